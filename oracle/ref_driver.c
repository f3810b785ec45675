/*
 * Test infrastructure only: a driver for the reference's ASSYMBLY hopping
 * matrix, compiled from the reference sources by oracle/Makefile into
 * oracle/_ref/ (never committed, never part of the product path).
 *
 * It follows the structure of tests/test_hopping.c (init_plqcd, amalloc of
 * pin/pout/ufield at :96-115, the eo-then-oe loop at :287-301) but reads its
 * inputs from a file instead of filling them with rand(), so that the
 * reference's output can be compared with other implementations on identical
 * inputs.
 *
 * usage:
 *   hopping_ref <input_file> hop <in.bin> <out.bin> <nmul>
 *       in.bin  = qin (V spinors) then u (4V su3), raw float64, even-odd order
 *       out.bin = qout after nmul applications of eo followed by oe
 *   hopping_ref <input_file> harness <in.bin> <out.bin>
 *       fills qin/qout/u exactly as tests/test_hopping.c:178-254 (srand(rank),
 *       rand()/RAND_MAX), writes the inputs and the output of one eo+oe
 *   hopping_ref <input_file> geometry <out.txt>
 *       print_params dump as tests/test_geometry.c:55
 *   hopping_ref <input_file> bench <seconds>
 *       harness fill, then eo+oe repeated until <seconds> of accumulated
 *       call time; prints "calls <n> seconds <t> sites <V>"
 */
#include "plqcd.h"
#include <stdlib.h>
#include <string.h>

static void die(const char *msg)
{
   fprintf(stderr, "hopping_ref: %s\n", msg);
   MPI_Abort(MPI_COMM_WORLD, 2);
   exit(2);
}

static void fill_like_harness(spinor *pin, spinor *pout, su3 *u, int V)
{
   /* same draw order as tests/test_hopping.c:193-254 */
   double rs[24], ru[18];
   srand(0u);
   for (int i = 0; i < V; i++) {
      double *p;
      for (int j = 0; j < 24; j++) rs[j] = rand() / (double)RAND_MAX;
      p = (double *)&pin[i];
      for (int j = 0; j < 24; j++) p[j] = rs[j];
      for (int j = 0; j < 24; j++) rs[j] = rand() / (double)RAND_MAX;
      p = (double *)&pout[i];
      for (int j = 0; j < 24; j++) p[j] = rs[j];
      for (int mu = 0; mu < 4; mu++) {
         for (int k = 0; k < 18; k++) ru[k] = rand() / (double)RAND_MAX;
         p = (double *)&u[4 * i + mu];
         for (int k = 0; k < 18; k++) p[k] = ru[k];
      }
   }
}

static void write_all(const char *path, const void *a, size_t na, const void *b, size_t nb)
{
   FILE *f = fopen(path, "wb");
   if (!f) die("cannot open output file");
   if (fwrite(a, 1, na, f) != na) die("short write");
   if (b && fwrite(b, 1, nb, f) != nb) die("short write");
   fclose(f);
}

int main(int argc, char **argv)
{
   if (argc < 3) {
      fprintf(stderr, "usage: %s <input_file> hop|harness|geometry|bench ...\n", argv[0]);
      return 2;
   }
   if (init_plqcd(argc, argv) != 0) die("init_plqcd failed");
   const int V = plqcd_g.VOLUME;
   const char *cmd = argv[2];

   if (strcmp(cmd, "geometry") == 0) {
      if (argc < 4) die("geometry needs <out.txt>");
      FILE *f = fopen(argv[3], "w");
      if (!f) die("cannot open geometry output");
      print_params(f);
      fclose(f);
      finalize_plqcd();
      return 0;
   }

   spinor *pin = (spinor *)amalloc(V * sizeof(spinor), plqcd_g.ALIGN);
   spinor *pout = (spinor *)amalloc(V * sizeof(spinor), plqcd_g.ALIGN);
   su3 *u = (su3 *)amalloc(4 * (size_t)V * sizeof(su3), plqcd_g.ALIGN);
   if (!pin || !pout || !u) die("allocation failed");

   if (strcmp(cmd, "hop") == 0) {
      if (argc < 6) die("hop needs <in.bin> <out.bin> <nmul>");
      FILE *f = fopen(argv[3], "rb");
      if (!f) die("cannot open input file");
      if (fread(pin, sizeof(spinor), V, f) != (size_t)V) die("short read (spinor)");
      if (fread(u, sizeof(su3), 4 * (size_t)V, f) != 4 * (size_t)V) die("short read (gauge)");
      fclose(f);
      memset(pout, 0, V * sizeof(spinor));
      int nmul = atoi(argv[5]);
      for (int n = 0; n < nmul; n++) {
         plqcd_hopping_matrix_eo_sse3_assymbly(pin, pout, u);
         plqcd_hopping_matrix_oe_sse3_assymbly(pin, pout, u);
      }
      write_all(argv[4], pout, V * sizeof(spinor), NULL, 0);
   } else if (strcmp(cmd, "harness") == 0) {
      if (argc < 5) die("harness needs <in.bin> <out.bin>");
      fill_like_harness(pin, pout, u, V);
      write_all(argv[3], pin, V * sizeof(spinor), u, 4 * (size_t)V * sizeof(su3));
      plqcd_hopping_matrix_eo_sse3_assymbly(pin, pout, u);
      plqcd_hopping_matrix_oe_sse3_assymbly(pin, pout, u);
      write_all(argv[4], pout, V * sizeof(spinor), NULL, 0);
   } else if (strcmp(cmd, "bench") == 0) {
      double target = argc > 3 ? atof(argv[3]) : 10.0;
      fill_like_harness(pin, pout, u, V);
      double total = 0.0;
      long calls = 0;
      /* one untimed warm-up pair */
      plqcd_hopping_matrix_eo_sse3_assymbly(pin, pout, u);
      plqcd_hopping_matrix_oe_sse3_assymbly(pin, pout, u);
      while (total < target) {
         total += plqcd_hopping_matrix_eo_sse3_assymbly(pin, pout, u);
         total += plqcd_hopping_matrix_oe_sse3_assymbly(pin, pout, u);
         calls++;
      }
      printf("calls %ld seconds %.6f sites %d threads %d\n", calls, total, V, plqcd_g.nthread);
   } else {
      die("unknown command");
   }
   afree(pin);
   afree(pout);
   afree(u);
   finalize_plqcd();
   return 0;
}
