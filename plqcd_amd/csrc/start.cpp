// Lattice geometry and initialisation with the API of lib/start.c for one
// process on one GPU.  The even-odd site order and the periodic neighbour
// tables are the reference's (start.c:406-526) because the fields that host
// code hands to the hopping matrix are laid out in that order.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "internal.h"

struct _PLQCD_GLOBAL_PARAMS plqcd_g;

namespace {

void free_tables()
{
   free(plqcd_g.ipt_eo);
   free(plqcd_g.ipt_lex);
   if (plqcd_g.iup) free(plqcd_g.iup[0]);
   if (plqcd_g.idn) free(plqcd_g.idn[0]);
   free(plqcd_g.iup);
   free(plqcd_g.idn);
   plqcd_g.ipt_eo = plqcd_g.ipt_lex = nullptr;
   plqcd_g.iup = plqcd_g.idn = nullptr;
}

// Validation shared by init_plqcd and plqcd_init_dims (start.c:118-140):
// global extents even and >= 4; one process per direction.
int check_params()
{
   const int g[4] = {plqcd_g.LX, plqcd_g.LY, plqcd_g.LZ, plqcd_g.LT};
   for (int mu = 0; mu < 4; mu++)
      if (g[mu] < 4 || (g[mu] % 2) != 0)
         return plqcd::set_error(2, "lattice extent L%d = %d: every extent must be even and at least 4", mu, g[mu]);
   for (int mu = 0; mu < 4; mu++)
      if (plqcd_g.nprocs[mu] != 1)
         return plqcd::set_error(3, "NPROC%d = %d: this library runs one process on one GPU, "
                                    "NPROC0..3 must be 1", mu, plqcd_g.nprocs[mu]);
   if ((long long)g[0] * g[1] * g[2] * g[3] > 0x7fffffffLL)
      return plqcd::set_error(2, "lattice volume exceeds 2^31-1 sites");
   if (plqcd_g.ALIGN < 0 || plqcd_g.ALIGN > 30)
      return plqcd::set_error(2, "ALIGN = %d out of range [0,30]", plqcd_g.ALIGN);
   return 0;
}

int finish_init()
{
   int rc = check_params();
   if (rc) return rc;
   for (int mu = 0; mu < 4; mu++) plqcd_g.cpr[mu] = 0;
   set_npr();
   rc = set_geometry();
   if (rc) return rc;
   for (int mu = 0; mu < 4; mu++) plqcd_g.phip[mu] = plqcd_g.phim[mu] = nullptr;
   plqcd_g.init = 1;
   return 0;
}

void set_defaults()
{
   for (int i = 0; i < 4; i++) plqcd_g.nprocs[i] = 1;
   plqcd_g.nthread = 1;
   plqcd_g.LX = plqcd_g.LY = plqcd_g.LZ = plqcd_g.LT = 4;
   plqcd_g.ALIGN = 16;
}

}  // namespace

extern "C" int init_plqcd(int argc, char *argv[])
{
   plqcd::clear_error();
   if (argc < 2 || argv == nullptr || argv[1] == nullptr) {
      printf("init_plqcd: the first argument must name the input file\n");
      return plqcd::set_error(-1, "init_plqcd: argv[1] must name the input file");
   }
   if (plqcd_g.init == 1) return 0;

   FILE *fp = fopen(argv[1], "r");
   if (!fp) {
      fprintf(stderr, "init_plqcd: cannot open input file '%s'\n", argv[1]);
      return plqcd::set_error(1, "init_plqcd: cannot open input file '%s'", argv[1]);
   }
   set_defaults();
   // whitespace-separated key/value pairs (start.c:84-107)
   char var[128], val[128];
   while (fscanf(fp, "%127s %127s", var, val) == 2) {
      const int v = atoi(val);
      if (!strcmp(var, "NPROC0")) plqcd_g.nprocs[0] = v;
      else if (!strcmp(var, "NPROC1")) plqcd_g.nprocs[1] = v;
      else if (!strcmp(var, "NPROC2")) plqcd_g.nprocs[2] = v;
      else if (!strcmp(var, "NPROC3")) plqcd_g.nprocs[3] = v;
      else if (!strcmp(var, "NTHREAD")) plqcd_g.nthread = v;
      else if (!strcmp(var, "L0")) plqcd_g.LX = v;
      else if (!strcmp(var, "L1")) plqcd_g.LY = v;
      else if (!strcmp(var, "L2")) plqcd_g.LZ = v;
      else if (!strcmp(var, "L3")) plqcd_g.LT = v;
      else if (!strcmp(var, "ALIGN")) plqcd_g.ALIGN = v;
   }
   fclose(fp);
   int rc = finish_init();
   if (rc) fprintf(stderr, "ERROR: %s\n", plqcd_hip_last_error());
   return rc;
}

extern "C" int plqcd_init_dims(int L0, int L1, int L2, int L3, int nthread, int align)
{
   plqcd::clear_error();
   if (plqcd_g.init == 1) finalize_plqcd();
   set_defaults();
   plqcd_g.LX = L0;
   plqcd_g.LY = L1;
   plqcd_g.LZ = L2;
   plqcd_g.LT = L3;
   plqcd_g.nthread = nthread > 0 ? nthread : 1;
   plqcd_g.ALIGN = align;
   return finish_init();
}

extern "C" void finalize_plqcd(void)
{
   plqcd::release_dropin_state();
   free_tables();
   plqcd_g.init = 0;
}

extern "C" int ipr(int n[])
{
   for (int i = 0; i < 4; i++)
      if (n[i] < 0 || n[i] >= plqcd_g.nprocs[i])
         return plqcd::set_error(-1, "ipr: process coordinate %d = %d out of range", i, n[i]);
   const int *p = plqcd_g.nprocs;
   return n[0] + p[0] * (n[1] + p[1] * (n[2] + p[2] * n[3]));
}

extern "C" void comp_cpr(int ip, int n[])
{
   const int *p = plqcd_g.nprocs;
   const int np = p[0] * p[1] * p[2] * p[3];
   if (ip < 0 || ip >= np) {
      printf("comp_cpr: process id %d is outside 0 .. %d\n", ip, np - 1);
      plqcd::set_error(-1, "comp_cpr: process id %d out of range", ip);
      return;
   }
   n[0] = ip % p[0];
   n[1] = (ip / p[0]) % p[1];
   n[2] = (ip / (p[0] * p[1])) % p[2];
   n[3] = (ip / (p[0] * p[1] * p[2])) % p[3];
}

extern "C" void set_npr(void)
{
   int n[4];
   for (int mu = 0; mu < 4; mu++) n[mu] = plqcd_g.cpr[mu];
   for (int mu = 0; mu < 4; mu++) {
      const int c = n[mu], P = plqcd_g.nprocs[mu];
      n[mu] = (c - 1 + P) % P;
      plqcd_g.npr[2 * mu] = ipr(n);
      n[mu] = (c + 1) % P;
      plqcd_g.npr[2 * mu + 1] = ipr(n);
      n[mu] = c;
   }
}

extern "C" void lex2cart(int ix, int xcart[])
{
   const int *L = plqcd_g.latdims;
   if (ix < 0 || ix >= plqcd_g.VOLUME) {
      printf("lex2cart: index %d is outside 0 .. VOLUME-1\n", ix);
      plqcd::set_error(-1, "lex2cart: index %d out of range", ix);
      return;
   }
   xcart[0] = ix % L[0];
   xcart[1] = (ix / L[0]) % L[1];
   xcart[2] = (ix / (L[0] * L[1])) % L[2];
   xcart[3] = (ix / (L[0] * L[1] * L[2])) % L[3];
}

extern "C" int cart2lex(int xcor[])
{
   const int *L = plqcd_g.latdims;
   for (int i = 0; i < 4; i++)
      if (xcor[i] < 0 || xcor[i] >= L[i]) {
         printf("cart2lex: coordinate %d = %d is outside the local lattice\n", i, xcor[i]);
         return plqcd::set_error(-1, "cart2lex: coordinate %d = %d out of range", i, xcor[i]);
      }
   return xcor[0] + L[0] * (xcor[1] + L[1] * (xcor[2] + L[2] * xcor[3]));
}

extern "C" int set_geometry(void)
{
   free_tables();
   const int g[4] = {plqcd_g.LX, plqcd_g.LY, plqcd_g.LZ, plqcd_g.LT};
   for (int mu = 0; mu < 4; mu++) {
      if (plqcd_g.nprocs[mu] != 1 || g[mu] < 2 || (g[mu] % 2) != 0)
         return plqcd::set_error(2, "set_geometry: invalid extent L%d = %d (NPROC%d = %d)", mu, g[mu], mu,
                                 plqcd_g.nprocs[mu]);
      plqcd_g.latdims[mu] = g[mu];
      plqcd_g.face[mu] = 0;  // no process boundaries
   }
   for (int k = 0; k < 8; k++) plqcd_g.nn_bnde[k] = plqcd_g.nn_bndo[k] = nullptr;
   const int *L = plqcd_g.latdims;
   const int V = L[0] * L[1] * L[2] * L[3];
   plqcd_g.VOLUME = V;

   plqcd_g.ipt_eo = static_cast<int *>(calloc(V, sizeof(int)));
   plqcd_g.ipt_lex = static_cast<int *>(calloc(V, sizeof(int)));
   plqcd_g.iup = static_cast<int **>(calloc(V, sizeof(int *)));
   plqcd_g.idn = static_cast<int **>(calloc(V, sizeof(int *)));
   int *up = static_cast<int *>(calloc(4 * (size_t)V, sizeof(int)));
   int *dn = static_cast<int *>(calloc(4 * (size_t)V, sizeof(int)));
   if (!plqcd_g.ipt_eo || !plqcd_g.ipt_lex || !plqcd_g.iup || !plqcd_g.idn || !up || !dn) {
      free(up);
      free(dn);
      if (plqcd_g.iup) plqcd_g.iup[0] = nullptr;
      if (plqcd_g.idn) plqcd_g.idn[0] = nullptr;
      free_tables();
      return plqcd::set_error(4, "set_geometry: insufficient memory for %d sites", V);
   }
   for (int i = 0; i < V; i++) {
      plqcd_g.iup[i] = up + 4 * (size_t)i;
      plqcd_g.idn[i] = dn + 4 * (size_t)i;
   }

   // Even sites first, then odd; lexicographic order (x fastest) inside each
   // parity.  The parity is that of the global coordinate sum.
   int n_even = 0, n_odd = V / 2;
   for (int lex = 0; lex < V; lex++) {
      int c[4];
      lex2cart(lex, c);
      const int odd = (c[0] + c[1] + c[2] + c[3]) & 1;
      const int eo = odd ? n_odd++ : n_even++;
      plqcd_g.ipt_eo[lex] = eo;
      plqcd_g.ipt_lex[eo] = lex;
   }
   // Periodic nearest neighbours in +mu / -mu.
   for (int i = 0; i < V; i++) {
      int c[4];
      lex2cart(plqcd_g.ipt_lex[i], c);
      for (int mu = 0; mu < 4; mu++) {
         const int s = c[mu];
         c[mu] = (s + 1) % L[mu];
         plqcd_g.iup[i][mu] = plqcd_g.ipt_eo[cart2lex(c)];
         c[mu] = (s - 1 + L[mu]) % L[mu];
         plqcd_g.idn[i][mu] = plqcd_g.ipt_eo[cart2lex(c)];
         c[mu] = s;
      }
   }
   return 0;
}

extern "C" void find_site(int xg[], int *pid, int *xeo)
{
   *pid = *xeo = -1;
   int proc[4], xc[4];
   for (int mu = 0; mu < 4; mu++) {
      if (xg[mu] < 0 || xg[mu] >= plqcd_g.nprocs[mu] * plqcd_g.latdims[mu]) {
         printf("find_site: global coordinate %d = %d is outside the lattice\n", mu, xg[mu]);
         plqcd::set_error(-1, "find_site: coordinate %d = %d out of range", mu, xg[mu]);
         return;
      }
      proc[mu] = xg[mu] / plqcd_g.latdims[mu];
      xc[mu] = xg[mu] - proc[mu] * plqcd_g.latdims[mu];
   }
   *xeo = plqcd_g.ipt_eo[cart2lex(xc)];
   *pid = ipr(proc);
}

// Same text as start.c:1351-1428 so that dumps can be compared line by line.
extern "C" void print_params(FILE *fp)
{
   const struct _PLQCD_GLOBAL_PARAMS &g = plqcd_g;
   fprintf(fp, "grid of processes: nprocs0 %d nprocs1 %d nprocs2 %d nprocs3 %d\n\n", g.nprocs[0], g.nprocs[1],
           g.nprocs[2], g.nprocs[3]);
   fprintf(fp, "Number of openmp threads %d\n\n", g.nthread);
   fprintf(fp, "Cartesian co-ordinates of my process on the grid of processes: %d %d %d %d\n\n", g.cpr[0],
           g.cpr[1], g.cpr[2], g.cpr[3]);
   fprintf(fp, "Nearst neighbours of my process (-0 +0 -1 +1 -2 +2 -3 +3):");
   for (int k = 0; k < 8; k++) fprintf(fp, " %10d", g.npr[k]);
   fprintf(fp, " \n\n");
   fprintf(fp, "Global lattice size: LX %d LY %d LZ %d LT %d\n\n", g.LX, g.LY, g.LZ, g.LT);
   fprintf(fp, "Local lattice size: lx %d ly %d lz %d lt %d, local volume %d\n\n", g.latdims[0], g.latdims[1],
           g.latdims[2], g.latdims[3], g.VOLUME);
   fprintf(fp, "Sizes of the boundaries: 0-boundary %d 1-boundary %d 2-boundary %d 3-boundary %d\n\n", g.face[0],
           g.face[1], g.face[2], g.face[3]);
   fprintf(fp, "printing the lattice sites w.r.t. the local lattice:\n\n");
   static const char *hdr[14] = {"x0",     "x1",     "x2",     "x3",     "ipt_eo", "ipt_lex", "iup[0]",
                                 "iup[1]", "iup[2]", "iup[3]", "idn[0]", "idn[1]",  "idn[2]", "idn[3]"};
   for (int k = 0; k < 14; k++) fprintf(fp, k ? " %8s" : "%8s", hdr[k]);
   fprintf(fp, "\n\n");
   for (int ix = 0; ix < g.VOLUME; ix++) {
      int x[4];
      const int ilex = g.ipt_lex[ix];
      lex2cart(ilex, x);
      const int row[14] = {x[0],         x[1],         x[2],         x[3],         ix,
                           ilex,         g.iup[ix][0], g.iup[ix][1], g.iup[ix][2], g.iup[ix][3],
                           g.idn[ix][0], g.idn[ix][1], g.idn[ix][2], g.idn[ix][3]};
      for (int k = 0; k < 14; k++) fprintf(fp, k ? " %8d" : "%8d", row[k]);
      fprintf(fp, "\n\n");
   }
   // boundary maps are empty with one process: only their headers are printed
   fprintf(fp, "the nn_bnde and nn_bndo arrays\n");
   for (int d = 0; d < 4; d++) {
      fprintf(fp, "nn_bnde[%d] \n", 2 * d);
      fprintf(fp, "nn_bndo[%d]\n", 2 * d);
      fprintf(fp, "nn_bnde[%d] \n", 2 * d + 1);
      fprintf(fp, "nn_bndo[%d]\n", 2 * d + 1);
   }
   fprintf(fp, "memory alignment in bytes: ALIGN %d\n\n", g.ALIGN);
   fprintf(fp, "initialization check parameter: init %d\n\n", g.init);
}
