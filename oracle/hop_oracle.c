/*
 * The oracle: plain-C restatement of PLQCD's ASSYMBLY hopping matrix.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library and
 * the timed CPU baseline ("port"); only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product never calls it.
 *
 * It restates, in four phases over half-spinor scratch fields exactly like the
 * reference (so that it is also a fair CPU baseline):
 *   P1 lib/dirac.c:79-116   phip[mu][x]   = upper((1 - gamma_mu) psi(x+mu))
 *   P2 lib/dirac.c:141-196  phim[mu][x]   = U_mu(x-mu)^+ upper((1 + gamma_mu) psi(x-mu))
 *   P3 lib/dirac.c:258-319  out(x)        = sum_mu recon_fwd(U_mu(x) phip[mu][x])
 *   P4 lib/dirac.c:363-399  out(x)       += sum_mu recon_bwd(phim[mu][x])
 * and the oe mirror lib/dirac.c:414-791.  The arithmetic follows the plain-C
 * macros of include/su3.h:111-306 in the same operation order (complex
 * products written out as (ac-bd, ad+bc), which is what the C99 operator
 * evaluates for finite operands), so that with -ffp-contract=off the result is
 * bitwise identical to the compiled reference.
 *
 * Two modes:
 *   ORACLE_MODE_COMPAT  what the compiled reference outputs.  It differs from
 *     Eq. 2.1 of plqcd.pdf in two places:
 *     (1) lib/dirac.c:373-374 and :756-757 use _vector_minus_assign
 *         (include/su3.h:135-138, r = -s) for the mu=0 backward lower spin
 *         components, discarding all forward contributions to s2 and s3;
 *     (2) the oe forward pass (lib/dirac.c:643-644) takes the links of odd
 *         output site j = i + V/2 from u[4*i], i.e. from the even site with
 *         the same index within its parity.
 *   ORACLE_MODE_WILSON  Eq. 2.1: both of the above corrected.
 *
 * Pinned by: tests/golden fixtures written by the compiled reference
 * (oracle/_ref), bitwise; for the wilson mode additionally by a dense
 * gamma-matrix construction, the free-field plane-wave closed form,
 * gamma5-hermiticity and gauge covariance (tests/test_oracle.py).
 */
#include <complex.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "hop_oracle.h"

typedef double _Complex cplx;
typedef struct { cplx c0, c1, c2; } vec3;          /* include/su3.h:57-60 */
typedef struct { vec3 s0, s1, s2, s3; } spin4;     /* include/su3.h:67-70 */
typedef struct { vec3 s0, s1; } spin2;             /* include/su3.h:72-75 */
typedef struct { cplx c00, c01, c02, c10, c11, c12, c20, c21, c22; } mat3; /* :52-55 */

/* ---------------------------------------------------------------- algebra */

static inline cplx cmul(cplx a, cplx b)
{
   double ar = creal(a), ai = cimag(a), br = creal(b), bi = cimag(b);
   return CMPLX(ar * br - ai * bi, ar * bi + ai * br);
}
static inline cplx cjmul(cplx a, cplx b) /* conj(a) * b */
{
   double ar = creal(a), ai = -cimag(a), br = creal(b), bi = cimag(b);
   return CMPLX(ar * br - ai * bi, ar * bi + ai * br);
}
static inline cplx ci(cplx a) { return CMPLX(-cimag(a), creal(a)); } /* i*a, exact */

static inline vec3 vadd(vec3 a, vec3 b) { vec3 r = {a.c0 + b.c0, a.c1 + b.c1, a.c2 + b.c2}; return r; }
static inline vec3 vsub(vec3 a, vec3 b) { vec3 r = {a.c0 - b.c0, a.c1 - b.c1, a.c2 - b.c2}; return r; }
static inline vec3 vneg(vec3 a) { vec3 r = {-a.c0, -a.c1, -a.c2}; return r; }
static inline vec3 viadd(vec3 a, vec3 b) { vec3 r = {a.c0 + ci(b.c0), a.c1 + ci(b.c1), a.c2 + ci(b.c2)}; return r; }
static inline vec3 visub(vec3 a, vec3 b) { vec3 r = {a.c0 - ci(b.c0), a.c1 - ci(b.c1), a.c2 - ci(b.c2)}; return r; }

/* r = U s  (include/su3.h:298-301) */
static inline vec3 mv(const mat3 *u, vec3 s)
{
   vec3 r;
   r.c0 = cmul(u->c00, s.c0) + cmul(u->c01, s.c1) + cmul(u->c02, s.c2);
   r.c1 = cmul(u->c10, s.c0) + cmul(u->c11, s.c1) + cmul(u->c12, s.c2);
   r.c2 = cmul(u->c20, s.c0) + cmul(u->c21, s.c1) + cmul(u->c22, s.c2);
   return r;
}
/* r = U^+ s  (include/su3.h:303-306) */
static inline vec3 mdv(const mat3 *u, vec3 s)
{
   vec3 r;
   r.c0 = cjmul(u->c00, s.c0) + cjmul(u->c10, s.c1) + cjmul(u->c20, s.c2);
   r.c1 = cjmul(u->c01, s.c0) + cjmul(u->c11, s.c1) + cjmul(u->c21, s.c2);
   r.c2 = cjmul(u->c02, s.c0) + cjmul(u->c12, s.c1) + cjmul(u->c22, s.c2);
   return r;
}

/* --------------------------------------------------------------- geometry */

/* lex = x + L0*(y + L1*(z + L2*t))   (lib/start.c:1278-1298) */
static inline int lex_of(const int L[4], const int c[4])
{
   return c[0] + L[0] * (c[1] + L[1] * (c[2] + L[2] * c[3]));
}
static inline void cart_of(const int L[4], int lex, int c[4])
{
   c[0] = lex % L[0];
   c[1] = (lex / L[0]) % L[1];
   c[2] = (lex / (L[0] * L[1])) % L[2];
   c[3] = lex / (L[0] * L[1] * L[2]);
}

int oracle_geometry(const int L[4], int *ipt_eo, int *ipt_lex, int *iup, int *idn)
{
   for (int mu = 0; mu < 4; mu++)
      if (L[mu] < 2 || (L[mu] & 1)) return -1;
   const int V = L[0] * L[1] * L[2] * L[3];
   /* even sites first, each parity in increasing lex order; the parity is
    * that of the coordinate sum (lib/start.c:406-433) */
   int ce = 0, co = V / 2;
   for (int lex = 0; lex < V; lex++) {
      int c[4];
      cart_of(L, lex, c);
      int par = (c[0] + c[1] + c[2] + c[3]) & 1;
      int eo = par ? co++ : ce++;
      ipt_eo[lex] = eo;
      ipt_lex[eo] = lex;
   }
   /* periodic nearest neighbours (lib/start.c:439-526, one process) */
   for (int i = 0; i < V; i++) {
      int c[4];
      cart_of(L, ipt_lex[i], c);
      for (int mu = 0; mu < 4; mu++) {
         int s = c[mu];
         c[mu] = (s + 1) % L[mu];
         iup[4 * i + mu] = ipt_eo[lex_of(L, c)];
         c[mu] = (s + L[mu] - 1) % L[mu];
         idn[4 * i + mu] = ipt_eo[lex_of(L, c)];
         c[mu] = s;
      }
   }
   return 0;
}

/* ----------------------------------------------------------- the operator */

typedef struct {
   int L[4], V;
   int *ipt_eo, *ipt_lex, *iup, *idn;
   spin2 *phip[4], *phim[4];
   int nthreads;
} octx;

static int octx_init(octx *c, const int L[4], int nthreads)
{
   memset(c, 0, sizeof(*c));
   for (int mu = 0; mu < 4; mu++) c->L[mu] = L[mu];
   c->V = L[0] * L[1] * L[2] * L[3];
   c->nthreads = nthreads;
   c->ipt_eo = malloc(sizeof(int) * c->V);
   c->ipt_lex = malloc(sizeof(int) * c->V);
   c->iup = malloc(sizeof(int) * 4 * (size_t)c->V);
   c->idn = malloc(sizeof(int) * 4 * (size_t)c->V);
   int bad = !c->ipt_eo || !c->ipt_lex || !c->iup || !c->idn;
   for (int mu = 0; mu < 4 && !bad; mu++) {
      c->phip[mu] = malloc(sizeof(spin2) * (size_t)(c->V / 2));
      c->phim[mu] = malloc(sizeof(spin2) * (size_t)(c->V / 2));
      bad = !c->phip[mu] || !c->phim[mu];
   }
   if (bad || oracle_geometry(L, c->ipt_eo, c->ipt_lex, c->iup, c->idn) != 0) return -1;
   return 0;
}

static void octx_free(octx *c)
{
   free(c->ipt_eo); free(c->ipt_lex); free(c->iup); free(c->idn);
   for (int mu = 0; mu < 4; mu++) { free(c->phip[mu]); free(c->phim[mu]); }
   memset(c, 0, sizeof(*c));
}

/* One half of H.  dst = 0: qout[even] from qin[odd] (eo); dst = 1: oe. */
static void half_hop(const octx *c, const spin4 *qin, spin4 *qout, const mat3 *u, int dst, int mode)
{
   const int V = c->V, Vh = V / 2;
   const int src0 = dst ? 0 : Vh;  /* first source site      */
   const int dst0 = dst ? Vh : 0;  /* first destination site */
   /* the two reference defects, separately switchable for the tests that
    * show each one is present in the compiled reference's output */
   const int mu0_assign = (mode == ORACLE_MODE_COMPAT || mode == ORACLE_MODE_TEST_NO_LINK_SHIFT);
   const int link_shift = (mode == ORACLE_MODE_COMPAT || mode == ORACLE_MODE_TEST_NO_MU0_ASSIGN);
   spin2 *const *phip = c->phip, *const *phim = c->phim;
   const int *iup = c->iup, *idn = c->idn;
   int nt = c->nthreads;
#ifdef _OPENMP
   if (nt <= 0) nt = omp_get_max_threads();
#pragma omp parallel num_threads(nt)
#endif
   {
      /* P1: forward projections scattered to the destination parity */
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
      for (int i = src0; i < src0 + Vh; i++) {
         const spin4 *s = &qin[i];
         spin2 *p;
         p = &phip[0][idn[4 * i + 0] - dst0]; p->s0 = vadd(s->s0, s->s2);  p->s1 = vadd(s->s1, s->s3);
         p = &phip[1][idn[4 * i + 1] - dst0]; p->s0 = viadd(s->s0, s->s3); p->s1 = viadd(s->s1, s->s2);
         p = &phip[2][idn[4 * i + 2] - dst0]; p->s0 = vadd(s->s0, s->s3);  p->s1 = vsub(s->s1, s->s2);
         p = &phip[3][idn[4 * i + 3] - dst0]; p->s0 = viadd(s->s0, s->s2); p->s1 = visub(s->s1, s->s3);
      }
      /* P2: backward projections times U^+ of the source site */
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
      for (int i = src0; i < src0 + Vh; i++) {
         const spin4 *s = &qin[i];
         const mat3 *ub = &u[4 * (size_t)i];
         spin2 *p;
         p = &phim[0][iup[4 * i + 0] - dst0];
         p->s0 = mdv(ub + 0, vsub(s->s0, s->s2));  p->s1 = mdv(ub + 0, vsub(s->s1, s->s3));
         p = &phim[1][iup[4 * i + 1] - dst0];
         p->s0 = mdv(ub + 1, visub(s->s0, s->s3)); p->s1 = mdv(ub + 1, visub(s->s1, s->s2));
         p = &phim[2][iup[4 * i + 2] - dst0];
         p->s0 = mdv(ub + 2, vsub(s->s0, s->s3));  p->s1 = mdv(ub + 2, vadd(s->s1, s->s2));
         p = &phim[3][iup[4 * i + 3] - dst0];
         p->s0 = mdv(ub + 3, visub(s->s0, s->s2)); p->s1 = mdv(ub + 3, viadd(s->s1, s->s3));
      }
      /* P3: forward terms; the compat oe pass reads the links of site i, not j */
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
      for (int i = 0; i < Vh; i++) {
         const int j = dst0 + i;
         const mat3 *ub = &u[4 * (size_t)((link_shift && dst) ? i : j)];
         vec3 q0, q1, q2, q3, a0, a1;
         q0 = mv(ub + 0, phip[0][i].s0); q1 = mv(ub + 0, phip[0][i].s1);
         q2 = q0; q3 = q1;
         a0 = mv(ub + 1, phip[1][i].s0); a1 = mv(ub + 1, phip[1][i].s1);
         q0 = vadd(q0, a0); q1 = vadd(q1, a1); q2 = visub(q2, a1); q3 = visub(q3, a0);
         a0 = mv(ub + 2, phip[2][i].s0); a1 = mv(ub + 2, phip[2][i].s1);
         q0 = vadd(q0, a0); q1 = vadd(q1, a1); q2 = vsub(q2, a1); q3 = vadd(q3, a0);
         a0 = mv(ub + 3, phip[3][i].s0); a1 = mv(ub + 3, phip[3][i].s1);
         q0 = vadd(q0, a0); q1 = vadd(q1, a1); q2 = visub(q2, a0); q3 = viadd(q3, a1);
         qout[j].s0 = q0; qout[j].s1 = q1; qout[j].s2 = q2; qout[j].s3 = q3;
      }
      /* P4: backward terms; compat assigns the mu=0 lower components */
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
      for (int i = 0; i < Vh; i++) {
         spin4 *o = &qout[dst0 + i];
         const spin2 *b;
         b = &phim[0][i];
         o->s0 = vadd(o->s0, b->s0); o->s1 = vadd(o->s1, b->s1);
         if (mu0_assign) { o->s2 = vneg(b->s0); o->s3 = vneg(b->s1); }
         else        { o->s2 = vsub(o->s2, b->s0); o->s3 = vsub(o->s3, b->s1); }
         b = &phim[1][i];
         o->s0 = vadd(o->s0, b->s0); o->s1 = vadd(o->s1, b->s1);
         o->s2 = viadd(o->s2, b->s1); o->s3 = viadd(o->s3, b->s0);
         b = &phim[2][i];
         o->s0 = vadd(o->s0, b->s0); o->s1 = vadd(o->s1, b->s1);
         o->s2 = vadd(o->s2, b->s1); o->s3 = vsub(o->s3, b->s0);
         b = &phim[3][i];
         o->s0 = vadd(o->s0, b->s0); o->s1 = vadd(o->s1, b->s1);
         o->s2 = viadd(o->s2, b->s0); o->s3 = visub(o->s3, b->s1);
      }
   }
}

static void run_halves(const double *qin, double *qout, const double *u, const int L[4],
                       int mode, int nthreads, int do_eo, int do_oe)
{
   octx c;
   if (octx_init(&c, L, nthreads) != 0) { octx_free(&c); return; }
   if (do_eo) half_hop(&c, (const spin4 *)qin, (spin4 *)qout, (const mat3 *)u, 0, mode);
   if (do_oe) half_hop(&c, (const spin4 *)qin, (spin4 *)qout, (const mat3 *)u, 1, mode);
   octx_free(&c);
}

void oracle_hopping_eo(const double *qin, double *qout, const double *u, const int L[4], int mode, int nthreads)
{
   run_halves(qin, qout, u, L, mode, nthreads, 1, 0);
}

void oracle_hopping_oe(const double *qin, double *qout, const double *u, const int L[4], int mode, int nthreads)
{
   run_halves(qin, qout, u, L, mode, nthreads, 0, 1);
}

void oracle_hopping_full(const double *qin, double *qout, const double *u, const int L[4], int mode, int nthreads)
{
   run_halves(qin, qout, u, L, mode, nthreads, 1, 1);
}

/* ------------------------------------------------------------ CPU timing */

static octx bench_ctx;
static int bench_ready = 0;

int oracle_bench_setup(const int L[4], int nthreads)
{
   if (bench_ready) oracle_bench_teardown();
   if (octx_init(&bench_ctx, L, nthreads) != 0) { octx_free(&bench_ctx); return -1; }
   bench_ready = 1;
   return 0;
}

double oracle_bench_run(const double *qin, double *qout, const double *u, int mode, int napply)
{
   if (!bench_ready) return -1.0;
   struct timeval t0, t1;
   gettimeofday(&t0, NULL);
   for (int n = 0; n < napply; n++) {
      half_hop(&bench_ctx, (const spin4 *)qin, (spin4 *)qout, (const mat3 *)u, 0, mode);
      half_hop(&bench_ctx, (const spin4 *)qin, (spin4 *)qout, (const mat3 *)u, 1, mode);
   }
   gettimeofday(&t1, NULL);
   return (t1.tv_sec - t0.tv_sec) + 1e-6 * (t1.tv_usec - t0.tv_usec);
}

void oracle_bench_teardown(void)
{
   if (bench_ready) octx_free(&bench_ctx);
   bench_ready = 0;
}
