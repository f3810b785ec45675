/*
 * The oracle: a plain-C restatement of PLQCD's ASSYMBLY hopping matrix
 * (lib/dirac.c:35-408 eo, :414-791 oe) and of its single-process geometry
 * (lib/start.c:406-526).  TEST INFRASTRUCTURE ONLY: it is the checker for the
 * HIP library and the timed CPU baseline, never a product code path.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * All fields use the reference's memory layout: sites in even-odd order
 * (even sites first), spinor = 4 spins x 3 colours of complex double
 * (include/su3.h:67-70), links u[4*i+mu] as row-major 3x3 complex
 * (include/su3.h:52-55).
 */
#ifndef HOP_ORACLE_H
#define HOP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* mode of the hopping matrix */
#define ORACLE_MODE_COMPAT 0 /* what the compiled reference computes */
#define ORACLE_MODE_WILSON 1 /* Eq. 2.1 of plqcd.pdf */
/* test-only: compat with one of its two defects removed (each must then
 * disagree with the compiled reference) */
#define ORACLE_MODE_TEST_NO_LINK_SHIFT 2
#define ORACLE_MODE_TEST_NO_MU0_ASSIGN 3

/* Geometry for one process with periodic boundaries, as lib/start.c:406-526.
 * L[4] = (L0,L1,L2,L3) = (x,y,z,t). iup/idn are V*4 ints, [i*4+mu].
 * Returns 0, or -1 if a dimension is odd or < 2. */
int oracle_geometry(const int L[4], int *ipt_eo, int *ipt_lex, int *iup, int *idn);

/* qout[even] = H_eo qin[odd]   (lib/dirac.c:35-408)                       */
void oracle_hopping_eo(const double *qin, double *qout, const double *u,
                       const int L[4], int mode, int nthreads);
/* qout[odd]  = H_oe qin[even]  (lib/dirac.c:414-791)                       */
void oracle_hopping_oe(const double *qin, double *qout, const double *u,
                       const int L[4], int mode, int nthreads);
/* eo then oe, as tests/test_hopping.c:296-297                              */
void oracle_hopping_full(const double *qin, double *qout, const double *u,
                         const int L[4], int mode, int nthreads);

/* Persistent-scratch variant for timing: geometry and half-spinor buffers are
 * built once by oracle_bench_setup, then oracle_bench_run applies eo+oe
 * `napply` times and returns the wall seconds.  nthreads<=0: OpenMP default. */
int oracle_bench_setup(const int L[4], int nthreads);
double oracle_bench_run(const double *qin, double *qout, const double *u, int mode, int napply);
void oracle_bench_teardown(void);

#ifdef __cplusplus
}
#endif
#endif
