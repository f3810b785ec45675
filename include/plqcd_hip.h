/*
 * plqcd_hip.h -- C ABI of the MI355X Wilson hopping-matrix library.
 *
 * Drop-in for PLQCD's hopping-matrix path.  Host C code that used
 *   #include "plqcd.h"          (include/plqcd.h:17-18 -> start.h, dirac.h)
 * includes this header instead and links libplqcd_hip.so.  The field types,
 * the global geometry struct `plqcd_g`, the start.c geometry API and the
 * utils.c helpers keep the reference's names, layouts and argument meanings;
 * the hopping routines gain a `_hip` suffix.
 *
 * Differences from the reference, all deliberate:
 *  - One process, one GPU: NPROC0..3 must be 1 (no MPI); face[] is 0.
 *  - Errors never exit(): functions return an error code (int), a negative
 *    time (double) or -1 index, and plqcd_hip_last_error() holds the message.
 *  - The ISA-specific split-layout tables of plqcd_g (ipt_eo_yzt,
 *    *_sse_split, *_avx_split, *_mic_split, phip512...) are not provided.
 */
#ifndef PLQCD_HIP_H
#define PLQCD_HIP_H

#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif
/* every declaration below is exported from libplqcd_hip.so */
#pragma GCC visibility push(default)

/* ------------------------------------------------------------------ types */
/* Same memory layout as include/su3.h:52-75 (complex = two doubles, re, im). */
#ifdef __cplusplus
typedef struct { double re, im; } plqcd_dcomplex;
#else
typedef _Complex double plqcd_dcomplex;
#endif

/* su3.h:52-55: row-major 3x3, c<row><col> */
typedef struct {
   plqcd_dcomplex c00, c01, c02, c10, c11, c12, c20, c21, c22;
} su3;
/* su3.h:57-60 */
typedef struct { plqcd_dcomplex c0, c1, c2; } su3_vector;
/* su3.h:67-70: spin-major, s0..s3 each c0..c2 */
typedef struct { su3_vector s0, s1, s2, s3; } spinor;
/* su3.h:72-75 */
typedef struct { su3_vector s0, s1; } halfspinor;

/* ------------------------------------------------------ global parameters */
/* Subset of struct _PLQCD_GLOBAL_PARAMS (include/plqcd_global_params.h:67-205)
 * with the same field names and meanings. */
struct _PLQCD_GLOBAL_PARAMS {
   int nprocs[4];   /* processes per direction: always 1 here              */
   int nthread;     /* NTHREAD from the input file (host threads)          */
   int cpr[4];      /* coordinates of this process: 0                      */
   int npr[8];      /* neighbour process ids: 0                            */
   int LX, LY, LZ, LT; /* global lattice extents (x,y,z,t)                  */
   int latdims[4];  /* local extents = global extents                      */
   int VOLUME;      /* latdims[0]*latdims[1]*latdims[2]*latdims[3]         */
   int face[4];     /* boundary sizes: 0 with one process                  */
   int *ipt_eo;     /* lexicographic -> even-odd index                     */
   int *ipt_lex;    /* even-odd -> lexicographic index                     */
   int *nn_bnde[8]; /* boundary maps: NULL with one process                */
   int *nn_bndo[8];
   int **iup;       /* iup[i][mu]: even-odd index of the +mu neighbour      */
   int **idn;       /* idn[i][mu]: even-odd index of the -mu neighbour      */
   halfspinor *phip[4]; /* host scratch of the CPU code: NULL (unused)      */
   halfspinor *phim[4];
   int ALIGN;       /* host allocations are aligned to 2^ALIGN bytes        */
   int init;        /* 1 once init_plqcd succeeded                          */
};
typedef struct _PLQCD_GLOBAL_PARAMS plqcd_global_params;
extern struct _PLQCD_GLOBAL_PARAMS plqcd_g;

/* ------------------------------------------- start.c (include/start.h) */
/* start.h:30 / start.c:19-232.  argv[1] names a key/value input file with the
 * keys L0..L3, NPROC0..3, NTHREAD, ALIGN (defaults 4,4,4,4 / 1 / 1 / 16).
 * Returns 0 on success (also when already initialised), -1 when argv[1] is
 * missing (as the reference), a positive code for an unreadable file or
 * invalid parameters (the reference exits instead). */
int init_plqcd(int argc, char *argv[]);
/* Programmatic form of init_plqcd (no input file). */
int plqcd_init_dims(int L0, int L1, int L2, int L3, int nthread, int align);
/* start.h:37 / start.c:237-247: releases the geometry tables and device state. */
void finalize_plqcd(void);
/* start.h:44 / start.c:1169-1183 */
int ipr(int n[]);
/* start.h:50 / start.c:1190-1212 */
void comp_cpr(int ip, int n[]);
/* start.h:61 / start.c:1227-1248 */
void set_npr(void);
/* start.h:67 / start.c:1256-1270; leaves xcor untouched on a bad index */
void lex2cart(int ix, int xcor[]);
/* start.h:72 / start.c:1278-1298; returns -1 on out-of-range coordinates */
int cart2lex(int xcor[]);
/* start.h:76 / start.c:255-526 (even-odd order and periodic neighbours).
 * Returns 0, or nonzero on invalid dimensions (the reference is void). */
int set_geometry(void);
/* start.h:84 / start.c:1308-1343; *pid = *xeo = -1 on bad coordinates */
void find_site(int xg[], int *pid, int *xeo);
/* start.h:88 / start.c:1351-1428: same text format */
void print_params(FILE *fp);

/* ---------------------------------------------- utils.c (include/utils.h) */
void *alloc(size_t size, size_t alignment);   /* utils.h:13 / utils.c:23-33 */
void *amalloc(size_t size, int p);            /* utils.h:31 / utils.c:35-65 */
void afree(void *addr);                       /* utils.h:32 / utils.c:68-90 */
double stop_watch(double t);                  /* utils.h:43 / utils.c:120-125 */

/* ------------------------------------------------------------- errors */
/* Message of the last failure in this thread ("" if none). */
const char *plqcd_hip_last_error(void);

/* ------------------------------------------- hopping matrix (dirac.h) */
#define PLQCD_HOP_COMPAT 0  /* the reference's compiled output, see below  */
#define PLQCD_HOP_WILSON 1  /* Eq. 2.1 of plqcd.pdf                        */

/* Mode used by the three drop-in calls below.  PLQCD_HOP_COMPAT (default)
 * reproduces lib/dirac.c's ASSYMBLY routines as compiled: the mu=0 backward
 * lower spin components are assigned, not accumulated (dirac.c:373-374,
 * :756-757), and the oe forward pass reads the links of site i-V/2
 * (dirac.c:643-644).  PLQCD_HOP_WILSON computes Eq. 2.1 exactly.
 * Returns 0, or -1 for an unknown mode. */
int plqcd_hip_set_mode(int mode);

/* Replaces plqcd_hopping_matrix_eo_sse3_assymbly (include/dirac.h:21,
 * lib/dirac.c:35-408): qout[0,V/2) = H_eo qin[V/2,V).  Host pointers in the
 * reference layout, even-odd ordered, after init_plqcd.  Returns the wall
 * seconds of the call including host<->device transfers, or a negative value
 * on failure.  qin and qout may alias (qin's odd half is read, qout's even
 * half written). */
double plqcd_hopping_matrix_eo_hip(spinor *qin, spinor *qout, su3 *u);
/* Replaces plqcd_hopping_matrix_oe_sse3_assymbly (include/dirac.h:22,
 * lib/dirac.c:414-791): qout[V/2,V) = H_oe qin[0,V/2). */
double plqcd_hopping_matrix_oe_hip(spinor *qin, spinor *qout, su3 *u);
/* eo followed by oe (tests/test_hopping.c:296-297) in one launch. */
double plqcd_hopping_matrix_hip(spinor *qin, spinor *qout, su3 *u);
/* When enabled, the drop-in calls skip the gauge upload if `u` is the same
 * pointer as in the previous call (the caller promises the links did not
 * change).  Disabled by default.  plqcd_hip_gauge_changed() forces the next
 * call to upload. */
void plqcd_hip_cache_gauge(int enable);
void plqcd_hip_gauge_changed(void);

/* ------------------------------------------------- device-resident API */
/* Storage precisions and link formats of device fields. */
#define PLQCD_PREC_F64 0
#define PLQCD_PREC_F32 1
#define PLQCD_PREC_F16 2  /* spinors: fp16 with a float scale per site      */
#define PLQCD_LINKS_18 18 /* full 3x3 complex                               */
#define PLQCD_LINKS_12 12 /* rows 0,1; row 2 = conj(row0 x row1) (SU(3) only) */

/* Output parity of a hopping application. */
#define PLQCD_PARITY_EVEN 0 /* eo: even outputs from odd inputs              */
#define PLQCD_PARITY_ODD 1  /* oe: odd outputs from even inputs              */
#define PLQCD_PARITY_ALL 2  /* full H                                        */

typedef struct plqcd_hip_lattice plqcd_hip_lattice;
typedef struct plqcd_hip_gauge plqcd_hip_gauge;
typedef struct plqcd_hip_field plqcd_hip_field;

/* A lattice with extents L[0..3] (x,y,z,t), each even and >= 2, with at most
 * 2^29 sites (the kernels' 32-bit link offsets), on the
 * current HIP device, with its own stream.  All calls taking a lattice
 * object are asynchronous on that stream unless stated.  Host transfers of
 * a lattice's fields and links share its staging buffers: one thread at a
 * time transfers to or from the fields of one lattice. */
int plqcd_hip_lattice_create(const int L[4], plqcd_hip_lattice **out);
void plqcd_hip_lattice_destroy(plqcd_hip_lattice *lat);
/* Use an existing hipStream_t (NULL: back to the lattice's own stream). */
int plqcd_hip_lattice_set_stream(plqcd_hip_lattice *lat, void *hip_stream);
void *plqcd_hip_lattice_stream(plqcd_hip_lattice *lat);
int plqcd_hip_lattice_sync(plqcd_hip_lattice *lat);
/* Event timer on the lattice stream: start, then stop returns milliseconds
 * between the two events (waits for the stop event). */
int plqcd_hip_timer_start(plqcd_hip_lattice *lat);
int plqcd_hip_timer_stop(plqcd_hip_lattice *lat, double *ms);

/* A gauge field in storage precision `prec` (PLQCD_PREC_F64/F32/F16) with
 * 18-real links, or 12-real links whose row 2 the kernels rebuild as
 * conj(row0 x row1) (exact for SU(3) links up to rounding). */
int plqcd_hip_gauge_create(plqcd_hip_lattice *lat, int prec, int links, plqcd_hip_gauge **out);
void plqcd_hip_gauge_destroy(plqcd_hip_gauge *g);
/* u: 4V links in the reference layout u[4*i+mu], even-odd site order.
 * Synchronous with respect to the host buffer. */
int plqcd_hip_gauge_upload(plqcd_hip_gauge *g, const su3 *u);
/* Device bytes of the gauge field. */
size_t plqcd_hip_gauge_bytes(const plqcd_hip_gauge *g);

/* A field of nrhs right-hand sides (1 <= nrhs <= plqcd_hip_max_rhs(L)). */
int plqcd_hip_field_create(plqcd_hip_lattice *lat, int prec, int nrhs, plqcd_hip_field **out);
/* The largest RHS count of one field on a lattice with extents L: the
 * kernels address a site inside a component block with a 32-bit byte offset,
 * so V/2 * nrhs * 16 bytes must not exceed 4 GiB (and nrhs <= 4096).  0 for
 * extents plqcd_hip_lattice_create rejects (odd, < 2, or V > 2^29).  Host
 * only; needs no device. */
int plqcd_hip_max_rhs(const int L[4]);
void plqcd_hip_field_destroy(plqcd_hip_field *f);
/* V spinors in the reference layout, even-odd order, for right-hand side
 * `rhs`.  Synchronous with respect to the host buffer. */
int plqcd_hip_field_upload(plqcd_hip_field *f, int rhs, const spinor *host);
int plqcd_hip_field_download(const plqcd_hip_field *f, int rhs, spinor *host);
size_t plqcd_hip_field_bytes(const plqcd_hip_field *f);

/* out = H in on the given output parity (the other parity of `out` is not
 * touched), mode PLQCD_HOP_COMPAT or PLQCD_HOP_WILSON.  in and out must be
 * distinct fields of the gauge's lattice and of the same precision and RHS
 * count.  Asynchronous. */
int plqcd_hip_hop(const plqcd_hip_gauge *g, const plqcd_hip_field *in, plqcd_hip_field *out,
                  int parity, int mode);

/* ------------------------------------------------------------- solvers */
/* The reference declares solver parameters but ships none
 * (include/plqcd_global_params.h:197-201); these are new, built on the
 * hopping matrix in wilson mode with M = 1 - kappa H, M^+ = gamma5 M gamma5. */

/* out = M in (dagger = 0) or M^+ in (dagger = 1), kappa_rhs[r] for RHS r
 * (host array of in->nrhs values).  Synchronous. */
int plqcd_hip_apply_m(const plqcd_hip_gauge *g, const plqcd_hip_field *in, plqcd_hip_field *out,
                      const double *kappa_rhs, int dagger);

/* plqcd_hip_apply_m without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops).  No
 * counterpart in the reference.  The per-RHS kappa goes into a static device
 * array of 64 doubles per GPU (filled by a kernel from launch arguments; no
 * host copy, no memset), so in->nrhs may be at most 64 (error 2 above).
 * Asynchronous on the lattice stream (plqcd_hip_apply_m synchronizes).
 * Calls that overlap on one GPU share that array and race. */
int plqcd_hip_apply_m_noalloc(const plqcd_hip_gauge *g, const plqcd_hip_field *in, plqcd_hip_field *out,
                              const double *kappa_rhs, int dagger);

/* Even-odd (Schur) preconditioned operator on one parity p (EVEN or ODD):
 * out_p = in_p - kappa_r^2 H_{p,q} H_{q,p} in_p  (q the other parity), built
 * from the eo/oe halves of dirac.h:21-22 in wilson mode (the second half
 * hop forms in_p - kappa^2 (H tmp) in its epilogue).  tmp receives
 * H_{q,p} in_p on parity q; the other parity of out is not touched.
 * Synchronous. */
int plqcd_hip_apply_schur(const plqcd_hip_gauge *g, const plqcd_hip_field *in, plqcd_hip_field *tmp,
                          plqcd_hip_field *out, const double *kappa_rhs, int parity);

typedef struct {
   double kappa;            /* M = 1 - kappa H                                  */
   double tol;              /* stop at ||b - M^+M x|| / ||b|| <= tol (FP64)     */
   int max_iter;            /* iterations of the inner CG                       */
   double delta;            /* reliable update when the iterated residual drops
                               by delta since the last one; <= 0: 1e-3 for F32,
                               1e-2 for F16 sloppy, only at convergence for F64 */
   int use_graph;           /* replay one captured hipGraph per iteration       */
   const double *kappa_rhs; /* optional per-RHS kappa (b->nrhs values)          */
   const double *tol_rhs;   /* optional per-RHS tolerance                       */
} plqcd_cg_params;

typedef struct {
   int iterations;        /* inner CG iterations of this RHS                    */
   int reliable_updates;  /* FP64 residual recomputations                       */
   int converged;         /* 1 when true_residual <= tol                        */
   double true_residual;  /* ||b - M^+M x|| / ||b|| computed in FP64            */
   double seconds;        /* wall time of the whole solve call                  */
   double hop_applications; /* applications of H spent on this RHS              */
} plqcd_cg_stats;

/* Solves M^+M x = b for every RHS of b (at most 32) with CG on the normal
 * equations.  g: F64 gauge (18- or 12-real links) used for the true residual;
 * g_sloppy: gauge of the inner iteration, any precision (NULL or g: plain FP64
 * CG).  x (F64, same RHS count) is overwritten, starting from 0.  stats:
 * b->nrhs entries.  Returns 0 when the solve ran (check stats[r].converged). */
int plqcd_hip_cg_solve(const plqcd_hip_gauge *g, const plqcd_hip_gauge *g_sloppy, const plqcd_hip_field *b,
                       plqcd_hip_field *x, const plqcd_cg_params *params, plqcd_cg_stats *stats);

/* ------------------------------------------------------------- serving */
/* A solve server with continuous batching: requests for CG solves of
 * M^+M x = b on any lattice are admitted by free HBM, share resident gauge
 * fields, and join running multi-RHS solves as slots free up.  New; the
 * reference has no serving layer. */
typedef struct plqcd_hip_server plqcd_hip_server;

typedef struct {
   int device;         /* HIP device                                            */
   int sloppy_prec;    /* inner CG precision: PLQCD_PREC_F64 / F32 / F16         */
   int sloppy_links;   /* 18 or 12                                               */
   int slots;          /* right-hand sides per engine, 1..32                     */
   size_t hbm_budget;  /* bytes the server may allocate (0: free HBM - 8 GiB)    */
   int use_graph;      /* replay captured CG iterations                          */
   int transfer_threads; /* concurrent host transfers of sources and solutions
                            (0: 8)                                               */
} plqcd_serve_config;

typedef struct {
   int L[4];             /* lattice extents                                     */
   uint64_t gauge_seed;  /* names the gauge field; generated as random SU(3)
                            with this seed unless `gauge` is given              */
   const su3 *gauge;     /* optional links (4V, reference layout) for the first
                            request naming this (L, seed)                       */
   double kappa, tol;
   int max_iter;
   int nrhs;
   const spinor *b;      /* nrhs*V sources; must stay valid until completion   */
   spinor *x;            /* nrhs*V solutions, written before completion        */
} plqcd_serve_request;

typedef struct {
   uint64_t id;
   int status;            /* 0 ok, -1 device error, -2 larger than the budget  */
   double submit_time, start_time, end_time;  /* seconds since server creation */
   int iterations;        /* largest over the request's RHS                    */
   double true_residual;  /* largest FP64 true relative residual               */
   int converged;         /* all RHS converged                                 */
} plqcd_serve_result;

int plqcd_hip_server_create(const plqcd_serve_config *cfg, plqcd_hip_server **out);
/* Thread-safe.  Returns at once; the request is served in the background. */
int plqcd_hip_server_submit(plqcd_hip_server *s, const plqcd_serve_request *req, uint64_t *id);
/* Moves up to `max` completed results into out[]; *n = how many. */
int plqcd_hip_server_poll(plqcd_hip_server *s, plqcd_serve_result *out, int max, int *n);
/* Waits until at least `count` results are ready to poll. */
int plqcd_hip_server_wait(plqcd_hip_server *s, int count, double timeout_s);
/* out = {budget bytes, bytes in use, cached (lattice, gauge) entries, internal
 * failures, engines created, cache entries evicted, host bytes transferred
 * (sources and solutions), seconds the transfer threads spent transferring}.
 * Serving runs one HIP stream per engine and per transfer thread: give the
 * process enough hardware queues (GPU_MAX_HW_QUEUES, e.g. 16) or unrelated
 * streams share queues and wait for each other. */
int plqcd_hip_server_info(plqcd_hip_server *s, double out[8]);
/* Finishes the queued requests, then frees everything. */
void plqcd_hip_server_destroy(plqcd_hip_server *s);

/* Launch shape of the hopping kernels for storage precision `prec` (-1: all
 * precisions), process-wide; the autotuner sets it.
 *   order 0: block b -> tile b; 1: each XCD takes a contiguous eighth of the
 *   lattice; 2: XCDs sweep the lattice together, XCD k taking runs of `group`
 *   consecutive tiles; block: threads per block (64, 128, 256 or 512);
 *   tmarch > 0: full-lattice wilson applications whose time extent is a
 *   multiple of tmarch use the t-march kernel (one thread walks tmarch
 *   t-slices of a column), otherwise the one-site-per-thread gather kernel;
 *   rb (1, 2 or 4): right-hand sides per thread of the gather kernel when it
 *   divides the field's RHS count (links are loaded once per rb RHS; the
 *   t-march kernel is used only with rb 1);
 *   tile_y, tile_z > 0: one-RHS t-march launches use the LDS-tiled kernel,
 *   whose blocks own tiles of L0 x tile_y x tile_z sites (both parities) and
 *   exchange neighbour spinors through LDS, when the tile divides the lattice
 *   and has 128 or 256 (or, below F64, 512) lanes = L0 tile_y tile_z
 *   (block is then ignored); 0, 0: no tile.
 * Each RHS's result is bitwise the same for every launch shape.  Returns 0,
 * or nonzero for bad values. */
int plqcd_hip_hop_tuning(int prec, int order, int group, int block, int tmarch, int rb, int tile_y, int tile_z);
/* Back to the built-in launch shapes of precision `prec` (-1: all), which
 * differ for one and for several right-hand sides; a shape set with
 * plqcd_hip_hop_tuning applies to every RHS count until reset. */
int plqcd_hip_hop_tuning_reset(int prec);
/* Current launch shape of precision `prec` (one RHS):
 * out = {order, group, block, tmarch, rb, tile_y, tile_z}. */
int plqcd_hip_hop_get_tuning(int prec, int out[7]);
/* Times every launch shape for the shape of (g, in) -- lattice, precision,
 * link format, RHS count -- with `reps` timed applications each, and makes
 * the fastest the launch shape of that shape from now on.  best = {order,
 * group, block, tmarch, rb, tile_y, tile_z}, *best_ms = its time per
 * application.  in/out are overwritten. */
int plqcd_hip_autotune(const plqcd_hip_gauge *g, const plqcd_hip_field *in, plqcd_hip_field *out, int reps,
                       int best[7], double *best_ms);
/* The same for the solvers' launches over all RHS of `in` (M = 1 - kappa H
 * with |Mp|^2, then M^+ fused with the CG residual update and |r|^2, and the residual
 * launch of reliable updates): times the pair of launches of one CG iteration
 * with `reps` timed pairs per shape, on scratch fields it allocates (`in`,
 * whose storage precision must be the gauge field's, is only read), and
 * makes the fastest the shape of those launches from now on; *best_ms is its
 * time per pair.  Plain applications of H keep their own shapes.  Returns 0,
 * or nonzero on bad arguments or a device error (the previous shape stays).
 * No counterpart in the reference, which has no solver and chooses among
 * its hopping variants by hand (tests/test_hopping.c:283-560). */
int plqcd_hip_autotune_solver(const plqcd_hip_gauge *g, const plqcd_hip_field *in, int reps, int best[7],
                              double *best_ms);
/* The launch shape cached for the shape of (g, in) -- lattice, precision,
 * link format, RHS count -- and operation `op`: 0 plain applications of H
 * (what plqcd_hip_autotune tunes), 1 the solvers' M / M^+ launches (what
 * plqcd_hip_autotune_solver tunes; built-in shapes until then).  Returns 1 and fills shape = {order, group,
 * block, tmarch, rb, tile_y, tile_z} when one is cached, 0 when none is, -1
 * on bad arguments.  No counterpart in the reference, which chooses among its
 * variants by hand (tests/test_hopping.c:283-560). */
int plqcd_hip_autotune_lookup(const plqcd_hip_gauge *g, const plqcd_hip_field *in, int op, int shape[7]);
/* Forgets every autotuned shape (back to the per-precision defaults). */
int plqcd_hip_autotune_clear(void);

/* One application of H (full lattice, given mode) on each of n independent
 * lattices -- each with its own gauge field and fields, any volume, up to 32
 * RHS -- in one launch per 16 lattices.  All items share storage
 * precision, link format and stream.  Asynchronous. */
int plqcd_hip_hop_batch(int n, const plqcd_hip_gauge *const *g, const plqcd_hip_field *const *in,
                        plqcd_hip_field *const *out, int mode);

/* HBM probe: achieved GB/s of a streaming kernel over `bytes` per buffer,
 * `reps` back-to-back launches timed together.  kind 0: copy (reads and
 * writes `bytes`); 1: read only; 2: write only; 3: the hopping matrix's
 * read:write mix of 4:1 (reads 4 x `bytes`, writes `bytes`; the best of a
 * sweep of grid sizes). */
int plqcd_hip_stream_bandwidth(int kind, size_t bytes, int reps, double *gbs);

/* ------------------------------------------------ host memory */
/* Page-locked host memory (hipHostMalloc).  The server moves requests'
 * sources and solutions that lie in such a buffer by DMA alone, without the
 * copy through its own pinned staging buffers (no counterpart in the
 * reference, which has no host/device transfers). */
int plqcd_hip_host_alloc(size_t bytes, void **out);
void plqcd_hip_host_free(void *p);

/* ------------------------------------------------ synthetic inputs */
/* Seeded, counter-based and thread-order independent: the value at each
 * (site, component) depends only on (seed, even-odd site index, component).
 * Random SU(3): rows 0,1 of a 3x3 matrix of N(0,1)+iN(0,1) entries are
 * Gram-Schmidt orthonormalised, row 2 = conj(row0 x row1).  V = volume. */
void plqcd_random_gauge_su3(su3 *u, int V, uint64_t seed);
void plqcd_random_spinor_gauss(spinor *s, int V, uint64_t seed);
/* u = 1 on every link (free field). */
void plqcd_unit_gauge(su3 *u, int V);

#pragma GCC visibility pop
#ifdef __cplusplus
}
#endif
#endif /* PLQCD_HIP_H */
