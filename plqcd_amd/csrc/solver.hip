// Conjugate gradient on the normal equations  M^+ M x = b,  M = 1 - kappa H
// (wilson mode, M^+ = gamma5 M gamma5), for all right-hand sides of a field
// at once, with per-RHS kappa and tolerance.
//
// The iteration runs in the "sloppy" precision of the sloppy gauge field
// (F64, F32 or F16 storage; F32/F64 arithmetic) and is corrected by reliable
// updates (Sleijpen and van der Vorst 1996, used for mixed precision by Clark
// et al. 2010): when the iterated residual of a right-hand side has dropped by
// `delta` since its last update, or looks converged, the partial solution is
// added to the FP64 solution, the true residual b - M^+M x is recomputed in
// FP64 and replaces the iterated one.  A right-hand side stops when its FP64
// true residual meets its tolerance.
//
// One iteration is a fixed sequence of kernels (beta, p update, M p with
// |Mp|^2, M^+, alpha, x/r update with |r|^2) whose scalars live on the
// device; it is captured once into a hipGraph and replayed.  Only |r|^2 comes
// back to the host each iteration, to decide reliable updates.  Every
// per-RHS quantity is computed from that RHS's data alone with fixed-order
// reductions, so a right-hand side follows bitwise the same iterates whatever
// else shares the launch.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "blas.h"
#include "dslash.h"
#include "reduce.h"

namespace {

constexpr int kMaxRhs = 32;

__global__ void k_alpha(const double *rr, const double *pap, const int *active, double *alpha, int R)
{
   const int r = threadIdx.x;
   if (r < R) alpha[r] = (active[r] && pap[r] > 0.0) ? rr[r] / pap[r] : 0.0;
}

__global__ void k_beta(double *rr, const double *rr_new, const int *active, double *beta, int R)
{
   const int r = threadIdx.x;
   if (r < R) {
      beta[r] = (active[r] && rr[r] > 0.0) ? rr_new[r] / rr[r] : 0.0;
      rr[r] = rr_new[r];
   }
}

struct Workspace {
   plqcd_hip_lattice *lat = nullptr;
   int R = 0;
   plqcd_hip_field *r = nullptr, *t1 = nullptr, *t2 = nullptr;             // F64
   plqcd_hip_field *rs = nullptr, *ps = nullptr, *qs = nullptr, *aps = nullptr, *xs = nullptr;  // sloppy
   double *partial = nullptr;
   double *scal = nullptr;  // kappa, rr, rr_new, pap, alpha, beta, tmp  (7 x kMaxRhs)
   int *imask = nullptr;    // active, update  (2 x kMaxRhs)
   double *host = nullptr;  // pinned: 2 x kMaxRhs
   hipGraphExec_t graph = nullptr;
   hipGraph_t graph_def = nullptr;

   double *kappa() { return scal; }
   double *rr() { return scal + kMaxRhs; }
   double *rr_new() { return scal + 2 * kMaxRhs; }
   double *pap() { return scal + 3 * kMaxRhs; }
   double *alpha() { return scal + 4 * kMaxRhs; }
   double *beta() { return scal + 5 * kMaxRhs; }
   double *tmp() { return scal + 6 * kMaxRhs; }
   int *active() { return imask; }
   int *update() { return imask + kMaxRhs; }

   ~Workspace()
   {
      if (graph) (void)hipGraphExecDestroy(graph);
      if (graph_def) (void)hipGraphDestroy(graph_def);
      const bool alias = (t1 == qs);
      for (plqcd_hip_field *f : {r, rs, ps, qs, aps, xs}) plqcd_hip_field_destroy(f);
      if (!alias) {
         plqcd_hip_field_destroy(t1);
         plqcd_hip_field_destroy(t2);
      }
      if (partial) (void)hipFree(partial);
      if (scal) (void)hipFree(scal);
      if (imask) (void)hipFree(imask);
      if (host) (void)hipHostFree(host);
   }
};

int make_field(plqcd_hip_lattice *lat, int prec, int R, plqcd_hip_field **f)
{
   return plqcd_hip_field_create(lat, prec, R, f);
}

int alloc_workspace(Workspace &w, plqcd_hip_lattice *lat, int sloppy, int R)
{
   w.lat = lat;
   w.R = R;
   int rc = make_field(lat, PLQCD_PREC_F64, R, &w.r);
   if (!rc) rc = make_field(lat, sloppy, R, &w.rs);
   if (!rc) rc = make_field(lat, sloppy, R, &w.ps);
   if (!rc) rc = make_field(lat, sloppy, R, &w.qs);
   if (!rc) rc = make_field(lat, sloppy, R, &w.aps);
   if (!rc) rc = make_field(lat, sloppy, R, &w.xs);
   if (rc) return rc;
   if (sloppy == PLQCD_PREC_F64) {  // the sloppy temporaries are free during a reliable update
      w.t1 = w.qs;
      w.t2 = w.aps;
   } else {
      rc = make_field(lat, PLQCD_PREC_F64, R, &w.t1);
      if (!rc) rc = make_field(lat, PLQCD_PREC_F64, R, &w.t2);
      if (rc) return rc;
   }
   PLQCD_HIP_TRY(hipMalloc(&w.partial, plqcd::partial_doubles(w.r) * sizeof(double)));
   PLQCD_HIP_TRY(hipMalloc(&w.scal, 7 * kMaxRhs * sizeof(double)));
   PLQCD_HIP_TRY(hipMalloc(&w.imask, 2 * kMaxRhs * sizeof(int)));
   PLQCD_HIP_TRY(hipHostMalloc(&w.host, 2 * kMaxRhs * sizeof(double), hipHostMallocDefault));
   return 0;
}

// One CG iteration on the sloppy fields (see the file comment).
int iteration(Workspace &w, const plqcd_hip_gauge *gs)
{
   hipStream_t st = w.lat->stream;
   const int nc = plqcd::reduce_chunks(w.lat->Vh);
   hipLaunchKernelGGL(k_beta, dim3(1), dim3(kMaxRhs), 0, st, w.rr(), w.rr_new(), w.active(), w.beta(), w.R);
   if (int rc = plqcd::blas_xpay(w.rs, w.beta(), w.ps)) return rc;
   plqcd::DslashOp m;
   m.kappa = w.kappa();
   m.partial = w.partial;
   if (int rc = plqcd::dslash(gs, w.ps, w.qs, PLQCD_PARITY_ALL, PLQCD_HOP_WILSON, m)) return rc;
   if (int rc = plqcd::reduce_partials(w.partial, 2 * nc, 0, 2 * nc, w.R, w.pap(), st)) return rc;
   plqcd::DslashOp md;
   md.kappa = w.kappa();
   md.dag = 1;
   if (int rc = plqcd::dslash(gs, w.qs, w.aps, PLQCD_PARITY_ALL, PLQCD_HOP_WILSON, md)) return rc;
   hipLaunchKernelGGL(k_alpha, dim3(1), dim3(kMaxRhs), 0, st, w.rr(), w.pap(), w.active(), w.alpha(), w.R);
   if (int rc = plqcd::blas_cg_update(w.xs, w.ps, w.rs, w.aps, w.alpha(), w.partial)) return rc;
   if (int rc = plqcd::reduce_partials(w.partial, 2 * nc, 0, 2 * nc, w.R, w.rr_new(), st)) return rc;
   PLQCD_HIP_TRY(hipMemcpyAsync(w.host, w.rr_new(), w.R * sizeof(double), hipMemcpyDeviceToHost, st));
   PLQCD_HIP_TRY(hipGetLastError());
   return 0;
}

// x += xs ; r = b - M^+M x ; rs = r ; tmp <- |r|^2, for the RHS in `update`.
int reliable_update(Workspace &w, const plqcd_hip_gauge *g, const plqcd_hip_field *b, plqcd_hip_field *x)
{
   hipStream_t st = w.lat->stream;
   const int nc = plqcd::reduce_chunks(w.lat->Vh);
   if (int rc = plqcd::blas_accum(x, w.xs, 1, w.update())) return rc;
   plqcd::DslashOp m;
   m.kappa = w.kappa();
   if (int rc = plqcd::dslash(g, x, w.t1, PLQCD_PARITY_ALL, PLQCD_HOP_WILSON, m)) return rc;
   m.dag = 1;
   if (int rc = plqcd::dslash(g, w.t1, w.t2, PLQCD_PARITY_ALL, PLQCD_HOP_WILSON, m)) return rc;
   if (int rc = plqcd::blas_residual(b, w.t2, w.r, w.partial, w.update())) return rc;
   if (int rc = plqcd::reduce_partials(w.partial, 2 * nc, 0, 2 * nc, w.R, w.tmp(), st)) return rc;
   if (int rc = plqcd::blas_copy(w.rs, w.r, nullptr, w.update())) return rc;
   PLQCD_HIP_TRY(hipMemcpyAsync(w.host + kMaxRhs, w.tmp(), w.R * sizeof(double), hipMemcpyDeviceToHost, st));
   PLQCD_HIP_TRY(hipStreamSynchronize(st));
   return 0;
}

}  // namespace

extern "C" int plqcd_hip_cg_solve(const plqcd_hip_gauge *g, const plqcd_hip_gauge *g_sloppy, const plqcd_hip_field *b,
                                  plqcd_hip_field *x, const plqcd_cg_params *prm, plqcd_cg_stats *stats)
{
   using clock = std::chrono::steady_clock;
   const auto t_start = clock::now();
   plqcd::clear_error();
   if (!g || !b || !x || !prm || !stats) return plqcd::set_error(1, "cg: null argument");
   const plqcd_hip_gauge *gs = g_sloppy ? g_sloppy : g;
   plqcd_hip_lattice *lat = g->lat;
   const int R = b->nrhs;
   if (g->prec != PLQCD_PREC_F64 || b->prec != PLQCD_PREC_F64 || x->prec != PLQCD_PREC_F64)
      return plqcd::set_error(1, "cg: the gauge field, b and x must be F64");
   if (gs->lat != lat || b->lat != lat || x->lat != lat) return plqcd::set_error(1, "cg: lattice mismatch");
   if (x->nrhs != R) return plqcd::set_error(1, "cg: b and x have different RHS counts");
   if (R > kMaxRhs) return plqcd::set_error(2, "cg: at most %d right-hand sides per solve", kMaxRhs);
   if (prm->max_iter <= 0) return plqcd::set_error(2, "cg: max_iter must be positive");
   const int sloppy = gs->prec;
   const bool mixed = gs != g;
   double delta = prm->delta;
   if (delta <= 0.0) delta = !mixed ? 0.0 : (sloppy == PLQCD_PREC_F16 ? 0.1 : 1e-3);

   Workspace w;
   if (int rc = alloc_workspace(w, lat, sloppy, R)) return rc;
   hipStream_t st = lat->stream;
   const int nc = plqcd::reduce_chunks(lat->Vh);

   std::vector<double> kap(kMaxRhs, 0.0), tol(R), bb(R), rmax(R), rr_new(R);
   std::vector<int> active(kMaxRhs, 0), update(kMaxRhs, 0);
   for (int r = 0; r < R; r++) {
      kap[r] = prm->kappa_rhs ? prm->kappa_rhs[r] : prm->kappa;
      tol[r] = prm->tol_rhs ? prm->tol_rhs[r] : prm->tol;
      stats[r] = plqcd_cg_stats{};
   }
   PLQCD_HIP_TRY(hipMemcpyAsync(w.kappa(), kap.data(), kMaxRhs * sizeof(double), hipMemcpyHostToDevice, st));

   // x = 0 ; r = b ; |b|^2 ; rs = r ; ps = 0 ; xs = 0 ; rr = rr_new = |b|^2
   int rc = plqcd::blas_zero(x, nullptr);
   if (!rc) rc = plqcd::blas_copy(w.r, b, w.partial, nullptr);
   if (!rc) rc = plqcd::reduce_partials(w.partial, 2 * nc, 0, 2 * nc, R, w.rr(), st);
   if (!rc) rc = plqcd::blas_copy(w.rs, w.r, nullptr, nullptr);
   if (!rc) rc = plqcd::blas_zero(w.ps, nullptr);
   if (!rc) rc = plqcd::blas_zero(w.xs, nullptr);
   if (rc) return rc;
   PLQCD_HIP_TRY(hipMemcpyAsync(w.rr_new(), w.rr(), R * sizeof(double), hipMemcpyDeviceToDevice, st));
   PLQCD_HIP_TRY(hipMemcpyAsync(w.host, w.rr(), R * sizeof(double), hipMemcpyDeviceToHost, st));
   PLQCD_HIP_TRY(hipStreamSynchronize(st));
   int nactive = 0;
   for (int r = 0; r < R; r++) {
      bb[r] = w.host[r];
      rmax[r] = std::sqrt(bb[r]);
      active[r] = bb[r] > 0.0;
      if (!active[r]) stats[r].converged = 1;  // b = 0: x = 0 is exact
      nactive += active[r];
   }
   PLQCD_HIP_TRY(hipMemcpyAsync(w.active(), active.data(), kMaxRhs * sizeof(int), hipMemcpyHostToDevice, st));

   // capture one iteration
   if (prm->use_graph) {
      PLQCD_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      rc = iteration(w, gs);
      hipGraph_t gdef = nullptr;
      hipError_t e = hipStreamEndCapture(st, &gdef);
      if (rc) {
         if (gdef) (void)hipGraphDestroy(gdef);
         return rc;
      }
      if (e != hipSuccess) return plqcd::set_error(10, "cg: graph capture failed: %s", hipGetErrorString(e));
      w.graph_def = gdef;
      PLQCD_HIP_TRY(hipGraphInstantiate(&w.graph, gdef, nullptr, nullptr, 0));
   }

   int iter = 0;
   while (nactive > 0 && iter < prm->max_iter) {
      if (w.graph) PLQCD_HIP_TRY(hipGraphLaunch(w.graph, st));
      else if ((rc = iteration(w, gs))) return rc;
      PLQCD_HIP_TRY(hipStreamSynchronize(st));
      iter++;
      const bool last = iter == prm->max_iter;
      int nupd = 0;
      for (int r = 0; r < R; r++) {
         update[r] = 0;
         if (!active[r]) continue;
         stats[r].iterations++;
         rr_new[r] = w.host[r];
         const double res = std::sqrt(rr_new[r]);
         if (res < delta * rmax[r] || res <= tol[r] * std::sqrt(bb[r]) || last || !std::isfinite(res)) {
            update[r] = 1;
            nupd++;
         }
      }
      if (nupd) {
         PLQCD_HIP_TRY(hipMemcpyAsync(w.update(), update.data(), kMaxRhs * sizeof(int), hipMemcpyHostToDevice, st));
         if ((rc = reliable_update(w, g, b, x))) return rc;
         for (int r = 0; r < R; r++) {
            if (!update[r]) continue;
            const double rt = w.host[kMaxRhs + r];
            stats[r].reliable_updates++;
            stats[r].true_residual = std::sqrt(rt / bb[r]);
            rr_new[r] = rt;
            rmax[r] = std::sqrt(rt);
            if (stats[r].true_residual <= tol[r]) {
               stats[r].converged = 1;
               active[r] = 0;
               nactive--;
            } else if (last || !std::isfinite(rt)) {
               active[r] = 0;
               nactive--;
            }
            w.host[r] = rr_new[r];
         }
         // the true |r|^2 replaces the iterated one before the next beta
         PLQCD_HIP_TRY(hipMemcpyAsync(w.rr_new(), w.host, R * sizeof(double), hipMemcpyHostToDevice, st));
         PLQCD_HIP_TRY(hipMemcpyAsync(w.active(), active.data(), kMaxRhs * sizeof(int), hipMemcpyHostToDevice, st));
         PLQCD_HIP_TRY(hipStreamSynchronize(st));
      }
   }
   const double secs = std::chrono::duration<double>(clock::now() - t_start).count();
   for (int r = 0; r < R; r++) {
      stats[r].seconds = secs;
      stats[r].hop_applications = 2.0 * (stats[r].iterations + stats[r].reliable_updates);
   }
   return 0;
}

extern "C" int plqcd_hip_apply_m(const plqcd_hip_gauge *g, const plqcd_hip_field *in, plqcd_hip_field *out,
                                 const double *kappa_rhs, int dagger)
{
   if (!g || !in || !out || !kappa_rhs) return plqcd::set_error(1, "apply_m: null argument");
   const int R = in->nrhs;
   if (R > kMaxRhs) return plqcd::set_error(2, "apply_m: at most %d right-hand sides", kMaxRhs);
   plqcd_hip_lattice *lat = g->lat;
   double *dk = nullptr;
   PLQCD_HIP_TRY(hipMallocAsync((void **)&dk, R * sizeof(double), lat->stream));
   hipError_t e = hipMemcpyAsync(dk, kappa_rhs, R * sizeof(double), hipMemcpyHostToDevice, lat->stream);
   int rc = e == hipSuccess ? 0 : plqcd::set_error(10, "apply_m: %s", hipGetErrorString(e));
   plqcd::DslashOp op;
   op.kappa = dk;
   op.dag = dagger ? 1 : 0;
   if (!rc) rc = plqcd::dslash(g, in, out, PLQCD_PARITY_ALL, PLQCD_HOP_WILSON, op);
   (void)hipFreeAsync(dk, lat->stream);
   if (!rc) PLQCD_HIP_TRY(hipStreamSynchronize(lat->stream));
   return rc;
}
