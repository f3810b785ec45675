// Conjugate gradient on the normal equations  M^+ M x = b,  M = 1 - kappa H
// (wilson mode, M^+ = gamma5 M gamma5), for a set of right-hand-side slots
// with per-slot kappa and tolerance.
//
// The iteration runs in the "sloppy" precision of the sloppy gauge field
// (F64, F32 or F16 storage; F32/F64 arithmetic) and is corrected by reliable
// updates (Sleijpen and van der Vorst 1996, used for mixed precision by Clark
// et al. 2010): when the iterated residual of a slot has dropped by `delta`
// since its last update, or looks converged, the partial solution is added to
// the FP64 solution, the true residual b - M^+M x is recomputed in FP64 and
// replaces the iterated one.  A slot stops when its FP64 true residual meets
// its tolerance.
//
// One iteration is a fixed sequence of kernels whose scalars live on the
// device: beta; one pass that applies the previous iteration's x += alpha p
// and forms p = r + beta p; M p with |Mp|^2; alpha; then M^+ (Mp) with the
// residual update r -= alpha Ap and |r|^2 fused into its epilogue, so Ap is
// never stored.  The iterate therefore trails by one step: a reliable update
// first applies the pending x += alpha p of the slots it updates and zeroes
// their alpha, so the next pass adds nothing twice.  The arithmetic of each
// update is unchanged, and so are the iterates.  The sequence is captured
// once into a hipGraph and replayed; only |r|^2 comes back to the host each
// iteration, to decide reliable updates.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <new>
#include <vector>

#include "blas.h"
#include "cg_engine.h"
#include "dslash.h"
#include "reduce.h"

namespace {

constexpr int kMaxSlots = 32;

__global__ void k_alpha(const double *rr, const double *pap, const int *active, double *alpha, int S)
{
   const int r = threadIdx.x;
   if (r < S) alpha[r] = (active[r] && pap[r] > 0.0) ? rr[r] / pap[r] : 0.0;
}

__global__ void k_beta(double *rr, const double *rr_new, const int *active, double *beta, int S)
{
   const int r = threadIdx.x;
   if (r < S) {
      beta[r] = (active[r] && rr[r] > 0.0) ? rr_new[r] / rr[r] : 0.0;
      rr[r] = rr_new[r];
   }
}

// rr[s] = rr_new[s] = src[s] and alpha[s] = 0 (no pending iterate update)
// for the slots in mask
__global__ void k_admit(const int *mask, const double *src, double *rr, double *rr_new, double *alpha, int S)
{
   const int r = threadIdx.x;
   if (r < S && mask[r]) {
      rr[r] = rr_new[r] = src[r];
      alpha[r] = 0.0;
   }
}

// v[s] = 0 for the slots in mask
__global__ void k_clear(const int *mask, double *v, int S)
{
   const int r = threadIdx.x;
   if (r < S && mask[r]) v[r] = 0.0;
}

}  // namespace

namespace plqcd {

// The smallest window of slots holding every slot with flag[s] != 0.
Window window_of(const std::vector<int> &flag)
{
   int lo = -1, hi = -1;
   for (int s = 0; s < (int)flag.size(); s++)
      if (flag[s]) {
         if (lo < 0) lo = s;
         hi = s;
      }
   return lo < 0 ? Window{0, 0} : Window{lo, hi - lo + 1};
}

size_t CgEngine::bytes_needed(const plqcd_hip_lattice *lat, int sloppy, int S)
{
   const size_t f64 = 2 * spinor_parity_bytes(PLQCD_PREC_F64, lat->Vh) * S;
   const size_t fs = 2 * spinor_parity_bytes(sloppy, lat->Vh) * S;
   const size_t nf64 = sloppy == PLQCD_PREC_F64 ? 2 : 3;  // b, x (+ t1)
   return nf64 * f64 + 4 * fs + (size_t)S * 2 * reduce_chunks(lat->Vh) * sizeof(double) + (1 << 20);
}

int CgEngine::create(const plqcd_hip_gauge *g, const plqcd_hip_gauge *gs, int S, double delta, bool use_graph,
                     plqcd_hip_field *b, plqcd_hip_field *x, CgEngine **out)
{
   *out = nullptr;
   if (!g || !gs) return set_error(1, "cg: null gauge");
   if (g->prec != PLQCD_PREC_F64) return set_error(1, "cg: the true-residual gauge field must be F64");
   if (gs->lat != g->lat) return set_error(1, "cg: gauge fields on different lattices");
   if (S < 1 || S > kMaxSlots) return set_error(2, "cg: 1 to %d right-hand sides per engine", kMaxSlots);
   if ((b == nullptr) != (x == nullptr)) return set_error(1, "cg: borrow both b and x or neither");
   if (b && (b->prec != PLQCD_PREC_F64 || x->prec != PLQCD_PREC_F64 || b->nrhs != S || x->nrhs != S ||
             b->lat != g->lat || x->lat != g->lat))
      return set_error(1, "cg: b and x must be F64 fields with the engine's RHS count on the gauge's lattice");
   CgEngine *e = new (std::nothrow) CgEngine();
   if (!e) return set_error(4, "out of host memory");
   e->lat_ = g->lat;
   e->g_ = g;
   e->gs_ = gs;
   e->S_ = S;
   e->use_graph_ = use_graph;
   const bool mixed = gs != g;
   e->delta_ = delta > 0 ? delta : (!mixed ? 0.0 : (gs->prec == PLQCD_PREC_F16 ? 1e-2 : 1e-3));
   e->b_ = b;
   e->x_ = x;
   e->own_bx_ = b == nullptr;
   e->slot_.assign(S, CgSlot{});
   e->graphs_.assign(S + 1, Graph{});
   e->active_.assign(S, 0);
   e->update_.assign(S, 0);
   if (int rc = e->alloc(gs->prec)) {
      delete e;
      return rc;
   }
   *out = e;
   return 0;
}

int CgEngine::alloc(int sloppy)
{
   if (own_bx_) {
      if (int rc = plqcd_hip_field_create(lat_, PLQCD_PREC_F64, S_, &b_)) return rc;
      if (int rc = plqcd_hip_field_create(lat_, PLQCD_PREC_F64, S_, &x_)) return rc;
   }
   for (plqcd_hip_field **f : {&rs_, &ps_, &qs_, &xs_})
      if (int rc = plqcd_hip_field_create(lat_, sloppy, S_, f)) return rc;
   if (sloppy == PLQCD_PREC_F64) {
      t1_ = qs_;  // free during a reliable update
   } else {
      if (int rc = plqcd_hip_field_create(lat_, PLQCD_PREC_F64, S_, &t1_)) return rc;
   }
   PLQCD_HIP_TRY(hipMalloc(&partial_, partial_doubles(b_) * sizeof(double)));
   PLQCD_HIP_TRY(hipMalloc(&scal_, 7 * S_ * sizeof(double)));
   PLQCD_HIP_TRY(hipMalloc(&imask_, 3 * S_ * sizeof(int)));
   PLQCD_HIP_TRY(hipHostMalloc(&host_, 2 * S_ * sizeof(double), hipHostMallocDefault));
   PLQCD_HIP_TRY(hipHostMalloc(&hmask_, 3 * S_ * sizeof(int), hipHostMallocDefault));
   PLQCD_HIP_TRY(hipEventCreateWithFlags(&ev_, hipEventDisableTiming));
   hipStream_t st = lat_->stream;
   std::vector<double> z(7 * S_, 0.0);
   PLQCD_HIP_TRY(hipMemcpyAsync(scal_, z.data(), z.size() * sizeof(double), hipMemcpyHostToDevice, st));
   std::vector<int> zi(3 * S_, 0);
   PLQCD_HIP_TRY(hipMemcpyAsync(imask_, zi.data(), zi.size() * sizeof(int), hipMemcpyHostToDevice, st));
   if (int rc = blas_zero(ps_, nullptr)) return rc;
   if (int rc = blas_zero(xs_, nullptr)) return rc;
   if (int rc = blas_zero(rs_, nullptr)) return rc;
   PLQCD_HIP_TRY(hipStreamSynchronize(st));
   return 0;
}

CgEngine::~CgEngine()
{
   if (lat_) (void)hipStreamSynchronize(lat_->stream);
   for (Graph &gr : graphs_) {
      if (gr.exec) (void)hipGraphExecDestroy(gr.exec);
      if (gr.def) (void)hipGraphDestroy(gr.def);
   }
   const bool alias = (t1_ == qs_);
   for (plqcd_hip_field *f : {rs_, ps_, qs_, xs_}) plqcd_hip_field_destroy(f);
   if (!alias) plqcd_hip_field_destroy(t1_);
   if (own_bx_) {
      plqcd_hip_field_destroy(b_);
      plqcd_hip_field_destroy(x_);
   }
   if (partial_) (void)hipFree(partial_);
   if (scal_) (void)hipFree(scal_);
   if (imask_) (void)hipFree(imask_);
   if (host_) (void)hipHostFree(host_);
   if (hmask_) (void)hipHostFree(hmask_);
   if (ev_) (void)hipEventDestroy(ev_);
}

int CgEngine::free_slot() const
{
   for (int s = 0; s < S_; s++)
      if (!slot_[s].active && !slot_[s].done) return s;
   return -1;
}

int CgEngine::active_count() const
{
   int n = 0;
   for (const CgSlot &s : slot_) n += s.active;
   return n;
}

int CgEngine::admit(const std::vector<int> &which, const double *kappa, const double *tol, const int *max_iter)
{
   hipStream_t st = lat_->stream;
   const int nc = reduce_chunks(lat_->Vh);
   std::vector<int> mask(S_, 0);
   for (size_t k = 0; k < which.size(); k++) {
      const int s = which[k];
      if (s < 0 || s >= S_ || slot_[s].active || slot_[s].done) return set_error(1, "cg: slot %d is not free", s);
      mask[s] = 1;
      CgSlot &sl = slot_[s];
      sl = CgSlot{};
      sl.kappa = kappa[k];
      sl.tol = tol[k];
      sl.max_iter = max_iter[k];
   }
   PLQCD_HIP_TRY(hipMemcpyAsync(update_d(), mask.data(), S_ * sizeof(int), hipMemcpyHostToDevice, st));
   const Window w = window_of(mask);
   if (w.nr == 0) return 0;
   // x = xs = ps = 0 ; rs = b ; |b|^2 (in FP64, before rounding to the sloppy precision)
   if (int rc = blas_zero(x_, update_d(), w)) return rc;
   if (int rc = blas_zero(xs_, update_d(), w)) return rc;
   if (int rc = blas_zero(ps_, update_d(), w)) return rc;
   if (int rc = blas_copy(rs_, b_, partial_, update_d(), w)) return rc;
   if (int rc = reduce_partials(partial_, 2 * nc, 0, 2 * nc, w.nr, tmp_d(), st, w.r0)) return rc;
   hipLaunchKernelGGL(k_admit, dim3(1), dim3(kMaxSlots), 0, st, update_d(), tmp_d(), rr_d(), rr_new_d(), alpha_d(),
                      S_);
   PLQCD_HIP_TRY(hipGetLastError());
   PLQCD_HIP_TRY(hipMemcpyAsync(host_ + S_, tmp_d(), S_ * sizeof(double), hipMemcpyDeviceToHost, st));
   PLQCD_HIP_TRY(hipStreamSynchronize(st));
   std::vector<double> kap(S_);
   for (int s = 0; s < S_; s++) kap[s] = slot_[s].kappa;
   for (int s : which) {
      CgSlot &sl = slot_[s];
      sl.bb = host_[S_ + s];
      sl.rmax = std::sqrt(sl.bb);
      if (sl.bb > 0.0) {
         sl.active = true;
      } else {  // b = 0: x = 0 is exact
         sl.done = true;
         sl.stats.converged = 1;
      }
      active_[s] = sl.active;
   }
   PLQCD_HIP_TRY(hipMemcpyAsync(kappa_d(), kap.data(), S_ * sizeof(double), hipMemcpyHostToDevice, st));
   PLQCD_HIP_TRY(hipMemcpyAsync(active_d(), active_.data(), S_ * sizeof(int), hipMemcpyHostToDevice, st));
   PLQCD_HIP_TRY(hipStreamSynchronize(st));
   return 0;
}

int CgEngine::admit_host(const std::vector<int> &which, const std::vector<const spinor *> &b_host, double kappa,
                         double tol, int max_iter)
{
   if (which.size() != b_host.size()) return set_error(1, "cg: one source per admitted slot");
   for (size_t k = 0; k < which.size(); k++) {
      if (which[k] < 0 || which[k] >= S_) return set_error(1, "cg: bad slot %d", which[k]);
      if (int rc = field_upload_range(b_, which[k], b_host[k], 0, lat_->V)) return rc;
   }
   const std::vector<double> kap(which.size(), kappa), tl(which.size(), tol);
   const std::vector<int> mi(which.size(), max_iter);
   return admit(which, kap.data(), tl.data(), mi.data());
}

// One CG iteration on the sloppy fields, for the slots 0 .. nr-1 (every
// active slot lies below nr; inactive ones inside get zero steps).  Device
// work only: this sequence is what the iteration graphs capture, and no graph
// holds a copy to or from host memory (launch() reads |r|^2 back after it).
int CgEngine::iteration(int nr)
{
   hipStream_t st = lat_->stream;
   const int nc = reduce_chunks(lat_->Vh);
   const Window w{0, nr};
   hipLaunchKernelGGL(k_beta, dim3(1), dim3(kMaxSlots), 0, st, rr_d(), rr_new_d(), active_d(), beta_d(), S_);
   // xs += alpha ps (alpha of the last iteration), then ps = rs + beta ps
   if (int rc = blas_cg_step(rs_, beta_d(), ps_, alpha_d(), xs_, w)) return rc;
   DslashOp m;
   m.nr = nr;
   m.kappa = kappa_d();
   m.partial = partial_;
   if (int rc = dslash(gs_, ps_, qs_, PLQCD_PARITY_ALL, PLQCD_HOP_WILSON, m)) return rc;
   if (int rc = reduce_partials(partial_, 2 * nc, 0, 2 * nc, nr, pap_d(), st)) return rc;
   hipLaunchKernelGGL(k_alpha, dim3(1), dim3(kMaxSlots), 0, st, rr_d(), pap_d(), active_d(), alpha_d(), S_);
   // Ap = M^+ q, fused with rs -= alpha Ap and |rs|^2
   DslashOp md;
   md.nr = nr;
   md.kappa = kappa_d();
   md.dag = 1;
   md.partial = partial_;
   md.alpha = alpha_d();
   if (int rc = dslash(gs_, qs_, rs_, PLQCD_PARITY_ALL, PLQCD_HOP_WILSON, md)) return rc;
   if (int rc = reduce_partials(partial_, 2 * nc, 0, 2 * nc, nr, rr_new_d(), st)) return rc;
   PLQCD_HIP_TRY(hipGetLastError());
   return 0;
}

int CgEngine::capture(int nr)
{
   hipStream_t st = lat_->stream;
   PLQCD_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
   const int rc = iteration(nr);
   hipGraph_t gdef = nullptr;
   const hipError_t e = hipStreamEndCapture(st, &gdef);
   if (rc) {
      if (gdef) (void)hipGraphDestroy(gdef);
      return rc;
   }
   if (e != hipSuccess) return set_error(10, "cg: graph capture failed: %s", hipGetErrorString(e));
   Graph &gr = graphs_[nr];
   gr.def = gdef;
   PLQCD_HIP_TRY(hipGraphInstantiate(&gr.exec, gdef, nullptr, nullptr, 0));
   return 0;
}

int CgEngine::capture_all()
{
   if (!use_graph_) return 0;
   for (int nr = 1; nr <= S_; nr++)
      if (!graphs_[nr].exec)
         if (int rc = capture(nr)) return rc;
   return 0;
}

void CgEngine::abort_active(std::vector<int> &finished)
{
   for (int s = 0; s < S_; s++) {
      CgSlot &sl = slot_[s];
      if (!sl.active) continue;
      sl.active = false;
      sl.done = true;
      sl.stats.converged = 0;
      active_[s] = 0;
      update_[s] = 0;
      finished.push_back(s);
   }
}

int CgEngine::launch()
{
   // slots 0 .. nr-1 hold every active slot (admission fills the lowest free slot)
   const Window act = window_of(active_);
   const int nr = act.r0 + act.nr;
   if (nr == 0) return 0;
   if (use_graph_) {
      if (!graphs_[nr].exec)
         if (int rc = capture(nr)) return rc;
      PLQCD_HIP_TRY(hipGraphLaunch(graphs_[nr].exec, lat_->stream));
   } else {
      if (int rc = iteration(nr)) return rc;
   }
   // |r|^2 of the iteration back to the host, on the stream after the graph
   // (a memcpy node into host memory inside the graph is executed by the
   // runtime at hipGraphLaunch; a host SIGSEGV was observed there once)
   PLQCD_HIP_TRY(hipMemcpyAsync(host_, rr_new_d(), S_ * sizeof(double), hipMemcpyDeviceToHost, lat_->stream));
   PLQCD_HIP_TRY(hipEventRecord(ev_, lat_->stream));
   return 0;
}

// xs += alpha ps (pending) ; x += xs ; r = b - M^+M x ; rs = r ; tmp <- |r|^2,
// for the slots in update.
// The FP64 operators run once over the list of slots being updated (the
// residual epilogue forms the residual of every RHS it covers, see EPI 3).
int CgEngine::reliable_update()
{
   hipStream_t st = lat_->stream;
   const int nc = reduce_chunks(lat_->Vh);
   const Window w = window_of(update_);
   if (w.nr == 0) return 0;
   // the pending xs += alpha ps of the last iteration, then alpha = 0 so
   // that the next iteration does not apply it again
   if (int rc = blas_x_step(alpha_d(), ps_, xs_, update_d(), w)) return rc;
   hipLaunchKernelGGL(k_clear, dim3(1), dim3(kMaxSlots), 0, st, update_d(), alpha_d(), S_);
   PLQCD_HIP_TRY(hipGetLastError());
   if (int rc = blas_accum(x_, xs_, 1, update_d(), w)) return rc;
   // every slot being updated in one launch of M and one of the residual, over
   // the list of those slots (begin_finish uploaded it): each link is loaded
   // once for all of them however the slots lie in the engine
   int nupd = 0;
   for (int s = w.r0; s < w.r0 + w.nr; s++) nupd += update_[s] ? 1 : 0;
   DslashOp m;
   m.rlist = ulist_d();
   m.nr = nupd;
   m.kappa = kappa_d();
   if (int rc = dslash(g_, x_, t1_, PLQCD_PARITY_ALL, PLQCD_HOP_WILSON, m)) return rc;
   // r = b - M^+ t1 in FP64, stored straight into the iterate's residual in
   // its own precision, with |r|^2
   m.dag = 1;
   m.partial = partial_;
   m.resid_b = b_;
   if (int rc = dslash(g_, t1_, rs_, PLQCD_PARITY_ALL, PLQCD_HOP_WILSON, m)) return rc;
   for (int s = w.r0; s < w.r0 + w.nr;) {  // |r|^2 per run of consecutive slots
      if (!update_[s]) {
         s++;
         continue;
      }
      int e = s;
      while (e < w.r0 + w.nr && update_[e]) e++;
      if (int rc = reduce_partials(partial_, 2 * nc, 0, 2 * nc, e - s, tmp_d(), st, s)) return rc;
      s = e;
   }
   PLQCD_HIP_TRY(hipMemcpyAsync(host_ + S_, tmp_d(), S_ * sizeof(double), hipMemcpyDeviceToHost, st));
   return 0;
}

int CgEngine::begin_finish(bool &updating)
{
   hipStream_t st = lat_->stream;
   updating = false;
   int nupd = 0;
   for (int s = 0; s < S_; s++) {
      update_[s] = 0;
      CgSlot &sl = slot_[s];
      if (!sl.active) continue;
      sl.stats.iterations++;
      const double res = std::sqrt(host_[s]);
      if (res < delta_ * sl.rmax || res <= sl.tol * std::sqrt(sl.bb) || sl.stats.iterations >= sl.max_iter ||
          !std::isfinite(res)) {
         update_[s] = 1;
         nupd++;
      }
   }
   if (!nupd) return 0;
   std::copy(update_.begin(), update_.end(), hmask_ + S_);
   for (int s = 0, k = 0; s < S_; s++)  // the slots being updated, in order
      if (update_[s]) hmask_[2 * S_ + k++] = s;
   PLQCD_HIP_TRY(hipMemcpyAsync(update_d(), hmask_ + S_, 2 * S_ * sizeof(int), hipMemcpyHostToDevice, st));
   if (int rc = reliable_update()) return rc;
   PLQCD_HIP_TRY(hipEventRecord(ev_, st));
   updating = true;
   return 0;
}

int CgEngine::end_finish(std::vector<int> &finished)
{
   hipStream_t st = lat_->stream;
   bool any = false;
   for (int s = 0; s < S_; s++) {
      if (!update_[s]) continue;
      any = true;
      CgSlot &sl = slot_[s];
      const double rt = host_[S_ + s];
      sl.stats.reliable_updates++;
      sl.stats.true_residual = std::sqrt(rt / sl.bb);
      sl.rmax = std::sqrt(rt);
      host_[s] = rt;  // the true |r|^2 replaces the iterated one before the next beta
      if (sl.stats.true_residual <= sl.tol) sl.stats.converged = 1;
      if (sl.stats.converged || sl.stats.iterations >= sl.max_iter || !std::isfinite(rt)) {
         sl.active = false;
         sl.done = true;
         active_[s] = 0;
         finished.push_back(s);
      }
      update_[s] = 0;
   }
   if (!any) return 0;
   // read by the next iteration, which the stream orders after these copies;
   // the host rewrites host_/hmask_ only after that iteration has completed
   std::copy(active_.begin(), active_.end(), hmask_);
   PLQCD_HIP_TRY(hipMemcpyAsync(rr_new_d(), host_, S_ * sizeof(double), hipMemcpyHostToDevice, st));
   PLQCD_HIP_TRY(hipMemcpyAsync(active_d(), hmask_, S_ * sizeof(int), hipMemcpyHostToDevice, st));
   return 0;
}

bool CgEngine::ready() const
{
   const hipError_t e = hipEventQuery(ev_);
   return e != hipErrorNotReady;  // errors surface at the next synchronizing call
}

int CgEngine::finish(std::vector<int> &finished)
{
   hipStream_t st = lat_->stream;
   PLQCD_HIP_TRY(hipStreamSynchronize(st));
   bool updating = false;
   if (int rc = begin_finish(updating)) return rc;
   if (updating) PLQCD_HIP_TRY(hipStreamSynchronize(st));
   if (int rc = end_finish(finished)) return rc;
   PLQCD_HIP_TRY(hipStreamSynchronize(st));
   return 0;
}

int CgEngine::retire(int s, spinor *x_host, plqcd_cg_stats *st)
{
   if (s < 0 || s >= S_ || !slot_[s].done) return set_error(1, "cg: slot %d has no result", s);
   if (x_host)
      if (int rc = field_download_range(x_, s, x_host, 0, lat_->V)) return rc;
   CgSlot &sl = slot_[s];
   sl.stats.hop_applications = 2.0 * (sl.stats.iterations + sl.stats.reliable_updates);
   if (st) *st = sl.stats;
   sl = CgSlot{};
   return 0;
}

}  // namespace plqcd

extern "C" int plqcd_hip_cg_solve(const plqcd_hip_gauge *g, const plqcd_hip_gauge *g_sloppy, const plqcd_hip_field *b,
                                  plqcd_hip_field *x, const plqcd_cg_params *prm, plqcd_cg_stats *stats)
{
   using clock = std::chrono::steady_clock;
   const auto t_start = clock::now();
   plqcd::clear_error();
   if (!g || !b || !x || !prm || !stats) return plqcd::set_error(1, "cg: null argument");
   const plqcd_hip_gauge *gs = g_sloppy ? g_sloppy : g;
   const int R = b->nrhs;
   if (x->nrhs != R) return plqcd::set_error(1, "cg: b and x have different RHS counts");
   if (prm->max_iter <= 0) return plqcd::set_error(2, "cg: max_iter must be positive");
   plqcd::CgEngine *e = nullptr;
   if (int rc = plqcd::CgEngine::create(g, gs, R, prm->delta, prm->use_graph != 0,
                                        const_cast<plqcd_hip_field *>(b), x, &e))
      return rc;
   std::vector<int> all(R);
   std::vector<double> kap(R), tol(R);
   std::vector<int> mi(R, prm->max_iter);
   for (int r = 0; r < R; r++) {
      all[r] = r;
      kap[r] = prm->kappa_rhs ? prm->kappa_rhs[r] : prm->kappa;
      tol[r] = prm->tol_rhs ? prm->tol_rhs[r] : prm->tol;
   }
   int rc = e->admit(all, kap.data(), tol.data(), mi.data());
   std::vector<int> finished;
   while (!rc && e->active_count() > 0) {
      rc = e->launch();
      if (!rc) rc = e->finish(finished);
   }
   const double secs = std::chrono::duration<double>(clock::now() - t_start).count();
   for (int r = 0; r < R && !rc; r++) {
      rc = e->retire(r, nullptr, &stats[r]);
      stats[r].seconds = secs;
   }
   delete e;
   return rc;
}

extern "C" int plqcd_hip_apply_m(const plqcd_hip_gauge *g, const plqcd_hip_field *in, plqcd_hip_field *out,
                                 const double *kappa_rhs, int dagger)
{
   if (!g || !in || !out || !kappa_rhs) return plqcd::set_error(1, "apply_m: null argument");
   const int R = in->nrhs;
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

extern "C" int plqcd_hip_apply_schur(const plqcd_hip_gauge *g, const plqcd_hip_field *in, plqcd_hip_field *tmp,
                                     plqcd_hip_field *out, const double *kappa_rhs, int parity)
{
   if (!g || !in || !tmp || !out || !kappa_rhs) return plqcd::set_error(1, "apply_schur: null argument");
   if (parity != PLQCD_PARITY_EVEN && parity != PLQCD_PARITY_ODD)
      return plqcd::set_error(1, "apply_schur: parity must be EVEN or ODD");
   if (in == tmp || in == out || tmp == out) return plqcd::set_error(1, "apply_schur: fields must be distinct");
   const int R = in->nrhs;
   plqcd_hip_lattice *lat = g->lat;
   std::vector<double> k2(R);
   for (int r = 0; r < R; r++) k2[r] = kappa_rhs[r] * kappa_rhs[r];
   double *dk = nullptr;
   PLQCD_HIP_TRY(hipMallocAsync((void **)&dk, R * sizeof(double), lat->stream));
   hipError_t e = hipMemcpyAsync(dk, k2.data(), R * sizeof(double), hipMemcpyHostToDevice, lat->stream);
   int rc = e == hipSuccess ? 0 : plqcd::set_error(10, "apply_schur: %s", hipGetErrorString(e));
   const int other = parity == PLQCD_PARITY_EVEN ? PLQCD_PARITY_ODD : PLQCD_PARITY_EVEN;
   // tmp(other) = H in(parity) ; out(parity) = in(parity) - kappa^2 H tmp(other),
   // the second half hop with the centre taken from in
   plqcd::DslashOp m;
   m.kappa = dk;
   m.centre = in;
   if (!rc) rc = plqcd_hip_hop(g, in, tmp, other, PLQCD_HOP_WILSON);
   if (!rc) rc = plqcd::dslash(g, tmp, out, parity, PLQCD_HOP_WILSON, m);
   (void)hipFreeAsync(dk, lat->stream);
   if (!rc) PLQCD_HIP_TRY(hipStreamSynchronize(lat->stream));
   return rc;
}
