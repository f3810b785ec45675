// Autotuners of the hopping kernels' launch shape.
//
// For one (lattice, storage precision, link format, RHS count) they time every
// candidate launch shape -- block->tile order (plain, contiguous per XCD, XCD
// runs of 8 or 16 tiles), threads per block, t-march length (gather kernel or
// T t-slices per thread, T dividing the time extent), right-hand sides per
// lane (RHS-blocked gather) and, for one RHS, the LDS tile of the t-march
// (y and z extents whose whole-x-row tiles give 128, 256 or 512 lanes) --
// with HIP events on the lattice stream, keep the fastest in the per-shape
// cache that every later launch of that shape consults, and return it.
// plqcd_hip_autotune times plain applications of H; plqcd_hip_autotune_solver
// times the pair of launches of one CG iteration (M p with |q|^2, then M^+ q
// fused with the residual update and |r|^2), whose shapes are cached apart.
#include <vector>

#include "blas.h"
#include "dslash.h"
#include "reduce.h"

namespace {

std::vector<HopTuning> candidates(const plqcd_hip_gauge *g, const plqcd_hip_field *in)
{
   const plqcd_hip_lattice *lat = g->lat;
   struct Cand {
      int order, group;
   };
   const std::vector<Cand> orders = {{0, 1}, {1, 1}, {2, 8}, {2, 16}};
   const std::vector<int> blocks = {64, 128, 256, 512};
   std::vector<int> marches = {0};
   for (int T : {2, 4, 8, 16, 32})
      if (lat->L[3] % T == 0) marches.push_back(T);
   std::vector<HopTuning> cands;
   for (int rb : {1, 2, 4}) {
      if (in->nrhs % rb) continue;
      for (int T : marches) {
         if (rb > 1 && T > 0) continue;  // the t-march kernels run one RHS per lane
         for (int blk : blocks)
            for (const Cand &o : orders) cands.push_back(HopTuning{o.order, o.group, blk, T, rb, 0, 0});
      }
   }
   if (in->nrhs == 1) {
      const int maxl = g->prec == PLQCD_PREC_F64 ? 256 : 512;
      for (int T : marches) {
         if (T == 0) continue;
         for (int ty = 1; ty <= lat->L[1]; ty++)
            for (int tz = 1; tz <= lat->L[2]; tz++) {
               const int lanes = 2 * lat->Lxh * ty * tz;
               if (lat->L[1] % ty || lat->L[2] % tz || (lanes != 128 && lanes != 256 && lanes != 512) || lanes > maxl)
                  continue;
               for (const Cand &o : orders) cands.push_back(HopTuning{o.order, o.group, 256, T, 1, ty, tz});
            }
      }
   }
   return cands;
}

void shape_out(const HopTuning &c, int best[7])
{
   if (!best) return;
   const int v[7] = {c.order, c.group, c.block, c.tmarch, c.rb, c.ty, c.tz};
   for (int k = 0; k < 7; k++) best[k] = v[k];
}

}  // namespace

extern "C" int plqcd_hip_autotune(const plqcd_hip_gauge *g, const plqcd_hip_field *in, plqcd_hip_field *out,
                                  int reps, int best[7], double *best_ms)
{
   plqcd::clear_error();
   if (!g || !in || !out || reps < 1) return plqcd::set_error(1, "autotune: bad arguments");
   plqcd_hip_lattice *lat = g->lat;
   // the shapes of plain H applications; the solver's launches keep their own
   const plqcd::TuneKey key{{lat->L[0], lat->L[1], lat->L[2], lat->L[3]}, g->prec, g->links, in->nrhs,
                            plqcd::kTuneOpHop};
   const std::vector<HopTuning> cands = candidates(g, in);
   HopTuning prev{};
   const bool had = plqcd::lookup_tuning(key, &prev);
   double best_t = 1e300;
   HopTuning best_c = plqcd::default_tuning(g->prec, g->links, in->nrhs);
   hipEvent_t e0 = nullptr, e1 = nullptr;
   if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
      if (e0) (void)hipEventDestroy(e0);
      return plqcd::set_error(10, "autotune: hipEventCreate");
   }
   int rc = 0;
   for (const HopTuning &c : cands) {
      plqcd::store_tuning(key, c);
      rc = plqcd_hip_hop(g, in, out, PLQCD_PARITY_ALL, PLQCD_HOP_WILSON);  // warm-up
      if (rc) break;
      (void)hipEventRecord(e0, lat->stream);
      for (int k = 0; k < reps && !rc; k++) rc = plqcd_hip_hop(g, in, out, PLQCD_PARITY_ALL, PLQCD_HOP_WILSON);
      (void)hipEventRecord(e1, lat->stream);
      if (rc) break;
      (void)hipEventSynchronize(e1);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (ms / reps < best_t) {
         best_t = ms / reps;
         best_c = c;
      }
   }
   (void)hipEventDestroy(e0);
   (void)hipEventDestroy(e1);
   if (rc) {
      if (had) plqcd::store_tuning(key, prev);
      else plqcd::clear_tuning(&key);
      return rc;
   }
   plqcd::store_tuning(key, best_c);
   shape_out(best_c, best);
   if (best_ms) *best_ms = best_t;
   return 0;
}

extern "C" int plqcd_hip_autotune_lookup(const plqcd_hip_gauge *g, const plqcd_hip_field *in, int op, int shape[7])
{
   plqcd::clear_error();
   if (!g || !in || !shape || (op != plqcd::kTuneOpHop && op != plqcd::kTuneOpSolver)) {
      plqcd::set_error(1, "autotune_lookup: bad arguments");
      return -1;
   }
   const plqcd_hip_lattice *lat = g->lat;
   const plqcd::TuneKey key{{lat->L[0], lat->L[1], lat->L[2], lat->L[3]}, g->prec, g->links, in->nrhs, op};
   HopTuning t{};
   if (!plqcd::lookup_tuning(key, &t)) return 0;
   const int v[7] = {t.order, t.group, t.block, t.tmarch, t.rb, t.ty, t.tz};
   for (int k = 0; k < 7; k++) shape[k] = v[k];
   return 1;
}

extern "C" int plqcd_hip_autotune_clear(void)
{
   plqcd::clear_tuning(nullptr);
   return 0;
}

extern "C" int plqcd_hip_autotune_solver(const plqcd_hip_gauge *g, const plqcd_hip_field *in, int reps, int best[7],
                                         double *best_ms)
{
   plqcd::clear_error();
   if (!g || !in || reps < 1 || in->lat != g->lat || in->prec != g->prec)
      return plqcd::set_error(1, "autotune_solver: bad arguments");
   plqcd_hip_lattice *lat = g->lat;
   const int R = in->nrhs;
   const plqcd::TuneKey key{{lat->L[0], lat->L[1], lat->L[2], lat->L[3]}, g->prec, g->links, R,
                            plqcd::kTuneOpSolver};
   // scratch: q = M p, the residual r of the fused update, kappa and alpha
   // per RHS, and the chunk sums; `in` plays p and is only read
   plqcd_hip_field *q = nullptr, *r = nullptr;
   double *dk = nullptr, *partial = nullptr;
   int rc = plqcd_hip_field_create(lat, in->prec, R, &q);
   if (!rc) rc = plqcd_hip_field_create(lat, in->prec, R, &r);
   if (!rc) {
      std::vector<double> ka(2 * size_t(R), 0.0);  // kappa 0.124 (a typical Wilson value), alpha 0
      for (int k = 0; k < R; k++) ka[k] = 0.124;
      hipError_t e = hipMalloc(&dk, ka.size() * sizeof(double));
      if (e == hipSuccess) e = hipMalloc(&partial, plqcd::partial_doubles(in) * sizeof(double));
      if (e == hipSuccess) e = hipMemcpy(dk, ka.data(), ka.size() * sizeof(double), hipMemcpyHostToDevice);
      if (e == hipSuccess) e = hipMemset(r->data, 0, r->bytes);
      if (e != hipSuccess) rc = plqcd::set_error(10, "autotune_solver: %s", hipGetErrorString(e));
   }
   plqcd::DslashOp m;  // q = M p and |q|^2
   m.kappa = dk;
   m.partial = partial;
   plqcd::DslashOp md;  // M^+ q fused with r -= alpha M^+ q and |r|^2
   md.kappa = dk;
   md.dag = 1;
   md.partial = partial;
   md.alpha = dk + R;
   auto pair = [&]() {
      int c = plqcd::dslash(g, in, q, PLQCD_PARITY_ALL, PLQCD_HOP_WILSON, m);
      if (!c) c = plqcd::dslash(g, q, r, PLQCD_PARITY_ALL, PLQCD_HOP_WILSON, md);
      return c;
   };
   HopTuning prev{};
   const bool had = plqcd::lookup_tuning(key, &prev);
   double best_t = 1e300;
   HopTuning best_c = plqcd::default_tuning(g->prec, g->links, R, true);
   hipEvent_t e0 = nullptr, e1 = nullptr;
   if (!rc && hipEventCreate(&e0) != hipSuccess) rc = plqcd::set_error(10, "autotune_solver: hipEventCreate");
   if (!rc && hipEventCreate(&e1) != hipSuccess) rc = plqcd::set_error(10, "autotune_solver: hipEventCreate");
   if (!rc) {
      for (const HopTuning &c : candidates(g, in)) {
         plqcd::store_tuning(key, c);
         rc = pair();  // warm-up
         if (rc == 2) {  // no kernel of this shape for the fused epilogues: not a candidate
            rc = 0;
            continue;
         }
         if (rc) break;
         (void)hipEventRecord(e0, lat->stream);
         for (int k = 0; k < reps && !rc; k++) rc = pair();
         (void)hipEventRecord(e1, lat->stream);
         if (rc) break;
         (void)hipEventSynchronize(e1);
         float ms = 0.f;
         (void)hipEventElapsedTime(&ms, e0, e1);
         if (ms / reps < best_t) {
            best_t = ms / reps;
            best_c = c;
         }
      }
      if (!rc && best_t == 1e300) rc = plqcd::set_error(2, "autotune_solver: no launch shape applies");
   }
   {
      // also after a failed launch: launches already queued still use the scratch freed below
      const hipError_t e = hipStreamSynchronize(lat->stream);
      if (!rc && e != hipSuccess) rc = plqcd::set_error(10, "autotune_solver: %s", hipGetErrorString(e));
   }
   if (e0) (void)hipEventDestroy(e0);
   if (e1) (void)hipEventDestroy(e1);
   if (dk) (void)hipFree(dk);
   if (partial) (void)hipFree(partial);
   plqcd_hip_field_destroy(q);
   plqcd_hip_field_destroy(r);
   if (rc) {
      if (had) plqcd::store_tuning(key, prev);
      else plqcd::clear_tuning(&key);
      return rc;
   }
   plqcd::store_tuning(key, best_c);
   shape_out(best_c, best);
   if (best_ms) *best_ms = best_t;
   return 0;
}
