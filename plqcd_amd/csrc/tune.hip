// Autotuner of the hopping kernels' launch shape.
//
// For one (lattice, storage precision, link format, RHS count) it times every
// candidate launch shape -- block->tile order (plain, contiguous per XCD, XCD
// runs of 8 or 16 tiles), threads per block, t-march length (gather kernel or
// T t-slices per thread, T dividing the time extent), right-hand sides per
// lane (RHS-blocked gather) and, for one RHS, the LDS tile of the t-march
// (y and z extents whose whole-x-row tiles give 128, 256 or 512 lanes) --
// with HIP events on the lattice stream, keeps the fastest in the per-shape
// cache that every later hopping launch of that shape consults, and returns it.
#include <vector>

#include "dslash.h"

extern "C" int plqcd_hip_autotune(const plqcd_hip_gauge *g, const plqcd_hip_field *in, plqcd_hip_field *out,
                                  int reps, int best[7], double *best_ms)
{
   plqcd::clear_error();
   if (!g || !in || !out || reps < 1) return plqcd::set_error(1, "autotune: bad arguments");
   plqcd_hip_lattice *lat = g->lat;
   // the shapes of plain H applications; the solver's launches keep their own
   const plqcd::TuneKey key{{lat->L[0], lat->L[1], lat->L[2], lat->L[3]}, g->prec, g->links, in->nrhs,
                            plqcd::kTuneOpHop};
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
   HopTuning prev{};
   const bool had = plqcd::lookup_tuning(key, &prev);
   double best_t = 1e300;
   HopTuning best_c = plqcd::default_tuning(g->prec, g->links, in->nrhs);
   hipEvent_t e0, e1;
   PLQCD_HIP_TRY(hipEventCreate(&e0));
   PLQCD_HIP_TRY(hipEventCreate(&e1));
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
   if (best) {
      best[0] = best_c.order;
      best[1] = best_c.group;
      best[2] = best_c.block;
      best[3] = best_c.tmarch;
      best[4] = best_c.rb;
      best[5] = best_c.ty;
      best[6] = best_c.tz;
   }
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
