// Host side of the Wilson hopping matrix: argument checks, dispatch to the
// per-format launchers (kernels in dslash_kernels.h, instantiated once per
// storage precision and link format in dslash_<precision>_l<links>.hip), and the
// process-wide launch shapes with the per-shape cache of the autotuner.
#include <map>
#include <mutex>

#include "dslash.h"

// Built-in launch shapes, from sweeps at 32^3 x 64.  One RHS, per storage
// format: the LDS-tiled t-march for F64 (tiles of 32 x 1 x 4 sites, 32
// t-slices: 400 us with 18-real links against 414 us for 32 x 4 x 1 since
// the x and one-plane z neighbour reads lost their per-lane test, and 434 us
// for 32 x 4 x 1 and 514 us untiled before that; 12-real links: 331 us) and
// F32 (32 x 1 x 4 sites
// and 32 t-slices with 18-real links: 237 against 240 us for 32 x 8 x 1; 32 x
// 8 x 2 and 8 t-slices with 12-real links: 168 against 176 us for 32 x 4 x
// 4), for F16 with 18-real links 32 x 8 x 1 tiles over 4 t-slices in XCD runs
// of 16 tiles (115 against 117 us for the untiled t-march), and with 12-real
// links the untiled t-march over 8 t-slices in 512-lane blocks (90-102 us,
// the gather kernel 116 us).  Shapes that do not
// fit a lattice fall back to the untiled t-march or the gather kernel.  Several
// RHS per launch, per precision: XCDs sweeping the lattice together, with
// t-marches of 8 slices for F32/F16 (F16 at 12 RHS: 80 us per RHS, the
// gather kernel 87 us) and 16 for F64 (which divides 48 as well as 32 and 64;
// longer ones gain under 1 % at 32^3 x 64).
namespace {
struct FormatShape {
   int prec, links;
   HopTuning t;
};
const FormatShape kOneRhs[6] = {{PLQCD_PREC_F64, 18, {1, 1, 256, 32, 1, 1, 4}},
                                {PLQCD_PREC_F64, 12, {1, 1, 256, 32, 1, 1, 4}},
                                {PLQCD_PREC_F32, 18, {1, 1, 256, 32, 1, 1, 4}},
                                {PLQCD_PREC_F32, 12, {1, 1, 256, 8, 1, 8, 2}},
                                {PLQCD_PREC_F16, 12, {1, 1, 512, 8, 1, 0, 0}},
                                {PLQCD_PREC_F16, 18, {2, 16, 256, 4, 1, 8, 1}}};
// One RHS with the solver epilogues (M with |Mp|^2, M^+ with the CG update),
// where they differ from plain H: F64 18-real keeps 32 x 4 x 1 tiles (one CG
// iteration's pair of launches 1.149 ms at 32^3 x 64; with 32 x 1 x 4 the
// FP64 one-RHS solve took 0.0806 against 0.0769 s).
const FormatShape kOneRhsSolver[1] = {{PLQCD_PREC_F64, 18, {1, 1, 256, 32, 1, 4, 1}}};
const HopTuning kMultiRhs[3] = {{2, 8, 256, 16, 1, 0, 0}, {2, 8, 256, 8, 1, 0, 0}, {2, 8, 256, 8, 1, 0, 0}};
// Several RHS with the solver epilogues (M, and M^+ with the residual
// update): short t-marches.  F16 with 4-slice marches: the 12-RHS mixed CG at
// 32^3 x 64 takes 0.2515 s against 0.256 s with the gather kernel, 1.24
// against 1.36-1.40 s at 48^3 x 96.
const HopTuning kMultiRhsSolver[3] = {{2, 8, 256, 16, 1, 0, 0}, {2, 8, 256, 4, 1, 0, 0}, {2, 8, 256, 4, 1, 0, 0}};
bool g_tuning_set[3] = {false, false, false};  // set by plqcd_hip_hop_tuning

HopTuning one_rhs_shape(int prec, int links)
{
   for (const FormatShape &f : kOneRhs)
      if (f.prec == prec && f.links == links) return f.t;
   for (const FormatShape &f : kOneRhs)
      if (f.prec == prec) return f.t;
   return kOneRhs[0].t;
}
}  // namespace

// The one-RHS shapes of the 18-real (F16: 12-real) formats, per precision;
// plqcd_hip_hop_tuning replaces them.
HopTuning g_tuning[3] = {one_rhs_shape(PLQCD_PREC_F64, 18), one_rhs_shape(PLQCD_PREC_F32, 18),
                         one_rhs_shape(PLQCD_PREC_F16, 12)};

HopTuning plqcd::default_tuning(int prec, int links, int R, bool solver)
{
   if (g_tuning_set[prec]) return g_tuning[prec];
   if (R <= 1 && solver)
      for (const FormatShape &f : kOneRhsSolver)
         if (f.prec == prec && f.links == links) return f.t;
   if (R <= 1) return one_rhs_shape(prec, links);
   return solver ? kMultiRhsSolver[prec] : kMultiRhs[prec];
}

namespace {
std::mutex g_tune_mu;
std::map<plqcd::TuneKey, HopTuning> g_tune_cache;
}  // namespace

bool plqcd::lookup_tuning(const TuneKey &k, HopTuning *t)
{
   std::lock_guard<std::mutex> lk(g_tune_mu);
   auto it = g_tune_cache.find(k);
   if (it == g_tune_cache.end()) return false;
   *t = it->second;
   return true;
}

void plqcd::store_tuning(const TuneKey &k, const HopTuning &t)
{
   std::lock_guard<std::mutex> lk(g_tune_mu);
   g_tune_cache[k] = t;
}

void plqcd::clear_tuning(const TuneKey *k)
{
   std::lock_guard<std::mutex> lk(g_tune_mu);
   if (k) g_tune_cache.erase(*k);
   else g_tune_cache.clear();
}

extern "C" int plqcd_hip_hop_batch(int n, const plqcd_hip_gauge *const *g, const plqcd_hip_field *const *in,
                                   plqcd_hip_field *const *out, int mode)
{
   plqcd::clear_error();
   if (n < 1 || !g || !in || !out) return plqcd::set_error(1, "hop_batch: bad arguments");
   for (int i = 0; i < n; i++) {
      if (!g[i] || !in[i] || !out[i] || in[i] == out[i] || in[i]->lat != g[i]->lat || out[i]->lat != g[i]->lat ||
          in[i]->nrhs != out[i]->nrhs || in[i]->prec != g[i]->prec || out[i]->prec != g[i]->prec ||
          g[i]->prec != g[0]->prec || g[i]->links != g[0]->links)
         return plqcd::set_error(1, "hop_batch: item %d is inconsistent (all items share precision and link format)", i);
      if (g[i]->lat->stream != g[0]->lat->stream)
         return plqcd::set_error(1, "hop_batch: all lattices must use the same stream (set_stream)");
   }
   if (mode != PLQCD_HOP_COMPAT && mode != PLQCD_HOP_WILSON) return plqcd::set_error(1, "hop_batch: bad mode");
   switch (g[0]->prec * 100 + g[0]->links) {
      case PLQCD_PREC_F64 * 100 + 18: return plqcd::hop_batch_launch<PLQCD_PREC_F64, 18>(n, g, in, out, mode);
      case PLQCD_PREC_F64 * 100 + 12: return plqcd::hop_batch_launch<PLQCD_PREC_F64, 12>(n, g, in, out, mode);
      case PLQCD_PREC_F32 * 100 + 18: return plqcd::hop_batch_launch<PLQCD_PREC_F32, 18>(n, g, in, out, mode);
      case PLQCD_PREC_F32 * 100 + 12: return plqcd::hop_batch_launch<PLQCD_PREC_F32, 12>(n, g, in, out, mode);
      case PLQCD_PREC_F16 * 100 + 12: return plqcd::hop_batch_launch<PLQCD_PREC_F16, 12>(n, g, in, out, mode);
      case PLQCD_PREC_F16 * 100 + 18: return plqcd::hop_batch_launch<PLQCD_PREC_F16, 18>(n, g, in, out, mode);
      default: return plqcd::set_error(1, "hop_batch: unsupported gauge format");
   }
}

HopArgs make_hop_args(const plqcd_hip_lattice *lat, int parity)
{
   HopArgs a{};
   for (int mu = 0; mu < 4; mu++) a.L[mu] = lat->L[mu];
   a.Lxh = lat->Lxh;
   a.Vh = lat->Vh;
   a.nparity = (parity == PLQCD_PARITY_ALL) ? 2 : 1;
   a.parity0 = (parity == PLQCD_PARITY_ODD) ? 1 : 0;
   a.order = 0;
   a.group = 1;
   a.S = lat->Vh / lat->L[3];
   a.T = 1;
   a.R = 1;
   a.spb = 256;
   return a;
}

int plqcd::dslash(const plqcd_hip_gauge *g, const plqcd_hip_field *in, plqcd_hip_field *out, int parity, int mode,
                  const DslashOp &op)
{
   if (!g || !in || !out) return plqcd::set_error(1, "hop: null argument");
   if (in == out) return plqcd::set_error(1, "hop: in and out must be distinct fields");
   if (in->lat != g->lat || out->lat != g->lat)
      return plqcd::set_error(1, "hop: fields and gauge belong to different lattices");
   if (in->prec != g->prec || (out->prec != g->prec && !op.resid_b))
      return plqcd::set_error(1, "hop: precision mismatch (gauge %d, in %d, out %d)", g->prec, in->prec, out->prec);
   if (in->nrhs != out->nrhs) return plqcd::set_error(1, "hop: RHS count mismatch");
   if (parity < PLQCD_PARITY_EVEN || parity > PLQCD_PARITY_ALL) return plqcd::set_error(1, "hop: bad parity %d", parity);
   if (mode != PLQCD_HOP_COMPAT && mode != PLQCD_HOP_WILSON) return plqcd::set_error(1, "hop: bad mode %d", mode);
   if (op.rlist && (op.nr < 1 || op.nr > in->nrhs))
      return plqcd::set_error(1, "hop: RHS list of %d entries for fields of %d", op.nr, in->nrhs);
   if (!op.rlist && (op.nr < 0 || op.r0 < 0 || (op.nr > 0 && op.r0 + op.nr > in->nrhs)))
      return plqcd::set_error(1, "hop: RHS window %d + %d outside the fields' %d", op.r0, op.nr, in->nrhs);
   if (op.centre) {
      const plqcd_hip_field *c = op.centre;
      if (c->lat != g->lat || c->prec != g->prec || c->nrhs != in->nrhs || c == out || !op.kappa || op.dag ||
          op.alpha || op.partial || op.resid_b)
         return plqcd::set_error(1, "hop: the centre field must be like in, distinct from out, with kappa and no "
                                    "other epilogue");
   }
   if (op.resid_b) {
      const plqcd_hip_field *b = op.resid_b;
      if (g->prec != PLQCD_PREC_F64 || b->lat != g->lat || b->prec != PLQCD_PREC_F64 || b->nrhs != in->nrhs ||
          b == in || b == out || op.alpha || !op.dag || !op.kappa || !op.partial)
         return plqcd::set_error(1, "hop: the residual epilogue needs an F64 gauge, an F64 b distinct from in and "
                                    "out, M^+, kappa and a norm");
   }
   switch (g->prec * 100 + g->links) {
      case PLQCD_PREC_F64 * 100 + 18: return plqcd::hop_launch<PLQCD_PREC_F64, 18>(g, in, out, parity, mode, op);
      case PLQCD_PREC_F64 * 100 + 12: return plqcd::hop_launch<PLQCD_PREC_F64, 12>(g, in, out, parity, mode, op);
      case PLQCD_PREC_F32 * 100 + 18: return plqcd::hop_launch<PLQCD_PREC_F32, 18>(g, in, out, parity, mode, op);
      case PLQCD_PREC_F32 * 100 + 12: return plqcd::hop_launch<PLQCD_PREC_F32, 12>(g, in, out, parity, mode, op);
      case PLQCD_PREC_F16 * 100 + 12: return plqcd::hop_launch<PLQCD_PREC_F16, 12>(g, in, out, parity, mode, op);
      case PLQCD_PREC_F16 * 100 + 18: return plqcd::hop_launch<PLQCD_PREC_F16, 18>(g, in, out, parity, mode, op);
      default: return plqcd::set_error(1, "hop: unsupported gauge format");
   }
}

extern "C" int plqcd_hip_hop(const plqcd_hip_gauge *g, const plqcd_hip_field *in, plqcd_hip_field *out, int parity,
                             int mode)
{
   return plqcd::dslash(g, in, out, parity, mode, plqcd::DslashOp{});
}

extern "C" int plqcd_hip_hop_tuning(int prec, int order, int group, int block, int tmarch, int rb, int tile_y,
                                    int tile_z)
{
   if (prec < -1 || prec > PLQCD_PREC_F16 || order < 0 || order > 2 || group < 1 || group > 4096 ||
       (block != 64 && block != 128 && block != 256 && block != 512) || tmarch < 0 || tmarch > 4096 ||
       (rb != 1 && rb != 2 && rb != 4) || tile_y < 0 || tile_z < 0 || tile_y > 4096 || tile_z > 4096 ||
       ((tile_y == 0) != (tile_z == 0)))
      return plqcd::set_error(1, "plqcd_hip_hop_tuning: bad parameters (prec %d, order %d, group %d, block %d, "
                                 "tmarch %d, rb %d, tile %d x %d)", prec, order, group, block, tmarch, rb, tile_y, tile_z);
   for (int p = 0; p < 3; p++)
      if (prec < 0 || p == prec) {
         g_tuning[p] = HopTuning{order, group, block, tmarch, rb, tile_y, tile_z};
         g_tuning_set[p] = true;
      }
   return 0;
}

extern "C" int plqcd_hip_hop_tuning_reset(int prec)
{
   if (prec < -1 || prec > PLQCD_PREC_F16) return plqcd::set_error(1, "plqcd_hip_hop_tuning_reset: bad precision");
   for (int p = 0; p < 3; p++)
      if (prec < 0 || p == prec) {
         g_tuning[p] = one_rhs_shape(p, p == PLQCD_PREC_F16 ? 12 : 18);
         g_tuning_set[p] = false;
      }
   return 0;
}

extern "C" int plqcd_hip_hop_get_tuning(int prec, int out[7])
{
   if (prec < 0 || prec > PLQCD_PREC_F16 || !out) return plqcd::set_error(1, "plqcd_hip_hop_get_tuning: bad arguments");
   out[0] = g_tuning[prec].order;
   out[1] = g_tuning[prec].group;
   out[2] = g_tuning[prec].block;
   out[3] = g_tuning[prec].tmarch;
   out[4] = g_tuning[prec].rb;
   out[5] = g_tuning[prec].ty;
   out[6] = g_tuning[prec].tz;
   return 0;
}
