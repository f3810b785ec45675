// Launch parameters of the hopping kernels and the internal entry point used
// by the solvers.
#pragma once

#include "device_common.h"

struct HopArgs {
   int L[4];     // extents x, y, z, t
   int Lxh;      // L[0] / 2: sites of one parity in an x row
   int Vh;       // sites per parity
   int nparity;  // 1 (eo or oe) or 2 (full H)
   int parity0;  // output parity when nparity == 1
   int order;    // block -> tile mapping (see logical_block)
   int group;    // blocks per XCD run for order 2
   int S;        // sites per parity in one t-slice (t-march)
   int T;        // t-slices per thread (t-march)
   int R;        // right-hand sides per launch: lanes per site
   int r0;       // RHS of the fields that lane 0 of a site handles (the launch covers r0 .. r0+R-1)
   const int *rlist;     // when set: the launch covers RHS rlist[0 .. R-1] instead (device array)
   int spb;      // sites per block
   int ty, tz;   // tile extents in y and z (LDS-tiled kernel)
   int nchunk;   // reduction chunks per parity
   const double *kappa;  // per-RHS kappa of M = 1 - kappa H (EPI = 1)
   double *partial;      // [R][2*nchunk] chunk sums of |out|^2 (NORM)
};

// Launch shape of the hopping kernels, process-wide.
struct HopTuning {
   int order, group, block;
   int tmarch;  // t-slices per thread of the t-march kernel (0: gather kernel)
   int rb;      // right-hand sides per lane (1, 2 or 4; RHS-blocked gather when > 1)
   int ty, tz;  // LDS tile of the t-march (y and z extents; 0: untiled)
};
extern HopTuning g_tuning[3];  // per storage precision

namespace plqcd {
// Launch shape for `R` right-hand sides of storage format (prec, links) before
// the per-shape cache: the precision's shape set by plqcd_hip_hop_tuning, else
// a built-in default for R = 1 (per format) or for R > 1 (per precision), and
// in both cases by whether the launch applies M = 1 - kappa H with or without
// the CG update (`solver`) where that differs.
HopTuning default_tuning(int prec, int links, int R, bool solver = false);
}  // namespace plqcd

HopArgs make_hop_args(const plqcd_hip_lattice *lat, int parity);

namespace plqcd {

// What a launch computes besides H in:
//   kappa   != null: out = in - kappa[r] * (H in)  (device array, one per RHS)
//   dag     != 0:    H is replaced by H^+ = gamma5 H gamma5
//   partial != null: chunk sums of |out|^2 into partial[r][2*nchunk]
//   alpha   != null: CG residual update (with dag, kappa and partial): Ap = in -
//                    kappa H^+ in is not stored; out -= alpha[r] Ap (out holds
//                    the residual), whose |out|^2 goes to partial
//   centre  != null: (with kappa; no dag, norm, alpha or resid_b) out = centre -
//                    kappa[r] * (H in) on the output parity, the centre read
//                    from that field rather than from in (gather kernel)
//   resid_b != null: residual (with dag, kappa and partial; F64 storage):
//                    out = resid_b - (in - kappa H^+ in) for every RHS of the
//                    launch, stored in out's own precision (any), and its |out|^2
//                    to partial
//   nr      > 0:     only the right-hand sides r0 .. r0+nr-1 of the fields (kappa,
//                    alpha and partial stay indexed by the field's RHS)
//   rlist   != null: only the nr right-hand sides rlist[0 .. nr-1] (a device
//                    array of distinct RHS of the fields; r0 is not used)
struct DslashOp {
   int r0 = 0, nr = 0;
   const int *rlist = nullptr;
   int dag = 0;
   const double *kappa = nullptr;
   double *partial = nullptr;
   const double *alpha = nullptr;
   const plqcd_hip_field *centre = nullptr;
   const plqcd_hip_field *resid_b = nullptr;
};

int dslash(const plqcd_hip_gauge *g, const plqcd_hip_field *in, plqcd_hip_field *out, int parity, int mode,
           const DslashOp &op);

// Launchers of one (storage precision, link format); arguments already checked.
template <int PREC, int LINKS>
int hop_launch(const plqcd_hip_gauge *g, const plqcd_hip_field *in, plqcd_hip_field *out, int parity, int mode,
               const DslashOp &op);
template <int PREC, int LINKS>
int hop_batch_launch(int n, const plqcd_hip_gauge *const *g, const plqcd_hip_field *const *in,
                     plqcd_hip_field *const *out, int mode);

// Launch shapes found by the autotuner, per (lattice, precision, links, RHS,
// operation).  `op` separates plain applications of H (0), which the
// autotuner times, from the solver's M / M^+ launches with their epilogues
// (1), whose built-in shapes differ and which a tuning of H must not change.
struct TuneKey {
   int L[4];
   int prec, links, R;
   int op;
   bool operator<(const TuneKey &o) const
   {
      for (int i = 0; i < 4; i++)
         if (L[i] != o.L[i]) return L[i] < o.L[i];
      if (prec != o.prec) return prec < o.prec;
      if (links != o.links) return links < o.links;
      if (R != o.R) return R < o.R;
      return op < o.op;
   }
};
constexpr int kTuneOpHop = 0, kTuneOpSolver = 1;
bool lookup_tuning(const TuneKey &k, HopTuning *t);
void store_tuning(const TuneKey &k, const HopTuning &t);
void clear_tuning(const TuneKey *k);  // null: all

}  // namespace plqcd
