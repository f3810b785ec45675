// Vector kernels of the solvers (device fields, per-RHS device scalars).
// All are asynchronous on the field's lattice stream.  `partial`, when not
// null, receives the chunk sums of the named norm (partial_doubles() doubles),
// which reduce_partials() (reduce.h) turns into one double per RHS.
#pragma once

#include "device_common.h"

namespace plqcd {

// A window of right-hand sides: RHS r0 .. r0+nr-1 of the fields (nr = 0: all).
struct Window {
   int r0 = 0, nr = 0;
};

size_t partial_doubles(const plqcd_hip_field *f);

// x += alpha p in the working precision of the storage (float for F32/F16,
// double for F64), then p = r + beta p in double, from one read of p
int blas_cg_step(const plqcd_hip_field *r, const double *beta, plqcd_hip_field *p, const double *alpha,
                 plqcd_hip_field *x, const Window &w = Window{});
// x += alpha p as in blas_cg_step, for the RHS with mask[r] != 0 only (mask required)
int blas_x_step(const double *alpha, const plqcd_hip_field *p, plqcd_hip_field *x, const int *mask,
                const Window &w = Window{});
// The masked kernels below touch only the RHS r with mask[r] != 0 (all when
// mask is null); the chunk sums of untouched RHS are written as 0.  With a
// window, RHS outside it are not touched at all (their chunk sums neither).
// y += x ; x = 0 if zero_x   (any precisions)
int blas_accum(plqcd_hip_field *y, plqcd_hip_field *x, int zero_x, const int *mask, const Window &w = Window{});
// y = x ; partial <- |y|^2   (any precisions)
int blas_copy(plqcd_hip_field *y, const plqcd_hip_field *x, double *partial, const int *mask,
              const Window &w = Window{});
// y = 0 for RHS r with mask[r] != 0 (every RHS when mask is null)
int blas_zero(plqcd_hip_field *y, const int *mask, const Window &w = Window{});

}  // namespace plqcd
