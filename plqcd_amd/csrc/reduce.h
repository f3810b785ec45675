// Deterministic reductions over lattice sites.
//
// Sums over a field are formed in two fixed stages that do not depend on the
// launch shape or on how many right-hand sides share a launch:
//   1. each chunk of REDUCE_CHUNK consecutive sites (within one parity) is
//      summed in a fixed binary tree, giving partial[r][chunk]
//      (chunk = parity * nchunk + j / REDUCE_CHUNK);
//   2. reduce_partials() sums the chunks of each RHS in an order fixed by the
//      chunk count (segments of 4096 chunks, then the segment sums).
// A right-hand side therefore gets bitwise the same sums alone or batched.
#pragma once

#include "device_common.h"

#define REDUCE_CHUNK 8

namespace plqcd {
inline int reduce_chunks(int Vh) { return (Vh + REDUCE_CHUNK - 1) / REDUCE_CHUNK; }

// out[r] = sum of partial[r][first .. first+n) for r0 <= r < r0+R (fixed
// order), on `st`.  The partials are overwritten.
int reduce_partials(double *partial, size_t stride, int first, int n, int R, double *out, hipStream_t st, int r0 = 0);
}  // namespace plqcd

// Stage 1 inside a block whose lanes are (js, r), tid = js * R + r, with
// REDUCE_CHUNK | (sites per block).  Every thread of the block must call it.
// The lane holding the first site of a chunk sums the chunk's 8 site values
// in the fixed tree ((v0 + v4) + (v2 + v6)) + ((v1 + v5) + (v3 + v7)) and
// writes the sum to out[chunk] when `write` is set.  One barrier per call:
// consecutive calls of a block alternate between two LDS buffers (`phase`
// counts the calls), so a call's writes cannot reach a buffer whose reads
// are still pending.
template <int BLOCK>
__device__ __forceinline__ void chunk_reduce(double v, int js, int r, int R, int chunk, bool write, double *out,
                                             int phase = 0)
{
   __shared__ double red[2][BLOCK];
   double *b = red[phase & 1];
   const int tid = threadIdx.x;
   b[tid] = v;
   __syncthreads();
   if ((js & (REDUCE_CHUNK - 1)) == 0 && write) {
      const double *c = b + tid;
      const double lo = (c[0] + c[4 * R]) + (c[2 * R] + c[6 * R]);
      const double hi = (c[R] + c[5 * R]) + (c[3 * R] + c[7 * R]);
      out[chunk] = lo + hi;
   }
}
