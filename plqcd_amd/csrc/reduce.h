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
// The lane holding the first site of a chunk writes the chunk sum to
// out[chunk] when `write` is set.
template <int BLOCK>
__device__ __forceinline__ void chunk_reduce(double v, int js, int r, int R, int chunk, bool write, double *out)
{
   __shared__ double red[BLOCK];
   const int tid = threadIdx.x;
   red[tid] = v;
   __syncthreads();
#pragma unroll
   for (int s = REDUCE_CHUNK / 2; s > 0; s >>= 1) {
      if ((js & (REDUCE_CHUNK - 1)) < s) red[tid] += red[tid + s * R];
      __syncthreads();
   }
   if ((js & (REDUCE_CHUNK - 1)) == 0 && write) out[chunk] = red[tid];
   __syncthreads();  // red is reused by the next call
}
