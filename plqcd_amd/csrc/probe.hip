// HBM bandwidth probes: plain streaming kernels of known byte counts, used to
// state the achievable bandwidth of the device beside the hopping kernels'
// roofline and to calibrate hardware counters (a copy of N bytes must read N
// and write N bytes).
#include <algorithm>

#include "device_common.h"

namespace {

// kind 0: b[i] = a[i] (non-temporal store); 1: sum of a[i]; 2: b[i] = const
template <int KIND>
__global__ __launch_bounds__(256) void k_stream(const d2 *__restrict__ a, d2 *__restrict__ b, size_t n, double *sink)
{
   const size_t T = (size_t)gridDim.x * blockDim.x;
   d2 acc = {0.0, 0.0};
   for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += T) {
      if (KIND == 0) __builtin_nontemporal_store(a[i], b + i);
      if (KIND == 1) acc += a[i];
      if (KIND == 2) b[i] = d2{1.0, 2.0};
   }
   if (KIND == 1 && acc.x == -1.2345e300) sink[0] = acc.y;  // keeps the loads alive
}

// The read:write mix of the hopping matrix (FP64: 768 B read, 192 B written
// per site): b[i] = a[i] + a[n + i] + a[2n + i] + a[3n + i], four streams
// read and one written with a non-temporal store.
__global__ __launch_bounds__(256) void k_mix41(const d2 *__restrict__ a, d2 *__restrict__ b, size_t n)
{
   const size_t T = (size_t)gridDim.x * blockDim.x;
   for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += T) {
      const d2 v = (a[i] + a[n + i]) + (a[2 * n + i] + a[3 * n + i]);
      __builtin_nontemporal_store(v, b + i);
   }
}

}  // namespace

extern "C" int plqcd_hip_stream_bandwidth(int kind, size_t bytes, int reps, double *gbs)
{
   plqcd::clear_error();
   if (kind < 0 || kind > 3 || reps < 1 || !gbs || bytes < 16) return plqcd::set_error(1, "stream_bandwidth: bad arguments");
   const size_t n = bytes / 16;
   d2 *a = nullptr, *b = nullptr;
   double *sink = nullptr;
   PLQCD_HIP_TRY(hipMalloc(&a, n * 16 * (kind == 3 ? 4 : 1)));
   hipError_t e = hipMalloc(&b, n * 16);
   if (e == hipSuccess) e = hipMalloc(&sink, 64);
   hipEvent_t e0 = nullptr, e1 = nullptr;
   if (e == hipSuccess) e = hipMemset(a, 0, n * 16 * (kind == 3 ? 4 : 1));
   if (e == hipSuccess) e = hipEventCreate(&e0);
   if (e == hipSuccess) e = hipEventCreate(&e1);
   float ms = 0.f;
   // copy, read and write: 1024 blocks, the best of a sweep of 1024..65536 on
   // MI355X; the 4:1 mix keeps the best of its own sweep, run here
   const int grids_plain[1] = {1024};
   const int grids_mix[6] = {1024, 2048, 4096, 8192, 16384, 32768};
   const int *grids = kind == 3 ? grids_mix : grids_plain;
   const int ngrid = kind == 3 ? 6 : 1;
   for (int gi = 0; gi < ngrid && e == hipSuccess; gi++) {
      const int grid = grids[gi];
      auto run = [&]() {
         if (kind == 0) hipLaunchKernelGGL(k_stream<0>, dim3(grid), dim3(256), 0, 0, a, b, n, sink);
         else if (kind == 1) hipLaunchKernelGGL(k_stream<1>, dim3(grid), dim3(256), 0, 0, a, b, n, sink);
         else if (kind == 2) hipLaunchKernelGGL(k_stream<2>, dim3(grid), dim3(256), 0, 0, a, b, n, sink);
         else hipLaunchKernelGGL(k_mix41, dim3(grid), dim3(256), 0, 0, a, b, n);
      };
      run();
      (void)hipEventRecord(e0, 0);
      for (int r = 0; r < reps; r++) run();
      (void)hipEventRecord(e1, 0);
      e = hipEventSynchronize(e1);
      float t = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
      if (e == hipSuccess && (gi == 0 || t < ms)) ms = t;
   }
   if (e0) (void)hipEventDestroy(e0);
   if (e1) (void)hipEventDestroy(e1);
   (void)hipFree(a);
   if (b) (void)hipFree(b);
   if (sink) (void)hipFree(sink);
   if (e != hipSuccess) return plqcd::set_error(10, "stream_bandwidth: %s", hipGetErrorString(e));
   const double moved = (double)n * 16 * (kind == 0 ? 2 : kind == 3 ? 5 : 1) * reps;
   *gbs = moved / (ms * 1e-3) / 1e9;
   return 0;
}
