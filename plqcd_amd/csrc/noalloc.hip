// plqcd_hip_apply_m_noalloc: out = M in or M^+ in (M = 1 - kappa H) like
// plqcd_hip_apply_m, for callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops).  plqcd_hip_apply_m keeps the
// per-RHS kappa its kernels read in a device array it allocates per call;
// here that array is the file-scope `scratch` below, filled by a one-block
// kernel from its launch arguments (no host copy, no memset), and the call is
// asynchronous on the lattice stream.
//
// `scratch` is a device symbol: each GPU holds its own copy.  Its address on
// a device is looked up once per device with hipGetSymbolAddress (with that
// device current) and kept in g_scratch_ptr; host code passes only that
// pointer to kernels.  Calls that overlap on one GPU share the array and race.
#include <mutex>

#include "dslash.h"

namespace {

constexpr int kNoAllocMaxRhs = 64;  // the per-launch RHS limit of the hopping kernels
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kNoAllocMaxRhs];

std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

struct KappaArgs {
   double k[kNoAllocMaxRhs];
};

__global__ void set_kappa(double *dst, KappaArgs v, int n)
{
   const int r = threadIdx.x;
   if (r < n) dst[r] = v.k[r];
}

// Address of `scratch` on `device`, looked up on the first call there.
int scratch_on(int device, double **out)
{
   if (device < 0 || device >= kMaxDevices) return plqcd::set_error(2, "apply_m_noalloc: device %d out of range", device);
   std::lock_guard<std::mutex> lk(g_scratch_mu);
   if (!g_scratch_ptr[device]) {
      int cur = -1;
      PLQCD_HIP_TRY(hipGetDevice(&cur));
      if (cur != device) PLQCD_HIP_TRY(hipSetDevice(device));
      void *p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (cur != device) (void)hipSetDevice(cur);
      if (e != hipSuccess || !p)
         return plqcd::set_error(10, "apply_m_noalloc: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
      g_scratch_ptr[device] = static_cast<double *>(p);
   }
   *out = g_scratch_ptr[device];
   return 0;
}

}  // namespace

extern "C" int plqcd_hip_apply_m_noalloc(const plqcd_hip_gauge *g, const plqcd_hip_field *in, plqcd_hip_field *out,
                                         const double *kappa_rhs, int dagger)
{
   plqcd::clear_error();
   if (!g || !in || !out || !kappa_rhs) return plqcd::set_error(1, "apply_m_noalloc: null argument");
   const int R = in->nrhs;
   if (R > kNoAllocMaxRhs)
      return plqcd::set_error(2, "apply_m_noalloc: %d right-hand sides exceed the scratch of %d", R, kNoAllocMaxRhs);
   plqcd_hip_lattice *lat = g->lat;
   double *dk = nullptr;
   if (int rc = scratch_on(lat->device, &dk)) return rc;
   KappaArgs v{};
   for (int r = 0; r < R; r++) v.k[r] = kappa_rhs[r];
   hipLaunchKernelGGL(set_kappa, dim3(1), dim3(kNoAllocMaxRhs), 0, lat->stream, dk, v, R);
   PLQCD_HIP_TRY(hipGetLastError());
   plqcd::DslashOp op;
   op.kappa = dk;
   op.dag = dagger ? 1 : 0;
   return plqcd::dslash(g, in, out, PLQCD_PARITY_ALL, PLQCD_HOP_WILSON, op);
}
