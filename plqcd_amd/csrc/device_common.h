// Device-side definitions shared by the HIP translation units: the objects
// behind the opaque C-ABI handles, the HBM layouts of fields, and complex
// helpers on ext_vector types (so that vector loads, stores and non-temporal
// builtins apply directly).
//
// HBM layouts (Vh = V/2 sites per parity; j = index of a site inside its
// parity, i.e. even-odd index minus parity*Vh, which keeps the reference's
// site order):
//   spinor F64  [parity][12 colour-spin][Vh][rhs]  double2   (192 B/site)
//   spinor F32  [parity][6 pairs][Vh][rhs]        float4    ( 96 B/site)
//   spinor F16  [parity][3 quads][Vh][rhs]        8 x fp16  ( 48 B/site)
//               + [parity][Vh][rhs] float max-abs scale     (  4 B/site)
//   gauge  F64  [parity][mu][9 or 6][Vh]          double2
//   gauge  F32  [parity][mu][9][Vh] float2  or  [parity][mu][3][Vh] float4
//   gauge  F16  12-real: [parity][mu][Vh] 8 x fp16 (row 0 + row1.c0) and
//               [parity][mu][Vh] 4 x fp16 (row1.c1, row1.c2);
//               18-real: [parity][mu][Vh] 8 x fp16 (k 0..3), [parity][mu][Vh]
//               8 x fp16 (k 4..7) and [parity][mu][Vh] 2 x fp16 (k 8)
// Component order inside a spinor: c = 3*spin + colour.  Inside a link:
// k = 3*row + col.  With 12-real links row 2 is rebuilt as conj(row0 x row1).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <cstddef>
#include <cstdint>

#include "internal.h"

typedef double d2 __attribute__((ext_vector_type(2)));
typedef float f2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));
typedef uint32_t u2 __attribute__((ext_vector_type(2)));

struct plqcd_hip_lattice {
   int L[4];
   int V, Vh, Lxh;
   int device;
   hipStream_t own_stream;
   hipStream_t stream;
   hipEvent_t ev_start, ev_stop;
   // host <-> device transfers, allocated on first use: two page-locked host
   // buffers and two device buffers of the same size, in which consecutive
   // chunks alternate, and an event per pair
   void *xpin[2];
   void *xdev[2];
   hipEvent_t xev[2];
};

struct plqcd_hip_gauge {
   plqcd_hip_lattice *lat;
   int prec;   // PLQCD_PREC_*
   int links;  // 18 or 12
   void *data;
   size_t bytes;
};

struct plqcd_hip_field {
   plqcd_hip_lattice *lat;
   int prec;
   int nrhs;
   void *data;
   size_t bytes;
};

namespace plqcd {

// Bytes of one parity of one RHS of a spinor field (data and, for F16, scales).
inline size_t spinor_parity_bytes(int prec, int Vh)
{
   switch (prec) {
      case PLQCD_PREC_F64: return size_t(Vh) * 12 * 16;
      case PLQCD_PREC_F32: return size_t(Vh) * 6 * 16;
      default: return size_t(Vh) * (3 * 16 + 4);
   }
}
// The kernels address a site inside one component block of a field with a
// 32-bit byte offset (j*R + r)*16, and a link with j*16 (fields.h), so a
// component block of Vh*R 16-byte chunks must not exceed 4 GiB.  The largest
// RHS count a field of this lattice may hold:
constexpr uint64_t kOffsetLimit = uint64_t(1) << 32;
inline int max_rhs_per_field(int Vh)
{
   const uint64_t n = kOffsetLimit / (uint64_t(Vh) * 16);
   return n > 4096 ? 4096 : (int)n;
}
inline size_t gauge_bytes(int prec, int links, int Vh)
{
   const size_t n = size_t(2) * 4 * Vh;  // links
   switch (prec) {
      case PLQCD_PREC_F64: return n * links / 2 * 16;
      case PLQCD_PREC_F32: return n * links * 4;
      default: return n * (links == 18 ? 16 + 16 + 4 : 16 + 8);
   }
}

#define PLQCD_HIP_TRY(expr)                                                                            \
   do {                                                                                                \
      hipError_t e_ = (expr);                                                                          \
      if (e_ != hipSuccess)                                                                            \
         return ::plqcd::set_error(10, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                                   __LINE__);                                                          \
   } while (0)

}  // namespace plqcd

// ------------------------------------------------------------- complex math

template <typename R>
struct cvec;
template <>
struct cvec<double> {
   typedef d2 t;
};
template <>
struct cvec<float> {
   typedef f2 t;
};

// c + a*b
template <typename C>
__device__ __forceinline__ C cmadd(C a, C b, C c)
{
   C r;
   r.x = fma(a.x, b.x, c.x);
   r.x = fma(-a.y, b.y, r.x);
   r.y = fma(a.x, b.y, c.y);
   r.y = fma(a.y, b.x, r.y);
   return r;
}
// c + conj(a)*b
template <typename C>
__device__ __forceinline__ C cjmadd(C a, C b, C c)
{
   C r;
   r.x = fma(a.x, b.x, c.x);
   r.x = fma(a.y, b.y, r.x);
   r.y = fma(a.x, b.y, c.y);
   r.y = fma(-a.y, b.x, r.y);
   return r;
}
// All multiply-adds are explicit fma (the HIP code is built with
// -ffp-contract=off), so a given site and RHS is computed by the same
// operation sequence in every kernel variant.
template <typename C>
__device__ __forceinline__ C cmul(C a, C b)
{
   C r;
   r.x = fma(a.x, b.x, -(a.y * b.y));
   r.y = fma(a.x, b.y, a.y * b.x);
   return r;
}
// c - k*a, componentwise
template <typename C, typename R>
__device__ __forceinline__ C cxmky(C c, R k, C a)
{
   C r;
   r.x = fma(-k, a.x, c.x);
   r.y = fma(-k, a.y, c.y);
   return r;
}
// i*a
template <typename C>
__device__ __forceinline__ C ci(C a)
{
   C r;
   r.x = -a.y;
   r.y = a.x;
   return r;
}
template <typename C>
__device__ __forceinline__ C cconj(C a)
{
   C r;
   r.x = a.x;
   r.y = -a.y;
   return r;
}

// Row 2 of an SU(3) matrix from rows 0 and 1: u2_a = conj(u0_b u1_c - u0_c u1_b).
template <typename C>
__device__ __forceinline__ void reconstruct_row2(C u[9])
{
   u[6] = cconj(cmul(u[1], u[5]) - cmul(u[2], u[4]));
   u[7] = cconj(cmul(u[2], u[3]) - cmul(u[0], u[5]));
   u[8] = cconj(cmul(u[0], u[4]) - cmul(u[1], u[3]));
}

__device__ __forceinline__ f2 h2_to_f2(uint32_t w)
{
   __half2 h = *reinterpret_cast<__half2 *>(&w);
   float2 f = __half22float2(h);
   return f2{f.x, f.y};
}
__device__ __forceinline__ uint32_t f2_to_h2(float a, float b)
{
   __half2 h = __floats2half2_rn(a, b);
   return *reinterpret_cast<uint32_t *>(&h);
}

namespace plqcd {

// Whether [p, p + n) lies inside one buffer from plqcd_hip_host_alloc.
bool host_pinned(const void *p, size_t n);

// Host <-> device transfers of one RHS of a field on a stream of their own,
// through two pinned host buffers and two device buffers: the CPU copy of one
// chunk overlaps the DMA and (un)packing of the previous one, and the field's
// lattice stream keeps running.  The caller keeps kernels on other streams
// away from the RHS being transferred.  Each call returns when its transfer
// is complete; one Stager serves one thread at a time.  The raw copies move a
// host range from plqcd_hip_host_alloc by DMA alone.  The host copies stay
// on the calling thread: the server already runs several stagers on transfer
// threads of their own (the shared host copy pool gave the same serving
// throughput: 7.71 against 7.78 solves/s at 2 requests/s over 60 s).
class Stager {
  public:
   static int create(int device, Stager **out);
   ~Stager();
   int upload(plqcd_hip_field *f, int rhs, const spinor *host);
   int download(const plqcd_hip_field *f, int rhs, spinor *host);
   // Plain copies of `bytes` between host memory and a device buffer (DMA
   // only, no kernels); download first waits (on the host) for `after` when
   // not null.
   int upload_raw(void *dev, const void *host, size_t bytes);
   int download_raw(void *host, const void *dev, size_t bytes, hipEvent_t after);

  private:
   Stager() = default;
   hipStream_t st_ = nullptr;
   void *pin_[2] = {nullptr, nullptr};
   void *dev_[2] = {nullptr, nullptr};
   hipEvent_t ev_[2] = {nullptr, nullptr};
};

// One RHS of a field from / to a device buffer of V spinors in the host
// (reference) layout, asynchronous on `st`.
int pack_rhs(plqcd_hip_field *f, int rhs, const void *dev_aos, hipStream_t st);
int unpack_rhs(const plqcd_hip_field *f, int rhs, void *dev_aos, hipStream_t st);

}  // namespace plqcd
