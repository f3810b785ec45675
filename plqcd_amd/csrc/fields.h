// Storage policies: how one site's spinor or one link is read from / written
// to HBM in each precision.  Compute happens in `real` (double for F64, float
// otherwise); loads are 16-byte vector loads on structure-of-arrays data so
// that consecutive lanes touch consecutive addresses.
#pragma once

#include "device_common.h"

// ------------------------------------------------------------------ spinors

template <int PREC>
struct SpinorIO;

template <>
struct SpinorIO<PLQCD_PREC_F64> {
   typedef double real;
   typedef d2 C;
   d2 *base;  // start of this RHS
   int Vh;
   __device__ __forceinline__ const d2 *site(int parity, int j) const { return base + (size_t)parity * 12 * Vh + j; }
   __device__ __forceinline__ void load(int parity, int j, C s[12]) const
   {
      const d2 *p = site(parity, j);
#pragma unroll
      for (int c = 0; c < 12; c++) s[c] = p[(size_t)c * Vh];
   }
   __device__ __forceinline__ void store(int parity, int j, const C s[12]) const
   {
      d2 *p = base + (size_t)parity * 12 * Vh + j;
#pragma unroll
      for (int c = 0; c < 12; c++) __builtin_nontemporal_store(s[c], p + (size_t)c * Vh);
   }
   __host__ __device__ static size_t rhs_elems(int Vh) { return (size_t)24 * Vh; }
};

template <>
struct SpinorIO<PLQCD_PREC_F32> {
   typedef float real;
   typedef f2 C;
   f4 *base;
   int Vh;
   __device__ __forceinline__ void load(int parity, int j, C s[12]) const
   {
      const f4 *p = base + (size_t)parity * 6 * Vh + j;
#pragma unroll
      for (int c = 0; c < 6; c++) {
         const f4 v = p[(size_t)c * Vh];
         s[2 * c] = f2{v.x, v.y};
         s[2 * c + 1] = f2{v.z, v.w};
      }
   }
   __device__ __forceinline__ void store(int parity, int j, const C s[12]) const
   {
      f4 *p = base + (size_t)parity * 6 * Vh + j;
#pragma unroll
      for (int c = 0; c < 6; c++)
         __builtin_nontemporal_store(f4{s[2 * c].x, s[2 * c].y, s[2 * c + 1].x, s[2 * c + 1].y},
                                     p + (size_t)c * Vh);
   }
};

// fp16 with one float scale per site: value = scale * half.  The scale is the
// site's largest |real component|, so every stored half lies in [-1, 1].
template <>
struct SpinorIO<PLQCD_PREC_F16> {
   typedef float real;
   typedef f2 C;
   u4 *base;  // [parity][3][Vh] u4, then [parity][Vh] float
   int Vh;
   __device__ __forceinline__ float *scales() const { return reinterpret_cast<float *>(base + (size_t)6 * Vh); }
   __device__ __forceinline__ void load(int parity, int j, C s[12]) const
   {
      const u4 *p = base + (size_t)parity * 3 * Vh + j;
      const float sc = scales()[(size_t)parity * Vh + j];
#pragma unroll
      for (int q = 0; q < 3; q++) {
         const u4 w = p[(size_t)q * Vh];
         s[4 * q + 0] = h2_to_f2(w.x) * sc;
         s[4 * q + 1] = h2_to_f2(w.y) * sc;
         s[4 * q + 2] = h2_to_f2(w.z) * sc;
         s[4 * q + 3] = h2_to_f2(w.w) * sc;
      }
   }
   __device__ __forceinline__ void store(int parity, int j, const C s[12]) const
   {
      float m = 0.f;
#pragma unroll
      for (int c = 0; c < 12; c++) m = fmaxf(m, fmaxf(fabsf(s[c].x), fabsf(s[c].y)));
      const float inv = m > 0.f ? 1.0f / m : 0.f;
      u4 *p = base + (size_t)parity * 3 * Vh + j;
#pragma unroll
      for (int q = 0; q < 3; q++) {
         u4 w;
         w.x = f2_to_h2(s[4 * q + 0].x * inv, s[4 * q + 0].y * inv);
         w.y = f2_to_h2(s[4 * q + 1].x * inv, s[4 * q + 1].y * inv);
         w.z = f2_to_h2(s[4 * q + 2].x * inv, s[4 * q + 2].y * inv);
         w.w = f2_to_h2(s[4 * q + 3].x * inv, s[4 * q + 3].y * inv);
         __builtin_nontemporal_store(w, p + (size_t)q * Vh);
      }
      __builtin_nontemporal_store(m, scales() + (size_t)parity * Vh + j);
   }
};

// -------------------------------------------------------------------- links

template <int PREC, int LINKS>
struct GaugeIO;

template <>
struct GaugeIO<PLQCD_PREC_F64, 18> {
   typedef d2 C;
   const d2 *base;
   int Vh;
   __device__ __forceinline__ void load(int parity, int mu, int j, C u[9]) const
   {
      const d2 *p = base + (size_t)(parity * 4 + mu) * 9 * Vh + j;
#pragma unroll
      for (int k = 0; k < 9; k++) u[k] = p[(size_t)k * Vh];
   }
};

template <>
struct GaugeIO<PLQCD_PREC_F64, 12> {
   typedef d2 C;
   const d2 *base;
   int Vh;
   __device__ __forceinline__ void load(int parity, int mu, int j, C u[9]) const
   {
      const d2 *p = base + (size_t)(parity * 4 + mu) * 6 * Vh + j;
#pragma unroll
      for (int k = 0; k < 6; k++) u[k] = p[(size_t)k * Vh];
      reconstruct_row2(u);
   }
};

template <>
struct GaugeIO<PLQCD_PREC_F32, 18> {
   typedef f2 C;
   const f2 *base;
   int Vh;
   __device__ __forceinline__ void load(int parity, int mu, int j, C u[9]) const
   {
      const f2 *p = base + (size_t)(parity * 4 + mu) * 9 * Vh + j;
#pragma unroll
      for (int k = 0; k < 9; k++) u[k] = p[(size_t)k * Vh];
   }
};

template <>
struct GaugeIO<PLQCD_PREC_F32, 12> {
   typedef f2 C;
   const f4 *base;
   int Vh;
   __device__ __forceinline__ void load(int parity, int mu, int j, C u[9]) const
   {
      const f4 *p = base + (size_t)(parity * 4 + mu) * 3 * Vh + j;
#pragma unroll
      for (int k = 0; k < 3; k++) {
         const f4 v = p[(size_t)k * Vh];
         u[2 * k] = f2{v.x, v.y};
         u[2 * k + 1] = f2{v.z, v.w};
      }
      reconstruct_row2(u);
   }
};

template <>
struct GaugeIO<PLQCD_PREC_F16, 12> {
   typedef f2 C;
   const u4 *base;  // [parity][mu][Vh] u4 (u00,u01,u02,u10), then [parity][mu][Vh] u2 (u11,u12)
   int Vh;
   __device__ __forceinline__ void load(int parity, int mu, int j, C u[9]) const
   {
      const size_t o = (size_t)(parity * 4 + mu) * Vh + j;
      const u4 a = base[o];
      const u2 b = reinterpret_cast<const u2 *>(base + (size_t)8 * Vh)[o];
      u[0] = h2_to_f2(a.x);
      u[1] = h2_to_f2(a.y);
      u[2] = h2_to_f2(a.z);
      u[3] = h2_to_f2(a.w);
      u[4] = h2_to_f2(b.x);
      u[5] = h2_to_f2(b.y);
      reconstruct_row2(u);
   }
};
