// Storage policies: how one site's spinor or one link is read from / written
// to HBM in each precision.  Compute happens in `real` (double for F64, float
// otherwise); loads are 16-byte vector loads on structure-of-arrays data so
// that consecutive lanes touch consecutive addresses.
#pragma once

#include "device_common.h"

// ------------------------------------------------------------------ spinors
//
// A field with R right-hand sides stores them innermost:
//   element (parity, chunk c, site j, rhs r) at ((parity*NC + c)*Vh + j)*R + r
// (NC 16-byte chunks per site).  Lanes that handle the R right-hand sides of
// one site therefore read one contiguous run, and when they read the same
// link the request is served once.  With R = 1 this is plain
// structure-of-arrays.  An accessor is bound to one RHS by at_rhs(r).

template <int PREC>
struct SpinorIO;

template <>
struct SpinorIO<PLQCD_PREC_F64> {
   typedef double real;
   typedef d2 C;
   static constexpr int NC = 12;
   d2 *base;
   int Vh, R;
   __device__ __forceinline__ size_t idx(int parity, int c, int j) const
   {
      return ((size_t)(parity * NC + c) * Vh + j) * R;
   }
   __device__ __forceinline__ SpinorIO at_rhs(int r) const
   {
      SpinorIO s = *this;
      s.base += r;
      return s;
   }
   __device__ __forceinline__ void load(int parity, int j, C s[12]) const
   {
#pragma unroll
      for (int c = 0; c < 12; c++) s[c] = base[idx(parity, c, j)];
   }
   __device__ __forceinline__ void store(int parity, int j, const C s[12]) const
   {
#pragma unroll
      for (int c = 0; c < 12; c++) __builtin_nontemporal_store(s[c], base + idx(parity, c, j));
   }
};

template <>
struct SpinorIO<PLQCD_PREC_F32> {
   typedef float real;
   typedef f2 C;
   static constexpr int NC = 6;
   f4 *base;
   int Vh, R;
   __device__ __forceinline__ size_t idx(int parity, int c, int j) const
   {
      return ((size_t)(parity * NC + c) * Vh + j) * R;
   }
   __device__ __forceinline__ SpinorIO at_rhs(int r) const
   {
      SpinorIO s = *this;
      s.base += r;
      return s;
   }
   __device__ __forceinline__ void load(int parity, int j, C s[12]) const
   {
#pragma unroll
      for (int c = 0; c < 6; c++) {
         const f4 v = base[idx(parity, c, j)];
         s[2 * c] = f2{v.x, v.y};
         s[2 * c + 1] = f2{v.z, v.w};
      }
   }
   __device__ __forceinline__ void store(int parity, int j, const C s[12]) const
   {
#pragma unroll
      for (int c = 0; c < 6; c++)
         __builtin_nontemporal_store(f4{s[2 * c].x, s[2 * c].y, s[2 * c + 1].x, s[2 * c + 1].y},
                                     base + idx(parity, c, j));
   }
};

// fp16 with one float scale per site: value = scale * half.  The scale is the
// site's largest |real component|, so every stored half lies in [-1, 1].
// Scales follow the data: [parity][Vh][R] floats.
template <>
struct SpinorIO<PLQCD_PREC_F16> {
   typedef float real;
   typedef f2 C;
   static constexpr int NC = 3;
   u4 *base;
   float *scale;
   int Vh, R;
   __device__ __forceinline__ size_t idx(int parity, int c, int j) const
   {
      return ((size_t)(parity * NC + c) * Vh + j) * R;
   }
   __device__ __forceinline__ size_t sidx(int parity, int j) const { return ((size_t)parity * Vh + j) * R; }
   __device__ __forceinline__ SpinorIO at_rhs(int r) const
   {
      SpinorIO s = *this;
      s.base += r;
      s.scale += r;
      return s;
   }
   __device__ __forceinline__ void load(int parity, int j, C s[12]) const
   {
      const float sc = scale[sidx(parity, j)];
#pragma unroll
      for (int q = 0; q < 3; q++) {
         const u4 w = base[idx(parity, q, j)];
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
#pragma unroll
      for (int q = 0; q < 3; q++) {
         u4 w;
         w.x = f2_to_h2(s[4 * q + 0].x * inv, s[4 * q + 0].y * inv);
         w.y = f2_to_h2(s[4 * q + 1].x * inv, s[4 * q + 1].y * inv);
         w.z = f2_to_h2(s[4 * q + 2].x * inv, s[4 * q + 2].y * inv);
         w.w = f2_to_h2(s[4 * q + 3].x * inv, s[4 * q + 3].y * inv);
         __builtin_nontemporal_store(w, base + idx(parity, q, j));
      }
      __builtin_nontemporal_store(m, scale + sidx(parity, j));
   }
};

// Accessor of a whole field (RHS 0; bind a lane with at_rhs).
template <int PREC>
inline SpinorIO<PREC> spinor_io(const plqcd_hip_field *f)
{
   SpinorIO<PREC> s;
   s.base = reinterpret_cast<decltype(s.base)>(f->data);
   s.Vh = f->lat->Vh;
   s.R = f->nrhs;
   if constexpr (PREC == PLQCD_PREC_F16)
      s.scale = reinterpret_cast<float *>(static_cast<char *>(f->data) + (size_t)2 * 3 * 16 * s.Vh * s.R);
   return s;
}

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
