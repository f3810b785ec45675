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

// Addressing: the wave-uniform part of an address (field base, parity,
// component, direction) is formed in scalar registers; each lane adds one
// 32-bit byte offset per site, (j*R + r)*16.  The backend then issues
// global loads in the saddr + 32-bit voffset form with no per-component
// vector arithmetic.  (A parity block of one component of a field must stay
// below 4 GiB: Vh*R*16 bytes.)

template <class T>
__device__ __forceinline__ T ld(const char *ubase, uint32_t off)
{
   return *reinterpret_cast<const T *>(ubase + off);
}
template <class T>
__device__ __forceinline__ void st_nt(char *ubase, uint32_t off, T v)
{
   __builtin_nontemporal_store(v, reinterpret_cast<T *>(ubase + off));
}

template <int PREC>
struct SpinorIO;

template <>
struct SpinorIO<PLQCD_PREC_F64> {
   typedef double real;
   typedef d2 C;
   static constexpr int NC = 12;
   char *base;        // field base
   size_t cstride;    // bytes between components (Vh*R*16)
   uint32_t R16;      // R*16: bytes between sites
   uint32_t roff;     // r*16 for the lane's RHS
   __device__ __forceinline__ uint32_t off(int j) const { return (uint32_t)j * R16 + roff; }
   __device__ __forceinline__ const char *cbase(int parity, int c) const
   {
      return base + (size_t)(parity * NC + c) * cstride;
   }
   __device__ __forceinline__ SpinorIO at_rhs(int r) const
   {
      SpinorIO s = *this;
      s.roff = (uint32_t)r * 16;
      return s;
   }
   __device__ __forceinline__ void load(int parity, int j, C s[12]) const
   {
      const uint32_t o = off(j);
#pragma unroll
      for (int c = 0; c < 12; c++) s[c] = ld<d2>(cbase(parity, c), o);
   }
   __device__ __forceinline__ void store(int parity, int j, const C s[12]) const
   {
      const uint32_t o = off(j);
#pragma unroll
      for (int c = 0; c < 12; c++) st_nt(const_cast<char *>(cbase(parity, c)), o, s[c]);
   }
};

template <>
struct SpinorIO<PLQCD_PREC_F32> {
   typedef float real;
   typedef f2 C;
   static constexpr int NC = 6;
   char *base;
   size_t cstride;
   uint32_t R16, roff;
   __device__ __forceinline__ uint32_t off(int j) const { return (uint32_t)j * R16 + roff; }
   __device__ __forceinline__ const char *cbase(int parity, int c) const
   {
      return base + (size_t)(parity * NC + c) * cstride;
   }
   __device__ __forceinline__ SpinorIO at_rhs(int r) const
   {
      SpinorIO s = *this;
      s.roff = (uint32_t)r * 16;
      return s;
   }
   __device__ __forceinline__ void load(int parity, int j, C s[12]) const
   {
      const uint32_t o = off(j);
#pragma unroll
      for (int c = 0; c < 6; c++) {
         const f4 v = ld<f4>(cbase(parity, c), o);
         s[2 * c] = f2{v.x, v.y};
         s[2 * c + 1] = f2{v.z, v.w};
      }
   }
   __device__ __forceinline__ void store(int parity, int j, const C s[12]) const
   {
      const uint32_t o = off(j);
#pragma unroll
      for (int c = 0; c < 6; c++)
         st_nt(const_cast<char *>(cbase(parity, c)), o, f4{s[2 * c].x, s[2 * c].y, s[2 * c + 1].x, s[2 * c + 1].y});
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
   char *base;
   char *sbase;       // scales
   size_t cstride;    // bytes between data components (Vh*R*16)
   size_t spstride;   // bytes between scale parities (Vh*R*4)
   uint32_t R16, roff, R4, sroff;
   __device__ __forceinline__ uint32_t off(int j) const { return (uint32_t)j * R16 + roff; }
   __device__ __forceinline__ uint32_t soff(int j) const { return (uint32_t)j * R4 + sroff; }
   __device__ __forceinline__ const char *cbase(int parity, int c) const
   {
      return base + (size_t)(parity * NC + c) * cstride;
   }
   __device__ __forceinline__ SpinorIO at_rhs(int r) const
   {
      SpinorIO s = *this;
      s.roff = (uint32_t)r * 16;
      s.sroff = (uint32_t)r * 4;
      return s;
   }
   __device__ __forceinline__ void load(int parity, int j, C s[12]) const
   {
      const uint32_t o = off(j);
      const float sc = ld<float>(sbase + parity * spstride, soff(j));
#pragma unroll
      for (int q = 0; q < 3; q++) {
         const u4 w = ld<u4>(cbase(parity, q), o);
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
      const uint32_t o = off(j);
#pragma unroll
      for (int q = 0; q < 3; q++) {
         u4 w;
         w.x = f2_to_h2(s[4 * q + 0].x * inv, s[4 * q + 0].y * inv);
         w.y = f2_to_h2(s[4 * q + 1].x * inv, s[4 * q + 1].y * inv);
         w.z = f2_to_h2(s[4 * q + 2].x * inv, s[4 * q + 2].y * inv);
         w.w = f2_to_h2(s[4 * q + 3].x * inv, s[4 * q + 3].y * inv);
         st_nt(const_cast<char *>(cbase(parity, q)), o, w);
      }
      st_nt(sbase + parity * spstride, soff(j), m);
   }
};

// Accessor of a whole field (RHS 0; bind a lane with at_rhs).
// A site's spinor of any storage precision converted to / from double.
template <int P>
__device__ __forceinline__ void load_d(const SpinorIO<P> &f, int par, int j, d2 s[12])
{
   typename SpinorIO<P>::C t[12];
   f.load(par, j, t);
#pragma unroll
   for (int c = 0; c < 12; c++) s[c] = d2{(double)t[c].x, (double)t[c].y};
}

template <int P>
__device__ __forceinline__ void store_d(const SpinorIO<P> &f, int par, int j, const d2 s[12])
{
   typedef typename SpinorIO<P>::C C;
   typedef typename SpinorIO<P>::real real;
   C t[12];
#pragma unroll
   for (int c = 0; c < 12; c++) t[c] = C{(real)s[c].x, (real)s[c].y};
   f.store(par, j, t);
}

// |s|^2 of a site in a fixed order
__device__ __forceinline__ double norm2(const d2 s[12])
{
   double n = 0.0;
#pragma unroll
   for (int c = 0; c < 12; c++) n = fma(s[c].y, s[c].y, fma(s[c].x, s[c].x, n));
   return n;
}

template <int PREC>
inline SpinorIO<PREC> spinor_io(const plqcd_hip_field *f)
{
   SpinorIO<PREC> s{};
   const size_t Vh = f->lat->Vh, R = f->nrhs;
   s.base = static_cast<char *>(f->data);
   s.cstride = Vh * R * 16;
   s.R16 = (uint32_t)(R * 16);
   s.roff = 0;
   if constexpr (PREC == PLQCD_PREC_F16) {
      s.sbase = s.base + 2 * 3 * s.cstride;
      s.spstride = Vh * R * 4;
      s.R4 = (uint32_t)(R * 4);
      s.sroff = 0;
   }
   return s;
}

// -------------------------------------------------------------------- links

template <int PREC, int LINKS>
struct GaugeIO;

template <>
struct GaugeIO<PLQCD_PREC_F64, 18> {
   typedef d2 C;
   const char *base;
   size_t kstride;  // Vh*16: bytes between link components
   __device__ __forceinline__ void load(int parity, int mu, int j, C u[9]) const
   {
      const char *b = base + (size_t)(parity * 4 + mu) * 9 * kstride;
      const uint32_t o = (uint32_t)j * 16;
#pragma unroll
      for (int k = 0; k < 9; k++) u[k] = ld<d2>(b + k * kstride, o);
   }
};

template <>
struct GaugeIO<PLQCD_PREC_F64, 12> {
   typedef d2 C;
   const char *base;
   size_t kstride;
   __device__ __forceinline__ void load(int parity, int mu, int j, C u[9]) const
   {
      const char *b = base + (size_t)(parity * 4 + mu) * 6 * kstride;
      const uint32_t o = (uint32_t)j * 16;
#pragma unroll
      for (int k = 0; k < 6; k++) u[k] = ld<d2>(b + k * kstride, o);
      reconstruct_row2(u);
   }
};

template <>
struct GaugeIO<PLQCD_PREC_F32, 18> {
   typedef f2 C;
   const char *base;
   size_t kstride;  // Vh*8
   __device__ __forceinline__ void load(int parity, int mu, int j, C u[9]) const
   {
      const char *b = base + (size_t)(parity * 4 + mu) * 9 * kstride;
      const uint32_t o = (uint32_t)j * 8;
#pragma unroll
      for (int k = 0; k < 9; k++) u[k] = ld<f2>(b + k * kstride, o);
   }
};

template <>
struct GaugeIO<PLQCD_PREC_F32, 12> {
   typedef f2 C;
   const char *base;
   size_t kstride;  // Vh*16
   __device__ __forceinline__ void load(int parity, int mu, int j, C u[9]) const
   {
      const char *b = base + (size_t)(parity * 4 + mu) * 3 * kstride;
      const uint32_t o = (uint32_t)j * 16;
#pragma unroll
      for (int k = 0; k < 3; k++) {
         const f4 v = ld<f4>(b + k * kstride, o);
         u[2 * k] = f2{v.x, v.y};
         u[2 * k + 1] = f2{v.z, v.w};
      }
      reconstruct_row2(u);
   }
};

template <>
struct GaugeIO<PLQCD_PREC_F16, 12> {
   typedef f2 C;
   const char *base;  // [parity][mu][Vh] u4 (u00,u01,u02,u10), then [parity][mu][Vh] u2 (u11,u12)
   size_t kstride;    // Vh*16
   __device__ __forceinline__ void load(int parity, int mu, int j, C u[9]) const
   {
      const size_t slot = (size_t)(parity * 4 + mu);
      const u4 a = ld<u4>(base + slot * kstride, (uint32_t)j * 16);
      const u2 b = ld<u2>(base + 8 * kstride + slot * (kstride / 2), (uint32_t)j * 8);
      u[0] = h2_to_f2(a.x);
      u[1] = h2_to_f2(a.y);
      u[2] = h2_to_f2(a.z);
      u[3] = h2_to_f2(a.w);
      u[4] = h2_to_f2(b.x);
      u[5] = h2_to_f2(b.y);
      reconstruct_row2(u);
   }
};

template <int PREC, int LINKS>
inline GaugeIO<PREC, LINKS> gauge_io(const plqcd_hip_gauge *g)
{
   GaugeIO<PREC, LINKS> io;
   io.base = static_cast<const char *>(g->data);
   const size_t Vh = g->lat->Vh;
   io.kstride = (PREC == PLQCD_PREC_F32 && LINKS == 18) ? Vh * 8 : Vh * 16;
   return io;
}
