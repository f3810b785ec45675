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
// component, direction) plus one 32-bit byte offset per lane and site,
// (j*R + r)*16.  The compiler often folds the lane offset into the
// per-component bases and issues 64-bit per-lane addresses (one v_mad_u64_u32
// per load).  Forcing every base into scalar registers (readfirstlane),
// which gives the saddr + voffset form, cut the FP64 tile kernel's vector
// instructions by a third but made it slower: 422 against 404 us at
// 32^3 x 64 (FP16 18-real tiles 112 against 119 us, the rest within 3 %).
// So the loads stay as written.  (A parity block of one component of a field
// must stay below 4 GiB: Vh*R*16 bytes.)

// NT: a non-temporal (streaming) load, for data that no other lane or block
// reads again soon; it leaves more of the L2 to data that is re-read.
template <class T, bool NT = false>
__device__ __forceinline__ T ld(const char *ubase, uint32_t off)
{
   if constexpr (NT)
      return __builtin_nontemporal_load(reinterpret_cast<const T *>(ubase + off));
   else
      return *reinterpret_cast<const T *>(ubase + off);
}
template <class T>
__device__ __forceinline__ void st_nt(char *ubase, uint32_t off, T v)
{
   __builtin_nontemporal_store(v, reinterpret_cast<T *>(ubase + off));
}

// ------------------------------------------------------------ held forms
//
// What a hopping kernel holds between loading an operand and using it (in
// registers or in LDS, as e[0..N-1] of 32- or 64-bit words):
//   spinor      F64/F32: 12 complex in the working precision (SpinV)
//               F16: the 12 stored half pairs and the site scale (SpinH)
//   link        F64/F32: 9 complex (LinkV); F16: 9 half pairs (LinkH)
//   half term   U (1 -+ gamma_mu) psi, 2 colour vectors: F64/F32 6 complex
//               (HalfV); F16 6 float complex and the scale of psi that
//               still multiplies them (HalfH)
typedef _Float16 hh2 __attribute__((ext_vector_type(2)));

template <class C>
struct SpinV {
   typedef C E;
   static constexpr int N = 12;
   C e[12];
};
struct SpinH {
   typedef uint32_t E;
   static constexpr int N = 13;
   uint32_t e[13];  // e[0..11]: (re, im) halves of component m; e[12]: the scale's bits
   __device__ __forceinline__ hh2 h(int m) const { return __builtin_bit_cast(hh2, e[m]); }
   __device__ __forceinline__ float s() const { return __builtin_bit_cast(float, e[12]); }
};
template <class C>
struct LinkV {
   typedef C E;
   C e[9];
};
struct LinkH {
   typedef uint32_t E;
   uint32_t e[9];
   __device__ __forceinline__ hh2 h(int k) const { return __builtin_bit_cast(hh2, e[k]); }
};
template <class C>
struct HalfV {
   typedef C E;
   static constexpr int N = 6;
   C e[6];
};
struct HalfH {
   typedef float E;
   static constexpr int N = 13;
   float e[13];  // e[2m], e[2m+1]: component m; e[12]: the scale
};

template <int PREC>
struct SpinorIO;

template <>
struct SpinorIO<PLQCD_PREC_F64> {
   typedef double real;
   typedef d2 C;
   typedef SpinV<C> Sp;
   typedef HalfV<C> Hv;
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
   __device__ __forceinline__ void load_sp(int parity, int j, Sp &s) const { load(parity, j, s.e); }
   static __device__ __forceinline__ void to_c(const Sp &s, C c[12])
   {
#pragma unroll
      for (int m = 0; m < 12; m++) c[m] = s.e[m];
   }
};

template <>
struct SpinorIO<PLQCD_PREC_F32> {
   typedef float real;
   typedef f2 C;
   typedef SpinV<C> Sp;
   typedef HalfV<C> Hv;
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
   __device__ __forceinline__ void load_sp(int parity, int j, Sp &s) const { load(parity, j, s.e); }
   static __device__ __forceinline__ void to_c(const Sp &s, C c[12])
   {
#pragma unroll
      for (int m = 0; m < 12; m++) c[m] = s.e[m];
   }
};

// fp16 with one float scale per site: value = scale * half.  The scale is the
// site's largest |real component|, so every stored half lies in [-1, 1].
// Scales follow the data: [parity][Vh][R] floats.
template <>
struct SpinorIO<PLQCD_PREC_F16> {
   typedef float real;
   typedef f2 C;
   typedef SpinH Sp;
   typedef HalfH Hv;
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
   __device__ __forceinline__ void load_sp(int parity, int j, Sp &s) const
   {
      const uint32_t o = off(j);
      s.e[12] = ld<uint32_t>(sbase + parity * spstride, soff(j));
#pragma unroll
      for (int q = 0; q < 3; q++) {
         const u4 w = ld<u4>(cbase(parity, q), o);
         s.e[4 * q + 0] = w.x;
         s.e[4 * q + 1] = w.y;
         s.e[4 * q + 2] = w.z;
         s.e[4 * q + 3] = w.w;
      }
   }
   static __device__ __forceinline__ void to_c(const Sp &s, C c[12])
   {
      const float sc = s.s();
#pragma unroll
      for (int m = 0; m < 12; m++) c[m] = h2_to_f2(s.e[m]) * sc;
   }
   __device__ __forceinline__ void load(int parity, int j, C s[12]) const
   {
      Sp h;
      load_sp(parity, j, h);
      to_c(h, s);
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
//
// load_u(parity, mu, j, u) reads the link U_mu of site j of `parity` into its
// held form; load_u<true> does the same with streaming (non-temporal) loads,
// for a link that only the loading lane reads.

template <int PREC, int LINKS>
struct GaugeIO;

template <>
struct GaugeIO<PLQCD_PREC_F64, 18> {
   typedef d2 C;
   typedef LinkV<C> U;
   const char *base;
   size_t kstride;  // Vh*16: bytes between link components
   template <bool NT = false>
   __device__ __forceinline__ void load(int parity, int mu, int j, C u[9]) const
   {
      const char *b = base + (size_t)(parity * 4 + mu) * 9 * kstride;
      const uint32_t o = (uint32_t)j * 16;
#pragma unroll
      for (int k = 0; k < 9; k++) u[k] = ld<d2, NT>(b + k * kstride, o);
   }
   template <bool NT = false>
   __device__ __forceinline__ void load_u(int parity, int mu, int j, U &u) const
   {
      load<NT>(parity, mu, j, u.e);
   }
};

template <>
struct GaugeIO<PLQCD_PREC_F64, 12> {
   typedef d2 C;
   typedef LinkV<C> U;
   const char *base;
   size_t kstride;
   template <bool NT = false>
   __device__ __forceinline__ void load(int parity, int mu, int j, C u[9]) const
   {
      const char *b = base + (size_t)(parity * 4 + mu) * 6 * kstride;
      const uint32_t o = (uint32_t)j * 16;
#pragma unroll
      for (int k = 0; k < 6; k++) u[k] = ld<d2, NT>(b + k * kstride, o);
      reconstruct_row2(u);
   }
   template <bool NT = false>
   __device__ __forceinline__ void load_u(int parity, int mu, int j, U &u) const
   {
      load<NT>(parity, mu, j, u.e);
   }
};

template <>
struct GaugeIO<PLQCD_PREC_F32, 18> {
   typedef f2 C;
   typedef LinkV<C> U;
   const char *base;
   size_t kstride;  // Vh*8
   template <bool NT = false>
   __device__ __forceinline__ void load(int parity, int mu, int j, C u[9]) const
   {
      const char *b = base + (size_t)(parity * 4 + mu) * 9 * kstride;
      const uint32_t o = (uint32_t)j * 8;
#pragma unroll
      for (int k = 0; k < 9; k++) u[k] = ld<f2, NT>(b + k * kstride, o);
   }
   template <bool NT = false>
   __device__ __forceinline__ void load_u(int parity, int mu, int j, U &u) const
   {
      load<NT>(parity, mu, j, u.e);
   }
};

template <>
struct GaugeIO<PLQCD_PREC_F32, 12> {
   typedef f2 C;
   typedef LinkV<C> U;
   const char *base;
   size_t kstride;  // Vh*16
   template <bool NT = false>
   __device__ __forceinline__ void load(int parity, int mu, int j, C u[9]) const
   {
      const char *b = base + (size_t)(parity * 4 + mu) * 3 * kstride;
      const uint32_t o = (uint32_t)j * 16;
#pragma unroll
      for (int k = 0; k < 3; k++) {
         const f4 v = ld<f4, NT>(b + k * kstride, o);
         u[2 * k] = f2{v.x, v.y};
         u[2 * k + 1] = f2{v.z, v.w};
      }
      reconstruct_row2(u);
   }
   template <bool NT = false>
   __device__ __forceinline__ void load_u(int parity, int mu, int j, U &u) const
   {
      load<NT>(parity, mu, j, u.e);
   }
};

// F16 links are held as half pairs, which the F16 hopping arithmetic multiplies
// with v_dot2_f32_f16 (c + a.lo * b.lo + a.hi * b.hi, products and sum in
// float) in its VOP3P form, written as inline assembly: the builtin selects
// the VOP2 form, which has neither operand modifiers (neg_lo / neg_hi negate
// a half of b) nor constant accumulators.  The hardware needs 3 wait states
// before a VALU instruction other than v_dot2_f32_f16 itself reads a
// v_dot2_f32_f16 result (the compiler inserts them for its own code, not
// for assembly), so each assembly block runs whole dependent chains of dot
// products and ends with s_nop 2; within a block a result only feeds the
// accumulator of the next v_dot2_f32_f16, which needs none.
#define PLQCD_D2Z(d, a, b, m) "v_dot2_f32_f16 " d ", " a ", " b ", 0" m "\n\t"
#define PLQCD_D2(d, a, b, m) "v_dot2_f32_f16 " d ", " a ", " b ", " d m "\n\t"
#define PLQCD_NLO " neg_lo:[0,1,0]"
#define PLQCD_NHI " neg_hi:[0,1,0]"
#define PLQCD_NBOTH " neg_lo:[0,1,0] neg_hi:[0,1,0]"

__device__ __forceinline__ hh2 hswap(hh2 a) { return hh2{a.y, a.x}; }

// Row 2 of a 12-real link, u2_a = conj(u0_b u1_c - u0_c u1_b), from rows 0
// and 1 as dot products of half pairs,
//   re conj(x y - z w) = x . (y.re, -y.im) + z . (-w.re, w.im)
//   im conj(x y - z w) = z . (w.im, w.re) - x . (y.im, y.re)
// rounded to half.
__device__ __forceinline__ void row2_h(uint32_t e[9])
{
   const hh2 u0 = __builtin_bit_cast(hh2, e[0]), u1 = __builtin_bit_cast(hh2, e[1]),
             u2 = __builtin_bit_cast(hh2, e[2]), u3 = __builtin_bit_cast(hh2, e[3]),
             u4 = __builtin_bit_cast(hh2, e[4]), u5 = __builtin_bit_cast(hh2, e[5]);
   const hh2 s3 = hswap(u3), s4 = hswap(u4), s5 = hswap(u5);
   float re6, im6, re7, im7, re8, im8;
   // u6: x u1, y u5, z u2, w u4; u7: x u2, y u3, z u0, w u5; u8: x u0, y u4, z u1, w u3
   asm(PLQCD_D2Z("%0", "%7", "%11", PLQCD_NHI)
       PLQCD_D2Z("%1", "%8", "%13", "")
       PLQCD_D2Z("%2", "%8", "%9", PLQCD_NHI)
       PLQCD_D2Z("%3", "%6", "%14", "")
       PLQCD_D2Z("%4", "%6", "%10", PLQCD_NHI)
       PLQCD_D2Z("%5", "%7", "%12", "")
       PLQCD_D2("%0", "%8", "%10", PLQCD_NLO)
       PLQCD_D2("%1", "%7", "%14", PLQCD_NBOTH)
       PLQCD_D2("%2", "%6", "%11", PLQCD_NLO)
       PLQCD_D2("%3", "%8", "%12", PLQCD_NBOTH)
       PLQCD_D2("%4", "%7", "%9", PLQCD_NLO)
       PLQCD_D2("%5", "%6", "%13", PLQCD_NBOTH)
       "s_nop 2"
       : "=&v"(re6), "=&v"(im6), "=&v"(re7), "=&v"(im7), "=&v"(re8), "=&v"(im8)
       : "v"(u0), "v"(u1), "v"(u2), "v"(u3), "v"(u4), "v"(u5), "v"(s3), "v"(s4), "v"(s5));
   e[6] = __builtin_bit_cast(uint32_t, hh2{(_Float16)re6, (_Float16)im6});
   e[7] = __builtin_bit_cast(uint32_t, hh2{(_Float16)re7, (_Float16)im7});
   e[8] = __builtin_bit_cast(uint32_t, hh2{(_Float16)re8, (_Float16)im8});
}

template <>
struct GaugeIO<PLQCD_PREC_F16, 12> {
   typedef f2 C;
   typedef LinkH U;
   const char *base;  // [parity][mu][Vh] u4 (u00,u01,u02,u10), then [parity][mu][Vh] u2 (u11,u12)
   size_t kstride;    // Vh*16
   template <bool NT = false>
   __device__ __forceinline__ void load_u(int parity, int mu, int j, U &u) const
   {
      const size_t slot = (size_t)(parity * 4 + mu);
      const u4 a = ld<u4, NT>(base + slot * kstride, (uint32_t)j * 16);
      const u2 b = ld<u2, NT>(base + 8 * kstride + slot * (kstride / 2), (uint32_t)j * 8);
      u.e[0] = a.x;
      u.e[1] = a.y;
      u.e[2] = a.z;
      u.e[3] = a.w;
      u.e[4] = b.x;
      u.e[5] = b.y;
      row2_h(u.e);
   }
};

template <>
struct GaugeIO<PLQCD_PREC_F16, 18> {
   typedef f2 C;
   typedef LinkH U;
   const char *base;  // [parity][mu][Vh] u4 (u00..u10), u4 (u11..u21), then [parity][mu][Vh] u32 (u22)
   size_t kstride;    // Vh*16
   template <bool NT = false>
   __device__ __forceinline__ void load_u(int parity, int mu, int j, U &u) const
   {
      const size_t slot = (size_t)(parity * 4 + mu);
      const u4 a = ld<u4, NT>(base + slot * kstride, (uint32_t)j * 16);
      const u4 b = ld<u4, NT>(base + (8 + slot) * kstride, (uint32_t)j * 16);
      u.e[0] = a.x;
      u.e[1] = a.y;
      u.e[2] = a.z;
      u.e[3] = a.w;
      u.e[4] = b.x;
      u.e[5] = b.y;
      u.e[6] = b.z;
      u.e[7] = b.w;
      u.e[8] = ld<uint32_t, NT>(base + 16 * kstride + slot * (kstride / 4), (uint32_t)j * 4);
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
