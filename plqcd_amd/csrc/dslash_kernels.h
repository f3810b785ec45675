// The Wilson hopping matrix on gfx950.
//
//   out(x) = sum_mu [ U_mu(x) (1 - gamma_mu) in(x+mu) + U_mu(x-mu)^+ (1 + gamma_mu) in(x-mu) ]
//
// (plqcd.pdf Eq. 2.1; lib/dirac.c:35-408 and :414-791 compute the two parity
// halves of it in four phases over half-spinor scratch fields).  Here it is a
// single gather pass: one thread per output site reads its 8 neighbour
// spinors and 8 links straight from HBM-resident structure-of-arrays fields,
// projects each neighbour to a half-spinor (2 colour vectors) before the SU(3)
// multiply, and writes its output once with a non-temporal store.  There is
// no scratch field, no halo and no MPI: periodic neighbours are computed from
// the site coordinates inside the kernel.
//
// Gamma matrices, projectors and reconstruction are the reference's chiral
// basis (lib/dirac.c:95-114, :163-195, :269-311, :371-398):
//   mu  upper((1-g)f)          upper((1+g)g)           s2 +=            s3 +=
//   0   (f0+f2,   f1+f3)       (g0-g2,   g1-g3)        a0 - b0          a1 - b1
//   1   (f0+i f3, f1+i f2)     (g0-i g3, g1-i g2)      -i a1 + i b1     -i a0 + i b0
//   2   (f0+f3,   f1-f2)       (g0-g3,   g1+g2)        -a1 + b1         a0 - b0
//   3   (f0+i f2, f1-i f3)     (g0-i g2, g1+i g3)      -i a0 + i b0     i a1 - i b1
// with a = U_mu(x) h, b = U_mu(x-mu)^+ k, s0 += a0 + b0, s1 += a1 + b1.
//
// MODE_COMPAT reproduces what the compiled reference outputs: forward terms
// do not reach s2/s3 (dirac.c:373-374 and :756-757 assign the mu=0 backward
// term over them) and odd output sites take their forward links from the even
// site with the same index inside its parity (dirac.c:643-644).
#pragma once

#include <algorithm>
#include <type_traits>
#include <vector>

#include "fields.h"
#include "dslash.h"
#include "reduce.h"

namespace {

constexpr int MODE_COMPAT = PLQCD_HOP_COMPAT;

// Site of thread (parity p, index j inside the parity) and the within-parity
// indices of its 8 neighbours (which all have the other parity).
struct Nbr {
   int up[4], dn[4];
};

__device__ __forceinline__ Nbr neighbours(const HopArgs &a, int p, int j)
{
   const int row = j / a.Lxh;  // row = y + L1*(z + L2*t)
   const int xh = j - row * a.Lxh;
   const int y = row % a.L[1];
   const int zt = row / a.L[1];
   const int z = zt % a.L[2];
   const int t = zt / a.L[2];
   const int x = 2 * xh + ((p + y + z + t) & 1);
   Nbr n;
   const int xp = (x + 1 == a.L[0]) ? 0 : x + 1;
   const int xm = (x == 0) ? a.L[0] - 1 : x - 1;
   n.up[0] = row * a.Lxh + (xp >> 1);
   n.dn[0] = row * a.Lxh + (xm >> 1);
   const int ryp = (y == a.L[1] - 1) ? row - (a.L[1] - 1) : row + 1;
   const int rym = (y == 0) ? row + (a.L[1] - 1) : row - 1;
   const int sz = a.L[1];
   const int rzp = (z == a.L[2] - 1) ? row - (a.L[2] - 1) * sz : row + sz;
   const int rzm = (z == 0) ? row + (a.L[2] - 1) * sz : row - sz;
   const int st = a.L[1] * a.L[2];
   const int rtp = (t == a.L[3] - 1) ? row - (a.L[3] - 1) * st : row + st;
   const int rtm = (t == 0) ? row + (a.L[3] - 1) * st : row - st;
   n.up[1] = ryp * a.Lxh + xh;
   n.dn[1] = rym * a.Lxh + xh;
   n.up[2] = rzp * a.Lxh + xh;
   n.dn[2] = rzm * a.Lxh + xh;
   n.up[3] = rtp * a.Lxh + xh;
   n.dn[3] = rtm * a.Lxh + xh;
   return n;
}

// h = upper((1 - gamma_mu) f): h[0..2] = first colour vector, h[3..5] = second
template <int MU, class C>
__device__ __forceinline__ void proj_minus(const C f[12], C h[6])
{
#pragma unroll
   for (int c = 0; c < 3; c++) {
      if (MU == 0) { h[c] = f[c] + f[6 + c];     h[3 + c] = f[3 + c] + f[9 + c]; }
      if (MU == 1) { h[c] = f[c] + ci(f[9 + c]); h[3 + c] = f[3 + c] + ci(f[6 + c]); }
      if (MU == 2) { h[c] = f[c] + f[9 + c];     h[3 + c] = f[3 + c] - f[6 + c]; }
      if (MU == 3) { h[c] = f[c] + ci(f[6 + c]); h[3 + c] = f[3 + c] - ci(f[9 + c]); }
   }
}
// k = upper((1 + gamma_mu) g)
template <int MU, class C>
__device__ __forceinline__ void proj_plus(const C g[12], C k[6])
{
#pragma unroll
   for (int c = 0; c < 3; c++) {
      if (MU == 0) { k[c] = g[c] - g[6 + c];     k[3 + c] = g[3 + c] - g[9 + c]; }
      if (MU == 1) { k[c] = g[c] - ci(g[9 + c]); k[3 + c] = g[3 + c] - ci(g[6 + c]); }
      if (MU == 2) { k[c] = g[c] - g[9 + c];     k[3 + c] = g[3 + c] + g[6 + c]; }
      if (MU == 3) { k[c] = g[c] - ci(g[6 + c]); k[3 + c] = g[3 + c] + ci(g[9 + c]); }
   }
}
// acc += forward reconstruction of a (lower components only if LOWER)
template <int MU, bool LOWER, class C>
__device__ __forceinline__ void recon_minus(C acc[12], const C a[6])
{
#pragma unroll
   for (int c = 0; c < 3; c++) {
      acc[c] += a[c];
      acc[3 + c] += a[3 + c];
      if (LOWER) {
         if (MU == 0) { acc[6 + c] += a[c];          acc[9 + c] += a[3 + c]; }
         if (MU == 1) { acc[6 + c] -= ci(a[3 + c]); acc[9 + c] -= ci(a[c]); }
         if (MU == 2) { acc[6 + c] -= a[3 + c];     acc[9 + c] += a[c]; }
         if (MU == 3) { acc[6 + c] -= ci(a[c]);     acc[9 + c] += ci(a[3 + c]); }
      }
   }
}
template <int MU, class C>
__device__ __forceinline__ void recon_plus(C acc[12], const C b[6])
{
#pragma unroll
   for (int c = 0; c < 3; c++) {
      acc[c] += b[c];
      acc[3 + c] += b[3 + c];
      if (MU == 0) { acc[6 + c] -= b[c];          acc[9 + c] -= b[3 + c]; }
      if (MU == 1) { acc[6 + c] += ci(b[3 + c]); acc[9 + c] += ci(b[c]); }
      if (MU == 2) { acc[6 + c] += b[3 + c];     acc[9 + c] -= b[c]; }
      if (MU == 3) { acc[6 + c] += ci(b[c]);     acc[9 + c] -= ci(b[3 + c]); }
   }
}

// r = U v for the two colour vectors of a half-spinor
template <class C>
__device__ __forceinline__ void mul_u(const C u[9], const C v[6], C r[6])
{
#pragma unroll
   for (int s = 0; s < 2; s++)
#pragma unroll
      for (int a = 0; a < 3; a++) {
         C t = cmul(u[3 * a], v[3 * s]);
         t = cmadd(u[3 * a + 1], v[3 * s + 1], t);
         r[3 * s + a] = cmadd(u[3 * a + 2], v[3 * s + 2], t);
      }
}
// r = U^+ v
template <class C>
__device__ __forceinline__ void mul_udag(const C u[9], const C v[6], C r[6])
{
#pragma unroll
   for (int s = 0; s < 2; s++)
#pragma unroll
      for (int a = 0; a < 3; a++) {
         C t = cmul(cconj(u[a]), v[3 * s]);
         t = cjmadd(u[3 + a], v[3 * s + 1], t);
         r[3 * s + a] = cjmadd(u[6 + a], v[3 * s + 2], t);
      }
}

// (1 + SIGN gamma_mu) projection and reconstruction, SIGN = -1 or +1
template <int MU, int SIGN, class C>
__device__ __forceinline__ void proj(const C f[12], C h[6])
{
   if (SIGN < 0) proj_minus<MU>(f, h);
   else proj_plus<MU>(f, h);
}
template <int MU, int SIGN, bool LOWER, class C>
__device__ __forceinline__ void recon(C acc[12], const C a[6])
{
   if (SIGN < 0) recon_minus<MU, LOWER>(acc, a);
   else recon_plus<MU>(acc, a);
}

// ---------------------------------------------------------------- terms
//
// term_u:    r = U (1 + SG gamma_mu) f        (a forward term, before recon)
// term_udag: r = U^+ (1 + SG gamma_mu) f      (a backward term)
// accum:     acc += reconstruction of r
// on the held forms of fields.h.  F64 and F32 project, multiply and add in
// the working precision.

template <int MU, int SG, class C>
__device__ __forceinline__ void term_u(const SpinV<C> &f, const LinkV<C> &u, HalfV<C> &r)
{
   C h[6];
   proj<MU, SG>(f.e, h);
   mul_u(u.e, h, r.e);
}
template <int MU, int SG, class C>
__device__ __forceinline__ void term_udag(const SpinV<C> &f, const LinkV<C> &u, HalfV<C> &r)
{
   C h[6];
   proj<MU, SG>(f.e, h);
   mul_udag(u.e, h, r.e);
}
template <int MU, int SG, bool LOWER, class C>
__device__ __forceinline__ void accum(C acc[12], const HalfV<C> &r)
{
   recon<MU, SG, LOWER>(acc, r.e);
}

// F16: the projection adds stored half pairs of one site (one v_pk_add_f16
// per half-spinor component; sign changes and the factor i are operand
// modifiers), the SU(3) multiply is v_dot2_f32_f16 on half pairs with float
// accumulation, and the site scale of the neighbour multiplies the float
// result as it is accumulated (acc += s * r, one fma per real).  Every kernel
// variant uses these functions, so a site's result does not depend on the
// variant.

// X g for X = 0: +1, 1: -1, 2: +i, 3: -i
template <int X>
__device__ __forceinline__ hh2 xmul(hh2 g)
{
   if (X == 0) return g;
   if (X == 1) return -g;
   if (X == 2) return hh2{-g.y, g.x};
   return hh2{g.y, -g.x};
}

// The projection of proj_minus / proj_plus as h[c] = f[c] + X0 f[P0 + c],
// h[3 + c] = f[3 + c] + X1 f[P1 + c]; the lower rows of the reconstruction
// (recon_minus / recon_plus) as acc[6 + c] += Y2 r[S2 + c], acc[9 + c] +=
// Y3 r[S3 + c] (factors coded as in xmul).
template <int MU, int SG>
struct SpinTab {
   static constexpr int P0 = (MU == 0 || MU == 3) ? 6 : 9;
   static constexpr int P1 = 15 - P0;
   static constexpr int X0 = ((MU == 0 || MU == 2) ? 0 : 2) ^ (SG > 0);
   static constexpr int X1 = (MU == 0 ? 0 : MU == 1 ? 2 : MU == 2 ? 1 : 3) ^ (SG > 0);
   static constexpr int S2 = (MU == 0 || MU == 3) ? 0 : 3;
   static constexpr int S3 = 3 - S2;
   static constexpr int Y2 = (MU == 0 ? 0 : MU == 1 ? 3 : MU == 2 ? 1 : 3) ^ (SG > 0);
   static constexpr int Y3 = (MU == 0 ? 0 : MU == 1 ? 3 : MU == 2 ? 0 : 2) ^ (SG > 0);
};

// h = upper((1 + SG gamma_mu) f) in halves, and its swapped pairs (h.im, h.re)
template <int MU, int SG>
__device__ __forceinline__ void proj_h(const SpinH &f, hh2 h[6], hh2 hs[6])
{
   typedef SpinTab<MU, SG> T;
#pragma unroll
   for (int c = 0; c < 3; c++) {
      h[c] = f.h(c) + xmul<T::X0>(f.h(T::P0 + c));
      h[3 + c] = f.h(3 + c) + xmul<T::X1>(f.h(T::P1 + c));
      hs[c] = hswap(h[c]);
      hs[3 + c] = hswap(h[3 + c]);
   }
}

// One colour vector (h, hs = hswap(h)) times U (DAG false) or U^+ (DAG
// true), r = (re, im) of the 3 result colours, as one assembly block of 18
// v_dot2_f32_f16 (see fields.h for why assembly and the closing s_nop):
//   U h:   re_a = sum_b U_ab . (h_b.re, -h_b.im)   im_a = sum_b U_ab . hs_b
//   U^+ h: re_a = sum_b U_ba . h_b                  im_a = sum_b U_ba . (hs_b.re, -hs_b.im)
template <bool DAG>
__device__ __forceinline__ void mat_h(const LinkH &u, const hh2 h[3], const hh2 hs[3], float r[6])
{
   // operands: %0..%5 re0 im0 re1 im1 re2 im2, %6..%14 u_k (k = 3 row + col),
   // %15..%17 h, %18..%20 hs
   if constexpr (!DAG) {
      asm(PLQCD_D2Z("%0", "%6", "%15", PLQCD_NHI)
          PLQCD_D2Z("%1", "%6", "%18", "")
          PLQCD_D2Z("%2", "%9", "%15", PLQCD_NHI)
          PLQCD_D2Z("%3", "%9", "%18", "")
          PLQCD_D2Z("%4", "%12", "%15", PLQCD_NHI)
          PLQCD_D2Z("%5", "%12", "%18", "")
          PLQCD_D2("%0", "%7", "%16", PLQCD_NHI)
          PLQCD_D2("%1", "%7", "%19", "")
          PLQCD_D2("%2", "%10", "%16", PLQCD_NHI)
          PLQCD_D2("%3", "%10", "%19", "")
          PLQCD_D2("%4", "%13", "%16", PLQCD_NHI)
          PLQCD_D2("%5", "%13", "%19", "")
          PLQCD_D2("%0", "%8", "%17", PLQCD_NHI)
          PLQCD_D2("%1", "%8", "%20", "")
          PLQCD_D2("%2", "%11", "%17", PLQCD_NHI)
          PLQCD_D2("%3", "%11", "%20", "")
          PLQCD_D2("%4", "%14", "%17", PLQCD_NHI)
          PLQCD_D2("%5", "%14", "%20", "")
          "s_nop 2"
          : "=&v"(r[0]), "=&v"(r[1]), "=&v"(r[2]), "=&v"(r[3]), "=&v"(r[4]), "=&v"(r[5])
          : "v"(u.h(0)), "v"(u.h(1)), "v"(u.h(2)), "v"(u.h(3)), "v"(u.h(4)), "v"(u.h(5)), "v"(u.h(6)),
            "v"(u.h(7)), "v"(u.h(8)), "v"(h[0]), "v"(h[1]), "v"(h[2]), "v"(hs[0]), "v"(hs[1]), "v"(hs[2]));
   } else {
      asm(PLQCD_D2Z("%0", "%6", "%15", "")
          PLQCD_D2Z("%1", "%6", "%18", PLQCD_NHI)
          PLQCD_D2Z("%2", "%7", "%15", "")
          PLQCD_D2Z("%3", "%7", "%18", PLQCD_NHI)
          PLQCD_D2Z("%4", "%8", "%15", "")
          PLQCD_D2Z("%5", "%8", "%18", PLQCD_NHI)
          PLQCD_D2("%0", "%9", "%16", "")
          PLQCD_D2("%1", "%9", "%19", PLQCD_NHI)
          PLQCD_D2("%2", "%10", "%16", "")
          PLQCD_D2("%3", "%10", "%19", PLQCD_NHI)
          PLQCD_D2("%4", "%11", "%16", "")
          PLQCD_D2("%5", "%11", "%19", PLQCD_NHI)
          PLQCD_D2("%0", "%12", "%17", "")
          PLQCD_D2("%1", "%12", "%20", PLQCD_NHI)
          PLQCD_D2("%2", "%13", "%17", "")
          PLQCD_D2("%3", "%13", "%20", PLQCD_NHI)
          PLQCD_D2("%4", "%14", "%17", "")
          PLQCD_D2("%5", "%14", "%20", PLQCD_NHI)
          "s_nop 2"
          : "=&v"(r[0]), "=&v"(r[1]), "=&v"(r[2]), "=&v"(r[3]), "=&v"(r[4]), "=&v"(r[5])
          : "v"(u.h(0)), "v"(u.h(1)), "v"(u.h(2)), "v"(u.h(3)), "v"(u.h(4)), "v"(u.h(5)), "v"(u.h(6)),
            "v"(u.h(7)), "v"(u.h(8)), "v"(h[0]), "v"(h[1]), "v"(h[2]), "v"(hs[0]), "v"(hs[1]), "v"(hs[2]));
   }
}

template <int MU, int SG>
__device__ __forceinline__ void term_u(const SpinH &f, const LinkH &u, HalfH &r)
{
   hh2 h[6], hs[6];
   proj_h<MU, SG>(f, h, hs);
   mat_h<false>(u, h, hs, r.e);
   mat_h<false>(u, h + 3, hs + 3, r.e + 6);
   r.e[12] = f.s();
}

template <int MU, int SG>
__device__ __forceinline__ void term_udag(const SpinH &f, const LinkH &u, HalfH &r)
{
   hh2 h[6], hs[6];
   proj_h<MU, SG>(f, h, hs);
   mat_h<true>(u, h, hs, r.e);
   mat_h<true>(u, h + 3, hs + 3, r.e + 6);
   r.e[12] = f.s();
}

// acc += s * Y (x + i y), Y coded as in xmul
template <int Y>
__device__ __forceinline__ void acc_add(f2 &acc, float s, float x, float y)
{
   if (Y == 0) { acc.x = fma(s, x, acc.x);  acc.y = fma(s, y, acc.y); }
   if (Y == 1) { acc.x = fma(-s, x, acc.x); acc.y = fma(-s, y, acc.y); }
   if (Y == 2) { acc.x = fma(-s, y, acc.x); acc.y = fma(s, x, acc.y); }
   if (Y == 3) { acc.x = fma(s, y, acc.x);  acc.y = fma(-s, x, acc.y); }
}

// recon_minus (SG < 0; lower rows only if LOWER) / recon_plus (SG > 0)
template <int MU, int SG, bool LOWER>
__device__ __forceinline__ void accum(f2 acc[12], const HalfH &r)
{
   typedef SpinTab<MU, SG> T;
   const float s = r.e[12];
#pragma unroll
   for (int c = 0; c < 3; c++) {
      acc_add<0>(acc[c], s, r.e[2 * c], r.e[2 * c + 1]);
      acc_add<0>(acc[3 + c], s, r.e[2 * (3 + c)], r.e[2 * (3 + c) + 1]);
      if (SG > 0 || LOWER) {
         acc_add<T::Y2>(acc[6 + c], s, r.e[2 * (T::S2 + c)], r.e[2 * (T::S2 + c) + 1]);
         acc_add<T::Y3>(acc[9 + c], s, r.e[2 * (T::S3 + c)], r.e[2 * (T::S3 + c) + 1]);
      }
   }
}

// Forward term of direction MU for output site (p, j); jnb = site of x+mu.
// H uses (1 - gamma_mu) forward and (1 + gamma_mu) backward; the adjoint
// H^+ = gamma5 H gamma5 has the signs swapped (gamma5 anticommutes with
// gamma_mu), so DAG only flips the projector signs.
template <int MU, int MODE, bool DAG, class SIO, class GIO>
__device__ __forceinline__ void hop_fwd(const SIO &in, const GIO &g, int p, int j, int jnb, typename SIO::C acc[12])
{
   constexpr int SG = DAG ? 1 : -1;
   typename SIO::Sp f;
   typename GIO::U u;
   typename SIO::Hv r;
   in.load_sp(1 - p, jnb, f);
   g.load_u(MODE == MODE_COMPAT ? 0 : p, MU, j, u);
   term_u<MU, SG>(f, u, r);
   accum<MU, SG, MODE != MODE_COMPAT>(acc, r);
}

// Backward term of direction MU; jnb = site of x-mu (its link U_mu(x-mu) too).
template <int MU, bool DAG, class SIO, class GIO>
__device__ __forceinline__ void hop_bwd(const SIO &in, const GIO &g, int p, int jnb, typename SIO::C acc[12])
{
   constexpr int SG = DAG ? -1 : 1;
   typename SIO::Sp f;
   typename GIO::U u;
   typename SIO::Hv r;
   in.load_sp(1 - p, jnb, f);
   g.load_u(1 - p, MU, jnb, u);
   term_udag<MU, SG>(f, u, r);
   accum<MU, SG, true>(acc, r);
}

template <int MU, int MODE, bool DAG, class SIO, class GIO>
__device__ __forceinline__ void hop_dir(const SIO &in, const GIO &g, int p, int j, const Nbr &n,
                                        typename SIO::C acc[12])
{
   hop_fwd<MU, MODE, DAG>(in, g, p, j, n.up[MU], acc);
   hop_bwd<MU, DAG>(in, g, p, n.dn[MU], acc);
}

// Maps the hardware block index to a logical tile.  Blocks are handed to
// the 8 XCDs round-robin (block b runs on XCD b % 8), and each XCD has its
// own L2, so the order decides which L2 sees which part of the lattice:
//   order 0: logical = b
//   order 1: XCD k takes the k-th contiguous eighth of the lattice
//   order 2: the lattice is cut into chunks of 8*group logical blocks; XCD k
//            takes the k-th run of `group` blocks of every chunk, so all XCDs
//            sweep the lattice together and neighbouring tiles share an L2
// The output parity alternates between consecutive logical blocks, so a link
// is read as a forward and as a backward link by nearby tiles.
__device__ __forceinline__ int logical_block(const HopArgs &a)
{
   const int b = blockIdx.x, nb = gridDim.x;
   if (a.order == 1) return (b & 7) * (nb >> 3) + (b >> 3);
   if (a.order == 2) {
      const int q = b >> 3, G = a.group;
      return (q / G) * 8 * G + (b & 7) * G + (q % G);
   }
   return b;
}

// The RHS of the fields that lane-RHS r of a launch handles: r0 + r, or the
// r-th entry of the launch's RHS list.
__device__ __forceinline__ int rhs_of(const HopArgs &a, int r)
{
   return a.rlist ? a.rlist[r] : a.r0 + r;
}

// Operands of the epilogues: the per-RHS step alpha (EPI = 2), the
// right-hand side b (EPI = 3) and the centre field (EPI = 4).
template <class SIO>
struct CgFuse {
   SIO p;  // EPI 3: b; EPI 4: the centre field
   const double *alpha;
};

// Epilogue: EPI = 0 stores H in; EPI = 1 stores in(x) - kappa_r * (H in)(x).
// EPI = 2 forms Ap = in(x) - kappa_r * (H in)(x) in registers and applies the
// CG residual update at the site instead of storing it: out -= alpha_r Ap (out
// holds the residual), in the working precision of the kernel (float for
// F32/F16 storage, double for F64).  The iterate's update x += alpha_r p is
// left to the solver's next pass over p.
// EPI = 3 (F64 storage) stores the residual b - Ap, with b in f.p, in the
// precision of `out`, for every RHS of the launch (the reliable update of the
// mixed-precision CG: the true residual goes straight into the low-precision
// iterate's residual; the solver launches it once per run of consecutive
// RHS being updated).  No per-RHS predicate: a data-dependent condition
// around the stores made the compiler keep both paths' values live (280
// registers, one wave per SIMD, against 144 and three waves).
// EPI = 4 is EPI = 1 with the centre taken from another field, f.p: out =
// c(x) - kappa_r * (H in)(x) (the even-odd operator's second half hop, whose
// input lives on the other parity; gather kernel only).
// With NORM the site's |out|^2 joins the deterministic chunk reduction.  `in`
// and `out` are positioned at the RHS `r`, whose kappa and alpha apply.
template <int EPI, bool NORM, class SIO, class SO>
__device__ __forceinline__ double epilogue(const SIO &in, const SO &out, const HopArgs &a, int r, int p, int j,
                                           typename SIO::C acc[12], const typename SIO::C *cen, const CgFuse<SIO> &f)
{
   typedef typename SIO::real real;
   typedef typename SIO::C C;
   if (EPI >= 1) {
      C c0[12];
      if constexpr (EPI == 4)
         f.p.at_rhs(r).load(p, j, c0);
      else if (cen == nullptr)
         in.load(p, j, c0);
      const real kap = (real)a.kappa[r];
#pragma unroll
      for (int m = 0; m < 12; m++) acc[m] = cxmky(cen ? cen[m] : c0[m], kap, acc[m]);
   }
   if constexpr (EPI == 3) {
      const SIO b = f.p.at_rhs(r);
      const uint32_t o = b.off(j);
      d2 rs[12];
#pragma unroll
      for (int m = 0; m < 12; m++) {
         const C bm = ld<C>(b.cbase(p, m), o);
         rs[m] = d2{(double)bm.x - (double)acc[m].x, (double)bm.y - (double)acc[m].y};
      }
      store_d(out, p, j, rs);
      return NORM ? norm2(rs) : 0.0;
   } else if constexpr (EPI == 2) {
      const real al = (real)f.alpha[r];
      C rs[12];
      out.load(p, j, rs);
      double nrm = 0.0;
#pragma unroll
      for (int m = 0; m < 12; m++) {
         rs[m] = C{fma(-al, acc[m].x, rs[m].x), fma(-al, acc[m].y, rs[m].y)};
         nrm = fma((double)rs[m].y, (double)rs[m].y, fma((double)rs[m].x, (double)rs[m].x, nrm));
      }
      out.store(p, j, rs);
      return NORM ? nrm : 0.0;
   } else {
      double nrm = 0.0;
      if (NORM) {
#pragma unroll
         for (int m = 0; m < 12; m++)
            nrm = fma((double)acc[m].y, (double)acc[m].y, fma((double)acc[m].x, (double)acc[m].x, nrm));
      }
      out.store(p, j, acc);
      return nrm;
   }
}

// One block of the gather kernel: logical block `blk` of the launch.
template <int MODE, bool DAG, int EPI, bool NORM, int BLOCK, class SIO, class GIO, class SO>
__device__ __forceinline__ void hop_block(const SIO &in, const SO &out, const GIO &g, const HopArgs &a, int blk,
                                          const CgFuse<SIO> &f)
{
   typedef typename SIO::C C;
   int p;
   if (a.nparity == 2) {
      p = blk & 1;
      blk >>= 1;
   } else {
      p = a.parity0;
   }
   const int r = threadIdx.x % a.R;
   const int js = threadIdx.x / a.R;
   const int rr = rhs_of(a, r);
   const int j = blk * a.spb + js;
   double nrm = 0.0;
   if (js < a.spb && j < a.Vh) {
      const SIO si = in.at_rhs(rr);
      const SO so = out.at_rhs(rr);
      const Nbr n = neighbours(a, p, j);
      C acc[12];
#pragma unroll
      for (int c = 0; c < 12; c++) acc[c] = C{0, 0};
      hop_dir<0, MODE, DAG>(si, g, p, j, n, acc);
      hop_dir<1, MODE, DAG>(si, g, p, j, n, acc);
      hop_dir<2, MODE, DAG>(si, g, p, j, n, acc);
      hop_dir<3, MODE, DAG>(si, g, p, j, n, acc);
      nrm = epilogue<EPI, NORM>(si, so, a, rr, p, j, acc, (const C *)nullptr, f);
   }
   if (NORM) chunk_reduce<BLOCK>(nrm, js, r, a.R, p * a.nchunk + (blk * a.spb + js) / REDUCE_CHUNK,
                                 j - (js & (REDUCE_CHUNK - 1)) < a.Vh, a.partial + (size_t)rr * 2 * a.nchunk);
}

template <int MODE, bool DAG, int EPI, bool NORM, int BLOCK, class SIO, class GIO, class SO>
__global__ __launch_bounds__(BLOCK) void hop_kernel(SIO in, SO out, GIO g, HopArgs a, CgFuse<SIO> f)
{
   hop_block<MODE, DAG, EPI, NORM, BLOCK>(in, out, g, a, logical_block(a), f);
}

// RHS-blocked gather: each lane handles RB consecutive right-hand sides of
// one site, so a link is loaded (and, with 12-real links, rebuilt) once per
// RB right-hand sides.  Lanes per site G = R / RB.  Every RHS goes through
// the same operation sequence as in hop_block.
template <int MODE, bool DAG, int EPI, bool NORM, int BLOCK, int RB, class SIO, class GIO>
__device__ __forceinline__ void hop_block_rb(const SIO &in, const SIO &out, const GIO &g, const HopArgs &a, int blk,
                                             const CgFuse<SIO> &f)
{
   typedef typename SIO::C C;
   constexpr int SF = DAG ? 1 : -1, SB = -SF;
   int p;
   if (a.nparity == 2) {
      p = blk & 1;
      blk >>= 1;
   } else {
      p = a.parity0;
   }
   const int G = a.R / RB;
   const int gi = threadIdx.x % G;
   const int js = threadIdx.x / G;
   const int j = blk * a.spb + js;
   double nrm[RB];
#pragma unroll
   for (int k = 0; k < RB; k++) nrm[k] = 0.0;
   if (js < a.spb && j < a.Vh) {
      const Nbr n = neighbours(a, p, j);
      C acc[RB][12];
#pragma unroll
      for (int k = 0; k < RB; k++)
#pragma unroll
         for (int c = 0; c < 12; c++) acc[k][c] = C{0, 0};
      const SIO si = in.at_rhs(rhs_of(a, gi * RB));
#define PLQCD_RB_DIR(MU)                                                                      \
   {                                                                                          \
      typename GIO::U u;                                                                      \
      g.load_u(MODE == MODE_COMPAT ? 0 : p, MU, j, u);                                        \
      _Pragma("unroll") for (int k = 0; k < RB; k++)                                          \
      {                                                                                       \
         typename SIO::Sp f;                                                                  \
         typename SIO::Hv r;                                                                  \
         si.at_rhs(rhs_of(a, gi * RB + k)).load_sp(1 - p, n.up[MU], f);                           \
         term_u<MU, SF>(f, u, r);                                                             \
         accum<MU, SF, MODE != MODE_COMPAT>(acc[k], r);                                       \
      }                                                                                       \
      g.load_u(1 - p, MU, n.dn[MU], u);                                                       \
      _Pragma("unroll") for (int k = 0; k < RB; k++)                                          \
      {                                                                                       \
         typename SIO::Sp f;                                                                  \
         typename SIO::Hv r;                                                                  \
         si.at_rhs(rhs_of(a, gi * RB + k)).load_sp(1 - p, n.dn[MU], f);                           \
         term_udag<MU, SB>(f, u, r);                                                          \
         accum<MU, SB, true>(acc[k], r);                                                      \
      }                                                                                       \
   }
      PLQCD_RB_DIR(0)
      PLQCD_RB_DIR(1)
      PLQCD_RB_DIR(2)
      PLQCD_RB_DIR(3)
#undef PLQCD_RB_DIR
#pragma unroll
      for (int k = 0; k < RB; k++)
         nrm[k] = epilogue<EPI, NORM>(in.at_rhs(rhs_of(a, gi * RB + k)), out.at_rhs(rhs_of(a, gi * RB + k)), a,
                                      rhs_of(a, gi * RB + k), p, j, acc[k],
                                      (const C *)nullptr, f);
   }
   if (NORM) {
#pragma unroll
      for (int k = 0; k < RB; k++)
         chunk_reduce<BLOCK>(nrm[k], js, gi, G, p * a.nchunk + (blk * a.spb + js) / REDUCE_CHUNK,
                             j - (js & (REDUCE_CHUNK - 1)) < a.Vh,
                             a.partial + (size_t)(rhs_of(a, gi * RB + k)) * 2 * a.nchunk, k);
   }
}

template <int MODE, bool DAG, int EPI, bool NORM, int BLOCK, int RB, class SIO, class GIO>
__global__ __launch_bounds__(BLOCK) void hop_rb_kernel(SIO in, SIO out, GIO g, HopArgs a, CgFuse<SIO> f)
{
   hop_block_rb<MODE, DAG, EPI, NORM, BLOCK, RB>(in, out, g, a, logical_block(a), f);
}

// One application of H on each of several independent lattices (each with its
// own gauge field and fields) in a single launch over the concatenated
// logical blocks of the items.
template <class SIO, class GIO>
struct BatchItem {
   SIO in, out;
   GIO g;
   HopArgs a;
   int first;  // first global block of this item
   int nb;     // blocks of this item
   int march;  // t-march blocks (a.T t-slices per lane) instead of gather blocks
};
// A launch takes up to kBatchItems items by value, as a kernel argument (no
// device copy of the list); longer batches take several launches.
constexpr int kBatchItems = 16;
constexpr int kBatchGroup = 8;  // consecutive logical blocks an XCD takes in a ragged batch
template <class SIO, class GIO>
struct BatchList {
   BatchItem<SIO, GIO> it[kBatchItems];
};

// t-march variant (full lattice, either mode): a thread owns one (x,y,z)
// column and walks T consecutive t-slices.  Along t each spinor and each U_t
// is loaded once: psi(t+1), read as the forward t-neighbour at step t, is kept
// (it is the centre site of step t+1) and its projection, multiplied by
// U_t(t+1)^+ at step t+1 (that step's forward link), gives the backward
// t-term of step t+2.  The spatial directions read their neighbours from L2
// as the gather kernel does.  The site parity alternates from step to step;
// the index of the site inside its parity advances by S = L0/2*L1*L2 per
// t-slice.  Terms are accumulated in the same order as in hop_kernel.
template <int MODE, bool DAG, int EPI, bool NORM, int BLOCK, class SIO, class GIO, class SO>
__device__ __forceinline__ void hop_tmarch_block(const SIO &in, const SO &out, const GIO &g, const HopArgs &a, int lb,
                                                 const CgFuse<SIO> &fuse)
{
   typedef typename SIO::C C;
   constexpr int SF = DAG ? 1 : -1;  // forward projector sign
   const int S = a.S;
   const int nbS = (S + a.spb - 1) / a.spb;
   const int q = lb & 1;  // parity of the column's first site
   lb >>= 1;
   const int sb = lb % nbS;
   const int chunk = lb / nbS;
   const int r = threadIdx.x % a.R;
   const int js = threadIdx.x / a.R;
   const int rr = rhs_of(a, r);
   const int c = sb * a.spb + js;
   const bool valid = js < a.spb && c < S;  // a batch block may have spare lanes
   const SIO si = in.at_rhs(rr);
   const SO so = out.at_rhs(rr);

   const int t0 = chunk * a.T;
   const int cc = valid ? c : 0;
   const int row = cc / a.Lxh;  // y + L1*z
   const int xh = cc - row * a.Lxh;
   const int y = row % a.L[1];
   const int z = row / a.L[1];
   const int x = 2 * xh + ((q + y + z + t0) & 1);
   int cu[3], cd[3];
   {
      const int xp = (x + 1 == a.L[0]) ? 0 : x + 1;
      const int xm = (x == 0) ? a.L[0] - 1 : x - 1;
      cu[0] = row * a.Lxh + (xp >> 1);
      cd[0] = row * a.Lxh + (xm >> 1);
      const int ryp = (y == a.L[1] - 1) ? row - (a.L[1] - 1) : row + 1;
      const int rym = (y == 0) ? row + (a.L[1] - 1) : row - 1;
      const int rzp = (z == a.L[2] - 1) ? row - (a.L[2] - 1) * a.L[1] : row + a.L[1];
      const int rzm = (z == 0) ? row + (a.L[2] - 1) * a.L[1] : row - a.L[1];
      cu[1] = ryp * a.Lxh + xh;
      cd[1] = rym * a.Lxh + xh;
      cu[2] = rzp * a.Lxh + xh;
      cd[2] = rzm * a.Lxh + xh;
   }
   const int Lt = a.L[3];
   // The lane's centre spinor and backward t term live across t-slices.  In
   // the FP64 residual epilogue they stay in LDS (each lane reads only its
   // own column, so no barrier), which frees 72 registers (226-254 instead of
   // 256 + 20-38 AGPRs) and lets two blocks share a CU: at 32^3 x 64 with 12
   // RHS it took 4.1 instead of 5.0 ms.  M (EPI 1) took the same time either
   // way but, with two blocks per CU sharing the L2, moved 2.77 instead of
   // 1.81 times its compulsory bytes, and plain H ran 2 % slower, so they and
   // the other epilogues keep the state in registers, as do FP32 and FP16.
   typedef typename SIO::Sp Sp;
   typedef typename SIO::Hv Hv;
   constexpr bool CL = std::is_same<typename SIO::real, double>::value && EPI == 3;
   __shared__ typename Sp::E lcen[CL ? Sp::N : 1][CL ? BLOCK : 1];
   __shared__ typename Hv::E lbt[CL ? Hv::N : 1][CL ? BLOCK : 1];
   const int tid = threadIdx.x;
   Sp cen;
   Hv b;
   auto put_cen = [&](const Sp &v) {
      if constexpr (CL) {
#pragma unroll
         for (int m = 0; m < Sp::N; m++) lcen[m][tid] = v.e[m];
      } else {
         cen = v;
      }
   };
   auto get_cen = [&](Sp &v) {
      if constexpr (CL) {
#pragma unroll
         for (int m = 0; m < Sp::N; m++) v.e[m] = lcen[m][tid];
      } else {
         v = cen;
      }
   };
   auto put_b = [&](const Hv &v) {
      if constexpr (CL) {
#pragma unroll
         for (int m = 0; m < Hv::N; m++) lbt[m][tid] = v.e[m];
      } else {
         b = v;
      }
   };
   auto get_b = [&](Hv &v) {
      if constexpr (CL) {
#pragma unroll
         for (int m = 0; m < Hv::N; m++) v.e[m] = lbt[m][tid];
      } else {
         v = b;
      }
   };
   // prologue: cen = psi(t0); b = U_t(t0-1)^+ (1 -+ g_t) psi(t0-1)
   if (valid) {
      Sp f, c0;
      Hv b0;
      typename GIO::U u;
      si.load_sp(q, c + t0 * S, c0);
      put_cen(c0);
      const int tm = (t0 == 0 ? Lt : t0) - 1;
      si.load_sp(q ^ 1, c + tm * S, f);
      g.load_u(q ^ 1, 3, c + tm * S, u);
      term_udag<3, -SF>(f, u, b0);
      put_b(b0);
   }
   for (int s = 0; s < a.T; s++) {
      const int t = t0 + s;
      const int p = q ^ (s & 1);
      const int base = t * S;
      const int j = cc + base;
      double nrm = 0.0;
      if (valid) {
         C acc[12];
#pragma unroll
         for (int m = 0; m < 12; m++) acc[m] = C{0, 0};
         hop_fwd<0, MODE, DAG>(si, g, p, j, cu[0] + base, acc);
         hop_bwd<0, DAG>(si, g, p, cd[0] + base, acc);
         hop_fwd<1, MODE, DAG>(si, g, p, j, cu[1] + base, acc);
         hop_bwd<1, DAG>(si, g, p, cd[1] + base, acc);
         hop_fwd<2, MODE, DAG>(si, g, p, j, cu[2] + base, acc);
         hop_bwd<2, DAG>(si, g, p, cd[2] + base, acc);
         Sp f, cs;
         {
            typename GIO::U u;
            Hv rr, bt;
            const int tp = (t + 1 == Lt) ? 0 : t + 1;
            si.load_sp(p ^ 1, c + tp * S, f);
            // U_t(t) serves this slice's forward and the next slice's
            // backward t term from registers and no other block reads it: a
            // streaming load (FP16 with one RHS 6-9 % faster at 32^3 x 64,
            // 12 RHS under 1 % slower; choosing by R at run time cost the
            // one-RHS kernel its gain).  Compat mode's forward term takes the
            // even site's link with the same index, a second load.
            g.template load_u<true>(p, 3, j, u);
            if constexpr (MODE == MODE_COMPAT) {
               typename GIO::U uf;
               g.load_u(0, 3, j, uf);
               term_u<3, SF>(f, uf, rr);
            } else {
               term_u<3, SF>(f, u, rr);
            }
            accum<3, SF, MODE != MODE_COMPAT>(acc, rr);
            get_b(bt);
            accum<3, -SF, true>(acc, bt);
            get_cen(cs);
            term_udag<3, -SF>(cs, u, bt);
            put_b(bt);
         }
         if (EPI >= 1) {
            C cv[12];
            SIO::to_c(cs, cv);
            nrm = epilogue<EPI, NORM>(si, so, a, rr, p, j, acc, cv, fuse);
         } else {
            nrm = epilogue<EPI, NORM>(si, so, a, rr, p, j, acc, (const C *)nullptr, fuse);
         }
         put_cen(f);
      }
      if (NORM) chunk_reduce<BLOCK>(nrm, js, r, a.R, p * a.nchunk + (sb * a.spb + js + base) / REDUCE_CHUNK, valid,
                                    a.partial + (size_t)rr * 2 * a.nchunk, s);
   }
}

template <int MODE, bool DAG, int EPI, bool NORM, int BLOCK, class SIO, class GIO, class SO>
__global__ __launch_bounds__(BLOCK) void hop_tmarch_kernel(SIO in, SO out, GIO g, HopArgs a, CgFuse<SIO> fuse)
{
   hop_tmarch_block<MODE, DAG, EPI, NORM, BLOCK>(in, out, g, a, logical_block(a), fuse);
}

// One launch over a ragged batch (see BatchItem): each item runs t-march or
// gather blocks.  Blocks go round-robin to the 8 XCDs; as in block order 2,
// the XCDs sweep the concatenated logical blocks together, XCD k (blockIdx %
// 8 == k) taking the k-th run of kBatchGroup consecutive blocks of every
// 8 kBatchGroup.  The grid is a multiple of 8; a block takes logical blocks
// at a stride of gridDim / 8 runs, so any such grid covers the batch (the
// launch gives each logical block a block of its own).  KIND 0: every item runs gather blocks, 1:
// every item t-march blocks, 2: each item its own kind (the kernel then holds
// both code paths and needs the registers of the larger).
template <int MODE, int KIND, int BLOCK, class SIO, class GIO>
__global__ __launch_bounds__(BLOCK) void hop_batch_kernel(const BatchList<SIO, GIO> items, int nitems, int total)
{
   const int xcd = blockIdx.x & 7, per = gridDim.x >> 3;
   const int span = ((total + 8 * kBatchGroup - 1) / (8 * kBatchGroup)) * kBatchGroup;  // blocks per XCD, rounded up
   for (int i = blockIdx.x >> 3; i < span; i += per) {
      const int b = ((i / kBatchGroup) * 8 + xcd) * kBatchGroup + i % kBatchGroup;
      if (b >= total) break;
      int it = 0;
      while (it + 1 < nitems && b >= items.it[it + 1].first) it++;
      const BatchItem<SIO, GIO> &I = items.it[it];
      if constexpr (KIND == 1) {
         hop_tmarch_block<MODE, false, 0, false, BLOCK>(I.in, I.out, I.g, I.a, b - I.first, CgFuse<SIO>{});
      } else if constexpr (KIND == 2) {
         if (I.march)
            hop_tmarch_block<MODE, false, 0, false, BLOCK>(I.in, I.out, I.g, I.a, b - I.first, CgFuse<SIO>{});
         else
            hop_block<MODE, false, 0, false, BLOCK>(I.in, I.out, I.g, I.a, b - I.first, CgFuse<SIO>{});
      } else {
         hop_block<MODE, false, 0, false, BLOCK>(I.in, I.out, I.g, I.a, b - I.first, CgFuse<SIO>{});
      }
   }
}

// LDS bytes of the tiled kernel for BLOCK lanes: two buffers of centre
// spinors, the backward t terms and, in FP64 blocks of up to 128 lanes, the
// forward x links of a slice (hop_tile_kernel).
template <class S, class G, int BLOCK>
constexpr bool tile_xsh()
{
   return BLOCK <= 128 && std::is_same<typename S::real, double>::value;
}
template <class S, class G, int BLOCK>
constexpr size_t tile_lds_bytes()
{
   return (2 * sizeof(typename S::Sp) + sizeof(typename S::Hv)) * BLOCK +
          (tile_xsh<S, G, BLOCK>() ? 9 * sizeof(typename G::U::E) * BLOCK : 0);
}

// One direction of the tiled kernel: neighbours inside the tile come from the
// LDS slots su / sd (>= 0), the others from memory; same operation sequence
// as hop_fwd + hop_bwd.  With XSH (x direction only) the lane's forward link
// is streamed from memory and published in ux, and the backward link -- the
// forward link of the x - 1 lane, slot sd -- is read from there after a
// barrier; every lane of the block calls this.  In compat mode the forward
// link is the even site's with the same index j (slot sx in the tile): with
// XSH an odd lane reads it from ux after the barrier, otherwise from memory.
// NB: where the neighbours are -- 0 always in the tile (x: whole rows), 2
// never (z with one-plane tiles), 1 either, by lane (slot >= 0: in the tile).
// With NB 1 the compiler merges the two reads into one flat load.
template <int MU, int NB, int MODE, bool DAG, int BLOCK, bool XSH, class SIO, class GIO>
__device__ __forceinline__ void tile_dir(const SIO &in, const GIO &g, const typename SIO::Sp::E (*buf)[BLOCK], int p,
                                         int j, int base, int cu, int cd, int su, int sd, int sx,
                                         typename SIO::C acc[12], typename GIO::U::E (*ux)[BLOCK])
{
   typedef typename SIO::Sp Sp;
   constexpr int SG = DAG ? 1 : -1;
   constexpr bool share = MU == 0 && XSH;
   constexpr bool compat = MODE == MODE_COMPAT;
   Sp f;
   typename GIO::U u;
   typename SIO::Hv r;
   auto nb_load = [&](int slot, int jnb) {
      if (NB == 0 || (NB == 1 && slot >= 0)) {
#pragma unroll
         for (int m = 0; m < Sp::N; m++) f.e[m] = buf[m][slot];
      } else {
         in.load_sp(1 - p, jnb + base, f);
      }
   };
   nb_load(su, cu);
   if constexpr (share) {
      g.template load_u<true>(p, MU, j, u);
#pragma unroll
      for (int k = 0; k < 9; k++) ux[k][threadIdx.x] = u.e[k];
      if constexpr (compat) {
         __syncthreads();
         if (p) {
#pragma unroll
            for (int k = 0; k < 9; k++) u.e[k] = ux[k][sx];
         }
      }
   } else {
      g.load_u(compat ? 0 : p, MU, j, u);
   }
   term_u<MU, SG>(f, u, r);
   accum<MU, SG, !compat>(acc, r);
   nb_load(sd, cd);
   if constexpr (share) {
      if constexpr (!compat) __syncthreads();
#pragma unroll
      for (int k = 0; k < 9; k++) u.e[k] = ux[k][sd];
   } else {
      g.load_u(1 - p, MU, cd + base, u);
   }
   term_udag<MU, -SG>(f, u, r);
   accum<MU, -SG, true>(acc, r);
}

// LDS-tiled t-march (full lattice, one RHS).  A block owns the
// columns of a tile of Lxh x ty x tz (x, y, z) sites -- whole x rows -- for
// both parities: lanes q = 0 and 1 of a column hold its two x-adjacent sites,
// and the block walks T t-slices.  The centre spinors of a t-slice are in LDS
// (two buffers, one barrier per slice), and the x, y and z
// neighbours inside the tile are read from there; only the tile's y/z halo and
// the t+1 spinor are loaded from memory.  The centre spinor and the backward
// t term stay in LDS rather than in registers.  Links and the t direction are
// handled as in hop_tmarch_kernel, except that FP64 blocks of up to 128 lanes
// pass the x links between the two lanes that use each one through LDS (a
// second barrier per slice).  Terms are accumulated in the order of hop_kernel.
// Compat mode takes the forward links of odd sites from the even site with
// the same index (its t link with a second load, since the true one still
// serves the next slice's backward term).  BLOCK = 2 Lxh ty tz lanes.
template <int MODE, bool DAG, int EPI, bool NORM, int BLOCK, class SIO, class GIO>
__global__ __launch_bounds__(BLOCK) void hop_tile_kernel(SIO in, SIO out, GIO g, HopArgs a, CgFuse<SIO> fuse)
{
   typedef typename SIO::C C;
   const int r0 = rhs_of(a, 0);  // the launch's one RHS
   in = in.at_rhs(r0);
   out = out.at_rhs(r0);
   constexpr int SF = DAG ? 1 : -1;  // forward projector sign
   typedef typename SIO::Sp Sp;
   typedef typename SIO::Hv Hv;
   __shared__ typename Sp::E sm[2][Sp::N][BLOCK];
   __shared__ typename Hv::E sb[Hv::N][BLOCK];  // the lane's backward t term b (read and written by that lane only)
   // x links through LDS in FP64 blocks of two waves: a CU then holds several
   // blocks, and one works while another waits at the extra barrier (FP64
   // 18-real, 32x4x1 tiles at 32^3 x 64: 407 against 427 us).  FP32 tiles ran
   // up to 27 % slower with it (12-real, 32x4x1, 16 slices: 220 against 173 us).
   constexpr bool XSH = tile_xsh<SIO, GIO, BLOCK>();
   constexpr bool SPLIT = std::is_same<typename SIO::real, double>::value;
   __shared__ typename GIO::U::E ux[XSH ? 9 : 1][XSH ? BLOCK : 1];  // the lanes' forward x links of this slice
   typedef typename GIO::U::E(*UxRow)[BLOCK];
   const int S = a.S;
   constexpr int ncol = BLOCK / 2;
   const int nty = a.L[1] / a.ty;
   const int ntile = nty * (a.L[2] / a.tz);
   const int lb = logical_block(a);
   const int tile = lb % ntile;
   const int chunk = lb / ntile;
   const int tid = threadIdx.x;
   const int q = tid >= ncol ? 1 : 0;  // parity of the lane's first site
   const int lc = tid - q * ncol;      // xh + Lxh * (ly + ty * lz)
   const int xh = lc % a.Lxh;
   const int lyz = lc / a.Lxh;
   const int ly = lyz % a.ty, lz = lyz / a.ty;
   const int y0 = (tile % nty) * a.ty, z0 = (tile / nty) * a.tz;
   const int y = y0 + ly, z = z0 + lz;
   const int row = y + a.L[1] * z;
   const int c = row * a.Lxh + xh;
   const int t0 = chunk * a.T;
   const int x = 2 * xh + ((q + y + z + t0) & 1);  // the same at every t-slice
   const int sx = (1 - q) * ncol + lc;              // the lane with the same index j, other parity
   int cu[3], cd[3], su[3], sd[3];
   {
      const int oth = (1 - q) * ncol;  // lanes of the other parity
      const int xp = (x + 1 == a.L[0]) ? 0 : x + 1;
      const int xm = (x == 0) ? a.L[0] - 1 : x - 1;
      cu[0] = row * a.Lxh + (xp >> 1);
      cd[0] = row * a.Lxh + (xm >> 1);
      su[0] = oth + lc - xh + (xp >> 1);
      sd[0] = oth + lc - xh + (xm >> 1);
      const int yp = (y + 1 == a.L[1]) ? 0 : y + 1;
      const int ym = (y == 0) ? a.L[1] - 1 : y - 1;
      const int zp = (z + 1 == a.L[2]) ? 0 : z + 1;
      const int zm = (z == 0) ? a.L[2] - 1 : z - 1;
      cu[1] = (yp + a.L[1] * z) * a.Lxh + xh;
      cd[1] = (ym + a.L[1] * z) * a.Lxh + xh;
      cu[2] = (y + a.L[1] * zp) * a.Lxh + xh;
      cd[2] = (y + a.L[1] * zm) * a.Lxh + xh;
      const int lyp = yp - y0, lym = ym - y0, lzp = zp - z0, lzm = zm - z0;
      su[1] = (lyp >= 0 && lyp < a.ty) ? oth + xh + a.Lxh * (lyp + a.ty * lz) : -1;
      sd[1] = (lym >= 0 && lym < a.ty) ? oth + xh + a.Lxh * (lym + a.ty * lz) : -1;
      su[2] = (lzp >= 0 && lzp < a.tz) ? oth + xh + a.Lxh * (ly + a.ty * lzp) : -1;
      sd[2] = (lzm >= 0 && lzm < a.tz) ? oth + xh + a.Lxh * (ly + a.ty * lzm) : -1;
   }
   const int Lt = a.L[3];
   // prologue: psi(t0) into the first buffer; b = U_t(t0-1)^+ (1 -+ g_t) psi(t0-1)
   {
      Sp f;
      typename GIO::U u;
      Hv b;
      in.load_sp(q, c + t0 * S, f);
#pragma unroll
      for (int m = 0; m < Sp::N; m++) sm[0][m][tid] = f.e[m];
      const int tm = (t0 == 0 ? Lt : t0) - 1;
      in.load_sp(q ^ 1, c + tm * S, f);
      g.load_u(q ^ 1, 3, c + tm * S, u);
      term_udag<3, -SF>(f, u, b);
#pragma unroll
      for (int m = 0; m < Hv::N; m++) sb[m][tid] = b.e[m];
   }
   for (int s = 0; s < a.T; s++) {
      const int t = t0 + s;
      const int p = q ^ (s & 1);
      const int base = t * S;
      const int j = c + base;
      // buf holds the centre spinors of this t-slice; the t+1 spinors go to
      // nxt, whose readers (the previous slice) have all passed this barrier
      typename Sp::E(*buf)[BLOCK] = sm[s & 1];
      typename Sp::E(*nxt)[BLOCK] = sm[(s + 1) & 1];
      __syncthreads();
      C acc[12];
#pragma unroll
      for (int m = 0; m < 12; m++) acc[m] = C{0, 0};
      // the scheduling barriers keep each direction's loads in its own
      // section, which holds the kernel within 256 VGPRs without spills
      // The x neighbours (whole rows) are read from LDS with no test, and in
      // FP64 the z neighbours of one-plane tiles from memory with none: each
      // neighbour read under a per-lane test became one flat load serving
      // both places, waited for with every global load in flight (FP64
      // 12-real 32x1x4 tiles at 32^3 x 64: 333 against 351 us; FP32 18-real
      // 32x8x1: 240 against 245 us with the x test gone, 2.5 % slower with
      // the z one gone too, so FP32 keeps that).
      tile_dir<0, 0, MODE, DAG, BLOCK, XSH>(in, g, buf, p, j, base, cu[0], cd[0], su[0], sd[0], sx, acc, (UxRow)ux);
      __builtin_amdgcn_sched_barrier(0);
      tile_dir<1, 1, MODE, DAG, BLOCK, XSH>(in, g, buf, p, j, base, cu[1], cd[1], su[1], sd[1], sx, acc, (UxRow)ux);
      __builtin_amdgcn_sched_barrier(0);
      if (SPLIT && a.tz == 1)
         tile_dir<2, 2, MODE, DAG, BLOCK, XSH>(in, g, buf, p, j, base, cu[2], cd[2], su[2], sd[2], sx, acc, (UxRow)ux);
      else
         tile_dir<2, 1, MODE, DAG, BLOCK, XSH>(in, g, buf, p, j, base, cu[2], cd[2], su[2], sd[2], sx, acc, (UxRow)ux);
      __builtin_amdgcn_sched_barrier(0);
      {
         Sp f;
         typename GIO::U u;
         Hv rr, b;
         const int tp = (t + 1 == Lt) ? 0 : t + 1;
         in.load_sp(p ^ 1, c + tp * S, f);
#pragma unroll
         for (int m = 0; m < Sp::N; m++) nxt[m][tid] = f.e[m];
         g.template load_u<true>(p, 3, j, u);  // read by this lane only (see hop_tmarch_kernel)
         if constexpr (MODE == MODE_COMPAT) {
            typename GIO::U uf;
            g.load_u(0, 3, j, uf);
            term_u<3, SF>(f, uf, rr);
         } else {
            term_u<3, SF>(f, u, rr);
         }
         accum<3, SF, MODE != MODE_COMPAT>(acc, rr);
#pragma unroll
         for (int m = 0; m < Hv::N; m++) b.e[m] = sb[m][tid];
         accum<3, -SF, true>(acc, b);
#pragma unroll
         for (int m = 0; m < Sp::N; m++) f.e[m] = buf[m][tid];
         term_udag<3, -SF>(f, u, b);
#pragma unroll
         for (int m = 0; m < Hv::N; m++) sb[m][tid] = b.e[m];
      }
      double nrm;
      if (EPI >= 1) {
         Sp cs;
         C cen[12];
#pragma unroll
         for (int m = 0; m < Sp::N; m++) cs.e[m] = buf[m][tid];
         SIO::to_c(cs, cen);
         nrm = epilogue<EPI, NORM>(in, out, a, r0, p, j, acc, cen, fuse);
      } else {
         nrm = epilogue<EPI, NORM>(in, out, a, r0, p, j, acc, (const C *)nullptr, fuse);
      }
      if (NORM)
         chunk_reduce<BLOCK>(nrm, lc, 0, 1, p * a.nchunk + j / REDUCE_CHUNK, true,
                             a.partial + (size_t)r0 * 2 * a.nchunk, s);
   }
}

// kernel variants: gather (1 RHS per lane), t-march, RHS-blocked gather, LDS tile
enum Variant { V_GATHER = 0, V_MARCH = 1, V_RB2 = 2, V_RB4 = 4, V_TILE = 8 };

// S: the input field's accessor; SO: the output's (differs only with EPI 3,
// which runs the gather kernel).  Returns 0 when the kernel was enqueued, and
// an error when the variant does not exist for these template arguments (the
// host-side checks of launch_t keep such launches out; this keeps a future
// change from silently launching nothing).
template <int MODE, bool DAG, int EPI, bool NORM, int BLOCK, class S, class G, class SO>
int launch_one(int variant, dim3 grid, dim3 block, hipStream_t st, S si, SO so, G gg, const HopArgs &a,
               const CgFuse<S> &f)
{
   constexpr bool same = std::is_same<S, SO>::value && EPI != 3;  // the tile and RHS-blocked kernels
   if constexpr (EPI == 4) {  // the external-centre epilogue exists for the gather kernel only
      if (variant == V_GATHER) {
         hipLaunchKernelGGL((hop_kernel<MODE, DAG, EPI, NORM, BLOCK, S, G, SO>), grid, block, 0, st, si, so, gg, a, f);
         return 0;
      }
      return plqcd::set_error(2, "hop: the external-centre epilogue runs the gather kernel only (variant %d)", variant);
   }
   if (variant == V_MARCH) {
      hipLaunchKernelGGL((hop_tmarch_kernel<MODE, DAG, EPI, NORM, BLOCK, S, G, SO>), grid, block, 0, st, si, so, gg, a,
                         f);
      return 0;
   } else if (variant == V_TILE) {
      // two LDS buffers of BLOCK spinors, BLOCK half-spinors and (FP64 blocks
      // of up to 128 lanes) BLOCK x links must fit the CU's 160 KiB
      if constexpr (same && tile_lds_bytes<S, G, BLOCK>() <= 136 * 1024) {
         hipLaunchKernelGGL((hop_tile_kernel<MODE, DAG, EPI, NORM, BLOCK, S, G>), grid, block, 0, st, si, so, gg, a, f);
         return 0;
      }
   } else if (variant == V_RB2) {
      if constexpr (same) {
         hipLaunchKernelGGL((hop_rb_kernel<MODE, DAG, EPI, NORM, BLOCK, 2, S, G>), grid, block, 0, st, si, so, gg, a, f);
         return 0;
      }
   } else if (variant == V_RB4) {
      if constexpr (same) {
         hipLaunchKernelGGL((hop_rb_kernel<MODE, DAG, EPI, NORM, BLOCK, 4, S, G>), grid, block, 0, st, si, so, gg, a, f);
         return 0;
      }
   } else if (variant == V_GATHER) {
      hipLaunchKernelGGL((hop_kernel<MODE, DAG, EPI, NORM, BLOCK, S, G, SO>), grid, block, 0, st, si, so, gg, a, f);
      return 0;
   }
   return plqcd::set_error(2, "hop: kernel variant %d with %d-lane blocks does not exist for this operation", variant,
                           BLOCK);
}

template <int MODE, bool DAG, int EPI, bool NORM, class S, class G, class SO>
int launch_block(int maxblock, int variant, dim3 grid, dim3 block, hipStream_t st, S si, SO so, G gg,
                 const HopArgs &a, const CgFuse<S> &f = CgFuse<S>{})
{
   if (maxblock <= 128) return launch_one<MODE, DAG, EPI, NORM, 128>(variant, grid, block, st, si, so, gg, a, f);
   if (maxblock <= 256) return launch_one<MODE, DAG, EPI, NORM, 256>(variant, grid, block, st, si, so, gg, a, f);
   return launch_one<MODE, DAG, EPI, NORM, 512>(variant, grid, block, st, si, so, gg, a, f);
}

template <int PREC, int LINKS>
int launch_t(const plqcd_hip_gauge *g, const plqcd_hip_field *in, plqcd_hip_field *out, int parity, int mode,
             const plqcd::DslashOp &op)
{
   const plqcd_hip_lattice *lat = g->lat;
   HopArgs a = make_hop_args(lat, parity);
   typedef SpinorIO<PREC> S;
   typedef GaugeIO<PREC, LINKS> G;
   const S si = spinor_io<PREC>(in);
   const G gg = gauge_io<PREC, LINKS>(g);
   const bool resid = op.resid_b != nullptr;  // EPI 3: out may have another precision
   const bool centre = op.centre != nullptr;  // EPI 4: the gather kernel only
   // lanes: the window of right-hand sides [op.r0, op.r0 + R) of the fields
   const int R = op.nr > 0 ? op.nr : in->nrhs;
   a.r0 = op.nr > 0 && !op.rlist ? op.r0 : 0;
   a.rlist = op.rlist;
   const bool solver = op.kappa != nullptr;
   HopTuning tu = plqcd::default_tuning(PREC, LINKS, R, solver);
   plqcd::lookup_tuning(plqcd::TuneKey{{lat->L[0], lat->L[1], lat->L[2], lat->L[3]}, PREC, LINKS, R,
                                       solver ? plqcd::kTuneOpSolver : plqcd::kTuneOpHop},
                        &tu);
   // RHS per lane (RHS-blocked gather) when it divides R; none for the
   // residual epilogue, and the external-centre one runs the gather kernel only
   const int rb = !resid && !centre && (tu.rb == 2 || tu.rb == 4) && R % tu.rb == 0 ? tu.rb : 1;
   const int lanes = R / rb;  // lanes per site
   // block shape: `lanes` lanes per site and a whole number of reduction chunks
   int target = tu.block;
   if (lanes * REDUCE_CHUNK > target) {
      target = 128;
      while (target < lanes * REDUCE_CHUNK) target *= 2;
   }
   if (target > 512)
      return plqcd::set_error(2, "hop: %d right-hand sides exceed the per-launch limit of %d", R, 512 / REDUCE_CHUNK);
   a.R = R;
   a.spb = REDUCE_CHUNK * (target / (REDUCE_CHUNK * lanes));
   const int threads = lanes * a.spb;
   const int maxblock = threads <= 128 ? 128 : threads <= 256 ? 256 : 512;
   a.kappa = op.kappa;
   a.partial = op.partial;
   a.nchunk = plqcd::reduce_chunks(lat->Vh);
   const bool dag = op.dag != 0;
   const int epi = resid ? 3 : op.alpha ? 2 : op.centre ? 4 : op.kappa ? 1 : 0;
   const bool norm = op.partial != nullptr;
   if ((dag || epi || norm) && mode != PLQCD_HOP_WILSON)
      return plqcd::set_error(2, "hop: M, M^+ and norms need the wilson mode");
   if (epi == 2 && (!dag || !norm || !op.kappa))
      return plqcd::set_error(2, "hop: the CG-update epilogue needs M^+, a norm and kappa");
   if (epi == 4 && (dag || norm || !op.kappa))
      return plqcd::set_error(2, "hop: the external-centre epilogue needs kappa, and no M^+ or norm");
   if (epi == 3 && (PREC != PLQCD_PREC_F64 || !dag || !norm || !op.kappa))
      return plqcd::set_error(2, "hop: the residual epilogue needs F64 storage, M^+, a norm and kappa");
   CgFuse<S> fu{};
   if (epi == 2) {
      fu.alpha = op.alpha;
   } else if (epi == 4) {
      fu.p = spinor_io<PREC>(op.centre);
   } else if (epi == 3) {
      fu.p = spinor_io<PREC>(op.resid_b);
   }
   if (norm && a.nparity != 2) return plqcd::set_error(2, "hop: norms are only fused into full-lattice launches");
   const int T = tu.tmarch;
   // the t-march kernel serves full-lattice applications (either mode) whose
   // time extent is a multiple of T (and, with a norm, whose t-slices hold
   // whole reduction chunks)
   const bool march = !centre && rb == 1 && T > 0 && a.nparity == 2 && lat->L[3] % T == 0 &&
                      (!norm || a.S % REDUCE_CHUNK == 0);
   // the LDS-tiled t-march: one RHS, a tile of whole x rows dividing the y and
   // z extents, 2 Lxh ty tz lanes per block (at most 256 with F64 storage for
   // the LDS), and, with a norm, reduction chunks that do not leave a tile row
   const int tlanes = 2 * lat->Lxh * tu.ty * tu.tz;
   const bool tiled = march && !resid && R == 1 && tu.ty > 0 && tu.tz > 0 && lat->L[1] % tu.ty == 0 &&
                      lat->L[2] % tu.tz == 0 &&
                      (tlanes == 128 || tlanes == 256 || (tlanes == 512 && PREC != PLQCD_PREC_F64)) &&
                      (!norm || (lat->Lxh * tu.ty) % REDUCE_CHUNK == 0);
   const int variant = tiled ? V_TILE : march ? V_MARCH : (rb == 2 ? V_RB2 : rb == 4 ? V_RB4 : V_GATHER);
   a.T = march ? T : 1;
   int nb, nthreads = threads, mblock = maxblock;
   if (tiled) {
      a.ty = tu.ty;
      a.tz = tu.tz;
      nb = (lat->L[1] / tu.ty) * (lat->L[2] / tu.tz) * (lat->L[3] / T);
      nthreads = mblock = tlanes;
   } else {
      const int nblk = march ? ((a.S + a.spb - 1) / a.spb) * (lat->L[3] / T) : (lat->Vh + a.spb - 1) / a.spb;
      nb = nblk * a.nparity;
   }
   // the remapped orders need whole rounds of 8 XCDs (and of 8*group blocks)
   a.order = tu.order;
   a.group = tu.group > 0 ? tu.group : 1;
   if ((a.order == 1 && nb % 8) || (a.order == 2 && nb % (8 * a.group))) a.order = 0;
   const dim3 grid(nb), block(nthreads);
   hipStream_t st = lat->stream;
   int lrc = 0;
   if (epi == 3) {
      if constexpr (PREC == PLQCD_PREC_F64) {
         if (out->prec == PLQCD_PREC_F64)
            lrc = launch_block<PLQCD_HOP_WILSON, true, 3, true>(mblock, variant, grid, block, st, si,
                                                                 spinor_io<PLQCD_PREC_F64>(out), gg, a, fu);
         else if (out->prec == PLQCD_PREC_F32)
            lrc = launch_block<PLQCD_HOP_WILSON, true, 3, true>(mblock, variant, grid, block, st, si,
                                                                 spinor_io<PLQCD_PREC_F32>(out), gg, a, fu);
         else
            lrc = launch_block<PLQCD_HOP_WILSON, true, 3, true>(mblock, variant, grid, block, st, si,
                                                                 spinor_io<PLQCD_PREC_F16>(out), gg, a, fu);
      }
      if (lrc) return lrc;
      PLQCD_HIP_TRY(hipGetLastError());
      return 0;
   }
   const S so = spinor_io<PREC>(out);
   if (mode == PLQCD_HOP_COMPAT) {
      lrc = launch_block<PLQCD_HOP_COMPAT, false, 0, false>(mblock, variant, grid, block, st, si, so, gg, a);
   } else if (!dag) {
      if (epi == 0) lrc = launch_block<PLQCD_HOP_WILSON, false, 0, false>(mblock, variant, grid, block, st, si, so, gg, a);
      else if (epi == 4)
         lrc = launch_block<PLQCD_HOP_WILSON, false, 4, false>(mblock, variant, grid, block, st, si, so, gg, a, fu);
      else if (!norm) lrc = launch_block<PLQCD_HOP_WILSON, false, 1, false>(mblock, variant, grid, block, st, si, so, gg, a);
      else lrc = launch_block<PLQCD_HOP_WILSON, false, 1, true>(mblock, variant, grid, block, st, si, so, gg, a);
   } else {
      if (epi == 0) lrc = launch_block<PLQCD_HOP_WILSON, true, 0, false>(mblock, variant, grid, block, st, si, so, gg, a);
      else if (epi == 2)
         lrc = launch_block<PLQCD_HOP_WILSON, true, 2, true>(mblock, variant, grid, block, st, si, so, gg, a, fu);
      else if (!norm) lrc = launch_block<PLQCD_HOP_WILSON, true, 1, false>(mblock, variant, grid, block, st, si, so, gg, a);
      else lrc = launch_block<PLQCD_HOP_WILSON, true, 1, true>(mblock, variant, grid, block, st, si, so, gg, a);
   }
   if (lrc) return lrc;
   PLQCD_HIP_TRY(hipGetLastError());
   return 0;
}


template <int PREC, int LINKS>
int launch_batch_t(int n, const plqcd_hip_gauge *const *g, const plqcd_hip_field *const *in,
                   plqcd_hip_field *const *out, int mode)
{
   typedef SpinorIO<PREC> S;
   typedef GaugeIO<PREC, LINKS> G;
   static_assert(sizeof(BatchList<S, G>) <= 4096, "kernel argument too large");
   for (int i = 0; i < n; i++)
      if (in[i]->nrhs * REDUCE_CHUNK > 256) return plqcd::set_error(2, "hop_batch: at most 32 right-hand sides per item");
   hipStream_t st = g[0]->lat->stream;
   for (int i0 = 0; i0 < n; i0 += kBatchItems) {
      const int m = std::min(kBatchItems, n - i0);
      // block size: the largest whole number of reduction chunks of sites
      // times the RHS count that fits 256 lanes, maximised over the items (a
      // batch of 12-RHS items runs 192-lane blocks with no idle lane)
      int block = 64;
      for (int i = 0; i < m; i++) {
         const int R = in[i0 + i]->nrhs;
         block = std::max(block, REDUCE_CHUNK * (256 / (REDUCE_CHUNK * R)) * R);
      }
      BatchList<S, G> L{};
      int total = 0, nmarch = 0;
      for (int i = 0; i < m; i++) {
         BatchItem<S, G> &I = L.it[i];
         I.in = spinor_io<PREC>(in[i0 + i]);
         I.out = spinor_io<PREC>(out[i0 + i]);
         I.g = gauge_io<PREC, LINKS>(g[i0 + i]);
         const plqcd_hip_lattice *lat = g[i0 + i]->lat;
         I.a = make_hop_args(lat, PLQCD_PARITY_ALL);
         I.a.R = in[i0 + i]->nrhs;
         I.a.spb = REDUCE_CHUNK * (block / (REDUCE_CHUNK * I.a.R));
         // items march along t with the precision's t-slice count when it
         // divides the time extent; the others use gather blocks
         const int T = plqcd::default_tuning(PREC, LINKS, I.a.R).tmarch;
         I.march = T > 0 && lat->L[3] % T == 0;
         I.a.T = I.march ? T : 1;
         I.nb = I.march ? 2 * ((I.a.S + I.a.spb - 1) / I.a.spb) * (lat->L[3] / T)
                        : 2 * ((I.a.Vh + I.a.spb - 1) / I.a.spb);
         I.first = total;
         total += I.nb;
         nmarch += I.march;
      }
      // one hardware block per logical block (8 XCDs x span, see
      // hop_batch_kernel): the hardware then balances the items' blocks over
      // the CUs.  A persistent grid of occupancy x CUs blocks, each taking a
      // fixed stride of logical blocks, was 2-9 % slower than this at
      // 16^3x32 + 24^3x48 + 32^3x64 with 12 RHS.
      auto go = [&](auto kern) {
         const int span = ((total + 8 * kBatchGroup - 1) / (8 * kBatchGroup)) * kBatchGroup;
         hipLaunchKernelGGL(kern, dim3(8 * span), dim3(block), 0, st, L, m, total);
      };
      if (mode == PLQCD_HOP_COMPAT) {
         if (nmarch == 0) go(hop_batch_kernel<PLQCD_HOP_COMPAT, 0, 256, S, G>);
         else if (nmarch == m) go(hop_batch_kernel<PLQCD_HOP_COMPAT, 1, 256, S, G>);
         else go(hop_batch_kernel<PLQCD_HOP_COMPAT, 2, 256, S, G>);
      } else if (nmarch == 0) {
         go(hop_batch_kernel<PLQCD_HOP_WILSON, 0, 256, S, G>);
      } else if (nmarch == m) {
         go(hop_batch_kernel<PLQCD_HOP_WILSON, 1, 256, S, G>);
      } else {
         go(hop_batch_kernel<PLQCD_HOP_WILSON, 2, 256, S, G>);
      }
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) return plqcd::set_error(10, "hop_batch: %s", hipGetErrorString(e));
   }
   return 0;
}

}  // namespace

// Entry points of one (precision, link format), explicitly instantiated in
// dslash_<precision>_l<links>.hip.
template <int PREC, int LINKS>
int plqcd::hop_launch(const plqcd_hip_gauge *g, const plqcd_hip_field *in, plqcd_hip_field *out, int parity, int mode,
                      const DslashOp &op)
{
   return launch_t<PREC, LINKS>(g, in, out, parity, mode, op);
}

template <int PREC, int LINKS>
int plqcd::hop_batch_launch(int n, const plqcd_hip_gauge *const *g, const plqcd_hip_field *const *in,
                            plqcd_hip_field *const *out, int mode)
{
   return launch_batch_t<PREC, LINKS>(n, g, in, out, mode);
}
