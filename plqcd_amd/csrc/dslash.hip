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
#include "fields.h"
#include "dslash.h"

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

template <int MU, int MODE, class SIO, class GIO>
__device__ __forceinline__ void hop_dir(const SIO &in, const GIO &g, int p, int j, const Nbr &n,
                                        typename SIO::C acc[12])
{
   typedef typename SIO::C C;
   const int q = 1 - p;
   {  // forward: U_mu(x) (1 - gamma_mu) in(x + mu)
      C f[12], h[6], u[9], r[6];
      in.load(q, n.up[MU], f);
      proj_minus<MU>(f, h);
      g.load(MODE == MODE_COMPAT ? 0 : p, MU, j, u);
      mul_u(u, h, r);
      recon_minus<MU, MODE != MODE_COMPAT>(acc, r);
   }
   {  // backward: U_mu(x - mu)^+ (1 + gamma_mu) in(x - mu)
      C f[12], k[6], u[9], r[6];
      in.load(q, n.dn[MU], f);
      proj_plus<MU>(f, k);
      g.load(q, MU, n.dn[MU], u);
      mul_udag(u, k, r);
      recon_plus<MU>(acc, r);
   }
}

template <int MODE, class SIO, class GIO>
__global__ __launch_bounds__(HOP_BLOCK) void hop_kernel(SIO in, SIO out, GIO g, HopArgs a)
{
   typedef typename SIO::C C;
   // Full-lattice launches interleave the two output parities block by block,
   // so that a link is read as a forward link and as a backward link by sites
   // processed close together in time (and mostly served from L2).
   int p, blk = blockIdx.x;
   if (a.nparity == 2) {
      p = blk & 1;
      blk >>= 1;
   } else {
      p = a.parity0;
   }
   const int j = blk * HOP_BLOCK + threadIdx.x;
   if (j >= a.Vh) return;
   const size_t roff = (size_t)blockIdx.y * a.rhs_stride;
   in.base += roff;
   out.base += roff;

   const Nbr n = neighbours(a, p, j);
   C acc[12];
#pragma unroll
   for (int c = 0; c < 12; c++) acc[c] = C{0, 0};
   hop_dir<0, MODE>(in, g, p, j, n, acc);
   hop_dir<1, MODE>(in, g, p, j, n, acc);
   hop_dir<2, MODE>(in, g, p, j, n, acc);
   hop_dir<3, MODE>(in, g, p, j, n, acc);
   out.store(p, j, acc);
}

template <int PREC, int LINKS>
int launch_t(const plqcd_hip_gauge *g, const plqcd_hip_field *in, plqcd_hip_field *out, int parity, int mode)
{
   const plqcd_hip_lattice *lat = g->lat;
   HopArgs a = make_hop_args(lat, parity);
   typedef SpinorIO<PREC> S;
   typedef GaugeIO<PREC, LINKS> G;
   S si, so;
   si.base = reinterpret_cast<decltype(si.base)>(in->data);
   so.base = reinterpret_cast<decltype(so.base)>(out->data);
   si.Vh = so.Vh = lat->Vh;
   G gg;
   gg.base = reinterpret_cast<decltype(gg.base)>(g->data);
   gg.Vh = lat->Vh;
   a.rhs_stride = in->rhs_bytes / sizeof(*si.base);
   const int nblk = (lat->Vh + HOP_BLOCK - 1) / HOP_BLOCK;
   dim3 grid(nblk * a.nparity, in->nrhs);
   if (mode == PLQCD_HOP_COMPAT)
      hipLaunchKernelGGL((hop_kernel<PLQCD_HOP_COMPAT, S, G>), grid, dim3(HOP_BLOCK), 0, lat->stream, si, so, gg, a);
   else
      hipLaunchKernelGGL((hop_kernel<PLQCD_HOP_WILSON, S, G>), grid, dim3(HOP_BLOCK), 0, lat->stream, si, so, gg, a);
   PLQCD_HIP_TRY(hipGetLastError());
   return 0;
}

}  // namespace

HopArgs make_hop_args(const plqcd_hip_lattice *lat, int parity)
{
   HopArgs a;
   for (int mu = 0; mu < 4; mu++) a.L[mu] = lat->L[mu];
   a.Lxh = lat->Lxh;
   a.Vh = lat->Vh;
   a.nparity = (parity == PLQCD_PARITY_ALL) ? 2 : 1;
   a.parity0 = (parity == PLQCD_PARITY_ODD) ? 1 : 0;
   a.rhs_stride = 0;
   return a;
}

extern "C" int plqcd_hip_hop(const plqcd_hip_gauge *g, const plqcd_hip_field *in, plqcd_hip_field *out, int parity,
                             int mode)
{
   if (!g || !in || !out) return plqcd::set_error(1, "plqcd_hip_hop: null argument");
   if (in == out) return plqcd::set_error(1, "plqcd_hip_hop: in and out must be distinct fields");
   if (in->lat != g->lat || out->lat != g->lat)
      return plqcd::set_error(1, "plqcd_hip_hop: fields and gauge belong to different lattices");
   if (in->prec != g->prec || out->prec != g->prec)
      return plqcd::set_error(1, "plqcd_hip_hop: precision mismatch (gauge %d, in %d, out %d)", g->prec, in->prec,
                              out->prec);
   if (in->nrhs != out->nrhs) return plqcd::set_error(1, "plqcd_hip_hop: RHS count mismatch");
   if (parity < PLQCD_PARITY_EVEN || parity > PLQCD_PARITY_ALL)
      return plqcd::set_error(1, "plqcd_hip_hop: bad parity %d", parity);
   if (mode != PLQCD_HOP_COMPAT && mode != PLQCD_HOP_WILSON)
      return plqcd::set_error(1, "plqcd_hip_hop: bad mode %d", mode);
   switch (g->prec * 100 + g->links) {
      case PLQCD_PREC_F64 * 100 + 18: return launch_t<PLQCD_PREC_F64, 18>(g, in, out, parity, mode);
      case PLQCD_PREC_F64 * 100 + 12: return launch_t<PLQCD_PREC_F64, 12>(g, in, out, parity, mode);
      case PLQCD_PREC_F32 * 100 + 18: return launch_t<PLQCD_PREC_F32, 18>(g, in, out, parity, mode);
      case PLQCD_PREC_F32 * 100 + 12: return launch_t<PLQCD_PREC_F32, 12>(g, in, out, parity, mode);
      case PLQCD_PREC_F16 * 100 + 12: return launch_t<PLQCD_PREC_F16, 12>(g, in, out, parity, mode);
      default: return plqcd::set_error(1, "plqcd_hip_hop: unsupported gauge format");
   }
}
