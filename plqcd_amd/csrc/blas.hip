// Site-wise vector kernels for the solvers: every operand may be stored in
// any precision; arithmetic is done in double, except the CG iterate update
// x += alpha p, which runs in the storage's working precision.  Scalars are per right-hand
// side and live in device memory, so a solver iteration needs no host round
// trip.  Sums use the deterministic chunk reduction of reduce.h.
#include "blas.h"
#include "fields.h"
#include "reduce.h"

namespace {

constexpr int kBlock = 256;

// Lanes (js, lane): tid = js * R + lane; a block covers spb sites of one
// parity, and lane k works on the field's RHS r = r0 + k.
struct BlasArgs {
   int Vh, R, spb, nbp;  // nbp: blocks per parity
   int nchunk;
   int r0;               // first RHS of the launch's window
   double *partial;      // [nrhs][2*nchunk] or null
};

__device__ __forceinline__ bool blas_site(const BlasArgs &a, int &p, int &j, int &js, int &r)
{
   const int blk = blockIdx.x;
   p = blk / a.nbp;
   r = a.r0 + threadIdx.x % a.R;
   js = threadIdx.x / a.R;
   j = (blk - p * a.nbp) * a.spb + js;
   return j < a.Vh;
}

__device__ __forceinline__ void blas_reduce(const BlasArgs &a, double v, int p, int j, int js, int r)
{
   chunk_reduce<kBlock>(v, js, r - a.r0, a.R, p * a.nchunk + j / REDUCE_CHUNK, j < a.Vh,
                        a.partial + (size_t)r * 2 * a.nchunk);
}

// x += alpha p at one site, in the storage's working precision (float for
// F32/F16, double for F64): the CG iterate update
template <int P>
__device__ __forceinline__ void x_step(const SpinorIO<P> &X, const typename SpinorIO<P>::C ps[12], double alpha, int p,
                                       int j)
{
   typedef typename SpinorIO<P>::C C;
   typedef typename SpinorIO<P>::real real;
   const real al = (real)alpha;
   C xs[12];
   X.load(p, j, xs);
#pragma unroll
   for (int m = 0; m < 12; m++) xs[m] = C{fma(al, ps[m].x, xs[m].x), fma(al, ps[m].y, xs[m].y)};
   X.store(p, j, xs);
}

// x += alpha p (working precision), then p = r + beta p (double), from one
// read of p: the iterate update of the last CG iteration, deferred to this
// pass, and the new search direction
template <int P>
__global__ __launch_bounds__(kBlock) void k_cg_step(SpinorIO<P> rv, const double *beta, SpinorIO<P> pv,
                                                    const double *alpha, SpinorIO<P> xv, BlasArgs a)
{
   int p, j, js, r;
   if (!blas_site(a, p, j, js, r)) return;
   const SpinorIO<P> Rr = rv.at_rhs(r), Pp = pv.at_rhs(r);
   typename SpinorIO<P>::C ps[12];
   Pp.load(p, j, ps);
   x_step(xv.at_rhs(r), ps, alpha[r], p, j);
   const double be = beta[r];
   d2 rs[12], pn[12];
   load_d(Rr, p, j, rs);
#pragma unroll
   for (int c = 0; c < 12; c++)
      pn[c] = d2{fma(be, (double)ps[c].x, rs[c].x), fma(be, (double)ps[c].y, rs[c].y)};
   store_d(Pp, p, j, pn);
}

// x += alpha p (working precision) for the right-hand sides with mask[r] != 0
template <int P>
__global__ __launch_bounds__(kBlock) void k_x_step(const double *alpha, SpinorIO<P> pv, SpinorIO<P> xv, const int *mask,
                                                   BlasArgs a)
{
   int p, j, js, r;
   if (!blas_site(a, p, j, js, r) || !mask[r]) return;
   typename SpinorIO<P>::C ps[12];
   pv.at_rhs(r).load(p, j, ps);
   x_step(xv.at_rhs(r), ps, alpha[r], p, j);
}

// y += x ; x = 0 (when zero_x) -- folds a low-precision partial solution into
// the high-precision one
template <int PY, int PX>
__global__ __launch_bounds__(kBlock) void k_accum(SpinorIO<PY> yv, SpinorIO<PX> xv, int zero_x, const int *mask,
                                                  BlasArgs a)
{
   int p, j, js, r;
   if (!blas_site(a, p, j, js, r)) return;
   if (mask && !mask[r]) return;
   d2 ys[12], xs[12];
   const SpinorIO<PY> Y = yv.at_rhs(r);
   const SpinorIO<PX> X = xv.at_rhs(r);
   load_d(Y, p, j, ys);
   load_d(X, p, j, xs);
#pragma unroll
   for (int c = 0; c < 12; c++) ys[c] += xs[c];
   store_d(Y, p, j, ys);
   if (zero_x) {
#pragma unroll
      for (int c = 0; c < 12; c++) xs[c] = d2{0.0, 0.0};
      store_d(X, p, j, xs);
   }
}

// y = x (precision conversion) ; |y|^2 of the converted values
template <int PY, int PX>
__global__ __launch_bounds__(kBlock) void k_copy(SpinorIO<PY> yv, SpinorIO<PX> xv, const int *mask, BlasArgs a)
{
   int p, j, js, r;
   double n = 0.0;
   if (blas_site(a, p, j, js, r) && (!mask || mask[r])) {
      d2 xs[12];
      load_d(xv.at_rhs(r), p, j, xs);
      store_d(yv.at_rhs(r), p, j, xs);
      n = norm2(xs);
   }
   if (a.partial) blas_reduce(a, n, p, j, js, r);
}

// y = 0 for the right-hand sides with mask[r] != 0 (all when mask is null)
template <int P>
__global__ __launch_bounds__(kBlock) void k_zero(SpinorIO<P> yv, const int *mask, BlasArgs a)
{
   int p, j, js, r;
   if (!blas_site(a, p, j, js, r)) return;
   if (mask && !mask[r]) return;
   d2 z[12];
#pragma unroll
   for (int c = 0; c < 12; c++) z[c] = d2{0.0, 0.0};
   store_d(yv.at_rhs(r), p, j, z);
}

// out[r] = sum of partial[r*stride + first + k], k < n, in an order fixed by n
// alone.  Stage 1: block (b, r) sums segment b of kSeg values (thread t adds
// elements t, t+256, ... of the segment, then a block tree) and stores the sum
// over the segment's first element, which no other block reads.  Stage 2: one
// block per RHS sums the segment results the same way.
constexpr int kSeg = 4096;

__device__ __forceinline__ double block_sum256(double s)
{
   __shared__ double red[256];
   const int t = threadIdx.x;
   red[t] = s;
   __syncthreads();
   for (int w = 128; w > 0; w >>= 1) {
      if (t < w) red[t] += red[t + w];
      __syncthreads();
   }
   return red[0];
}

__global__ __launch_bounds__(256) void k_reduce_seg(double *partial, size_t stride, int first, int n)
{
   const int b = blockIdx.x, r = blockIdx.y, t = threadIdx.x;
   double *p = partial + r * stride + first + (size_t)b * kSeg;
   const int m = min(kSeg, n - b * kSeg);
   double s = 0.0;
   for (int k = t; k < m; k += 256) s += p[k];
   const double tot = block_sum256(s);  // every read of the segment precedes this
   if (t == 0) p[0] = tot;
}

__global__ __launch_bounds__(256) void k_reduce_fin(const double *partial, size_t stride, int first, int nseg,
                                                   double *out)
{
   const int r = blockIdx.x, t = threadIdx.x;
   const double *p = partial + r * stride + first;
   double s = 0.0;
   for (int k = t; k < nseg; k += 256) s += p[(size_t)k * kSeg];
   const double tot = block_sum256(s);
   if (t == 0) out[r] = tot;
}

BlasArgs make_args(const plqcd_hip_field *f, double *partial, const plqcd::Window &w)
{
   BlasArgs a;
   a.Vh = f->lat->Vh;
   a.R = w.nr > 0 ? w.nr : f->nrhs;
   a.r0 = w.nr > 0 ? w.r0 : 0;
   a.spb = REDUCE_CHUNK * (kBlock / (REDUCE_CHUNK * a.R));
   a.nbp = (a.Vh + a.spb - 1) / a.spb;
   a.nchunk = plqcd::reduce_chunks(a.Vh);
   a.partial = partial;
   return a;
}

int check_pair(const plqcd_hip_field *a, const plqcd_hip_field *b, const plqcd::Window &w)
{
   if (!a || !b) return plqcd::set_error(1, "blas: null field");
   if (a->lat != b->lat || a->nrhs != b->nrhs) return plqcd::set_error(1, "blas: fields of different shapes");
   if (a->nrhs * REDUCE_CHUNK > kBlock) return plqcd::set_error(2, "blas: at most %d right-hand sides", kBlock / REDUCE_CHUNK);
   if (w.nr < 0 || w.r0 < 0 || (w.nr > 0 && w.r0 + w.nr > a->nrhs))
      return plqcd::set_error(1, "blas: RHS window %d + %d outside the fields' %d", w.r0, w.nr, a->nrhs);
   return 0;
}

template <class F>
int dispatch1(int prec, F f)
{
   switch (prec) {
      case PLQCD_PREC_F64: f(std::integral_constant<int, PLQCD_PREC_F64>{}); break;
      case PLQCD_PREC_F32: f(std::integral_constant<int, PLQCD_PREC_F32>{}); break;
      default: f(std::integral_constant<int, PLQCD_PREC_F16>{});
   }
   PLQCD_HIP_TRY(hipGetLastError());
   return 0;
}

template <class F>
int dispatch2(int p1, int p2, F f)
{
   return dispatch1(p1, [&](auto a) { dispatch1(p2, [&](auto b) { f(a, b); }); });
}

dim3 grid_of(const BlasArgs &a) { return dim3(2 * a.nbp); }
dim3 block_of(const BlasArgs &a) { return dim3(a.R * a.spb); }

}  // namespace

namespace plqcd {

int reduce_partials(double *partial, size_t stride, int first, int n, int R, double *out, hipStream_t st, int r0)
{
   if (n < 1 || R < 1 || r0 < 0) return set_error(1, "reduce_partials: empty reduction");
   const int nseg = (n + kSeg - 1) / kSeg;
   partial += (size_t)r0 * stride;
   out += r0;
   hipLaunchKernelGGL(k_reduce_seg, dim3(nseg, R), dim3(256), 0, st, partial, stride, first, n);
   hipLaunchKernelGGL(k_reduce_fin, dim3(R), dim3(256), 0, st, partial, stride, first, nseg, out);
   PLQCD_HIP_TRY(hipGetLastError());
   return 0;
}

size_t partial_doubles(const plqcd_hip_field *f) { return (size_t)f->nrhs * 2 * reduce_chunks(f->lat->Vh); }

int blas_cg_step(const plqcd_hip_field *r, const double *beta, plqcd_hip_field *p, const double *alpha,
                 plqcd_hip_field *x, const Window &w)
{
   if (int rc = check_pair(r, p, w) | check_pair(r, x, w)) return rc;
   if (r->prec != p->prec || r->prec != x->prec) return set_error(1, "blas_cg_step: precision mismatch");
   if (p == x || p == r || r == x) return set_error(1, "blas_cg_step: r, p and x must be distinct fields");
   const BlasArgs a = make_args(r, nullptr, w);
   hipStream_t st = r->lat->stream;
   return dispatch1(r->prec, [&](auto P) {
      constexpr int PP = decltype(P)::value;
      hipLaunchKernelGGL(k_cg_step<PP>, grid_of(a), block_of(a), 0, st, spinor_io<PP>(r), beta, spinor_io<PP>(p),
                         alpha, spinor_io<PP>(x), a);
   });
}

int blas_x_step(const double *alpha, const plqcd_hip_field *p, plqcd_hip_field *x, const int *mask, const Window &w)
{
   if (int rc = check_pair(p, x, w)) return rc;
   if (p->prec != x->prec) return set_error(1, "blas_x_step: precision mismatch");
   if (p == x || !mask) return set_error(1, "blas_x_step: distinct p and x and a mask expected");
   const BlasArgs a = make_args(p, nullptr, w);
   hipStream_t st = p->lat->stream;
   return dispatch1(p->prec, [&](auto P) {
      constexpr int PP = decltype(P)::value;
      hipLaunchKernelGGL(k_x_step<PP>, grid_of(a), block_of(a), 0, st, alpha, spinor_io<PP>(p), spinor_io<PP>(x),
                         mask, a);
   });
}

int blas_accum(plqcd_hip_field *y, plqcd_hip_field *x, int zero_x, const int *mask, const Window &w)
{
   if (int rc = check_pair(y, x, w)) return rc;
   const BlasArgs a = make_args(y, nullptr, w);
   hipStream_t st = y->lat->stream;
   return dispatch2(y->prec, x->prec, [&](auto PY, auto PX) {
      constexpr int A = decltype(PY)::value, B = decltype(PX)::value;
      hipLaunchKernelGGL((k_accum<A, B>), grid_of(a), block_of(a), 0, st, spinor_io<A>(y), spinor_io<B>(x), zero_x, mask, a);
   });
}

int blas_copy(plqcd_hip_field *y, const plqcd_hip_field *x, double *partial, const int *mask, const Window &w)
{
   if (int rc = check_pair(y, x, w)) return rc;
   const BlasArgs a = make_args(y, partial, w);
   hipStream_t st = y->lat->stream;
   return dispatch2(y->prec, x->prec, [&](auto PY, auto PX) {
      constexpr int A = decltype(PY)::value, B = decltype(PX)::value;
      hipLaunchKernelGGL((k_copy<A, B>), grid_of(a), block_of(a), 0, st, spinor_io<A>(y), spinor_io<B>(x), mask, a);
   });
}

int blas_zero(plqcd_hip_field *y, const int *mask, const Window &w)
{
   if (int rc = check_pair(y, y, w)) return rc;
   const BlasArgs a = make_args(y, nullptr, w);
   hipStream_t st = y->lat->stream;
   return dispatch1(y->prec, [&](auto P) {
      constexpr int PP = decltype(P)::value;
      hipLaunchKernelGGL(k_zero<PP>, grid_of(a), block_of(a), 0, st, spinor_io<PP>(y), mask, a);
   });
}

}  // namespace plqcd
