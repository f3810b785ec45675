// Device objects behind the C-ABI handles (lattice, gauge, spinor fields) and
// the kernels that repack the reference's host layout (array of structures in
// even-odd site order, include/su3.h:52-75) into the structure-of-arrays HBM
// layouts described in device_common.h, converting precision on the way.
#include <sys/mman.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <type_traits>

#include "fields.h"

namespace {

constexpr size_t kXferBytes = size_t(32) << 20;  // one chunk of a host <-> device transfer
constexpr int kPackBlock = 256;

// --------------------------------------------------------- spinor repacking

// aos: n sites starting at even-odd index i0, 12 complex doubles each.
template <int PREC>
__global__ void pack_spinor(const d2 *__restrict__ aos, SpinorIO<PREC> f, int Vh, int rhs, int i0, int n)
{
   const int t = blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= n) return;
   f = f.at_rhs(rhs);
   const int i = i0 + t;
   const int p = i >= Vh;
   const int j = i - p * Vh;
   typename SpinorIO<PREC>::C s[12];
#pragma unroll
   for (int c = 0; c < 12; c++) {
      const d2 v = aos[(size_t)t * 12 + c];
      s[c].x = v.x;
      s[c].y = v.y;
   }
   f.store(p, j, s);
}

template <int PREC>
__global__ void unpack_spinor(d2 *__restrict__ aos, SpinorIO<PREC> f, int Vh, int rhs, int i0, int n)
{
   const int t = blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= n) return;
   f = f.at_rhs(rhs);
   const int i = i0 + t;
   const int p = i >= Vh;
   const int j = i - p * Vh;
   typename SpinorIO<PREC>::C s[12];
   f.load(p, j, s);
#pragma unroll
   for (int c = 0; c < 12; c++) aos[(size_t)t * 12 + c] = d2{(double)s[c].x, (double)s[c].y};
}

// ----------------------------------------------------------- link repacking

// aos: n links starting at link index l0 = 4*i0 (u[4*i+mu]), 9 complex each.
__global__ void pack_gauge(const d2 *__restrict__ aos, void *dst, int prec, int links, int Vh, int l0, int n)
{
   const int t = blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= n) return;
   const int l = l0 + t;
   const int i = l >> 2, mu = l & 3;
   const int p = i >= Vh;
   const int j = i - p * Vh;
   const d2 *u = aos + (size_t)t * 9;
   const size_t lk = (size_t)(p * 4 + mu);  // link slot [parity][mu]
   if (prec == PLQCD_PREC_F64) {
      const int nk = links / 2;
      d2 *o = static_cast<d2 *>(dst) + lk * nk * Vh + j;
      for (int k = 0; k < nk; k++) o[(size_t)k * Vh] = u[k];
   } else if (prec == PLQCD_PREC_F32 && links == 18) {
      f2 *o = static_cast<f2 *>(dst) + lk * 9 * Vh + j;
      for (int k = 0; k < 9; k++) o[(size_t)k * Vh] = f2{(float)u[k].x, (float)u[k].y};
   } else if (prec == PLQCD_PREC_F32) {
      f4 *o = static_cast<f4 *>(dst) + lk * 3 * Vh + j;
      for (int k = 0; k < 3; k++)
         o[(size_t)k * Vh] = f4{(float)u[2 * k].x, (float)u[2 * k].y, (float)u[2 * k + 1].x, (float)u[2 * k + 1].y};
   } else if (links == 18) {  // F16: u4 (k 0..3), u4 (k 4..7), then u32 (k 8)
      u4 *a = static_cast<u4 *>(dst) + lk * Vh + j;
      u4 *b = static_cast<u4 *>(dst) + (8 + lk) * Vh + j;
      uint32_t *c = reinterpret_cast<uint32_t *>(static_cast<u4 *>(dst) + (size_t)16 * Vh) + lk * Vh + j;
      u4 wa, wb;
      wa.x = f2_to_h2((float)u[0].x, (float)u[0].y);
      wa.y = f2_to_h2((float)u[1].x, (float)u[1].y);
      wa.z = f2_to_h2((float)u[2].x, (float)u[2].y);
      wa.w = f2_to_h2((float)u[3].x, (float)u[3].y);
      wb.x = f2_to_h2((float)u[4].x, (float)u[4].y);
      wb.y = f2_to_h2((float)u[5].x, (float)u[5].y);
      wb.z = f2_to_h2((float)u[6].x, (float)u[6].y);
      wb.w = f2_to_h2((float)u[7].x, (float)u[7].y);
      *a = wa;
      *b = wb;
      *c = f2_to_h2((float)u[8].x, (float)u[8].y);
   } else {
      u4 *a = static_cast<u4 *>(dst) + lk * Vh + j;
      u2 *b = reinterpret_cast<u2 *>(static_cast<u4 *>(dst) + (size_t)8 * Vh) + lk * Vh + j;
      u4 wa;
      wa.x = f2_to_h2((float)u[0].x, (float)u[0].y);
      wa.y = f2_to_h2((float)u[1].x, (float)u[1].y);
      wa.z = f2_to_h2((float)u[2].x, (float)u[2].y);
      wa.w = f2_to_h2((float)u[3].x, (float)u[3].y);
      u2 wb;
      wb.x = f2_to_h2((float)u[4].x, (float)u[4].y);
      wb.y = f2_to_h2((float)u[5].x, (float)u[5].y);
      *a = wa;
      *b = wb;
   }
}

#ifndef MADV_POPULATE_WRITE
#define MADV_POPULATE_WRITE 23
#endif
// Maps the whole pages of a host destination in one call before they are
// written: the first write to a fresh page otherwise takes a fault and a
// zeroing per page (a fresh 2 GB array filled at ~1.5-3.5 GB/s instead of
// ~9 GB/s).  The data is not touched; kernels without the call return an
// error, which changes nothing.
void prefault(void *p, size_t n)
{
   const uintptr_t pg = 4096;
   const uintptr_t a = (reinterpret_cast<uintptr_t>(p) + pg - 1) & ~(pg - 1);
   const uintptr_t b = (reinterpret_cast<uintptr_t>(p) + n) & ~(pg - 1);
   if (b > a) (void)madvise(reinterpret_cast<void *>(a), b - a, MADV_POPULATE_WRITE);
}

template <class F>
void dispatch_prec(int prec, F f)
{
   switch (prec) {
      case PLQCD_PREC_F64: f(std::integral_constant<int, PLQCD_PREC_F64>{}); break;
      case PLQCD_PREC_F32: f(std::integral_constant<int, PLQCD_PREC_F32>{}); break;
      default: f(std::integral_constant<int, PLQCD_PREC_F16>{});
   }
}

int ensure_xfer(plqcd_hip_lattice *lat)
{
   for (int b = 0; b < 2; b++) {
      if (!lat->xdev[b]) PLQCD_HIP_TRY(hipMalloc(&lat->xdev[b], kXferBytes));
      if (!lat->xpin[b]) PLQCD_HIP_TRY(hipHostMalloc(&lat->xpin[b], kXferBytes, hipHostMallocDefault));
      if (!lat->xev[b]) PLQCD_HIP_TRY(hipEventCreateWithFlags(&lat->xev[b], hipEventDisableTiming));
   }
   return 0;
}

// Host -> device transfer of `count` elements of `esize` bytes, from element
// `first` of `host`, in chunks that alternate between the lattice's two
// buffer pairs.  Each chunk is copied into a page-locked buffer (by several
// threads) while the previous one is in DMA and being repacked, or goes by
// DMA alone when the host range is page-locked (plqcd_hip_host_alloc).
// pack(device chunk, first element, elements) enqueues the repacking kernel
// on the lattice stream.  Returns when the last chunk has been repacked.
template <class Pack>
int upload_chunks_(plqcd_hip_lattice *lat, const void *host, size_t esize, size_t first, size_t count, Pack pack)
{
   hipStream_t st = lat->stream;
   const size_t per = kXferBytes / esize;
   const char *h = static_cast<const char *>(host);
   const bool pinned = plqcd::host_pinned(h + first * esize, count * esize);
   bool used[2] = {false, false};
   for (size_t i0 = first, c = 0; i0 < first + count; i0 += per, c++) {
      const int b = (int)(c & 1);
      const size_t n = std::min(per, first + count - i0);
      const void *src = h + i0 * esize;
      if (!pinned) {
         if (used[b]) PLQCD_HIP_TRY(hipEventSynchronize(lat->xev[b]));  // its previous chunk has left
         plqcd::host_copy(lat->xpin[b], src, n * esize);
         src = lat->xpin[b];
      }
      PLQCD_HIP_TRY(hipMemcpyAsync(lat->xdev[b], src, n * esize, hipMemcpyHostToDevice, st));
      pack(lat->xdev[b], i0, n);
      PLQCD_HIP_TRY(hipGetLastError());
      PLQCD_HIP_TRY(hipEventRecord(lat->xev[b], st));
      used[b] = true;
   }
   PLQCD_HIP_TRY(hipStreamSynchronize(st));
   return 0;
}

// Device -> host transfer, the reverse of upload_chunks: unpack(device chunk,
// first element, elements) enqueues the kernel that writes a chunk in the host
// layout; each chunk's DMA overlaps the copy-out of the previous one.
template <class Unpack>
int download_chunks_(plqcd_hip_lattice *lat, void *host, size_t esize, size_t first, size_t count, Unpack unpack)
{
   hipStream_t st = lat->stream;
   const size_t per = kXferBytes / esize;
   char *h = static_cast<char *>(host);
   const bool pinned = plqcd::host_pinned(h + first * esize, count * esize);
   if (!pinned) prefault(h + first * esize, count * esize);
   size_t pend_i0 = 0, pend_n = 0;
   int pend_b = -1;  // chunk whose DMA is in flight
   for (size_t i0 = first, c = 0; i0 < first + count; i0 += per, c++) {
      const int b = (int)(c & 1);
      const size_t n = std::min(per, first + count - i0);
      unpack(lat->xdev[b], i0, n);
      PLQCD_HIP_TRY(hipGetLastError());
      void *dst = pinned ? static_cast<void *>(h + i0 * esize) : lat->xpin[b];
      PLQCD_HIP_TRY(hipMemcpyAsync(dst, lat->xdev[b], n * esize, hipMemcpyDeviceToHost, st));
      PLQCD_HIP_TRY(hipEventRecord(lat->xev[b], st));
      if (!pinned && pend_b >= 0) {
         PLQCD_HIP_TRY(hipEventSynchronize(lat->xev[pend_b]));
         plqcd::host_copy(h + pend_i0 * esize, lat->xpin[pend_b], pend_n * esize);
      }
      pend_i0 = i0;
      pend_n = n;
      pend_b = b;
   }
   if (!pinned && pend_b >= 0) {
      PLQCD_HIP_TRY(hipEventSynchronize(lat->xev[pend_b]));
      plqcd::host_copy(h + pend_i0 * esize, lat->xpin[pend_b], pend_n * esize);
   }
   PLQCD_HIP_TRY(hipStreamSynchronize(st));
   return 0;
}

// A transfer that fails part-way leaves no DMA in flight from the staging
// buffers, which the next transfer of the lattice reuses.
template <class Pack>
int upload_chunks(plqcd_hip_lattice *lat, const void *host, size_t esize, size_t first, size_t count, Pack pack)
{
   if (int rc = ensure_xfer(lat)) return rc;
   const int rc = upload_chunks_(lat, host, esize, first, count, pack);
   if (rc) (void)hipStreamSynchronize(lat->stream);
   return rc;
}
template <class Unpack>
int download_chunks(plqcd_hip_lattice *lat, void *host, size_t esize, size_t first, size_t count, Unpack unpack)
{
   if (int rc = ensure_xfer(lat)) return rc;
   const int rc = download_chunks_(lat, host, esize, first, count, unpack);
   if (rc) (void)hipStreamSynchronize(lat->stream);
   return rc;
}

int check_rhs(const plqcd_hip_field *f, int rhs)
{
   if (!f) return plqcd::set_error(1, "null field");
   if (rhs < 0 || rhs >= f->nrhs) return plqcd::set_error(1, "rhs %d out of range [0,%d)", rhs, f->nrhs);
   return 0;
}

}  // namespace

// ================================================================== lattice

extern "C" int plqcd_hip_lattice_create(const int L[4], plqcd_hip_lattice **out)
{
   plqcd::clear_error();
   if (!L || !out) return plqcd::set_error(1, "plqcd_hip_lattice_create: null argument");
   *out = nullptr;
   long long V = 1;
   for (int mu = 0; mu < 4; mu++) {
      if (L[mu] < 2 || (L[mu] & 1))
         return plqcd::set_error(2, "plqcd_hip_lattice_create: L[%d] = %d must be even and >= 2", mu, L[mu]);
      V *= L[mu];
   }
   // a link component block of Vh 16-byte entries must stay within the
   // kernels' 32-bit offsets: V/2 * 16 <= 4 GiB
   if (V > (1LL << 29))
      return plqcd::set_error(2, "plqcd_hip_lattice_create: volume %lld exceeds 2^29 sites", V);
   plqcd_hip_lattice *lat = new (std::nothrow) plqcd_hip_lattice();
   if (!lat) return plqcd::set_error(4, "out of host memory");
   for (int mu = 0; mu < 4; mu++) lat->L[mu] = L[mu];
   lat->V = (int)V;
   lat->Vh = (int)(V / 2);
   lat->Lxh = L[0] / 2;
   hipError_t e = hipGetDevice(&lat->device);
   if (e == hipSuccess) e = hipStreamCreateWithFlags(&lat->own_stream, hipStreamNonBlocking);
   if (e == hipSuccess) e = hipEventCreate(&lat->ev_start);
   if (e == hipSuccess) e = hipEventCreate(&lat->ev_stop);
   if (e != hipSuccess) {
      delete lat;
      return plqcd::set_error(10, "plqcd_hip_lattice_create: %s", hipGetErrorString(e));
   }
   lat->stream = lat->own_stream;
   *out = lat;
   return 0;
}

extern "C" void plqcd_hip_lattice_destroy(plqcd_hip_lattice *lat)
{
   if (!lat) return;
   (void)hipStreamSynchronize(lat->stream);
   for (int b = 0; b < 2; b++) {
      if (lat->xdev[b]) (void)hipFree(lat->xdev[b]);
      if (lat->xpin[b]) (void)hipHostFree(lat->xpin[b]);
      if (lat->xev[b]) (void)hipEventDestroy(lat->xev[b]);
   }
   (void)hipEventDestroy(lat->ev_start);
   (void)hipEventDestroy(lat->ev_stop);
   (void)hipStreamDestroy(lat->own_stream);
   delete lat;
}

extern "C" int plqcd_hip_lattice_set_stream(plqcd_hip_lattice *lat, void *stream)
{
   if (!lat) return plqcd::set_error(1, "null lattice");
   lat->stream = stream ? static_cast<hipStream_t>(stream) : lat->own_stream;
   return 0;
}

extern "C" void *plqcd_hip_lattice_stream(plqcd_hip_lattice *lat) { return lat ? (void *)lat->stream : nullptr; }

extern "C" int plqcd_hip_lattice_sync(plqcd_hip_lattice *lat)
{
   if (!lat) return plqcd::set_error(1, "null lattice");
   PLQCD_HIP_TRY(hipStreamSynchronize(lat->stream));
   return 0;
}

extern "C" int plqcd_hip_timer_start(plqcd_hip_lattice *lat)
{
   if (!lat) return plqcd::set_error(1, "null lattice");
   PLQCD_HIP_TRY(hipEventRecord(lat->ev_start, lat->stream));
   return 0;
}

extern "C" int plqcd_hip_timer_stop(plqcd_hip_lattice *lat, double *ms)
{
   if (!lat || !ms) return plqcd::set_error(1, "null argument");
   PLQCD_HIP_TRY(hipEventRecord(lat->ev_stop, lat->stream));
   PLQCD_HIP_TRY(hipEventSynchronize(lat->ev_stop));
   float f = 0.f;
   PLQCD_HIP_TRY(hipEventElapsedTime(&f, lat->ev_start, lat->ev_stop));
   *ms = f;
   return 0;
}

// ==================================================================== gauge

extern "C" int plqcd_hip_gauge_create(plqcd_hip_lattice *lat, int prec, int links, plqcd_hip_gauge **out)
{
   plqcd::clear_error();
   if (!lat || !out) return plqcd::set_error(1, "plqcd_hip_gauge_create: null argument");
   *out = nullptr;
   const bool ok = (prec == PLQCD_PREC_F64 && (links == 18 || links == 12)) ||
                   (prec == PLQCD_PREC_F32 && (links == 18 || links == 12)) ||
                   (prec == PLQCD_PREC_F16 && (links == 18 || links == 12));
   if (!ok) return plqcd::set_error(2, "plqcd_hip_gauge_create: unsupported precision %d with %d-real links", prec, links);
   plqcd_hip_gauge *g = new (std::nothrow) plqcd_hip_gauge();
   if (!g) return plqcd::set_error(4, "out of host memory");
   g->lat = lat;
   g->prec = prec;
   g->links = links;
   g->bytes = plqcd::gauge_bytes(prec, links, lat->Vh);
   hipError_t e = hipMalloc(&g->data, g->bytes);
   if (e != hipSuccess) {
      delete g;
      return plqcd::set_error(5, "plqcd_hip_gauge_create: hipMalloc(%zu): %s", plqcd::gauge_bytes(prec, links, lat->Vh),
                              hipGetErrorString(e));
   }
   *out = g;
   return 0;
}

extern "C" void plqcd_hip_gauge_destroy(plqcd_hip_gauge *g)
{
   if (!g) return;
   (void)hipStreamSynchronize(g->lat->stream);
   (void)hipFree(g->data);
   delete g;
}

extern "C" size_t plqcd_hip_gauge_bytes(const plqcd_hip_gauge *g) { return g ? g->bytes : 0; }

extern "C" int plqcd_hip_gauge_upload(plqcd_hip_gauge *g, const su3 *u)
{
   if (!g || !u) return plqcd::set_error(1, "plqcd_hip_gauge_upload: null argument");
   plqcd_hip_lattice *lat = g->lat;
   return upload_chunks(lat, u, sizeof(su3), 0, size_t(4) * lat->V, [&](void *d, size_t l0, size_t n) {
      hipLaunchKernelGGL(pack_gauge, dim3((unsigned)((n + kPackBlock - 1) / kPackBlock)), dim3(kPackBlock), 0,
                         lat->stream, static_cast<const d2 *>(d), g->data, g->prec, g->links, lat->Vh, (int)l0,
                         (int)n);
   });
}

// =================================================================== fields

extern "C" int plqcd_hip_field_create(plqcd_hip_lattice *lat, int prec, int nrhs, plqcd_hip_field **out)
{
   plqcd::clear_error();
   if (!lat || !out) return plqcd::set_error(1, "plqcd_hip_field_create: null argument");
   *out = nullptr;
   if (prec < PLQCD_PREC_F64 || prec > PLQCD_PREC_F16)
      return plqcd::set_error(2, "plqcd_hip_field_create: bad precision %d", prec);
   if (nrhs < 1 || nrhs > 4096) return plqcd::set_error(2, "plqcd_hip_field_create: bad RHS count %d", nrhs);
   if (nrhs > plqcd::max_rhs_per_field(lat->Vh))
      return plqcd::set_error(2,
                              "plqcd_hip_field_create: %d right-hand sides of %d sites per parity exceed the "
                              "4 GiB component block of the kernels' 32-bit offsets (at most %d)",
                              nrhs, lat->Vh, plqcd::max_rhs_per_field(lat->Vh));
   plqcd_hip_field *f = new (std::nothrow) plqcd_hip_field();
   if (!f) return plqcd::set_error(4, "out of host memory");
   f->lat = lat;
   f->prec = prec;
   f->nrhs = nrhs;
   f->bytes = (2 * plqcd::spinor_parity_bytes(prec, lat->Vh) * nrhs + 255) & ~size_t(255);
   hipError_t e = hipMalloc(&f->data, f->bytes);
   if (e != hipSuccess) {
      delete f;
      return plqcd::set_error(5, "plqcd_hip_field_create: hipMalloc: %s", hipGetErrorString(e));
   }
   *out = f;
   return 0;
}

extern "C" int plqcd_hip_max_rhs(const int L[4])
{
   if (!L) return 0;
   long long V = 1;
   for (int mu = 0; mu < 4; mu++) {
      if (L[mu] < 2 || (L[mu] & 1)) return 0;
      V *= L[mu];
   }
   if (V > (1LL << 29)) return 0;
   return plqcd::max_rhs_per_field((int)(V / 2));
}

extern "C" void plqcd_hip_field_destroy(plqcd_hip_field *f)
{
   if (!f) return;
   (void)hipStreamSynchronize(f->lat->stream);
   (void)hipFree(f->data);
   delete f;
}

extern "C" size_t plqcd_hip_field_bytes(const plqcd_hip_field *f) { return f ? f->bytes : 0; }

int plqcd::field_upload_range(plqcd_hip_field *f, int rhs, const spinor *host, int first, int last)
{
   if (int rc = check_rhs(f, rhs)) return rc;
   if (!host) return plqcd::set_error(1, "plqcd_hip_field_upload: null host pointer");
   plqcd_hip_lattice *lat = f->lat;
   if (first < 0 || last > lat->V || first > last) return plqcd::set_error(1, "bad site range");
   return upload_chunks(lat, host, sizeof(spinor), first, last - first, [&](void *d, size_t i0, size_t n) {
      const dim3 grid((unsigned)((n + kPackBlock - 1) / kPackBlock)), block(kPackBlock);
      dispatch_prec(f->prec, [&](auto P) {
         constexpr int PP = decltype(P)::value;
         hipLaunchKernelGGL(pack_spinor<PP>, grid, block, 0, lat->stream, static_cast<const d2 *>(d),
                            spinor_io<PP>(f), lat->Vh, rhs, (int)i0, (int)n);
      });
   });
}

int plqcd::field_download_range(const plqcd_hip_field *f, int rhs, spinor *host, int first, int last)
{
   if (int rc = check_rhs(f, rhs)) return rc;
   if (!host) return plqcd::set_error(1, "plqcd_hip_field_download: null host pointer");
   plqcd_hip_lattice *lat = f->lat;
   if (first < 0 || last > lat->V || first > last) return plqcd::set_error(1, "bad site range");
   return download_chunks(lat, host, sizeof(spinor), first, last - first, [&](void *d, size_t i0, size_t n) {
      const dim3 grid((unsigned)((n + kPackBlock - 1) / kPackBlock)), block(kPackBlock);
      dispatch_prec(f->prec, [&](auto P) {
         constexpr int PP = decltype(P)::value;
         hipLaunchKernelGGL(unpack_spinor<PP>, grid, block, 0, lat->stream, static_cast<d2 *>(d), spinor_io<PP>(f),
                            lat->Vh, rhs, (int)i0, (int)n);
      });
   });
}

namespace {
constexpr size_t kStagerChunk = 32u << 20;  // bytes per pinned / device buffer
}  // namespace

int plqcd::Stager::create(int device, Stager **out)
{
   *out = nullptr;
   Stager *g = new (std::nothrow) Stager();
   if (!g) return set_error(4, "out of host memory");
   hipError_t e = hipSetDevice(device);
   if (e == hipSuccess) e = hipStreamCreateWithFlags(&g->st_, hipStreamNonBlocking);
   for (int b = 0; b < 2 && e == hipSuccess; b++) {
      e = hipHostMalloc(&g->pin_[b], kStagerChunk, hipHostMallocDefault);
      if (e == hipSuccess) e = hipMalloc(&g->dev_[b], kStagerChunk);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&g->ev_[b], hipEventDisableTiming);
   }
   if (e != hipSuccess) {
      delete g;
      return set_error(10, "stager: %s", hipGetErrorString(e));
   }
   *out = g;
   return 0;
}

plqcd::Stager::~Stager()
{
   if (st_) (void)hipStreamSynchronize(st_);
   for (int b = 0; b < 2; b++) {
      if (ev_[b]) (void)hipEventDestroy(ev_[b]);
      if (dev_[b]) (void)hipFree(dev_[b]);
      if (pin_[b]) (void)hipHostFree(pin_[b]);
   }
   if (st_) (void)hipStreamDestroy(st_);
}

int plqcd::Stager::upload(plqcd_hip_field *f, int rhs, const spinor *host)
{
   if (int rc = check_rhs(f, rhs)) return rc;
   if (!host) return set_error(1, "stager: null host pointer");
   const int V = f->lat->V, Vh = f->lat->Vh;
   const int per = (int)(kStagerChunk / sizeof(spinor));
   bool used[2] = {false, false};
   for (int i0 = 0, c = 0; i0 < V; i0 += per, c++) {
      const int b = c & 1, n = std::min(per, V - i0);
      if (used[b]) PLQCD_HIP_TRY(hipEventSynchronize(ev_[b]));  // buffer b's previous chunk is packed
      std::memcpy(pin_[b], host + i0, (size_t)n * sizeof(spinor));
      PLQCD_HIP_TRY(hipMemcpyAsync(dev_[b], pin_[b], (size_t)n * sizeof(spinor), hipMemcpyHostToDevice, st_));
      const d2 *src = static_cast<const d2 *>(dev_[b]);
      const dim3 grid((n + kPackBlock - 1) / kPackBlock), block(kPackBlock);
      dispatch_prec(f->prec, [&](auto P) {
         constexpr int PP = decltype(P)::value;
         hipLaunchKernelGGL(pack_spinor<PP>, grid, block, 0, st_, src, spinor_io<PP>(f), Vh, rhs, i0, n);
      });
      PLQCD_HIP_TRY(hipGetLastError());
      PLQCD_HIP_TRY(hipEventRecord(ev_[b], st_));
      used[b] = true;
   }
   PLQCD_HIP_TRY(hipStreamSynchronize(st_));
   return 0;
}

int plqcd::Stager::download(const plqcd_hip_field *f, int rhs, spinor *host)
{
   if (int rc = check_rhs(f, rhs)) return rc;
   if (!host) return set_error(1, "stager: null host pointer");
   const int V = f->lat->V, Vh = f->lat->Vh;
   const int per = (int)(kStagerChunk / sizeof(spinor));
   int pend_i0 = -1, pend_n = 0, pend_b = 0;  // chunk whose D2H is in flight
   for (int i0 = 0, c = 0; i0 < V; i0 += per, c++) {
      const int b = c & 1, n = std::min(per, V - i0);
      d2 *dst = static_cast<d2 *>(dev_[b]);
      const dim3 grid((n + kPackBlock - 1) / kPackBlock), block(kPackBlock);
      dispatch_prec(f->prec, [&](auto P) {
         constexpr int PP = decltype(P)::value;
         hipLaunchKernelGGL(unpack_spinor<PP>, grid, block, 0, st_, dst, spinor_io<PP>(f), Vh, rhs, i0, n);
      });
      PLQCD_HIP_TRY(hipGetLastError());
      PLQCD_HIP_TRY(hipMemcpyAsync(pin_[b], dev_[b], (size_t)n * sizeof(spinor), hipMemcpyDeviceToHost, st_));
      PLQCD_HIP_TRY(hipEventRecord(ev_[b], st_));
      if (pend_i0 >= 0) {  // copy out the previous chunk while this one is in flight
         PLQCD_HIP_TRY(hipEventSynchronize(ev_[pend_b]));
         std::memcpy(host + pend_i0, pin_[pend_b], (size_t)pend_n * sizeof(spinor));
      }
      pend_i0 = i0;
      pend_n = n;
      pend_b = b;
   }
   if (pend_i0 >= 0) {
      PLQCD_HIP_TRY(hipEventSynchronize(ev_[pend_b]));
      std::memcpy(host + pend_i0, pin_[pend_b], (size_t)pend_n * sizeof(spinor));
   }
   return 0;
}

// ------------------------------------------------------- pinned host memory
//
// Host buffers from plqcd_hip_host_alloc are page-locked: the stager moves
// them by DMA alone, without the copy through its own pinned buffers.  The
// allocations are recorded so that a transfer can tell whether its whole
// host range lies in one of them.
namespace {
std::mutex g_pin_mu;
std::map<uintptr_t, size_t> g_pinned;  // base -> bytes
}  // namespace

bool plqcd::host_pinned(const void *p, size_t n)
{
   const uintptr_t a = reinterpret_cast<uintptr_t>(p);
   std::lock_guard<std::mutex> lk(g_pin_mu);
   auto it = g_pinned.upper_bound(a);
   if (it == g_pinned.begin()) return false;
   --it;
   return a >= it->first && a + n <= it->first + it->second;
}

extern "C" int plqcd_hip_host_alloc(size_t bytes, void **out)
{
   plqcd::clear_error();
   if (!out || bytes == 0) return plqcd::set_error(1, "plqcd_hip_host_alloc: bad arguments");
   *out = nullptr;
   void *p = nullptr;
   const hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
   if (e != hipSuccess) return plqcd::set_error(4, "plqcd_hip_host_alloc(%zu): %s", bytes, hipGetErrorString(e));
   std::lock_guard<std::mutex> lk(g_pin_mu);
   g_pinned[reinterpret_cast<uintptr_t>(p)] = bytes;
   *out = p;
   return 0;
}

extern "C" void plqcd_hip_host_free(void *p)
{
   if (!p) return;
   {
      std::lock_guard<std::mutex> lk(g_pin_mu);
      g_pinned.erase(reinterpret_cast<uintptr_t>(p));
   }
   (void)hipHostFree(p);
}

int plqcd::Stager::upload_raw(void *dev, const void *host, size_t bytes)
{
   if (!dev || !host) return set_error(1, "stager: null pointer");
   if (host_pinned(host, bytes)) {  // DMA straight from the caller's buffer
      PLQCD_HIP_TRY(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, st_));
      PLQCD_HIP_TRY(hipStreamSynchronize(st_));
      return 0;
   }
   bool used[2] = {false, false};
   for (size_t o = 0, c = 0; o < bytes; o += kStagerChunk, c++) {
      const int b = (int)(c & 1);
      const size_t n = std::min(kStagerChunk, bytes - o);
      if (used[b]) PLQCD_HIP_TRY(hipEventSynchronize(ev_[b]));  // buffer b's previous chunk has left
      std::memcpy(pin_[b], static_cast<const char *>(host) + o, n);
      PLQCD_HIP_TRY(hipMemcpyAsync(static_cast<char *>(dev) + o, pin_[b], n, hipMemcpyHostToDevice, st_));
      PLQCD_HIP_TRY(hipEventRecord(ev_[b], st_));
      used[b] = true;
   }
   PLQCD_HIP_TRY(hipStreamSynchronize(st_));
   return 0;
}

int plqcd::Stager::download_raw(void *host, const void *dev, size_t bytes, hipEvent_t after)
{
   if (!dev || !host) return set_error(1, "stager: null pointer");
   if (host_pinned(host, bytes)) {  // DMA straight into the caller's buffer
      if (after) PLQCD_HIP_TRY(hipEventSynchronize(after));
      PLQCD_HIP_TRY(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, st_));
      PLQCD_HIP_TRY(hipStreamSynchronize(st_));
      return 0;
   }
   prefault(host, bytes);  // while the solution may still be unpacking
   // a host-side wait: a stream wait on `after` would tie this stream to the
   // other stream, which fails while that stream is being captured
   if (after) PLQCD_HIP_TRY(hipEventSynchronize(after));
   size_t pend_o = 0, pend_n = 0;
   int pend_b = -1;
   for (size_t o = 0, c = 0; o < bytes; o += kStagerChunk, c++) {
      const int b = (int)(c & 1);
      const size_t n = std::min(kStagerChunk, bytes - o);
      PLQCD_HIP_TRY(hipMemcpyAsync(pin_[b], static_cast<const char *>(dev) + o, n, hipMemcpyDeviceToHost, st_));
      PLQCD_HIP_TRY(hipEventRecord(ev_[b], st_));
      if (pend_b >= 0) {  // copy out the previous chunk while this one is in flight
         PLQCD_HIP_TRY(hipEventSynchronize(ev_[pend_b]));
         std::memcpy(static_cast<char *>(host) + pend_o, pin_[pend_b], pend_n);
      }
      pend_o = o;
      pend_n = n;
      pend_b = b;
   }
   if (pend_b >= 0) {
      PLQCD_HIP_TRY(hipEventSynchronize(ev_[pend_b]));
      std::memcpy(static_cast<char *>(host) + pend_o, pin_[pend_b], pend_n);
   }
   return 0;
}

int plqcd::pack_rhs(plqcd_hip_field *f, int rhs, const void *dev_aos, hipStream_t st)
{
   if (int rc = check_rhs(f, rhs)) return rc;
   const int V = f->lat->V, Vh = f->lat->Vh;
   const d2 *src = static_cast<const d2 *>(dev_aos);
   const dim3 grid((V + kPackBlock - 1) / kPackBlock), block(kPackBlock);
   dispatch_prec(f->prec, [&](auto P) {
      constexpr int PP = decltype(P)::value;
      hipLaunchKernelGGL(pack_spinor<PP>, grid, block, 0, st, src, spinor_io<PP>(f), Vh, rhs, 0, V);
   });
   PLQCD_HIP_TRY(hipGetLastError());
   return 0;
}

int plqcd::unpack_rhs(const plqcd_hip_field *f, int rhs, void *dev_aos, hipStream_t st)
{
   if (int rc = check_rhs(f, rhs)) return rc;
   const int V = f->lat->V, Vh = f->lat->Vh;
   d2 *dst = static_cast<d2 *>(dev_aos);
   const dim3 grid((V + kPackBlock - 1) / kPackBlock), block(kPackBlock);
   dispatch_prec(f->prec, [&](auto P) {
      constexpr int PP = decltype(P)::value;
      hipLaunchKernelGGL(unpack_spinor<PP>, grid, block, 0, st, dst, spinor_io<PP>(f), Vh, rhs, 0, V);
   });
   PLQCD_HIP_TRY(hipGetLastError());
   return 0;
}

extern "C" int plqcd_hip_field_upload(plqcd_hip_field *f, int rhs, const spinor *host)
{
   if (!f) return plqcd::set_error(1, "null field");
   return plqcd::field_upload_range(f, rhs, host, 0, f->lat->V);
}

extern "C" int plqcd_hip_field_download(const plqcd_hip_field *f, int rhs, spinor *host)
{
   if (!f) return plqcd::set_error(1, "null field");
   return plqcd::field_download_range(f, rhs, host, 0, f->lat->V);
}

// Copies the sites [first, last) of one RHS between two fields of the same
// lattice and precision (device to device, asynchronous).
int plqcd::field_copy_range(const plqcd_hip_field *src, plqcd_hip_field *dst, int rhs, int parity)
{
   if (src->prec != dst->prec || src->lat != dst->lat) return plqcd::set_error(1, "field copy: mismatch");
   const size_t pb = plqcd::spinor_parity_bytes(src->prec, src->lat->Vh);
   if (src->prec == PLQCD_PREC_F16 || src->nrhs != 1 || dst->nrhs != 1 || rhs != 0)
      return plqcd::set_error(1, "field copy: only single-RHS F64/F32 parity copies are supported");
   const char *s = static_cast<const char *>(src->data) + parity * pb;
   char *d = static_cast<char *>(dst->data) + parity * pb;
   PLQCD_HIP_TRY(hipMemcpyAsync(d, s, pb, hipMemcpyDeviceToDevice, src->lat->stream));
   return 0;
}
