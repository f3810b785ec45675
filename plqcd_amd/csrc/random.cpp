// Seeded synthetic inputs.  Counter-based (a keyed SplitMix64 hash of
// (seed, stream, counter)), so every value depends only on its site and
// component and the fill can run on any number of threads.
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "internal.h"

namespace {

inline uint64_t mix64(uint64_t z)
{
   z += 0x9e3779b97f4a7c15ULL;
   z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
   z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
   return z ^ (z >> 31);
}

struct Stream {
   uint64_t key;
   Stream(uint64_t seed, uint64_t tag) : key(mix64(seed ^ mix64(tag))) {}
   // uniform in (0,1), 53 random bits, never 0
   double uniform(uint64_t counter) const
   {
      return ((mix64(key ^ mix64(counter)) >> 11) + 0.5) * 0x1.0p-53;
   }
   // Box-Muller pair from counters 2c, 2c+1
   void gauss2(uint64_t c, double &a, double &b) const
   {
      const double u1 = uniform(2 * c), u2 = uniform(2 * c + 1);
      const double r = std::sqrt(-2.0 * std::log(u1));
      const double th = 6.283185307179586476925 * u2;
      a = r * std::cos(th);
      b = r * std::sin(th);
   }
};

constexpr uint64_t kGaugeTag = 1, kSpinorTag = 2;

template <class F>
void parallel_sites(int V, F f)
{
   unsigned nt = std::thread::hardware_concurrency();
   if (nt == 0) nt = 1;
   if (nt > 16) nt = 16;
   if (V < 4096) nt = 1;
   std::vector<std::thread> pool;
   const int chunk = (V + nt - 1) / nt;
   for (unsigned t = 0; t < nt; t++) {
      const int lo = t * chunk, hi = std::min(V, lo + chunk);
      if (lo >= hi) break;
      try {
         pool.emplace_back([=] {
            for (int i = lo; i < hi; i++) f(i);
         });
      } catch (...) {  // no thread: the chunk runs here (each site's value depends on its index only)
         for (int i = lo; i < hi; i++) f(i);
      }
   }
   for (auto &th : pool) th.join();
}

// Random SU(3) from a Gaussian complex 3x3 matrix m[row][col][re/im]:
// Gram-Schmidt on rows 0 and 1, row 2 = conj(row0 x row1) (det = +1).
void su3_from_gauss(double m[3][3][2], double *out18)
{
   double *r0 = &m[0][0][0], *r1 = &m[1][0][0];
   double n0 = 0;
   for (int k = 0; k < 6; k++) n0 += r0[k] * r0[k];
   n0 = 1.0 / std::sqrt(n0);
   for (int k = 0; k < 6; k++) r0[k] *= n0;
   // <r0, r1> = sum conj(r0_k) r1_k
   double pr = 0, pi = 0;
   for (int k = 0; k < 3; k++) {
      pr += r0[2 * k] * r1[2 * k] + r0[2 * k + 1] * r1[2 * k + 1];
      pi += r0[2 * k] * r1[2 * k + 1] - r0[2 * k + 1] * r1[2 * k];
   }
   for (int k = 0; k < 3; k++) {  // r1 -= <r0,r1> r0
      r1[2 * k] -= pr * r0[2 * k] - pi * r0[2 * k + 1];
      r1[2 * k + 1] -= pr * r0[2 * k + 1] + pi * r0[2 * k];
   }
   double n1 = 0;
   for (int k = 0; k < 6; k++) n1 += r1[k] * r1[k];
   n1 = 1.0 / std::sqrt(n1);
   for (int k = 0; k < 6; k++) r1[k] *= n1;
   // r2_a = conj(r0_b r1_c - r0_c r1_b), (a,b,c) cyclic
   for (int a = 0; a < 3; a++) {
      const int b = (a + 1) % 3, c = (a + 2) % 3;
      const double re = (r0[2 * b] * r1[2 * c] - r0[2 * b + 1] * r1[2 * c + 1]) -
                        (r0[2 * c] * r1[2 * b] - r0[2 * c + 1] * r1[2 * b + 1]);
      const double im = (r0[2 * b] * r1[2 * c + 1] + r0[2 * b + 1] * r1[2 * c]) -
                        (r0[2 * c] * r1[2 * b + 1] + r0[2 * c + 1] * r1[2 * b]);
      m[2][a][0] = re;
      m[2][a][1] = -im;
   }
   for (int k = 0; k < 18; k++) out18[k] = (&m[0][0][0])[k];
}

}  // namespace

extern "C" void plqcd_random_gauge_su3(su3 *u, int V, uint64_t seed)
{
   const Stream st(seed, kGaugeTag);
   double *base = reinterpret_cast<double *>(u);
   parallel_sites(V, [&](int i) {
      for (int mu = 0; mu < 4; mu++) {
         const uint64_t c0 = (uint64_t(i) * 4 + mu) * 9;
         double m[3][3][2];
         for (int k = 0; k < 9; k++) st.gauss2(c0 + k, m[k / 3][k % 3][0], m[k / 3][k % 3][1]);
         su3_from_gauss(m, base + (uint64_t(i) * 4 + mu) * 18);
      }
   });
}

extern "C" void plqcd_random_spinor_gauss(spinor *s, int V, uint64_t seed)
{
   const Stream st(seed, kSpinorTag);
   double *base = reinterpret_cast<double *>(s);
   parallel_sites(V, [&](int i) {
      double *p = base + uint64_t(i) * 24;
      for (int k = 0; k < 12; k++) st.gauss2(uint64_t(i) * 12 + k, p[2 * k], p[2 * k + 1]);
   });
}

extern "C" void plqcd_unit_gauge(su3 *u, int V)
{
   double *base = reinterpret_cast<double *>(u);
   parallel_sites(V, [&](int i) {
      for (int mu = 0; mu < 4; mu++) {
         double *p = base + (uint64_t(i) * 4 + mu) * 18;
         for (int k = 0; k < 18; k++) p[k] = 0.0;
         p[0] = p[8] = p[16] = 1.0;  // c00, c11, c22
      }
   });
}
