// Drop-in hopping-matrix entry points with the signature of
// include/dirac.h:21-22: host fields in the reference layout go in and out,
// the work runs on the GPU.  Device state (lattice, F64 gauge, two F64
// spinor fields) is created on first use for the geometry in plqcd_g and
// kept until finalize_plqcd() or a geometry change.
#include <mutex>

#include "internal.h"

namespace {

struct DropIn {
   plqcd_hip_lattice *lat = nullptr;
   plqcd_hip_gauge *gauge = nullptr;
   plqcd_hip_field *in = nullptr, *out = nullptr;
   int L[4] = {0, 0, 0, 0};
   const su3 *cached_u = nullptr;  // last uploaded link pointer
   bool cache_enabled = false;
   bool gauge_valid = false;
   int mode = PLQCD_HOP_COMPAT;
};

DropIn g_state;
std::mutex g_mu;

void release_locked()
{
   plqcd_hip_field_destroy(g_state.in);
   plqcd_hip_field_destroy(g_state.out);
   plqcd_hip_gauge_destroy(g_state.gauge);
   plqcd_hip_lattice_destroy(g_state.lat);
   g_state.in = g_state.out = nullptr;
   g_state.gauge = nullptr;
   g_state.lat = nullptr;
   g_state.gauge_valid = false;
   g_state.cached_u = nullptr;
}

int ensure_state()
{
   if (plqcd_g.init != 1) return plqcd::set_error(20, "plqcd hopping matrix called before init_plqcd");
   const int *L = plqcd_g.latdims;
   if (g_state.lat && L[0] == g_state.L[0] && L[1] == g_state.L[1] && L[2] == g_state.L[2] && L[3] == g_state.L[3])
      return 0;
   release_locked();
   int rc = plqcd_hip_lattice_create(L, &g_state.lat);
   if (!rc) rc = plqcd_hip_gauge_create(g_state.lat, PLQCD_PREC_F64, PLQCD_LINKS_18, &g_state.gauge);
   if (!rc) rc = plqcd_hip_field_create(g_state.lat, PLQCD_PREC_F64, 1, &g_state.in);
   if (!rc) rc = plqcd_hip_field_create(g_state.lat, PLQCD_PREC_F64, 1, &g_state.out);
   if (rc) {
      release_locked();
      return rc;
   }
   for (int mu = 0; mu < 4; mu++) g_state.L[mu] = L[mu];
   return 0;
}

int upload_gauge(const su3 *u)
{
   if (g_state.cache_enabled && g_state.gauge_valid && u == g_state.cached_u) return 0;
   if (int rc = plqcd_hip_gauge_upload(g_state.gauge, u)) return rc;
   g_state.cached_u = u;
   g_state.gauge_valid = true;
   return 0;
}

// parity: output parity (EVEN = eo, ODD = oe, ALL = eo then oe)
double apply(spinor *qin, spinor *qout, su3 *u, int parity)
{
   const double t0 = stop_watch(0.0);
   plqcd::clear_error();
   if (!qin || !qout || !u) return plqcd::set_error(1, "hopping matrix: null argument"), -1.0;
   std::lock_guard<std::mutex> lk(g_mu);
   if (ensure_state() || upload_gauge(u)) return -1.0;
   const int V = plqcd_g.VOLUME, Vh = V / 2;
   int rc = 0;
   if (parity == PLQCD_PARITY_EVEN) {
      rc = plqcd::field_upload_range(g_state.in, 0, qin, Vh, V);
      if (!rc) rc = plqcd_hip_hop(g_state.gauge, g_state.in, g_state.out, PLQCD_PARITY_EVEN, g_state.mode);
      if (!rc) rc = plqcd::field_download_range(g_state.out, 0, qout, 0, Vh);
   } else if (parity == PLQCD_PARITY_ODD) {
      rc = plqcd::field_upload_range(g_state.in, 0, qin, 0, Vh);
      if (!rc) rc = plqcd_hip_hop(g_state.gauge, g_state.in, g_state.out, PLQCD_PARITY_ODD, g_state.mode);
      if (!rc) rc = plqcd::field_download_range(g_state.out, 0, qout, Vh, V);
   } else if (qin != qout) {
      rc = plqcd::field_upload_range(g_state.in, 0, qin, 0, V);
      if (!rc) rc = plqcd_hip_hop(g_state.gauge, g_state.in, g_state.out, PLQCD_PARITY_ALL, g_state.mode);
      if (!rc) rc = plqcd::field_download_range(g_state.out, 0, qout, 0, V);
   } else {
      // in place: eo overwrites the even half that oe then reads, as the
      // sequence eo(q,q,u); oe(q,q,u) does on the host
      rc = plqcd::field_upload_range(g_state.in, 0, qin, 0, V);
      if (!rc) rc = plqcd_hip_hop(g_state.gauge, g_state.in, g_state.out, PLQCD_PARITY_EVEN, g_state.mode);
      if (!rc) rc = plqcd::field_copy_range(g_state.out, g_state.in, 0, 0);
      if (!rc) rc = plqcd_hip_hop(g_state.gauge, g_state.in, g_state.out, PLQCD_PARITY_ODD, g_state.mode);
      if (!rc) rc = plqcd::field_download_range(g_state.out, 0, qout, 0, V);
   }
   if (rc) return -1.0;
   return stop_watch(t0);
}

}  // namespace

void plqcd::release_dropin_state()
{
   std::lock_guard<std::mutex> lk(g_mu);
   release_locked();
}

extern "C" int plqcd_hip_set_mode(int mode)
{
   if (mode != PLQCD_HOP_COMPAT && mode != PLQCD_HOP_WILSON)
      return plqcd::set_error(-1, "plqcd_hip_set_mode: unknown mode %d", mode);
   std::lock_guard<std::mutex> lk(g_mu);
   g_state.mode = mode;
   return 0;
}

extern "C" void plqcd_hip_cache_gauge(int enable)
{
   std::lock_guard<std::mutex> lk(g_mu);
   g_state.cache_enabled = enable != 0;
   g_state.gauge_valid = false;
}

extern "C" void plqcd_hip_gauge_changed(void)
{
   std::lock_guard<std::mutex> lk(g_mu);
   g_state.gauge_valid = false;
}

extern "C" double plqcd_hopping_matrix_eo_hip(spinor *qin, spinor *qout, su3 *u)
{
   return apply(qin, qout, u, PLQCD_PARITY_EVEN);
}

extern "C" double plqcd_hopping_matrix_oe_hip(spinor *qin, spinor *qout, su3 *u)
{
   return apply(qin, qout, u, PLQCD_PARITY_ODD);
}

extern "C" double plqcd_hopping_matrix_hip(spinor *qin, spinor *qout, su3 *u)
{
   return apply(qin, qout, u, PLQCD_PARITY_ALL);
}
