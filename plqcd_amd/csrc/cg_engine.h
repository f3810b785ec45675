// Slot-based CG engine on the normal equations M^+M x = b, M = 1 - kappa H.
//
// An engine owns (or borrows) F64 fields b and x with S right-hand-side
// slots, the F64 temporaries of the reliable updates and the sloppy-precision
// iteration fields.  Right-hand sides are admitted into free slots at any
// time between iterations and retired as soon as they converge, so one
// engine serves a stream of solves (continuous batching).  Each slot's
// arithmetic depends only on its own data (per-slot scalars, masked kernels,
// fixed-order reductions): a solve gives bitwise the same result whatever
// occupies the other slots.
#pragma once

#include <vector>

#include "blas.h"

namespace plqcd {

struct CgSlot {
   bool active = false;     // iterating
   bool done = false;       // converged or stopped, result not yet retired
   double kappa = 0, tol = 0, bb = 0, rmax = 0;
   int max_iter = 0;
   plqcd_cg_stats stats{};
};

class CgEngine {
  public:
   // g: F64 gauge (true residual); gs: sloppy gauge (== g for plain FP64 CG).
   // b/x: borrowed F64 fields with S RHS, or null to allocate them.
   static int create(const plqcd_hip_gauge *g, const plqcd_hip_gauge *gs, int S, double delta, bool use_graph,
                     plqcd_hip_field *b, plqcd_hip_field *x, CgEngine **out);
   ~CgEngine();

   // Device bytes an engine with S slots needs (b and x included).
   static size_t bytes_needed(const plqcd_hip_lattice *lat, int sloppy_prec, int S);

   int slots() const { return S_; }
   int free_slot() const;
   int active_count() const;
   const CgSlot &slot(int s) const { return slot_[s]; }
   plqcd_hip_field *b() { return b_; }
   plqcd_hip_field *x() { return x_; }
   plqcd_hip_lattice *lattice() { return lat_; }

   // Admits the slots listed in `which` (their b already in the b field);
   // kappa/tol/max_iter per listed slot.  Asynchronous except for one |b|^2
   // readback.
   int admit(const std::vector<int> &which, const double *kappa, const double *tol, const int *max_iter);
   // Uploads host sources into the listed slots and admits them together.
   int admit_host(const std::vector<int> &which, const std::vector<const spinor *> &b_host, double kappa, double tol,
                  int max_iter);
   // Enqueues one iteration of all active slots (asynchronous).  The kernels
   // cover the slots up to the highest active one; each such width has its
   // own captured graph.
   int launch();
   // Waits for the iteration, runs the reliable updates it calls for and
   // appends the slots that finished to `finished`.
   int finish(std::vector<int> &finished);
   // The same without blocking, for a loop that drives several engines:
   //   launch(); ... ready() ... begin_finish(updating) (after the iteration
   //   completed; enqueues the reliable updates it calls for and sets
   //   `updating`); ... ready() ... end_finish(finished) (after those updates,
   //   or at once when there were none).  Slots are admitted only between
   //   end_finish and the next launch.
   bool ready() const;  // the engine's last enqueued work has completed
   // Captures the iteration graphs of every window width now (with
   // use_graph), so that no stream capture happens later: another thread may
   // then wait on events of the engine's stream at any time.
   int capture_all();
   // Stops every active slot unconverged (after a failed launch or update)
   // and lists it in `finished`.
   void abort_active(std::vector<int> &finished);
   int begin_finish(bool &updating);
   int end_finish(std::vector<int> &finished);
   // Downloads x of a finished slot into host memory (may be null), returns
   // its stats and frees the slot.
   int retire(int s, spinor *x_host, plqcd_cg_stats *st);

  private:
   CgEngine() = default;
   int alloc(int sloppy);
   int iteration(int nr);
   int reliable_update();
   int capture(int nr);

   plqcd_hip_lattice *lat_ = nullptr;
   const plqcd_hip_gauge *g_ = nullptr, *gs_ = nullptr;
   int S_ = 0;
   double delta_ = 0;
   bool use_graph_ = true;
   bool own_bx_ = false;
   plqcd_hip_field *b_ = nullptr, *x_ = nullptr;
   plqcd_hip_field *t1_ = nullptr;                                                  // F64
   plqcd_hip_field *rs_ = nullptr, *ps_ = nullptr, *qs_ = nullptr, *xs_ = nullptr;  // sloppy
   double *partial_ = nullptr;
   double *scal_ = nullptr;  // kappa, rr, rr_new, pap, alpha, beta, tmp (7 x S)
   int *imask_ = nullptr;    // active, update, the list of slots being updated (3 x S)
   double *host_ = nullptr;  // pinned: rr_new, tmp (2 x S)
   int *hmask_ = nullptr;    // pinned: active, update, update list (3 x S), the sources of their uploads
   hipEvent_t ev_ = nullptr; // recorded after each enqueued iteration / reliable update
   struct Graph {
      hipGraph_t def = nullptr;
      hipGraphExec_t exec = nullptr;
   };
   std::vector<Graph> graphs_;  // by window width 1..S
   std::vector<CgSlot> slot_;
   std::vector<int> active_, update_;

   double *kappa_d() { return scal_; }
   double *rr_d() { return scal_ + S_; }
   double *rr_new_d() { return scal_ + 2 * S_; }
   double *pap_d() { return scal_ + 3 * S_; }
   double *alpha_d() { return scal_ + 4 * S_; }
   double *beta_d() { return scal_ + 5 * S_; }
   double *tmp_d() { return scal_ + 6 * S_; }
   int *active_d() { return imask_; }
   int *update_d() { return imask_ + S_; }
   int *ulist_d() { return imask_ + 2 * S_; }
};

}  // namespace plqcd
