// Solve server: continuous batching of CG requests on one GPU.
//
// Requests name a lattice and a gauge field (by seed) and bring 1..n
// right-hand sides.  A worker thread owns the GPU:
//  * gauge fields (F64 and the sloppy copy) stay resident in a cache keyed by
//    (lattice, seed) and are evicted least-recently-used when HBM is needed;
//  * each (lattice, gauge) has one CG engine with S slots on its own stream;
//    right-hand sides enter free slots between iterations and leave as soon
//    as they converge, so requests that arrive while others run join them;
//  * engines of different lattices iterate concurrently and independently:
//    each is a small state machine (ready -> iteration in flight -> reliable
//    update in flight -> ready) driven by an event on its own stream, so a
//    small lattice's engine is not held back by a large one's;
//  * admission is by HBM: a new gauge or engine is created only if its bytes
//    fit the budget after evicting idle cache entries; otherwise the request
//    waits;
//  * host transfers run on a pool of transfer threads, each with its own
//    stream and pinned staging (plqcd::Stager), and are DMA only: a source
//    lands in the slot's device buffer in the host layout while the engine
//    iterates on its other slots, and is packed into the slot on the engine's
//    own stream when the slot is admitted; a finished slot's solution is
//    unpacked into that buffer on the engine stream and downloaded while the
//    engine goes on.  (Packing on the transfer streams would wait for compute
//    units behind the engines' kernels.)  A slot moves free -> uploading ->
//    uploaded -> active -> downloading -> free; it joins the engine's
//    iteration only between iterations.
// A right-hand side's iterates do not depend on the other slots (see
// cg_engine.h), so its result is the same alone and in the mix.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <string>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "cg_engine.h"
#include "internal.h"

namespace {

using clock_t_ = std::chrono::steady_clock;
constexpr int kTransferThreads = 8;  // default concurrent host transfers (CPU copies into pinned staging)

struct Request {
   uint64_t id = 0;
   plqcd_serve_request req{};
   double t_submit = 0, t_start = -1, t_end = 0;
   int remaining = 0;  // RHS not finished
   int next_rhs = 0;   // next RHS to admit
   int iterations = 0;
   double residual = 0;
   int converged = 1;
   int status = 0;
   int inflight = 0;  // transfers of this request not yet finished
};

struct Key {
   int L[4];
   uint64_t seed;
   bool operator<(const Key &o) const
   {
      for (int i = 0; i < 4; i++)
         if (L[i] != o.L[i]) return L[i] < o.L[i];
      return seed < o.seed;
   }
};

struct Entry {
   Key key{};
   plqcd_hip_lattice *lat = nullptr;
   plqcd_hip_gauge *g = nullptr, *gs = nullptr;
   std::unique_ptr<plqcd::CgEngine> eng;
   enum Phase { READY, ITERATING, UPDATING } phase = READY;
   size_t gauge_bytes = 0, engine_bytes = 0;
   double last_used = 0;
   // slot -> (request, rhs) and the slot's state
   std::vector<std::pair<Request *, int>> owner;
   enum SlotState { FREE, UPLOADING, UPLOADED, ACTIVE, DOWNLOADING };
   std::vector<int> sstate;
   std::vector<int> parts, part_rc;     // per slot: transfer parts still out, their combined status
   char *io = nullptr;                  // per slot: V spinors in the host layout (device)
   std::vector<hipEvent_t> io_ev;       // per slot: the solution is in io
   char *io_slot(int c) const { return io + (size_t)c * lat->V * sizeof(spinor); }
   bool idle() const
   {
      for (int st : sstate)
         if (st != FREE) return false;
      return phase == READY && (!eng || eng->active_count() == 0);
   }
   ~Entry()
   {
      eng.reset();  // waits for the engine's stream
      for (hipEvent_t ev : io_ev) (void)hipEventDestroy(ev);
      if (io) (void)hipFree(io);
      plqcd_hip_gauge_destroy(gs != g ? gs : nullptr);
      plqcd_hip_gauge_destroy(g);
      plqcd_hip_lattice_destroy(lat);
   }
};

}  // namespace

// A host transfer job: the source of `rhs` of request r into slot `slot`
// of the entry's engine (up), or that slot's solution into the request (down).
// One transfer job: bytes [off, off + len) of one RHS between the request's
// host array and the slot's device buffer (both in the host layout).  A large
// RHS is split into parts that several transfer threads move at once: a
// solution lands in host pages the caller has just allocated, and the kernel
// zeroes each page on its first write at a few GB/s per core.
struct Xfer {
   bool down = false;
   Entry *e = nullptr;
   int slot = 0;
   Request *r = nullptr;
   int rhs = 0;
   int rc = 0;
   size_t off = 0, len = 0;
};
constexpr size_t kXferPart = size_t(256) << 20;  // split above this many bytes per RHS

struct plqcd_hip_server {
   plqcd_serve_config cfg{};
   size_t budget = 0, used = 0;
   clock_t_::time_point t0;
   std::thread worker;
   std::mutex mu;
   std::condition_variable cv;
   bool stop = false;
   uint64_t next_id = 1;
   std::deque<std::unique_ptr<Request>> incoming;   // submitted, not yet seen by the worker
   std::vector<std::unique_ptr<Request>> waiting;   // worker-owned, not complete
   std::vector<plqcd_serve_result> results;         // completed, not yet polled
   std::map<Key, std::unique_ptr<Entry>> cache;     // worker-owned
   std::atomic<size_t> used_pub{0}, entries_pub{0};
   std::atomic<int> failed{0};
   std::atomic<long> engines_created{0}, evictions{0};
   std::mutex logmu;
   std::set<std::string> logged;  // failure messages already reported
   std::atomic<double> xfer_bytes{0.0}, xfer_seconds{0.0};
   // transfer pool
   std::vector<std::thread> xfer;
   std::mutex xmu;
   std::condition_variable xcv;
   std::deque<Xfer> xjobs, xdone;
   bool xstop = false;

   double now() const { return std::chrono::duration<double>(clock_t_::now() - t0).count(); }
};

namespace {

// Counts an internal failure and reports each distinct message once on stderr
// (the request concerned gets status -1).
void note_failure(plqcd_hip_server *s, const char *where)
{
   s->failed++;
   char msg[512];
   std::snprintf(msg, sizeof(msg), "%s: %s", where, plqcd_hip_last_error());
   std::lock_guard<std::mutex> lk(s->logmu);
   if (s->logged.insert(msg).second) std::fprintf(stderr, "plqcd server: %s\n", msg);
}

void atomic_add(std::atomic<double> &a, double v)
{
   double cur = a.load();
   while (!a.compare_exchange_weak(cur, cur + v)) {
   }
}

size_t gauge_bytes_of(const plqcd_hip_lattice *lat, int prec, int links)
{
   return plqcd::gauge_bytes(prec, links, lat->Vh);
}

int slots_for(plqcd_hip_server *s, const plqcd_hip_lattice *lat)
{
   const size_t per = plqcd::CgEngine::bytes_needed(lat, s->cfg.sloppy_prec, 1) + (size_t)lat->V * sizeof(spinor);
   size_t fit = (s->budget / 3) / std::max<size_t>(per, 1);
   int S = (int)std::min<size_t>(fit, (size_t)s->cfg.slots);
   // an engine's fields hold S right-hand sides within the kernels' 32-bit offsets
   S = std::min(S, plqcd::max_rhs_per_field(lat->Vh));
   return std::max(S, 1);
}

// Frees idle cache entries, least recently used first, until `need` bytes fit.
bool make_room(plqcd_hip_server *s, size_t need, const Key *keep)
{
   while (s->used + need > s->budget) {
      Entry *victim = nullptr;
      for (auto &kv : s->cache) {
         Entry *e = kv.second.get();
         if (keep && !(e->key < *keep) && !(*keep < e->key)) continue;
         if (!e->idle()) continue;
         bool busy = false;
         for (auto &o : e->owner) busy |= o.first != nullptr;
         if (busy) continue;
         if (!victim || e->last_used < victim->last_used) victim = e;
      }
      if (!victim) return false;
      s->used -= victim->gauge_bytes + victim->engine_bytes;
      s->cache.erase(victim->key);
      s->evictions++;
   }
   return true;
}

// Finds or creates the cache entry and engine of a request; null when it has
// to wait for memory.
Entry *entry_for(plqcd_hip_server *s, Request *r)
{
   Key k;
   std::memcpy(k.L, r->req.L, sizeof(k.L));
   k.seed = r->req.gauge_seed;
   auto it = s->cache.find(k);
   Entry *e = it == s->cache.end() ? nullptr : it->second.get();
   if (!e) {
      plqcd_hip_lattice *lat = nullptr;
      if (plqcd_hip_lattice_create(r->req.L, &lat)) {
         note_failure(s, "lattice");
         r->status = -1;
         return nullptr;
      }
      size_t gb = gauge_bytes_of(lat, PLQCD_PREC_F64, 18);
      if (s->cfg.sloppy_prec != PLQCD_PREC_F64) gb += gauge_bytes_of(lat, s->cfg.sloppy_prec, s->cfg.sloppy_links);
      if (gb + plqcd::CgEngine::bytes_needed(lat, s->cfg.sloppy_prec, 1) > s->budget) {
         r->status = -2;  // can never fit
         plqcd_hip_lattice_destroy(lat);
         return nullptr;
      }
      if (!make_room(s, gb, nullptr)) {
         plqcd_hip_lattice_destroy(lat);
         return nullptr;
      }
      auto ne = std::make_unique<Entry>();
      ne->key = k;
      ne->lat = lat;
      std::vector<su3> u;
      const su3 *src = r->req.gauge;
      if (!src) {
         u.resize((size_t)4 * lat->V);
         plqcd_random_gauge_su3(u.data(), lat->V, k.seed);
         src = u.data();
      }
      if (plqcd_hip_gauge_create(lat, PLQCD_PREC_F64, 18, &ne->g) || plqcd_hip_gauge_upload(ne->g, src)) {
         note_failure(s, "gauge field");
         r->status = -1;
         return nullptr;
      }
      ne->gs = ne->g;
      if (s->cfg.sloppy_prec != PLQCD_PREC_F64) {
         if (plqcd_hip_gauge_create(lat, s->cfg.sloppy_prec, s->cfg.sloppy_links, &ne->gs) ||
             plqcd_hip_gauge_upload(ne->gs, src)) {
            note_failure(s, "sloppy gauge field");
            r->status = -1;
            return nullptr;
         }
      }
      ne->gauge_bytes = gb;
      s->used += gb;
      e = ne.get();
      s->cache[k] = std::move(ne);
   }
   e->last_used = s->now();
   if (!e->eng) {
      const int S = slots_for(s, e->lat);
      const size_t iob = (size_t)S * e->lat->V * sizeof(spinor);
      const size_t eb = plqcd::CgEngine::bytes_needed(e->lat, s->cfg.sloppy_prec, S) + iob;
      if (!make_room(s, eb, &e->key)) return nullptr;
      plqcd::CgEngine *eng = nullptr;
      if (plqcd::CgEngine::create(e->g, e->gs, S, 0.0, s->cfg.use_graph != 0, nullptr, nullptr, &eng)) {
         note_failure(s, "engine");
         r->status = -1;
         return nullptr;
      }
      // transfer threads wait on events of the engine's stream: no capture
      // may start on it later
      if (eng->capture_all()) {
         note_failure(s, "engine graphs");
         delete eng;
         r->status = -1;
         return nullptr;
      }
      e->eng.reset(eng);
      hipError_t he = hipMalloc(&e->io, iob);
      e->io_ev.assign(S, nullptr);
      for (int c = 0; c < S && he == hipSuccess; c++) he = hipEventCreateWithFlags(&e->io_ev[c], hipEventDisableTiming);
      if (he != hipSuccess) {
         plqcd::set_error(10, "%zu bytes of transfer buffers: %s", iob, hipGetErrorString(he));
         note_failure(s, "engine");
         r->status = -1;
         e->eng.reset();
         for (hipEvent_t &ev : e->io_ev)
            if (ev) (void)hipEventDestroy(ev);
         e->io_ev.clear();
         if (e->io) (void)hipFree(e->io);
         e->io = nullptr;
         return nullptr;
      }
      s->engines_created++;
      e->engine_bytes = eb;
      e->owner.assign(S, {nullptr, 0});
      e->sstate.assign(S, Entry::FREE);
      e->parts.assign(S, 0);
      e->part_rc.assign(S, 0);
      s->used += eb;
   }
   return e;
}

void complete(plqcd_hip_server *s, Request *r)
{
   r->t_end = s->now();
   plqcd_serve_result res{};
   res.id = r->id;
   res.status = r->status;
   res.submit_time = r->t_submit;
   res.start_time = r->t_start;
   res.end_time = r->t_end;
   res.iterations = r->iterations;
   res.true_residual = r->residual;
   res.converged = r->converged;
   std::lock_guard<std::mutex> lk(s->mu);
   s->results.push_back(res);
   s->cv.notify_all();
}

// Queues the transfer of one RHS of slot `slot` as up to one part per
// transfer thread (parts of at least kXferPart bytes, whole spinors).
void push_xfer(plqcd_hip_server *s, bool down, Entry *e, int slot, Request *r, int rhs)
{
   const size_t bytes = (size_t)e->lat->V * sizeof(spinor);
   const size_t nt = std::max<size_t>(1, s->xfer.size());
   const size_t k = std::min(nt, std::max<size_t>(1, bytes / kXferPart));
   const size_t per = ((bytes / k + sizeof(spinor) - 1) / sizeof(spinor)) * sizeof(spinor);
   e->parts[slot] = 0;
   e->part_rc[slot] = 0;
   std::lock_guard<std::mutex> lk(s->xmu);
   for (size_t o = 0; o < bytes; o += per) {
      Xfer j{down, e, slot, r, rhs, 0, o, std::min(per, bytes - o)};
      e->parts[slot]++;
      s->xjobs.push_back(j);
   }
   s->xcv.notify_all();
}

// Starts the uploads of the pending right-hand sides of waiting requests into
// free slots (the engine may be iterating on its other slots meanwhile).
void admit_pending(plqcd_hip_server *s)
{
   for (auto &up : s->waiting) {
      Request *r = up.get();
      if (r->status != 0 || r->next_rhs >= r->req.nrhs) continue;
      Entry *e = entry_for(s, r);
      if (!e) continue;
      for (int c = 0; c < e->eng->slots() && r->next_rhs < r->req.nrhs; c++) {
         if (e->sstate[c] != Entry::FREE || e->eng->slot(c).active || e->eng->slot(c).done) continue;
         e->sstate[c] = Entry::UPLOADING;
         e->owner[c] = {r, r->next_rhs};
         if (r->t_start < 0) r->t_start = s->now();
         r->inflight++;
         push_xfer(s, false, e, c, r, r->next_rhs);
         r->next_rhs++;
      }
   }
}

// Applies the transfers the pool has finished.
void collect_xfers(plqcd_hip_server *s)
{
   std::deque<Xfer> done;
   {
      std::lock_guard<std::mutex> lk(s->xmu);
      done.swap(s->xdone);
   }
   for (const Xfer &j : done) {
      Entry *e = j.e;
      if (j.rc) e->part_rc[j.slot] = j.rc;
      if (--e->parts[j.slot] > 0) continue;  // the slot's other parts are still moving
      const int rc = e->part_rc[j.slot];
      if (j.r) j.r->inflight--;
      if (rc) {
         if (j.r) j.r->status = -1;  // reported by the transfer thread
      }
      if (!j.down) {
         if (rc || !j.r || j.r->status != 0) {
            e->sstate[j.slot] = Entry::FREE;
            e->owner[j.slot] = {nullptr, 0};
         } else {
            e->sstate[j.slot] = Entry::UPLOADED;
         }
      } else {
         e->sstate[j.slot] = Entry::FREE;
         e->owner[j.slot] = {nullptr, 0};
         if (j.r) j.r->remaining--;
      }
   }
}

void xfer_main(plqcd_hip_server *s)
{
   plqcd::Stager *g = nullptr;
   const int rc0 = plqcd::Stager::create(s->cfg.device, &g);
   while (true) {
      Xfer j;
      {
         std::unique_lock<std::mutex> lk(s->xmu);
         s->xcv.wait(lk, [&] { return s->xstop || !s->xjobs.empty(); });
         if (s->xjobs.empty()) break;
         j = s->xjobs.front();
         s->xjobs.pop_front();
      }
      const size_t V = (size_t)j.e->lat->V;
      const auto t0 = clock_t_::now();
      char *dev = j.e->io_slot(j.slot) + j.off;
      if (rc0) j.rc = rc0;  // reported below
      else if (j.down)
         j.rc = j.r ? g->download_raw(reinterpret_cast<char *>(j.r->req.x + j.rhs * V) + j.off, dev, j.len,
                                      j.e->io_ev[j.slot])
                    : 0;
      else j.rc = g->upload_raw(dev, reinterpret_cast<const char *>(j.r->req.b + j.rhs * V) + j.off, j.len);
      if (!j.rc && (j.r || !j.down)) {
         atomic_add(s->xfer_bytes, (double)j.len);
         atomic_add(s->xfer_seconds, std::chrono::duration<double>(clock_t_::now() - t0).count());
      }
      if (j.rc) note_failure(s, j.down ? "solution download" : "source upload");
      {
         std::lock_guard<std::mutex> lk(s->xmu);
         s->xdone.push_back(j);
      }
      std::lock_guard<std::mutex> lk(s->mu);
      s->cv.notify_all();  // wakes the worker
   }
   delete g;
}

// Frees the finished slots in the engine and starts their downloads.
void retire_finished(plqcd_hip_server *s, Entry *e, const std::vector<int> &fin)
{
   for (int slot : fin) {
      Request *r = e->owner[slot].first;
      plqcd_cg_stats st{};
      if (e->eng->retire(slot, nullptr, &st)) note_failure(s, "retire");
      if (!r) {  // its request failed: nothing to deliver
         e->sstate[slot] = Entry::FREE;
         e->owner[slot] = {nullptr, 0};
         continue;
      }
      r->iterations = std::max(r->iterations, st.iterations);
      r->residual = std::max(r->residual, st.true_residual);
      r->converged &= st.converged;
      // the solution into the slot's buffer, on the engine's stream (x of a
      // retired slot is not written until the slot is admitted again)
      if (plqcd::unpack_rhs(e->eng->x(), slot, e->io_slot(slot), e->lat->stream) ||
          hipEventRecord(e->io_ev[slot], e->lat->stream) != hipSuccess) {
         note_failure(s, "solution unpack");
         r->status = -1;
         e->sstate[slot] = Entry::FREE;
         e->owner[slot] = {nullptr, 0};
         continue;
      }
      e->sstate[slot] = Entry::DOWNLOADING;
      r->inflight++;
      push_xfer(s, true, e, slot, r, e->owner[slot].second);
   }
}

// Admits the uploaded slots of an engine that is between iterations.
void admit_uploaded(plqcd_hip_server *s, Entry *e)
{
   std::vector<int> which, mi;
   std::vector<double> kap, tol;
   for (int c = 0; c < e->eng->slots(); c++) {
      if (e->sstate[c] != Entry::UPLOADED) continue;
      const Request *r = e->owner[c].first;
      which.push_back(c);
      kap.push_back(r->req.kappa);
      tol.push_back(r->req.tol);
      mi.push_back(r->req.max_iter);
   }
   if (which.empty()) return;
   int rc = 0;
   for (int c : which)  // the sources into the b field, on the engine's stream
      if (!rc) rc = plqcd::pack_rhs(e->eng->b(), c, e->io_slot(c), e->lat->stream);
   if (rc || e->eng->admit(which, kap.data(), tol.data(), mi.data())) {
      note_failure(s, "admission");
      for (int c : which) {
         if (Request *r = e->owner[c].first) r->status = -1;
         e->sstate[c] = Entry::FREE;
         e->owner[c] = {nullptr, 0};
      }
      return;
   }
   std::vector<int> zero;  // zero sources are finished at admission
   for (int c : which) {
      e->sstate[c] = Entry::ACTIVE;
      if (e->eng->slot(c).done) zero.push_back(c);
   }
   retire_finished(s, e, zero);
}

// Stops an engine's active slots after a failure; their requests fail.
void fail_engine(plqcd_hip_server *s, Entry *e, const char *where)
{
   note_failure(s, where);
   (void)hipStreamSynchronize(e->lat->stream);
   (void)hipGetLastError();
   std::vector<int> fin;
   e->eng->abort_active(fin);
   for (int slot : fin) {
      if (Request *r = e->owner[slot].first) r->status = -1;
      plqcd_cg_stats st{};
      (void)e->eng->retire(slot, nullptr, &st);
      e->sstate[slot] = Entry::FREE;
      e->owner[slot] = {nullptr, 0};
   }
   e->phase = Entry::READY;
}

// Advances every engine whose work in flight has completed; returns whether
// anything happened.
bool run_engines(plqcd_hip_server *s)
{
   bool progressed = false;
   for (auto &kv : s->cache) {
      Entry *e = kv.second.get();
      if (!e->eng) continue;
      if (e->phase == Entry::ITERATING && e->eng->ready()) {
         bool updating = false;
         progressed = true;
         if (e->eng->begin_finish(updating)) {
            fail_engine(s, e, "iteration");
            continue;
         }
         if (updating) {
            e->phase = Entry::UPDATING;
         } else {
            std::vector<int> fin;
            if (e->eng->end_finish(fin)) {
               fail_engine(s, e, "iteration end");
               continue;
            }
            retire_finished(s, e, fin);
            e->phase = Entry::READY;
         }
      } else if (e->phase == Entry::UPDATING && e->eng->ready()) {
         std::vector<int> fin;
         progressed = true;
         if (e->eng->end_finish(fin)) {
            fail_engine(s, e, "reliable update");
            continue;
         }
         retire_finished(s, e, fin);
         e->phase = Entry::READY;
      }
      if (e->phase == Entry::READY) admit_uploaded(s, e);
      if (e->phase == Entry::READY && e->eng->active_count() > 0) {
         progressed = true;
         e->last_used = s->now();
         if (e->eng->launch()) {
            fail_engine(s, e, "launch");
            continue;
         }
         e->phase = Entry::ITERATING;
      }
   }
   return progressed;
}

void worker_main(plqcd_hip_server *s)
{
   (void)hipSetDevice(s->cfg.device);
   while (true) {
      {
         std::unique_lock<std::mutex> lk(s->mu);
         bool busy = false;
         for (auto &kv : s->cache) busy |= !kv.second->idle();
         if (!busy && s->incoming.empty() && !s->stop) {
            bool pending = false;
            for (auto &r : s->waiting) pending |= r->next_rhs < r->req.nrhs && r->status == 0;
            if (!pending) s->cv.wait_for(lk, std::chrono::milliseconds(2));
         }
         if (s->stop && s->incoming.empty() && s->waiting.empty()) break;
         while (!s->incoming.empty()) {
            s->waiting.push_back(std::move(s->incoming.front()));
            s->incoming.pop_front();
         }
      }
      collect_xfers(s);
      admit_pending(s);
      const bool progressed = run_engines(s);
      s->used_pub = s->used;
      s->entries_pub = s->cache.size();
      // completed or failed requests leave the waiting list once no transfer
      // of theirs is in flight
      for (auto it = s->waiting.begin(); it != s->waiting.end();) {
         Request *r = it->get();
         if (r->status != 0) {
            // release what a failed request still holds: uploaded slots, and
            // the owner of slots that go on iterating until they stop
            for (auto &kv : s->cache) {
               Entry *e = kv.second.get();
               for (size_t c = 0; c < e->owner.size(); c++)
                  if (e->owner[c].first == r && e->sstate[c] != Entry::UPLOADING &&
                      e->sstate[c] != Entry::DOWNLOADING) {
                     if (e->sstate[c] == Entry::UPLOADED) e->sstate[c] = Entry::FREE;
                     e->owner[c].first = nullptr;
                  }
            }
         }
         const bool ok_done = r->remaining == 0 && r->next_rhs == r->req.nrhs;
         if ((r->status != 0 || ok_done) && r->inflight == 0) {
            complete(s, r);
            it = s->waiting.erase(it);
         } else {
            ++it;
         }
      }
      if (!progressed) {  // wait for work in flight or for a finished transfer
         std::unique_lock<std::mutex> lk(s->mu);
         s->cv.wait_for(lk, std::chrono::microseconds(50));
      }
   }
   {
      std::lock_guard<std::mutex> lk(s->xmu);
      s->xstop = true;
      s->xcv.notify_all();
   }
   for (std::thread &t : s->xfer)
      if (t.joinable()) t.join();
   s->cache.clear();
}

}  // namespace

extern "C" int plqcd_hip_server_create(const plqcd_serve_config *cfg, plqcd_hip_server **out)
{
   plqcd::clear_error();
   if (!cfg || !out) return plqcd::set_error(1, "server: null argument");
   *out = nullptr;
   if (cfg->slots < 1 || cfg->slots > 32) return plqcd::set_error(2, "server: slots must be 1..32");
   if (cfg->transfer_threads < 0 || cfg->transfer_threads > 64)
      return plqcd::set_error(2, "server: transfer_threads must be 0..64");
   const bool ok = cfg->sloppy_prec == PLQCD_PREC_F64 ||
                   (cfg->sloppy_prec == PLQCD_PREC_F32 && (cfg->sloppy_links == 18 || cfg->sloppy_links == 12)) ||
                   (cfg->sloppy_prec == PLQCD_PREC_F16 && (cfg->sloppy_links == 18 || cfg->sloppy_links == 12));
   if (!ok) return plqcd::set_error(2, "server: unsupported sloppy precision");
   plqcd_hip_server *s = new (std::nothrow) plqcd_hip_server();
   if (!s) return plqcd::set_error(4, "out of host memory");
   s->cfg = *cfg;
   // every failure below stops the threads already started and frees s;
   // nothing may throw through the C ABI
   auto fail = [s](int rc) {
      {
         std::lock_guard<std::mutex> lk(s->xmu);
         s->xstop = true;
         s->xcv.notify_all();
      }
      for (std::thread &t : s->xfer)
         if (t.joinable()) t.join();
      delete s;
      return rc;
   };
   if (hipError_t e = hipSetDevice(cfg->device); e != hipSuccess) {
      (void)hipGetLastError();  // not left for a later, unrelated check
      return fail(plqcd::set_error(10, "server: hipSetDevice(%d): %s", cfg->device, hipGetErrorString(e)));
   }
   // one stream per engine and per transfer thread: with HIP's default of 4
   // hardware queues per process unrelated streams share queues and wait for
   // each other (a 60 s request stream took 53.6 s of GPU work instead of
   // 36.4 s); the variable is read when HIP initialises, so only warn here
   {
      const char *q = std::getenv("GPU_MAX_HW_QUEUES");
      const int nq = q ? std::atoi(q) : 4;
      if (nq < 8)
         std::fprintf(stderr,
                      "plqcd server: GPU_MAX_HW_QUEUES is %d; set it to 16 in the environment before the process "
                      "starts using the GPU, or engine and transfer streams share hardware queues\n",
                      nq);
   }
   size_t free_b = 0, total_b = 0;
   if (hipError_t e = hipMemGetInfo(&free_b, &total_b); e != hipSuccess)
      return fail(plqcd::set_error(10, "server: hipMemGetInfo: %s", hipGetErrorString(e)));
   const size_t margin = (size_t)8 << 30;  // the runtime's own allocations, transfer stagers, fragmentation
   s->budget = cfg->hbm_budget ? cfg->hbm_budget : (free_b > margin ? free_b - margin : free_b / 2);
   if (s->budget > free_b) s->budget = free_b;
   s->t0 = clock_t_::now();
   const int nx = cfg->transfer_threads > 0 ? cfg->transfer_threads : kTransferThreads;
   try {
      for (int i = 0; i < nx; i++) s->xfer.emplace_back(xfer_main, s);
      s->worker = std::thread(worker_main, s);
   } catch (...) {
      return fail(plqcd::set_error(4, "server: cannot start its threads"));
   }
   *out = s;
   return 0;
}

extern "C" int plqcd_hip_server_submit(plqcd_hip_server *s, const plqcd_serve_request *req, uint64_t *id)
{
   if (!s || !req || !id) return plqcd::set_error(1, "server: null argument");
   long long V = 1;
   for (int i = 0; i < 4; i++) {
      if (req->L[i] < 2 || (req->L[i] & 1)) return plqcd::set_error(2, "server: bad lattice extent %d", req->L[i]);
      V *= req->L[i];
   }
   if (V > (1LL << 29)) return plqcd::set_error(2, "server: volume %lld exceeds 2^29 sites", V);
   if (req->nrhs < 1 || !req->b || !req->x || req->max_iter < 1 || !(req->tol > 0))
      return plqcd::set_error(2, "server: bad request (nrhs %d, max_iter %d)", req->nrhs, req->max_iter);
   auto r = std::make_unique<Request>();
   r->req = *req;
   r->remaining = req->nrhs;
   std::lock_guard<std::mutex> lk(s->mu);
   r->id = s->next_id++;
   r->t_submit = s->now();
   *id = r->id;
   s->incoming.push_back(std::move(r));
   s->cv.notify_all();
   return 0;
}

extern "C" int plqcd_hip_server_poll(plqcd_hip_server *s, plqcd_serve_result *out, int max, int *n)
{
   if (!s || !n || (max > 0 && !out)) return plqcd::set_error(1, "server: null argument");
   std::lock_guard<std::mutex> lk(s->mu);
   const int k = std::min<int>(max, (int)s->results.size());
   for (int i = 0; i < k; i++) out[i] = s->results[i];
   s->results.erase(s->results.begin(), s->results.begin() + k);
   *n = k;
   return 0;
}

extern "C" int plqcd_hip_server_wait(plqcd_hip_server *s, int count, double timeout_s)
{
   if (!s) return plqcd::set_error(1, "server: null argument");
   std::unique_lock<std::mutex> lk(s->mu);
   const bool ok = s->cv.wait_for(lk, std::chrono::duration<double>(timeout_s),
                                  [&] { return (int)s->results.size() >= count; });
   return ok ? 0 : plqcd::set_error(3, "server: timed out waiting for %d results", count);
}

extern "C" int plqcd_hip_server_info(plqcd_hip_server *s, double out[8])
{
   if (!s || !out) return plqcd::set_error(1, "server: null argument");
   out[0] = (double)s->budget;
   out[1] = (double)s->used_pub;
   out[2] = (double)s->entries_pub;
   out[3] = (double)s->failed;
   out[4] = (double)s->engines_created;
   out[5] = (double)s->evictions;
   out[6] = s->xfer_bytes;
   out[7] = s->xfer_seconds;
   return 0;
}

extern "C" void plqcd_hip_server_destroy(plqcd_hip_server *s)
{
   if (!s) return;
   {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stop = true;
      s->cv.notify_all();
   }
   if (s->worker.joinable()) s->worker.join();
   delete s;
}
