// Solve server: continuous batching of CG requests on one GPU.
//
// Requests name a lattice and a gauge field (by seed) and bring 1..n
// right-hand sides.  A worker thread owns the GPU:
//  * gauge fields (F64 and the sloppy copy) stay resident in a cache keyed by
//    (lattice, seed) and are evicted least-recently-used when HBM is needed;
//  * each (lattice, gauge) has one CG engine with S slots on its own stream;
//    right-hand sides enter free slots between iterations and leave as soon
//    as they converge, so requests that arrive while others run join them;
//  * engines of different lattices iterate concurrently (their iteration
//    launches are issued back to back on separate streams before any of them
//    is waited for);
//  * admission is by HBM: a new gauge or engine is created only if its bytes
//    fit the budget after evicting idle cache entries; otherwise the request
//    waits.
// A right-hand side's iterates do not depend on the other slots (see
// cg_engine.h), so its result is the same alone and in the mix.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "cg_engine.h"

namespace {

using clock_t_ = std::chrono::steady_clock;

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
   size_t gauge_bytes = 0, engine_bytes = 0;
   double last_used = 0;
   // slot -> (request, rhs)
   std::vector<std::pair<Request *, int>> owner;
   ~Entry()
   {
      eng.reset();
      plqcd_hip_gauge_destroy(gs != g ? gs : nullptr);
      plqcd_hip_gauge_destroy(g);
      plqcd_hip_lattice_destroy(lat);
   }
};

}  // namespace

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

   double now() const { return std::chrono::duration<double>(clock_t_::now() - t0).count(); }
};

namespace {

size_t gauge_bytes_of(const plqcd_hip_lattice *lat, int prec, int links)
{
   return plqcd::gauge_bytes(prec, links, lat->Vh);
}

int slots_for(plqcd_hip_server *s, const plqcd_hip_lattice *lat)
{
   const size_t per = plqcd::CgEngine::bytes_needed(lat, s->cfg.sloppy_prec, 1);
   size_t fit = (s->budget / 4) / std::max<size_t>(per, 1);
   int S = (int)std::min<size_t>(fit, (size_t)s->cfg.slots);
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
         if (e->eng && (e->eng->active_count() > 0)) continue;
         bool busy = false;
         for (auto &o : e->owner) busy |= o.first != nullptr;
         if (busy) continue;
         if (!victim || e->last_used < victim->last_used) victim = e;
      }
      if (!victim) return false;
      s->used -= victim->gauge_bytes + victim->engine_bytes;
      s->cache.erase(victim->key);
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
         r->status = -1;
         return nullptr;
      }
      ne->gs = ne->g;
      if (s->cfg.sloppy_prec != PLQCD_PREC_F64) {
         if (plqcd_hip_gauge_create(lat, s->cfg.sloppy_prec, s->cfg.sloppy_links, &ne->gs) ||
             plqcd_hip_gauge_upload(ne->gs, src)) {
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
      const size_t eb = plqcd::CgEngine::bytes_needed(e->lat, s->cfg.sloppy_prec, S);
      if (!make_room(s, eb, &e->key)) return nullptr;
      plqcd::CgEngine *eng = nullptr;
      if (plqcd::CgEngine::create(e->g, e->gs, S, 0.0, s->cfg.use_graph != 0, nullptr, nullptr, &eng)) {
         r->status = -1;
         return nullptr;
      }
      e->eng.reset(eng);
      e->engine_bytes = eb;
      e->owner.assign(S, {nullptr, 0});
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

// Admits the pending right-hand sides of waiting requests into free slots.
void admit_pending(plqcd_hip_server *s)
{
   for (auto &up : s->waiting) {
      Request *r = up.get();
      if (r->status != 0 || r->next_rhs >= r->req.nrhs) continue;
      Entry *e = entry_for(s, r);
      if (!e) continue;
      const size_t V = (size_t)e->lat->V;
      while (r->next_rhs < r->req.nrhs) {
         const int slot = e->eng->free_slot();
         if (slot < 0) break;
         const int k = r->next_rhs;
         if (e->eng->admit_host(slot, r->req.b + k * V, r->req.kappa, r->req.tol, r->req.max_iter)) {
            r->status = -1;
            break;
         }
         if (r->t_start < 0) r->t_start = s->now();
         r->next_rhs++;
         if (e->eng->slot(slot).done) {  // zero source: finished at admission
            plqcd_cg_stats st;
            e->eng->retire(slot, r->req.x + k * V, &st);
            r->remaining--;
         } else {
            e->owner[slot] = {r, k};
         }
      }
   }
}

void run_engines(plqcd_hip_server *s)
{
   std::vector<Entry *> live;
   for (auto &kv : s->cache)
      if (kv.second->eng && kv.second->eng->active_count() > 0) live.push_back(kv.second.get());
   // issue every engine's iteration before waiting for any, so they overlap
   for (Entry *e : live)
      if (e->eng->launch()) s->failed++;
   for (Entry *e : live) {
      std::vector<int> fin;
      if (e->eng->finish(fin)) {
         s->failed++;
         continue;
      }
      e->last_used = s->now();
      const size_t V = (size_t)e->lat->V;
      for (int slot : fin) {
         Request *r = e->owner[slot].first;
         const int k = e->owner[slot].second;
         plqcd_cg_stats st{};
         if (e->eng->retire(slot, r ? r->req.x + k * V : nullptr, &st)) s->failed++;
         e->owner[slot] = {nullptr, 0};
         if (!r) continue;
         r->iterations = std::max(r->iterations, st.iterations);
         r->residual = std::max(r->residual, st.true_residual);
         r->converged &= st.converged;
         r->remaining--;
      }
   }
}

void worker_main(plqcd_hip_server *s)
{
   (void)hipSetDevice(s->cfg.device);
   while (true) {
      {
         std::unique_lock<std::mutex> lk(s->mu);
         bool busy = false;
         for (auto &kv : s->cache) busy |= kv.second->eng && kv.second->eng->active_count() > 0;
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
      admit_pending(s);
      run_engines(s);
      s->used_pub = s->used;
      s->entries_pub = s->cache.size();
      // completed or failed requests leave the waiting list
      for (auto it = s->waiting.begin(); it != s->waiting.end();) {
         Request *r = it->get();
         if (r->status != 0 || (r->remaining == 0 && r->next_rhs == r->req.nrhs)) {
            if (r->status != 0) {
               // release the slots a failed request still holds
               for (auto &kv : s->cache)
                  for (auto &o : kv.second->owner)
                     if (o.first == r) o.first = nullptr;
            }
            complete(s, r);
            it = s->waiting.erase(it);
         } else {
            ++it;
         }
      }
   }
   s->cache.clear();
}

}  // namespace

extern "C" int plqcd_hip_server_create(const plqcd_serve_config *cfg, plqcd_hip_server **out)
{
   plqcd::clear_error();
   if (!cfg || !out) return plqcd::set_error(1, "server: null argument");
   *out = nullptr;
   if (cfg->slots < 1 || cfg->slots > 32) return plqcd::set_error(2, "server: slots must be 1..32");
   const bool ok = cfg->sloppy_prec == PLQCD_PREC_F64 ||
                   (cfg->sloppy_prec == PLQCD_PREC_F32 && (cfg->sloppy_links == 18 || cfg->sloppy_links == 12)) ||
                   (cfg->sloppy_prec == PLQCD_PREC_F16 && cfg->sloppy_links == 12);
   if (!ok) return plqcd::set_error(2, "server: unsupported sloppy precision");
   plqcd_hip_server *s = new (std::nothrow) plqcd_hip_server();
   if (!s) return plqcd::set_error(4, "out of host memory");
   s->cfg = *cfg;
   PLQCD_HIP_TRY(hipSetDevice(cfg->device));
   size_t free_b = 0, total_b = 0;
   PLQCD_HIP_TRY(hipMemGetInfo(&free_b, &total_b));
   const size_t margin = (size_t)4 << 30;
   s->budget = cfg->hbm_budget ? cfg->hbm_budget : (free_b > margin ? free_b - margin : free_b / 2);
   if (s->budget > free_b) s->budget = free_b;
   s->t0 = clock_t_::now();
   s->worker = std::thread(worker_main, s);
   *out = s;
   return 0;
}

extern "C" int plqcd_hip_server_submit(plqcd_hip_server *s, const plqcd_serve_request *req, uint64_t *id)
{
   if (!s || !req || !id) return plqcd::set_error(1, "server: null argument");
   for (int i = 0; i < 4; i++)
      if (req->L[i] < 2 || (req->L[i] & 1)) return plqcd::set_error(2, "server: bad lattice extent %d", req->L[i]);
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

extern "C" int plqcd_hip_server_info(plqcd_hip_server *s, double out[4])
{
   if (!s || !out) return plqcd::set_error(1, "server: null argument");
   out[0] = (double)s->budget;
   out[1] = (double)s->used_pub;
   out[2] = (double)s->entries_pub;
   out[3] = (double)s->failed;
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
