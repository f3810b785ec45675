// Multi-threaded host memcpy for the transfers between caller memory (pageable,
// in the reference's layout) and the library's page-locked staging buffers.
// One thread copies some 8-12 GB/s here, a PCIe 5 x16 link moves about twice
// that, so a transfer that copies on one thread is bound by the copy.  A small
// pool of workers, started on first use and kept for the life of the process,
// splits each large copy into page-aligned pieces; the calling thread takes a
// piece too and returns when every piece is done.  Copies from several
// threads take turns, so paths that already copy on many threads at once
// (the server's stagers) keep their own memcpy.
#include <sched.h>
#include <unistd.h>

#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "internal.h"

namespace {

constexpr size_t kMinParallel = size_t(4) << 20;  // smaller copies stay on the calling thread
constexpr size_t kAlign = 4096;

int pool_threads()
{
   cpu_set_t set;
   int n = 0;
   if (sched_getaffinity(0, sizeof(set), &set) == 0) n = CPU_COUNT(&set);
   if (n <= 0) n = (int)std::thread::hardware_concurrency();
   // half the CPUs the process may use, at most 8: host memory bandwidth, not
   // the core count, bounds a copy
   return std::max(1, std::min(8, n / 2));
}

class CopyPool {
  public:
   // never destroyed: the workers stay blocked on the pool's condition
   // variable until the process ends (joining them from a static destructor
   // would run while other threads may still copy, and a forked child holds
   // thread objects of threads it does not have)
   static CopyPool *get()
   {
      static CopyPool *pool = new (std::nothrow) CopyPool();
      return pool;  // null if out of memory: the caller copies on its own thread
   }

   void copy(void *dst, const void *src, size_t n)
   {
      const int parts = nworkers_ + 1;
      // a child forked after the pool started has none of its threads
      if (n < kMinParallel || parts == 1 || getpid() != owner_) {
         std::memcpy(dst, src, n);
         return;
      }
      std::lock_guard<std::mutex> one(call_mu_);  // one parallel copy at a time
      size_t piece = (n + parts - 1) / parts;
      piece = (piece + kAlign - 1) / kAlign * kAlign;
      {
         std::lock_guard<std::mutex> lk(mu_);
         dst_ = static_cast<char *>(dst);
         src_ = static_cast<const char *>(src);
         n_ = n;
         piece_ = piece;
         pending_ = nworkers_;
         gen_++;
      }
      cv_.notify_all();
      run_piece(0);  // the calling thread's share
      std::unique_lock<std::mutex> lk(mu_);
      done_cv_.wait(lk, [&] { return pending_ == 0; });
   }

  private:
   CopyPool()
   {
      owner_ = getpid();
      const int want = pool_threads() - 1;
      // a thread that cannot be started leaves a smaller pool (no workers:
      // plain memcpy); nothing may throw through the C ABI
      for (int k = 0; k < want; k++) {
         try {
            th_.emplace_back([this, k] { worker(k + 1); });
         } catch (...) {
            break;
         }
         nworkers_ = k + 1;
      }
   }
   void run_piece(int k)
   {
      const size_t o = size_t(k) * piece_;
      if (o >= n_) return;
      std::memcpy(dst_ + o, src_ + o, std::min(piece_, n_ - o));
   }

   void worker(int k)
   {
      uint64_t seen = 0;
      for (;;) {
         {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return gen_ != seen; });
            seen = gen_;
         }
         run_piece(k);
         {
            std::lock_guard<std::mutex> lk(mu_);
            if (--pending_ == 0) done_cv_.notify_one();
         }
      }
   }

   pid_t owner_ = 0;
   int nworkers_ = 0;
   std::vector<std::thread> th_;
   std::mutex call_mu_, mu_;
   std::condition_variable cv_, done_cv_;
   uint64_t gen_ = 0;
   int pending_ = 0;
   char *dst_ = nullptr;
   const char *src_ = nullptr;
   size_t n_ = 0, piece_ = 0;
};

}  // namespace

void plqcd::host_copy(void *dst, const void *src, size_t n)
{
   if (CopyPool *pool = CopyPool::get())
      pool->copy(dst, src, n);
   else
      std::memcpy(dst, src, n);
}
