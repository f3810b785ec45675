// Host utilities with the semantics of lib/utils.c: aligned allocation that
// remembers the true address (utils.c:35-90), wall-clock stop watch
// (utils.c:120-125), plus the library's error state.
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <sys/time.h>

#include "internal.h"

namespace plqcd {

static thread_local char g_error[512] = "";

int set_error(int code, const char *fmt, ...)
{
   va_list ap;
   va_start(ap, fmt);
   vsnprintf(g_error, sizeof(g_error), fmt, ap);
   va_end(ap);
   return code;
}

void clear_error() { g_error[0] = '\0'; }

}  // namespace plqcd

extern "C" const char *plqcd_hip_last_error(void) { return plqcd::g_error; }

// ---------------------------------------------------------------- allocation

extern "C" void *alloc(size_t size, size_t alignment)
{
   void *ptr = nullptr;
   if (posix_memalign(&ptr, alignment, size) != 0 || ptr == nullptr) {
      plqcd::set_error(1, "alloc(): posix_memalign of %zu bytes failed", size);
      fprintf(stderr, "alloc(): could not allocate %zu bytes\n", size);
      return nullptr;
   }
   return ptr;
}

namespace {
// Aligned block -> address returned by malloc.  A mutex-protected list, like
// the reference's singly linked list but safe to use from several threads.
struct Block {
   char *addr;
   char *true_addr;
   Block *next;
};
Block *g_blocks = nullptr;
std::mutex g_blocks_mu;
}  // namespace

extern "C" void *amalloc(size_t size, int p)
{
   if (size == 0 || p < 0 || p > 30) return nullptr;
   const size_t shift = size_t(1) << p;
   char *true_addr = static_cast<char *>(malloc(size + shift));
   Block *b = static_cast<Block *>(malloc(sizeof(Block)));
   if (!true_addr || !b) {
      free(true_addr);
      free(b);
      return nullptr;
   }
   // first multiple of 2^p strictly above true_addr, as utils.c:55
   uintptr_t a = (reinterpret_cast<uintptr_t>(true_addr) + shift) & ~(uintptr_t)(shift - 1);
   b->addr = reinterpret_cast<char *>(a);
   b->true_addr = true_addr;
   std::lock_guard<std::mutex> lk(g_blocks_mu);
   b->next = g_blocks;
   g_blocks = b;
   return b->addr;
}

extern "C" void afree(void *addr)
{
   std::lock_guard<std::mutex> lk(g_blocks_mu);
   Block *prev = nullptr;
   for (Block *b = g_blocks; b; prev = b, b = b->next) {
      if (b->addr == addr) {
         (prev ? prev->next : g_blocks) = b->next;
         free(b->true_addr);
         free(b);
         return;
      }
   }
   // unknown address: nothing happens (utils.c:68-90)
}

extern "C" double stop_watch(double t)
{
   struct timeval tv;
   gettimeofday(&tv, nullptr);
   return (tv.tv_sec + tv.tv_usec * 1e-6) - t;
}
