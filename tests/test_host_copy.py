"""CPU test of the library's multi-threaded host copy (plqcd_amd/csrc/hostcopy.cpp),
the copy that moves caller memory into the page-locked staging buffers of every
host <-> device transfer: compiled with g++ beside a small driver, it must copy
exactly (sizes below and above the parallel threshold, unaligned ends and
offsets), from several threads at once, and in a child forked after the pool
started (which has none of the pool's threads)."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]

DRIVER = r"""
#include <sys/wait.h>
#include <unistd.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>
#include "internal.h"

static int check(size_t n, size_t off, unsigned seed)
{
   std::vector<unsigned char> a(n + off + 7), b(n + off + 7, 0xAB);
   for (size_t i = 0; i < a.size(); i++) a[i] = (unsigned char)((i * 2654435761u + seed) >> 13);
   plqcd::host_copy(b.data() + off, a.data() + off, n);
   if (std::memcmp(a.data() + off, b.data() + off, n) != 0) return 1;
   for (size_t i = 0; i < off; i++) if (b[i] != 0xAB) return 2;                  // nothing before
   for (size_t i = off + n; i < b.size(); i++) if (b[i] != 0xAB) return 3;       // nothing after
   return 0;
}

int main()
{
   const size_t sizes[] = {0, 1, 4095, 4 << 20, (4 << 20) + 1, (64 << 20) + 12345, 201326592};
   for (size_t n : sizes)
      for (size_t off : {0, 3})
         if (int rc = check(n, off, (unsigned)n)) { std::printf("fail n=%zu off=%zu rc=%d\n", n, off, rc); return 1; }
   int bad = 0;
   std::vector<std::thread> th;
   for (int k = 0; k < 4; k++)
      th.emplace_back([&bad, k] { for (int it = 0; it < 3; it++) if (check((48 << 20) + k, k, 7 * k + it)) bad = 1; });
   for (auto &t : th) t.join();
   if (bad) { std::printf("fail: concurrent copies\n"); return 1; }
   pid_t pid = fork();
   if (pid == 0) _exit(check((32 << 20) + 5, 1, 99));
   int st = 0;
   waitpid(pid, &st, 0);
   if (!WIFEXITED(st) || WEXITSTATUS(st) != 0) { std::printf("fail: forked child\n"); return 1; }
   std::printf("ok\n");
   return 0;
}
"""


@pytest.mark.timeout(300)
def test_host_copy_exact_concurrent_and_after_fork(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    src = tmp_path / "drv.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "drv"
    csrc = ROOT / "plqcd_amd" / "csrc"
    subprocess.run([gxx, "-O2", "-std=c++17", "-pthread", f"-I{csrc}", f"-I{ROOT / 'include'}",
                    str(src), str(csrc / "hostcopy.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
