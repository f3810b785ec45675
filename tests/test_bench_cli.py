"""The bench.py contract on the CPU: the one-GPU guard, the suites on the
command line, and the per-launch traffic the bench line reports, read from
the committed rocprofv3 summary of the headline kernel."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def run(*args):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_exits_with_a_message():
    r = run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_help_lists_the_suites():
    r = run("--help")
    assert r.returncode == 0
    for suite in ("hop", "solver", "serve", "calib", "cg"):
        assert suite in r.stdout


def test_headline_traffic_comes_from_the_profile():
    sys.path.insert(0, str(ROOT))
    import bench
    t = bench.read_traffic("hop_f64_l18_32x32x32x64")
    compulsory = bench.BYTES_PER_SITE[(0, 18)] * 32 * 32 * 32 * 64
    assert t is not None and 1.0 <= t / compulsory < 2.0
    assert bench.read_traffic("no_such_kernel") is None


def test_cpu_baseline_uses_every_available_core():
    """cpu_baseline runs on every CPU this process may use (the physical cores
    of the benchmark hosts), and says how many."""
    import os
    sys.path.insert(0, str(ROOT))
    import bench
    cores, desc = bench.host_cores()
    avail = len(os.sched_getaffinity(0))
    assert 1 <= cores <= avail
    assert f"{avail} CPUs available" in desc
