#!/usr/bin/env python3
"""Benchmark of the Wilson hopping matrix H on one MI355X.

Workload (one step = one application of the full H, one kernel launch):
32^3 x 64 lattice (V = 2,097,152 sites), FP64 spinors and 18-real FP64 links,
seeded random SU(3) gauge field and Gaussian source resident in HBM, periodic
boundaries, Eq. 2.1 (wilson mode).  Prints one JSON line.

  value     = 1320 flop/site x V x steps / (event time)      [GFLOP/s]
  roofline  = compulsory bytes (960 B/site: spinor in 192, links 576,
              spinor out 192) x V / (average launch time) against 8 TB/s
  cpu_baseline = the reference's CPU hopping matrix (compiled ASSYMBLY eo+oe,
              OpenMP) or, where that binary is absent, the oracle's C/OpenMP
              restatement, timed on this host on a bounded sample
  formats   = the same H in every storage format: FP64 12-real links, FP32
              and FP16 with 18/12-real links, FP64 18-real (200 launches each)
  compat    = the reference's own operator (compat mode), device-resident
  schur     = the even-odd operator 1 - kappa^2 H_eo H_oe on the even sites
  dropin    = wall time per call of the drop-in eo then oe on host arrays
              (transfers included), links uploaded per call or cached, and
              of the fused call
  solver    = plain FP64 CG on one RHS against the mixed-precision CG on 12
              RHS (FP16 storage, FP64 reliable updates), true residual 1e-10
  cg_48c96  = the same comparison at 48^3 x 96, one RHS each
  multi_rhs = H on 1, 4, 12 and 24 RHS sharing one gauge field (FP64 18-real,
              FP16 12-real), us per application and per RHS
  ragged_batch = one launch over 16^3x32 + 24^3x48 + 32^3x64 with 12 RHS each
              against three separate launches
  serve     = a short Poisson serving run (--serve-seconds, 2 requests/s):
              p50/p99 latency and solves/s beside one-at-a-time FP64 solves
The headline is timed last, after the other GPU measurements, so that its
warm-up launches start on a GPU at its steady clocks.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
# The serving suite runs one HIP stream per solver engine and per transfer
# thread; with HIP's default of 4 hardware queues per process unrelated
# streams share queues and wait for each other.  Read at HIP initialisation.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")

METRIC = "hopping GFLOP/s & %HBM roofline at 32^3x64; batched mixed vs FP64 solves/sec @1e-10"
HBM_PEAK_GBS = 8000.0
FLOP_PER_SITE = 1320
BYTES_PER_SITE = {(0, 18): 960, (0, 12): 768, (1, 18): 480, (1, 12): 384, (2, 12): 200, (2, 18): 248}
PREC_NAMES = {0: "f64", 1: "f32", 2: "f16"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic(kernel_tag: str):
    """L2-to-memory bytes per launch measured with rocprofv3 --pmc, from profiles/."""
    f = ROOT / "profiles" / f"{kernel_tag}_pmc.txt"
    if not f.exists():
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", f.read_text(), flags=re.M)
    return float(m.group(1)) if m else None


class HopCase:
    """Resident inputs of one application of H: seeded links and source in
    HBM (the setup is not timed)."""

    def __init__(self, L, prec, links, mode):
        from plqcd_amd import hopping as hp
        self.hp, self.mode = hp, mode
        V = int(np.prod(L))
        t0 = time.time()
        u = hp.random_gauge_su3(V, 1)
        q = hp.random_spinor(V, 2)
        log(f"[bench] inputs generated in {time.time() - t0:.1f}s")
        self.lat = hp.Lattice(L)
        self.g = hp.Gauge(self.lat, prec, links).upload(u)
        self.fi = hp.Field(self.lat, prec).upload(q)
        self.fo = hp.Field(self.lat, prec)
        self.lat.sync()

    def run(self, steps, warmup):
        """`warmup` untimed applications, then milliseconds of `steps` applications
        between two events on the lattice stream (host-synchronised on both sides)."""
        hp, lat = self.hp, self.lat
        for _ in range(warmup):
            hp.hop(self.g, self.fi, self.fo, hp.PARITY_ALL, self.mode)
        lat.sync()
        lat.timer_start()
        for _ in range(steps):
            hp.hop(self.g, self.fi, self.fo, hp.PARITY_ALL, self.mode)
        ms = lat.timer_stop()
        lat.sync()
        return ms

    def close(self):
        self.fo.close(); self.fi.close(); self.g.close(); self.lat.close()


def gpu_run(L, prec, links, steps, warmup, mode):
    c = HopCase(L, prec, links, mode)
    try:
        return c.run(steps, warmup)
    finally:
        c.close()


def host_cores():
    """(threads to use, description): every CPU this process may run on, which
    on the benchmark hosts is every physical core (lscpu: sockets x cores per
    socket, one thread per core)."""
    avail = len(os.sched_getaffinity(0))
    phys = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        f = {k.strip(): v.strip() for k, v in (ln.split(":", 1) for ln in out.splitlines() if ":" in ln)}
        phys = int(f["Socket(s)"]) * int(f["Core(s) per socket"])
        tpc = int(f.get("Thread(s) per core", "1"))
    except Exception:  # lscpu missing: the affinity mask alone
        tpc = 1
    cores = min(avail, phys) if phys else avail
    desc = f"{cores} threads on {avail} CPUs available" + (
        f", {phys} physical cores ({tpc} thread(s) per core, lscpu)" if phys else "")
    return cores, desc


def cpu_baseline(L, seconds: float):
    """The reference's own CPU hopping matrix where its binary was built, else
    the oracle, on every host core (OMP_NUM_THREADS = cores, as the reference's
    NTHREAD, lib/start.c:110-114)."""
    import oracle
    V = int(np.prod(L))
    cores, cores_desc = host_cores()
    if oracle.ref_available():
        try:
            with tempfile.TemporaryDirectory() as td:
                inp = Path(td) / "bench.input"
                oracle.write_input_file(inp, L, nthread=cores, align=6)
                env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close", OMP_PLACES="cores")
                r = subprocess.run([str(oracle.REF_BIN), str(inp), "bench", str(seconds)], cwd=td, env=env,
                                   capture_output=True, text=True, timeout=seconds * 4 + 120)
                m = re.search(r"calls (\d+) seconds ([0-9.]+)", r.stdout)
                if r.returncode == 0 and m:
                    calls, sec = int(m.group(1)), float(m.group(2))
                    return {"value": round(calls * V * FLOP_PER_SITE / sec / 1e9, 3), "unit": "GFLOP/s",
                            "cores": cores, "kind": "reference",
                            "sample": f"reference ASSYMBLY eo+oe (gcc -O3 -fopenmp, {cores_desc}), "
                                      f"32^3x64 FP64, {calls} applications of H in {sec:.1f}s, "
                                      f"harness rand() fill"}
                log(f"[bench] reference CPU binary failed (rc {r.returncode}); using the restatement")
        except Exception as e:  # the port below is the documented fallback
            log(f"[bench] reference CPU binary unusable: {e}")
    from plqcd_amd import hopping as hp
    u = hp.random_gauge_su3(V, 1)
    q = hp.random_spinor(V, 2)
    t1 = oracle.bench(q, u, L, 1, oracle.MODE_WILSON, cores)
    n = max(1, int(seconds / max(t1, 1e-3)))
    t = oracle.bench(q, u, L, n, oracle.MODE_WILSON, cores)
    return {"value": round(n * V * FLOP_PER_SITE / t / 1e9, 3), "unit": "GFLOP/s", "cores": cores, "kind": "port",
            "sample": f"C/OpenMP restatement eo+oe ({cores_desc}), 32^3x64 FP64, {n} applications of H in {t:.1f}s"}


def formats_run(L, mode, steps=200, warmup=10, headline_case=None):
    """The same application of H in every storage format (FP64 with 12-real
    links, FP32 and FP16 with 18- and 12-real links, and last FP64 with
    18-real links, the headline's own format, on the headline's resident
    inputs when given), default launch shapes; per format: time per launch,
    GFLOP/s and compulsory GB/s."""
    V = int(np.prod(L))
    out = {}
    for prec, links in ((0, 12), (1, 18), (1, 12), (2, 18), (2, 12), (0, 18)):
        if (prec, links) == (0, 18) and headline_case is not None:
            ms = headline_case.run(steps, warmup)
        else:
            ms = gpu_run(L, prec, links, steps, warmup, mode)
        t = ms / 1e3 / steps
        gbs = BYTES_PER_SITE[(prec, links)] * V / t / 1e9
        out[f"{PREC_NAMES[prec]}_l{links}"] = {"us_per_launch": round(t * 1e6, 2),
                                               "gflops": round(FLOP_PER_SITE * V / t / 1e9, 1),
                                               "compulsory_bytes_per_site": BYTES_PER_SITE[(prec, links)],
                                               "compulsory_GBps": round(gbs, 1),
                                               "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4)}
    return out


def mix_ceiling(L, reps=20):
    """GB/s of a streaming kernel with the FP64 hopping matrix's byte counts
    and read:write mix at this lattice: four streams of V x 192 B read, one of
    V x 192 B written (plqcd_hip_stream_bandwidth kind 3); the practical
    memory ceiling for H's traffic, measured in the same run."""
    import ctypes as C
    from plqcd_amd import capi
    V = int(np.prod(L))
    gbs = C.c_double()
    capi.check(capi.lib().plqcd_hip_stream_bandwidth(3, V * 192, reps, C.byref(gbs)), "stream_bandwidth")
    return round(gbs.value, 1)


def compat_run(L, steps=200, warmup=10):
    """The reference's own operator (compat mode: lib/dirac.c as compiled,
    tests/test_hopping.c:287-301), FP64 18-real, device-resident fields."""
    V = int(np.prod(L))
    ms = gpu_run(L, 0, 18, steps, warmup, 0)
    t = ms / 1e3 / steps
    return {"us_per_launch": round(t * 1e6, 2), "gflops": round(FLOP_PER_SITE * V / t / 1e9, 1),
            "frac_of_hbm_peak": round(960 * V / t / 1e9 / HBM_PEAK_GBS, 4)}


def schur_run(L, steps=100, warmup=10):
    """The even-odd (Schur) operator out_e = in_e - kappa^2 H_eo H_oe in_e
    (plqcd_hip_apply_schur: two half-lattice hops, each reading every link, the
    second forming in_e - kappa^2 (H tmp) in its epilogue; each call
    synchronises), FP64 18-real, one RHS, kappa 0.124.  Compulsory bytes per
    application, per site of the lattice: 2 x (576 link + 96 spinor in + 96
    spinor out) for the hops and 96 for the centre in_e = 1632 B."""
    from plqcd_amd import hopping as hp, solver as sv
    V = int(np.prod(L))
    lat = hp.Lattice(L)
    g = hp.Gauge(lat).upload(hp.random_gauge_su3(V, 1))
    fi = hp.Field(lat).upload(hp.random_spinor(V, 2))
    tmp, fo = hp.Field(lat), hp.Field(lat)
    for _ in range(warmup):
        sv.apply_schur(g, fi, tmp, fo, 0.124, 0)
    lat.timer_start()
    for _ in range(steps):
        sv.apply_schur(g, fi, tmp, fo, 0.124, 0)
    t = lat.timer_stop() / 1e3 / steps  # seconds per application
    for f in (fi, tmp, fo, g):
        f.close()
    lat.close()
    nbytes = 1632 * V
    return {"us_per_application": round(t * 1e6, 1), "compulsory_bytes_per_site": 1632,
            "compulsory_GBps": round(nbytes / t / 1e9, 1), "frac_of_hbm_peak": round(nbytes / t / 1e9 / HBM_PEAK_GBS, 4),
            "parity": "even", "steps": steps}


def dropin_run(L, calls=5):
    """Wall seconds per call of the drop-in entry points on host arrays in the
    reference layout, compat mode (tests/test_hopping.c:287-301 times the eo
    then oe pair on host arrays): the pair as the harness calls it, with the
    links uploaded on every call (the default) and with the link cache on
    (plqcd_hip_cache_gauge: uploaded once, reused while the pointer is the
    same), and the fused plqcd_hopping_matrix_hip.  Each call includes its
    transfers."""
    from plqcd_amd import capi, hopping as hp
    V = int(np.prod(L))
    u = hp.random_gauge_su3(V, 1)
    q = hp.random_spinor(V, 2)
    out = hp.spinor_field(V)
    hp.init_dims(L)
    hp.set_mode(hp.HOP_COMPAT)
    lib = capi.lib()

    def pair():
        return hp.hopping_matrix_eo(q, out, u) + hp.hopping_matrix_oe(q, out, u)

    def timed(fn):
        fn()  # device buffers and first-touch costs
        return float(np.median([fn() for _ in range(calls)]))

    t_pair = timed(pair)
    lib.plqcd_hip_cache_gauge(1)
    t_cached = timed(pair)
    lib.plqcd_hip_cache_gauge(0)
    t_fused = timed(lambda: hp.hopping_matrix(q, out, u))
    hp.finalize_plqcd()
    spin, links = V * 192, V * 576
    rec = {"calls": calls}
    for key, t, host_bytes in (("eo_oe_pair", t_pair, 2 * links + 2 * spin),
                               ("eo_oe_pair_links_cached", t_cached, 2 * spin),
                               ("fused", t_fused, links + 2 * spin)):
        rec[key] = {"seconds_per_call": round(t, 4), "gflops": round(FLOP_PER_SITE * V / t / 1e9, 2),
                    "host_bytes_per_call": host_bytes, "host_GBps": round(host_bytes / t / 1e9, 2)}
    return rec


def cg48_compare(kappa=0.124, tol=1e-10):
    """48^3 x 96 CG to a 1e-10 true residual: plain FP64 against the mixed
    FP16/12-real inner solve with FP64 reliable updates (one RHS each; the
    second of two solves timed)."""
    from plqcd_amd import hopping as hp, solver as sv
    L = (48, 48, 48, 96)
    V = int(np.prod(L))
    u = hp.random_gauge_su3(V, 1)
    lat = hp.Lattice(L)
    g = hp.Gauge(lat).upload(u)
    gs = hp.Gauge(lat, 2, 12).upload(u)
    del u
    b = hp.Field(lat).upload(hp.random_spinor(V, 2))
    x = hp.Field(lat)
    out = {"lattice": list(L), "kappa": kappa, "tol": tol}
    for label, sl in (("fp64", None), ("mixed_f16_l12", gs)):
        sv.cg_solve(g, b, x, kappa, tol, sloppy=sl)
        lat.sync()
        t0 = time.perf_counter()
        st = sv.cg_solve(g, b, x, kappa, tol, sloppy=sl)[0]
        t = time.perf_counter() - t0
        out[label] = {"seconds": round(t, 4), "iterations": st["iterations"],
                      "reliable_updates": st["reliable_updates"], "true_residual": st["true_residual"],
                      "converged": bool(st["converged"])}
    out["speedup"] = round(out["fp64"]["seconds"] / out["mixed_f16_l12"]["seconds"], 3)
    x.close(); b.close(); gs.close(); g.close(); lat.close()
    return out


def solve_compare(L, nrhs=12, sloppy=(2, 12), kappa=0.124, tol=1e-10, reps=3, tune=False):
    """Plain FP64 CG on one RHS vs the mixed-precision CG on a batch of nrhs,
    same gauge field and sources, true residual <= tol; each solve timed
    `reps` times after one untimed solve, the median reported.  With `tune`,
    both solvers' launch shapes are autotuned first (plqcd_hip_autotune_solver
    on the iterating gauge field and RHS count of each).  Returns a dict."""
    import ctypes as C
    from plqcd_amd import capi, hopping as hp, solver as sv
    V = int(np.prod(L))
    u = hp.random_gauge_su3(V, 1)
    lat = hp.Lattice(L)
    g = hp.Gauge(lat).upload(u)
    gs = hp.Gauge(lat, *sloppy).upload(u)
    out = {}
    for label, R, sl in (("fp64_1rhs", 1, None), (f"mixed_{PREC_NAMES[sloppy[0]]}_{nrhs}rhs", nrhs, gs)):
        b = hp.Field(lat, 0, R)
        x = hp.Field(lat, 0, R)
        for r in range(R):
            b.upload(hp.random_spinor(V, 100 + r), r)
        shape = None
        if tune:
            gi = sl if sl is not None else g
            fsl = hp.Field(lat, gi.prec, R)
            best = (C.c_int * 7)()
            bms = C.c_double()
            capi.check(capi.lib().plqcd_hip_autotune_solver(gi.h, fsl.h, 5, best, C.byref(bms)), "autotune_solver")
            fsl.close()
            shape = {"shape": list(best), "ms_per_iteration_pair": round(bms.value, 4)}
        sv.cg_solve(g, b, x, kappa, tol, sloppy=sl)  # graph capture and first-touch costs
        ts = []
        for _ in range(reps):
            lat.sync()
            t0 = time.perf_counter()
            st = sv.cg_solve(g, b, x, kappa, tol, sloppy=sl)
            ts.append(time.perf_counter() - t0)
        t = float(np.median(ts))
        out[label] = {"rhs": R, "seconds": round(t, 4), "seconds_min_max": [round(min(ts), 4), round(max(ts), 4)],
                      "solves_per_s": round(R / t, 3),
                      "iterations": max(s_["iterations"] for s_ in st),
                      "reliable_updates": max(s_["reliable_updates"] for s_ in st),
                      "max_true_residual": max(s_["true_residual"] for s_ in st),
                      "all_converged": all(s_["converged"] for s_ in st)}
        if shape is not None:
            out[label]["tuned"] = shape
        b.close(); x.close()
    keys = list(out)
    out["ratio_solves_per_s"] = round(out[keys[1]]["solves_per_s"] / out[keys[0]]["solves_per_s"], 3)
    out["lattice"] = list(L)
    out["kappa"], out["tol"] = kappa, tol
    if tune:
        capi.lib().plqcd_hip_autotune_clear()
    gs.close(); g.close(); lat.close()
    return out


def emit(rec):
    print(json.dumps(rec), flush=True)


def time_hop(hp, g, fi, fo, n=20, warmup=10):
    """us per application of H over n launches after `warmup` untimed ones
    (fresh fields and a GPU coming out of other work need a few)."""
    lat = g.lat
    for _ in range(warmup):
        hp.hop(g, fi, fo, hp.PARITY_ALL, hp.HOP_WILSON)
    lat.timer_start()
    for _ in range(n):
        hp.hop(g, fi, fo, hp.PARITY_ALL, hp.HOP_WILSON)
    return lat.timer_stop() / n


def multi_rhs_run(L, formats=((0, 18), (2, 12)), rhs=(1, 4, 12, 24)):
    """H on R right-hand sides sharing one gauge field: us per
    application and per RHS, default launch shapes, 100 timed applications
    after 10 untimed (20 read about 3 % high while the clocks settle).
    Returns {format: {R: {...}}}."""
    from plqcd_amd import hopping as hp
    V = int(np.prod(L))
    u = hp.random_gauge_su3(V, 1)
    lat = hp.Lattice(L)
    out = {}
    for prec, links in formats:
        g = hp.Gauge(lat, prec, links).upload(u)
        rec = {}
        for R in rhs:
            fi, fo = hp.Field(lat, prec, R), hp.Field(lat, prec, R)
            for r in range(R):
                fi.upload(hp.random_spinor(V, 10 + r), r)
            ms = time_hop(hp, g, fi, fo, n=100)
            rec[f"R{R}"] = {"us_per_application": round(ms * 1e3, 1), "us_per_rhs": round(ms * 1e3 / R, 1),
                            "gflops": round(FLOP_PER_SITE * V * R / ms / 1e6, 1)}
            fi.close(); fo.close()
        out[f"{PREC_NAMES[prec]}_l{links}"] = rec
        g.close()
    lat.close()
    return out


def ragged_run(formats=((1, 12), (2, 12), (0, 18))):
    """One launch of H over the ragged batch 16^3x32 + 24^3x48 + 32^3x64 with
    its own gauge field and 12 RHS each against three separate
    launches: medians of three alternating rounds of 20.  Returns {format: {...}}."""
    import ctypes as C
    from plqcd_amd import capi, hopping as hp
    lib = capi.lib()
    shapes = [(16, 16, 16, 32), (24, 24, 24, 48), (32, 32, 32, 64)]
    sources = [[hp.random_spinor(int(np.prod(Ls)), 50 + r) for r in range(12)] for Ls in shapes]
    gauges = [hp.random_gauge_su3(int(np.prod(Ls)), 5 + k) for k, Ls in enumerate(shapes)]
    out = {}
    for prec, links in formats:
        lats, gs_, fis, fos = [], [], [], []
        for k, Ls in enumerate(shapes):
            la = hp.Lattice(Ls)
            if lats:  # one stream for the whole batch
                la.set_stream(lib.plqcd_hip_lattice_stream(lats[0].h))
            lats.append(la)
            gs_.append(hp.Gauge(la, prec, links).upload(gauges[k]))
            fi, fo = hp.Field(la, prec, 12), hp.Field(la, prec, 12)
            for r in range(12):
                fi.upload(sources[k][r], r)
            fis.append(fi); fos.append(fo)
        n = len(shapes)
        garr = (C.c_void_p * n)(*[x_.h.value for x_ in gs_])
        iarr = (C.c_void_p * n)(*[x_.h.value for x_ in fis])
        oarr = (C.c_void_p * n)(*[x_.h.value for x_ in fos])
        for _ in range(3):
            capi.check(lib.plqcd_hip_hop_batch(n, garr, iarr, oarr, 1), "hop_batch")
        tb, ts = [], []
        lat = lats[0]
        for _ in range(3):
            lat.timer_start()
            for _ in range(20):
                capi.check(lib.plqcd_hip_hop_batch(n, garr, iarr, oarr, 1), "hop_batch")
            tb.append(lat.timer_stop() / 20)
            lat.timer_start()
            for _ in range(20):
                for k in range(n):
                    hp.hop(gs_[k], fis[k], fos[k], hp.PARITY_ALL, hp.HOP_WILSON)
            ts.append(lat.timer_stop() / 20)
        t_batch, t_sep = float(np.median(tb)), float(np.median(ts))
        sites = sum(int(np.prod(Ls)) for Ls in shapes) * 12
        out[f"{PREC_NAMES[prec]}_l{links}"] = {"us_one_launch": round(t_batch * 1e3, 1),
                                               "us_separate_launches": round(t_sep * 1e3, 1),
                                               "gflops_one_launch": round(FLOP_PER_SITE * sites / t_batch / 1e6, 1)}
        for k in reversed(range(n)):
            fis[k].close(); fos[k].close(); gs_[k].close(); lats[k].close()
    return out


def solver_suite():
    """configs[2] (48^3x96 CG: FP64 vs mixed), configs[3] (multi-RHS hop and
    CG at 32^3x64, ragged batch in one launch, autotuner).  One JSON line each."""
    import ctypes as C
    from plqcd_amd import capi, hopping as hp, solver as sv
    lib = capi.lib()
    # --- single-RHS CG at 48^3 x 96, kappa 0.124, 1e-10
    L = (48, 48, 48, 96)
    V = int(np.prod(L))
    u = hp.random_gauge_su3(V, 1)
    lat = hp.Lattice(L)
    g = hp.Gauge(lat).upload(u)
    b = hp.Field(lat).upload(hp.random_spinor(V, 2))
    x = hp.Field(lat)
    for label, sl in (("fp64", None), ("mixed_f32_l12", (1, 12)), ("mixed_f16_l12", (2, 12))):
        gs = hp.Gauge(lat, *sl).upload(u) if sl else None
        sv.cg_solve(g, b, x, 0.124, 1e-10, sloppy=gs)
        t0 = time.perf_counter()
        st = sv.cg_solve(g, b, x, 0.124, 1e-10, sloppy=gs)[0]
        t = time.perf_counter() - t0
        emit({"suite": "solver", "case": f"cg_48c96_{label}", "seconds": round(t, 4), "iterations": st["iterations"],
              "reliable_updates": st["reliable_updates"], "true_residual": st["true_residual"],
              "converged": st["converged"]})
        if gs:
            gs.close()
    x.close(); b.close(); g.close(); lat.close()
    del u
    # --- multi-RHS at 32^3 x 64
    L = (32, 32, 32, 64)
    V = int(np.prod(L))
    for fmt, rec in multi_rhs_run(L, ((0, 18), (1, 12), (2, 18), (2, 12))).items():
        for rk, r in rec.items():
            emit({"suite": "solver", "case": f"hop_32c64_{fmt}_{rk}", **r})
    for R in (1, 4, 12, 24):
        for sl in ((1, 12), (2, 12)):
            rec = solve_compare(L, nrhs=R, sloppy=sl)
            rec.update({"suite": "solver", "case": f"cg_32c64_mixed_{PREC_NAMES[sl[0]]}_R{R}_vs_fp64_R1"})
            emit(rec)
    # both solvers with their launch shapes autotuned first (FP16 12-real, 12 RHS)
    rec = solve_compare(L, nrhs=12, sloppy=(2, 12), tune=True)
    rec.update({"suite": "solver", "case": "cg_32c64_mixed_f16_R12_vs_fp64_R1_solver_autotuned"})
    emit(rec)
    # ragged batch: 16^3x32, 24^3x48, 32^3x64, each 12 RHS, one launch
    for fmt, r in ragged_run().items():
        emit({"suite": "solver", "case": f"ragged_batch_{fmt}_3lattices_12rhs", **r})
    u = hp.random_gauge_su3(V, 1)
    lat = hp.Lattice(L)
    # autotuner at 32^3 x 64
    for prec, links, R in ((0, 18, 1), (1, 12, 12), (2, 12, 12)):
        g = hp.Gauge(lat, prec, links).upload(u)
        fi, fo = hp.Field(lat, prec, R), hp.Field(lat, prec, R)
        for r in range(R):
            fi.upload(hp.random_spinor(V, 10 + r), r)
        t_def = time_hop(hp, g, fi, fo)
        best = (C.c_int * 7)()
        bms = C.c_double()
        capi.check(lib.plqcd_hip_autotune(g.h, fi.h, fo.h, 5, best, C.byref(bms)), "autotune")
        t_tuned = time_hop(hp, g, fi, fo)
        emit({"suite": "solver", "case": f"autotune_32c64_{PREC_NAMES[prec]}_l{links}_R{R}",
              "default_us": round(t_def * 1e3, 1), "tuned_us": round(t_tuned * 1e3, 1),
              "best": {"order": best[0], "group": best[1], "block": best[2], "tmarch": best[3],
                       "rb": best[4], "tile_y": best[5], "tile_z": best[6]}})
        fi.close(); fo.close(); g.close()
    lib.plqcd_hip_autotune_clear()
    lat.close()


def calib_suite():
    """HBM probes of known byte counts (copy, read, write and the 4:1 mix over
    1 GiB, the mix at H's own sizes, Infinity-Cache- and HBM-sized working
    sets), then two applications of H per storage format at 32^3 x 64 with the
    default launch shapes, M on 12 RHS (FP64) and H on 12 RHS (FP64 18-real,
    FP16 12-real, and FP16 under two other block orders): one run holds the
    counter calibration and the kernels."""
    import ctypes as C
    from plqcd_amd import capi, hopping as hp
    lib = capi.lib()
    for kind, name in ((0, "copy"), (1, "read"), (2, "write"), (3, "mix41")):
        gbs = C.c_double()
        capi.check(lib.plqcd_hip_stream_bandwidth(kind, 1 << 30, 10, C.byref(gbs)), "stream_bandwidth")
        emit({"suite": "calib", "case": f"stream_{name}_1GiB", "bytes_per_buffer": 1 << 30, "timed_launches": 10,
              "GBps": round(gbs.value, 1)})
    # the FP64 hopping matrix's own byte counts at 32^3 x 64: 4 x 402 MB read, 402 MB written
    gbs = C.c_double()
    capi.check(lib.plqcd_hip_stream_bandwidth(3, 2097152 * 192, 20, C.byref(gbs)), "stream_bandwidth")
    emit({"suite": "calib", "case": "stream_mix41_hop_f64_32x32x32x64", "bytes_per_buffer": 2097152 * 192,
          "timed_launches": 20, "GBps": round(gbs.value, 1)})
    # the same probes over working sets that fit the 256 MB Infinity Cache
    # (read 128 MiB, copy 2 x 128 MiB) and that do not (4 GiB)
    for kind, name, mib, reps in ((1, "read", 128, 100), (0, "copy", 128, 100), (1, "read", 4096, 10),
                                  (0, "copy", 4096, 10)):
        gbs = C.c_double()
        capi.check(lib.plqcd_hip_stream_bandwidth(kind, mib << 20, reps, C.byref(gbs)), "stream_bandwidth")
        emit({"suite": "calib", "case": f"stream_{name}_{mib}MiB", "bytes_per_buffer": mib << 20,
              "timed_launches": reps, "GBps": round(gbs.value, 1)})
    L = (32, 32, 32, 64)
    V = int(np.prod(L))
    u = hp.random_gauge_su3(V, 1)
    q = hp.random_spinor(V, 2)
    lat = hp.Lattice(L)
    for prec, links in ((0, 18), (1, 18), (1, 12), (2, 18), (2, 12)):
        g = hp.Gauge(lat, prec, links).upload(u)
        fi, fo = hp.Field(lat, prec).upload(q), hp.Field(lat, prec)
        for _ in range(2):
            hp.hop(g, fi, fo, hp.PARITY_ALL, hp.HOP_WILSON)
        lat.sync()
        emit({"suite": "calib", "case": f"hop_{PREC_NAMES[prec]}_l{links}_32x32x32x64", "applications": 2,
              "compulsory_bytes": BYTES_PER_SITE[(prec, links)] * V})
        fi.close(); fo.close(); g.close()
    # FP64 M = 1 - kappa H on 12 RHS (the reliable update's operator), twice
    from plqcd_amd import solver as sv
    g = hp.Gauge(lat, 0, 18).upload(u)
    fi, fo = hp.Field(lat, 0, 12), hp.Field(lat, 0, 12)
    for r in range(12):
        fi.upload(hp.random_spinor(V, 10 + r), r)
    for _ in range(2):
        sv.apply_m(g, fi, fo, 0.124)
    lat.sync()
    emit({"suite": "calib", "case": "apply_m_f64_l18_32x32x32x64_R12", "applications": 2,
          "compulsory_bytes": (12 * (192 + 192) + 576) * V})
    fi.close(); fo.close(); g.close()
    # 12 RHS: H with the multi-RHS default shapes (FP64 18-real, FP16 12-real),
    # then the FP16 H under the other block orders (plain, contiguous per XCD)
    from plqcd_amd import capi as _capi
    for prec, links, tune, tag in ((0, 18, None, "default"), (2, 12, None, "default"),
                                   (2, 12, (0, 1, 256, 8, 1, 0, 0), "order0_T8"),
                                   (2, 12, (1, 1, 256, 8, 1, 0, 0), "order1_T8")):
        if tune:
            _capi.check(lib.plqcd_hip_hop_tuning(prec, *tune), "hop_tuning")
        g = hp.Gauge(lat, prec, links).upload(u)
        fi, fo = hp.Field(lat, prec, 12), hp.Field(lat, prec, 12)
        for r in range(12):
            fi.upload(hp.random_spinor(V, 10 + r), r)
        for _ in range(2):
            hp.hop(g, fi, fo, hp.PARITY_ALL, hp.HOP_WILSON)
        lat.sync()
        link_b = BYTES_PER_SITE[(prec, links)] - (384 if prec == 0 else 104)
        emit({"suite": "calib", "case": f"hop_{PREC_NAMES[prec]}_l{links}_32x32x32x64_R12_{tag}", "applications": 2,
              "compulsory_bytes": (12 * (BYTES_PER_SITE[(prec, links)] - link_b) + link_b) * V})
        fi.close(); fo.close(); g.close()
        lib.plqcd_hip_hop_tuning_reset(prec)
    lat.close()


def serve_suite(duration: float, rate: float, seed: int = 7, pinned: bool = False):
    """Poisson request stream for `duration` seconds; volumes from
    {16^4, 24^3x48, 32^3x64, 48^3x96}, 1..12 RHS, tol in {1e-8,1e-10,1e-12},
    kappa in [0.120, 0.1245], two gauge fields per volume.  Returns p50/p99
    latency and solves/s beside one-at-a-time plain FP64 solves.  With `pinned`
    the sources live in page-locked host memory (moved by DMA alone)."""
    from plqcd_amd import hopping as hp, solver as sv
    from plqcd_amd.serve import Server
    rng = np.random.default_rng(seed)
    vols = [(16, 16, 16, 16), (24, 24, 24, 48), (32, 32, 32, 64), (48, 48, 48, 96)]
    pool = {L: np.stack([hp.random_spinor(int(np.prod(L)), 1000 + k) for k in range(12)]) for L in vols}
    if pinned:
        from plqcd_amd.serve import pinned_empty
        for L in vols:
            pp = pinned_empty(pool[L].shape)
            pp[...] = pool[L]
            pool[L] = pp
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=int(rate * duration * 2) + 10))
    arrivals = arrivals[arrivals < duration]
    reqs = []
    for t in arrivals:
        L = vols[rng.integers(len(vols))]
        reqs.append({"t": float(t), "L": L, "nrhs": int(rng.integers(1, 13)), "tol": float(rng.choice([1e-8, 1e-10, 1e-12])),
                     "kappa": float(rng.uniform(0.120, 0.1245)), "seed": 11 + 100 * vols.index(L) + int(rng.integers(2))})
    lat_res = {}
    with Server(sloppy_prec=2, sloppy_links=12, slots=12) as srv:
        t0 = time.perf_counter()
        submitted = {}
        k = 0
        done = {}
        while k < len(reqs) or len(done) < len(reqs):
            now = time.perf_counter() - t0
            while k < len(reqs) and reqs[k]["t"] <= now:
                r = reqs[k]
                rid = srv.submit(r["L"], r["seed"], pool[r["L"]][: r["nrhs"]], r["kappa"], r["tol"])
                submitted[rid] = r
                k += 1
            for d in srv.poll():
                done[d["id"]] = d
                d.pop("x", None)
            time.sleep(0.002)
        wall = time.perf_counter() - t0
        info = srv.info()
    lat_s = np.array([d["end_time"] - d["submit_time"] for d in done.values()])
    nrhs = sum(submitted[i]["nrhs"] for i in done)
    ok = all(d["status"] == 0 and d["converged"] for d in done.values())
    # one-at-a-time plain FP64 baseline, per volume: seconds per single-RHS
    # solve with the fields resident, and with the host transfers of its
    # source and solution (what the server's requests include)
    per_rhs, per_rhs_io = {}, {}
    for L in vols:
        V = int(np.prod(L))
        la = hp.Lattice(L)
        g = hp.Gauge(la).upload(hp.random_gauge_su3(V, 11 + 100 * vols.index(L)))
        b = hp.Field(la).upload(pool[L][0])
        x = hp.Field(la)
        sv.cg_solve(g, b, x, 0.1225, 1e-10)
        t1 = time.perf_counter()
        sv.cg_solve(g, b, x, 0.1225, 1e-10)
        per_rhs[L] = time.perf_counter() - t1
        xh = np.empty_like(pool[L][0])
        t1 = time.perf_counter()
        b.upload(pool[L][1])
        sv.cg_solve(g, b, x, 0.1225, 1e-10)
        xh[...] = x.download()
        per_rhs_io[L] = time.perf_counter() - t1
        del xh
        x.close(); b.close(); g.close(); la.close()
    base_seconds = sum(per_rhs[submitted[i]["L"]] * submitted[i]["nrhs"] for i in done)
    base_io_seconds = sum(per_rhs_io[submitted[i]["L"]] * submitted[i]["nrhs"] for i in done)
    return {"suite": "serve", "case": "poisson_mixed_volumes", "duration_s": duration, "rate_per_s": rate,
          "pinned_sources": pinned,
          "requests": len(done), "rhs_solved": nrhs, "all_converged": ok, "wall_s": round(wall, 2),
          "p50_latency_s": round(float(np.percentile(lat_s, 50)), 4),
          "p99_latency_s": round(float(np.percentile(lat_s, 99)), 4),
          "solves_per_s": round(nrhs / wall, 3),
          "fp64_one_at_a_time_solves_per_s": round(nrhs / base_seconds, 3),
          "fp64_one_at_a_time_with_host_io_solves_per_s": round(nrhs / base_io_seconds, 3),
          "fp64_seconds_per_rhs": {"x".join(map(str, L)): round(v, 4) for L, v in per_rhs.items()},
          "fp64_seconds_per_rhs_with_host_io": {"x".join(map(str, L)): round(v, 4) for L, v in per_rhs_io.items()},
          "hbm_budget_gb": round(info["budget"] / 2**30, 1), "failed": info["failed"],
          "engines_created": info["engines_created"], "evictions": info["evictions"],
          "host_transfer_gb": round(info["transfer_bytes"] / 1e9, 1),
          "host_transfer_thread_seconds": round(info["transfer_seconds"], 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--prec", choices=["f64", "f32", "f16"], default="f64")
    ap.add_argument("--links", type=int, default=None)
    ap.add_argument("--lattice", default="32,32,32,64")
    ap.add_argument("--mode", choices=["wilson", "compat"], default="wilson")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solver", action="store_true")
    ap.add_argument("--no-formats", action="store_true", help="skip the FP32/FP16 storage-format timings")
    ap.add_argument("--serve-seconds", type=float, default=10.0,
                    help="hop line: seconds of the short Poisson serving run (0: none)")
    ap.add_argument("--suite", choices=["hop", "solver", "serve", "calib", "cg"], default="hop",
                    help="hop: the headline line (default); solver/serve: the extended studies, one JSON line "
                         "each; cg: the bench line's solver comparison alone")
    ap.add_argument("--duration", type=float, default=60.0, help="serve suite: seconds of Poisson arrivals")
    ap.add_argument("--rate", type=float, default=2.0, help="serve suite: requests per second")
    ap.add_argument("--pinned", action="store_true", help="serve suite: sources in page-locked host memory")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the hopping matrix here is a "
              "single-GPU kernel; multi-GPU sharding is a design, not a build", file=sys.stderr)
        sys.exit(2)

    if args.suite == "solver":
        solver_suite()
        return
    if args.suite == "serve":
        emit(serve_suite(args.duration, args.rate, pinned=args.pinned))
        return
    if args.suite == "calib":
        calib_suite()
        return
    if args.suite == "cg":
        rec = solve_compare(tuple(int(v) for v in args.lattice.split(",")))
        rec.update({"suite": "cg"})
        emit(rec)
        return

    L = tuple(int(x) for x in args.lattice.split(","))
    V = int(np.prod(L))
    prec = {"f64": 0, "f32": 1, "f16": 2}[args.prec]
    links = args.links or (18 if prec < 2 else 12)
    mode = 1 if args.mode == "wilson" else 0

    headline = prec == 0 and links == 18 and mode == 1 and L == (32, 32, 32, 64)
    # The CPU baseline first (the GPU idles meanwhile), then the GPU figures
    # with the headline's inputs already resident, and the headline last, right
    # after the storage formats (the last of which is this same kernel, 200
    # launches): after about 2 s without work the GPU needs some 20 ms of load
    # to return to its steady clocks (20 applications after an idle gap take
    # 448 us each, 413 us once warm), which a short timed region would
    # otherwise measure instead of the kernel.
    cpu = None if args.no_cpu else cpu_baseline(L, args.cpu_seconds)
    case = HopCase(L, prec, links, mode)
    compat = compat_run(L) if headline and not args.no_formats else None
    schur = schur_run(L) if headline and not args.no_formats else None
    dropin = dropin_run(L) if headline and not args.no_formats else None
    solver = None if args.no_solver else solve_compare(L)
    cg48 = cg48_compare() if headline and not args.no_solver else None
    multi = multi_rhs_run(L) if headline and not args.no_solver else None
    ragged = ragged_run(((1, 12), (0, 18))) if headline and not args.no_solver else None
    serve = serve_suite(args.serve_seconds, 2.0) if headline and args.serve_seconds > 0 else None
    ceiling = mix_ceiling(L) if headline else None
    formats = None if (args.no_formats or prec != 0) else formats_run(
        L, mode, headline_case=case if (prec, links) == (0, 18) else None)
    ms = case.run(args.steps, args.warmup)
    case.close()
    t_launch = ms / 1e3 / args.steps
    gflops = FLOP_PER_SITE * V / t_launch / 1e9
    bytes_launch = BYTES_PER_SITE[(prec, links)] * V
    achieved = bytes_launch / t_launch / 1e9
    tag = f"hop_{PREC_NAMES[prec]}_l{links}_{'x'.join(map(str, L))}"
    traffic = read_traffic(tag)
    traffic_src = (f"profiles/{tag}_pmc.txt: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of this kernel at this "
                   f"shape, 2 x FETCH + WRITE (copy-kernel calibration), committed, not measured in this run"
                   if traffic else None)
    rec = {
        "metric": METRIC,
        "value": round(gflops, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": PREC_NAMES[prec],
        "data": "synthetic (seeded random SU(3) links, Gaussian spinor)",
        "config": {"workload": f"Wilson hopping matrix H (full lattice, {args.mode} mode), "
                               f"{'x'.join(map(str, L))}, {PREC_NAMES[prec]} spinors, {links}-real links",
                   "lattice": list(L), "sites": V, "flop_per_site": FLOP_PER_SITE,
                   "bytes_per_site": BYTES_PER_SITE[(prec, links)], "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": traffic_src,
                     "measured_ceiling": None if ceiling is None else {
                         "probe": "stream kernel with H's byte counts: 4 x V x 192 B read, V x 192 B written",
                         "GBps": ceiling, "frac_of_ceiling": round(achieved / ceiling, 4)}},
        "cpu_baseline": cpu,
        "formats": formats,
        "compat": compat,
        "schur": schur,
        "dropin": dropin,
        "solver": solver,
        "cg_48c96": cg48,
        "multi_rhs": multi,
        "ragged_batch": ragged,
        "serve": serve,
    }
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
