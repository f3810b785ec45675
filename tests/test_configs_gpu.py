"""GPU tests at the workload shapes of the benchmarked configurations, through
the C ABI, against the oracle:

  * 24 right-hand sides at 32^3 x 64 (FP64, and FP16 with 12-real links): every
    RHS bitwise equal to its own one-RHS launch, two of them against the FP64
    oracle (FP64 1e-13, FP16 4e-3 relative max-norm per site);
  * FP16 storage (12- and 18-real links) and FP64 with 12-real links at
    32^3 x 64 against the FP64 oracle at every site;
  * the ragged batch 16^3 x 32 + 24^3 x 48 + 32^3 x 64 with 12 RHS each, one
    launch, bitwise equal to separate launches and against the oracle;
  * CG on M^+M at 48^3 x 96, kappa 0.124, plain FP64 and mixed FP16/12-real
    sloppy: the FP64 true residual recomputed by the oracle <= 1.05e-10;
  * the autotuner on 8^3 x 16: tuned output bitwise equal to the default
    shape's, the tuned shape cached and reused, the solver's launch shapes
    untouched, a failed tuning restoring the previous entry.

The shapes follow the reference's benchmark input (tests/test.input:1-19,
32^3 x 64) and its field allocation (tests/test_hopping.c:96-115)."""
import ctypes as C

import numpy as np
import pytest

from conftest import site_rel_err

pytestmark = pytest.mark.gpu

TOL64, TOL32, TOL16 = 1e-13, 2e-6, 4e-3
L32 = (32, 32, 32, 64)
KAPPA = 0.124


def _close(*objs):
    for o in objs:
        o.close()


# ----------------------------------------------------------- 24 RHS at 32^3 x 64

@pytest.mark.parametrize("prec,links,tol", [(0, 18, TOL64), (2, 12, TOL16)])
def test_24_rhs_at_32c64(hp, orc, prec, links, tol):
    R = 24
    V = int(np.prod(L32))
    u = hp.random_gauge_su3(V, 1)
    lat = hp.Lattice(L32)
    g = hp.Gauge(lat, prec, links).upload(u)
    fi, fo = hp.Field(lat, prec, R), hp.Field(lat, prec, R)
    f1i, f1o = hp.Field(lat, prec), hp.Field(lat, prec)
    seed = lambda r: 500 + r  # noqa: E731
    try:
        for r in range(R):
            fi.upload(hp.random_spinor(V, seed(r)), r)
        for mode, omode, checked in ((hp.HOP_WILSON, orc.MODE_WILSON, range(R)),
                                     (hp.HOP_COMPAT, orc.MODE_COMPAT, (0, R - 1))):
            hp.hop(g, fi, fo, hp.PARITY_ALL, mode)
            for r in checked:
                q = hp.random_spinor(V, seed(r))
                f1i.upload(q)
                hp.hop(g, f1i, f1o, hp.PARITY_ALL, mode)
                multi = fo.download(r)
                assert np.array_equal(multi, f1o.download()), (mode, r)
                if r in (0, R - 1):
                    err = site_rel_err(multi, orc.hopping_full(q, u, L32, omode))
                    print(f"24 RHS prec {prec}/{links} mode {mode} rhs {r}: {err:.3e}")
                    assert err < tol, (mode, r)
    finally:
        _close(f1o, f1i, fo, fi, g, lat)


# ------------------------------------------- reduced-precision storage at 32^3 x 64

@pytest.mark.parametrize("prec,links,tol", [(2, 12, TOL16), (2, 18, TOL16), (0, 12, 1e-12)])
def test_storage_formats_at_32c64_vs_f64_oracle(hp, orc, prec, links, tol):
    V = int(np.prod(L32))
    u = hp.random_gauge_su3(V, 1)
    q = hp.random_spinor(V, 2)
    ref = orc.hopping_full(q, u, L32, orc.MODE_WILSON)
    lat = hp.Lattice(L32)
    g = hp.Gauge(lat, prec, links).upload(u)
    fi, fo = hp.Field(lat, prec).upload(q), hp.Field(lat, prec)
    try:
        hp.hop(g, fi, fo, hp.PARITY_ALL, hp.HOP_WILSON)
        err = site_rel_err(fo.download(), ref)
        print(f"32^3x64 prec {prec} {links}-real links: {err:.3e}")
        assert err < tol
    finally:
        _close(fo, fi, g, lat)


# -------------------------------------------------------------- ragged batch

RAGGED_SHAPES = [(16, 16, 16, 32), (24, 24, 24, 48), (32, 32, 32, 64)]


@pytest.mark.parametrize("prec,links,tol", [(0, 18, TOL64), (1, 12, TOL32), (2, 12, TOL16)])
def test_ragged_batch_benchmark_shapes(hp, orc, prec, links, tol):
    """16^3x32 + 24^3x48 + 32^3x64, each with its own gauge field and 12 RHS,
    in one launch: every item and RHS bitwise equal to the item's own launch;
    the first and last RHS of each item against the oracle."""
    from plqcd_amd import capi
    R = 12
    lib = capi.lib()
    lats, gs, fis, fos, fsep, meta = [], [], [], [], [], []
    try:
        for k, L in enumerate(RAGGED_SHAPES):
            V = int(np.prod(L))
            lat = hp.Lattice(L)
            if lats:
                lat.set_stream(lib.plqcd_hip_lattice_stream(lats[0].h))
            lats.append(lat)
            u = hp.random_gauge_su3(V, 700 + k)
            gs.append(hp.Gauge(lat, prec, links).upload(u))
            fi, fo, fs = hp.Field(lat, prec, R), hp.Field(lat, prec, R), hp.Field(lat, prec, R)
            fis.append(fi); fos.append(fo); fsep.append(fs)
            for r in range(R):
                fi.upload(hp.random_spinor(V, 800 + 31 * k + r), r)
            meta.append((L, V, u))
        hp.hop_batch(gs, fis, fos, hp.HOP_WILSON)
        for k in range(len(RAGGED_SHAPES)):
            hp.hop(gs[k], fis[k], fsep[k], hp.PARITY_ALL, hp.HOP_WILSON)
        lats[0].sync()
        for k, (L, V, u) in enumerate(meta):
            for r in range(R):
                out = fos[k].download(r)
                assert np.array_equal(out, fsep[k].download(r)), (L, r)
                if r in (0, R - 1):
                    q = hp.random_spinor(V, 800 + 31 * k + r)
                    err = site_rel_err(out, orc.hopping_full(q, u, L, orc.MODE_WILSON))
                    assert err < tol, (L, r, err)
    finally:
        for objs in (fis, fos, fsep, gs):
            for o in objs:
                o.close()
        for lat in reversed(lats):
            lat.close()


# ------------------------------------------------------------- CG at 48^3 x 96

def test_cg_48c96_fp64_and_mixed_f16(hp):
    """48^3 x 96, kappa 0.124, true relative residual 1e-10, plain
    FP64 and FP16 storage with 12-real links plus FP64 reliable updates; the
    residual of each solution recomputed on the CPU by the oracle."""
    from oracle import solver as osv
    from plqcd_amd import solver as sv
    L = (48, 48, 48, 96)
    V = int(np.prod(L))
    u = hp.random_gauge_su3(V, 41)
    b = hp.random_spinor(V, 42)
    lat = hp.Lattice(L)
    g = hp.Gauge(lat).upload(u)
    gs = hp.Gauge(lat, 2, 12).upload(u)
    bf, x = hp.Field(lat).upload(b), hp.Field(lat)
    try:
        iters = {}
        for name, sloppy in (("fp64", None), ("mixed_f16_l12", gs)):
            st = sv.cg_solve(g, bf, x, KAPPA, 1e-10, sloppy=sloppy)[0]
            res = osv.true_residual(b, x.download(), u, L, KAPPA)
            print(f"48^3x96 {name}: {st['iterations']} iterations, {st['reliable_updates']} reliable updates, "
                  f"{st['seconds']:.3f} s, GPU residual {st['true_residual']:.3e}, oracle residual {res:.3e}")
            assert st["converged"] and st["true_residual"] <= 1e-10
            assert res <= 1.05e-10, (name, res)
            iters[name] = st["iterations"]
        assert iters["mixed_f16_l12"] <= 1.2 * iters["fp64"]
    finally:
        _close(x, bf, gs, g, lat)


def test_field_rhs_limit_at_48c96(hp):
    """A field of 48^3 x 96 holds at most 50 RHS (32-bit offsets); 51 is
    rejected before any allocation, and so is a CG on such a field."""
    from plqcd_amd import capi
    L = (48, 48, 48, 96)
    lib = capi.lib()
    assert lib.plqcd_hip_max_rhs(capi.i4(*L)) == 50
    lat = hp.Lattice(L)
    try:
        h = C.c_void_p()
        assert lib.plqcd_hip_field_create(lat.h, 2, 51, C.byref(h)) == 2
        assert "32-bit offsets" in capi.last_error() and not h.value
    finally:
        lat.close()


# ---------------------------------------------------------------- autotuner

def test_autotune_bitwise_cached_and_solver_untouched(hp):
    from plqcd_amd import capi
    from plqcd_amd import solver as sv
    lib = capi.lib()
    L = (8, 8, 8, 16)
    V = int(np.prod(L))
    u = hp.random_gauge_su3(V, 51)
    q = hp.random_spinor(V, 52)
    lat = hp.Lattice(L)
    g = hp.Gauge(lat, 1, 12).upload(u)
    g64 = hp.Gauge(lat).upload(u)
    fi, fo = hp.Field(lat, 1).upload(q), hp.Field(lat, 1)
    b, x = hp.Field(lat).upload(q), hp.Field(lat)
    shape = (C.c_int * 7)()
    best = (C.c_int * 7)()
    ms = C.c_double()
    try:
        assert lib.plqcd_hip_autotune_clear() == 0
        assert lib.plqcd_hip_autotune_lookup(g.h, fi.h, 2, shape) == -1  # no such operation
        assert lib.plqcd_hip_autotune_lookup(g.h, fi.h, 0, shape) == 0
        hp.hop(g, fi, fo, hp.PARITY_ALL, hp.HOP_WILSON)
        default_out = fo.download()
        st0 = sv.cg_solve(g64, b, x, KAPPA, 1e-10, sloppy=g)[0]
        x0 = x.download()

        assert lib.plqcd_hip_autotune(g.h, fi.h, fo.h, 3, best, C.byref(ms)) == 0, capi.last_error()
        assert ms.value > 0
        assert lib.plqcd_hip_autotune_lookup(g.h, fi.h, 0, shape) == 1
        assert list(shape) == list(best)
        # the solver's M / M^+ launches keep their built-in shapes
        assert lib.plqcd_hip_autotune_lookup(g.h, fi.h, 1, shape) == 0
        # later launches of this shape use the tuned one; the output is unchanged
        fo.upload(np.zeros_like(q))
        hp.hop(g, fi, fo, hp.PARITY_ALL, hp.HOP_WILSON)
        assert np.array_equal(fo.download(), default_out)
        st1 = sv.cg_solve(g64, b, x, KAPPA, 1e-10, sloppy=g)[0]
        assert st1["iterations"] == st0["iterations"] and st1["true_residual"] == st0["true_residual"]
        assert np.array_equal(x.download(), x0)

        # a tuning that fails (output of another precision) restores the entry
        f64o = hp.Field(lat)
        try:
            assert lib.plqcd_hip_autotune(g.h, fi.h, f64o.h, 1, None, None) != 0
        finally:
            f64o.close()
        assert lib.plqcd_hip_autotune_lookup(g.h, fi.h, 0, shape) == 1
        assert list(shape) == list(best)
        assert lib.plqcd_hip_autotune_clear() == 0
        assert lib.plqcd_hip_autotune_lookup(g.h, fi.h, 0, shape) == 0
    finally:
        lib.plqcd_hip_autotune_clear()
        _close(x, b, fo, fi, g64, g, lat)


def test_autotune_solver_bitwise_and_cached_apart(hp):
    """The solver tuner on 4 RHS with FP16 / 18-real sloppy fields: the tuned
    shape is cached under the solver operation only, and a mixed-precision CG
    afterwards takes the same iterations to the same bits."""
    from plqcd_amd import capi
    from plqcd_amd import solver as sv
    lib = capi.lib()
    L = (8, 8, 8, 16)
    R = 4
    V = int(np.prod(L))
    u = hp.random_gauge_su3(V, 61)
    lat = hp.Lattice(L)
    g64 = hp.Gauge(lat).upload(u)
    g16 = hp.Gauge(lat, 2, 18).upload(u)
    b, x = hp.Field(lat, 0, R), hp.Field(lat, 0, R)
    fs = hp.Field(lat, 2, R)
    shape = (C.c_int * 7)()
    best = (C.c_int * 7)()
    ms = C.c_double()
    try:
        for r in range(R):
            b.upload(hp.random_spinor(V, 62 + r), r)
        assert lib.plqcd_hip_autotune_clear() == 0
        kap = [KAPPA, 0.12, 0.11, 0.1]
        st0 = sv.cg_solve(g64, b, x, kap, 1e-10, sloppy=g16)
        x0 = [x.download(r) for r in range(R)]

        assert lib.plqcd_hip_autotune_solver(g16.h, fs.h, 2, best, C.byref(ms)) == 0, capi.last_error()
        assert ms.value > 0
        assert lib.plqcd_hip_autotune_lookup(g16.h, fs.h, 1, shape) == 1
        assert list(shape) == list(best)
        assert lib.plqcd_hip_autotune_lookup(g16.h, fs.h, 0, shape) == 0  # plain H untouched

        st1 = sv.cg_solve(g64, b, x, kap, 1e-10, sloppy=g16)
        for s0, s1 in zip(st0, st1):
            assert s1["converged"] and s1["iterations"] == s0["iterations"]
            assert s1["true_residual"] == s0["true_residual"]
        for r in range(R):
            assert np.array_equal(x.download(r), x0[r])

        # bad arguments (a field of another precision than the gauge) leave the entry
        assert lib.plqcd_hip_autotune_solver(g16.h, b.h, 1, None, None) != 0
        assert lib.plqcd_hip_autotune_lookup(g16.h, fs.h, 1, shape) == 1
        assert list(shape) == list(best)
    finally:
        lib.plqcd_hip_autotune_clear()
        _close(fs, x, b, g16, g64, lat)
