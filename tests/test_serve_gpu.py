"""Solve server on the GPU: continuous batching across volumes, admission by
HBM, and each request's result the same alone and in the mix (bitwise, as
each slot's iterates depend only on its own data; the documented tolerance is
1e-12 relative in case a kernel variant changes between runs).  Residuals are
recomputed on the CPU by the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def source(hp, L, nrhs, seed):
    V = int(np.prod(L))
    return np.stack([hp.random_spinor(V, seed + k) for k in range(nrhs)])


def test_alone_equals_in_mix(hp):
    from oracle import solver as osv
    from plqcd_amd.serve import Server
    reqs = [((8, 8, 8, 8), 11, 2, 0.124, 1e-10), ((4, 4, 8, 8), 12, 3, 0.12, 1e-8),
            ((8, 4, 4, 8), 13, 1, 0.1245, 1e-12), ((8, 8, 8, 8), 11, 1, 0.122, 1e-10)]
    srcs = [source(hp, L, n, 100 * i) for i, (L, _, n, _, _) in enumerate(reqs)]
    alone = []
    for (L, seed, n, k, tol), b in zip(reqs, srcs):
        with Server(slots=4) as s:
            s.submit(L, seed, b, k, tol)
            res = s.run_all()
        alone.append(next(iter(res.values())))
    with Server(slots=4) as s:
        ids = [s.submit(L, seed, b, k, tol) for (L, seed, n, k, tol), b in zip(reqs, srcs)]
        mixed = s.run_all()
        info = s.info()
    assert info["failed"] == 0
    for i, ((L, seed, n, k, tol), b) in enumerate(zip(reqs, srcs)):
        a, m = alone[i], mixed[ids[i]]
        assert a["status"] == 0 and m["status"] == 0 and a["converged"] and m["converged"]
        assert a["iterations"] == m["iterations"]
        assert np.abs(a["x"] - m["x"]).max() <= 1e-12 * np.abs(a["x"]).max()
        assert np.array_equal(a["x"], m["x"])
        V = int(np.prod(L))
        u = hp.random_gauge_su3(V, seed)
        for r in range(n):
            assert osv.true_residual(b[r], m["x"][r], u, L, k) <= tol * 1.05


def test_admission_by_budget_and_reuse(hp):
    from plqcd_amd.serve import Server
    L1, L2 = (8, 8, 8, 8), (8, 8, 8, 16)
    # a budget that holds about one (gauge, engine) pair of the larger lattice
    with Server(slots=2, hbm_budget=64 << 20) as s:
        ids = [s.submit(L1 if i % 2 else L2, 20 + i % 3, source(hp, L1 if i % 2 else L2, 1, i), 0.12, 1e-9)
               for i in range(6)]
        res = s.run_all()
        info = s.info()
    assert info["used"] <= info["budget"]
    assert all(res[i]["status"] == 0 and res[i]["converged"] for i in ids)
    with Server(slots=2, hbm_budget=8 << 20) as s:
        rid = s.submit((16, 16, 16, 16), 1, source(hp, (16, 16, 16, 16), 1, 0), 0.12, 1e-9)
        res = s.run_all()
    assert res[rid]["status"] == -2


def test_large_rhs_moves_in_parts(hp):
    """A right-hand side of more than 256 MiB (32^3x96: 604 MB) is moved by
    several transfer threads at once, each a part of it; the solution is the
    same as with one thread moving it whole."""
    from plqcd_amd.serve import Server
    L = (32, 32, 32, 96)
    b = source(hp, L, 2, 7)
    out = []
    for threads in (1, 8):
        with Server(slots=2, transfer_threads=threads) as s:
            rid = s.submit(L, 3, b, 0.12, 1e-6)
            res = s.run_all()[rid]
            info = s.info()
        assert info["failed"] == 0 and res["status"] == 0 and res["converged"]
        assert info["transfer_bytes"] == 2 * b.nbytes
        out.append(res["x"])
    assert np.array_equal(out[0], out[1])
    assert np.isfinite(out[1]).all() and np.abs(out[1]).max() > 0


def test_pinned_buffers_move_by_dma(hp):
    """Sources and solutions in page-locked host memory (pinned_empty) take
    the DMA-only path; the solution equals the one through pageable memory."""
    from plqcd_amd.serve import Server, pinned_empty
    L = (8, 8, 8, 16)
    b = source(hp, L, 3, 21)
    bp = pinned_empty(b.shape)
    bp[...] = b
    xp = pinned_empty(b.shape)
    with Server(slots=4) as s:
        r1 = s.submit(L, 5, b, 0.123, 1e-9)
        r2 = s.submit(L, 5, bp, 0.123, 1e-9, out=xp)
        res = s.run_all()
    assert res[r1]["status"] == 0 and res[r2]["status"] == 0
    assert res[r2]["x"] is xp
    assert np.array_equal(res[r1]["x"], xp)
    with pytest.raises(ValueError):
        with Server(slots=1) as s:
            s.submit(L, 5, b, 0.123, 1e-9, out=np.empty((1,) + b.shape[1:], np.complex128))


def test_server_create_rejects_bad_config(hp):
    """Creation errors come back as error codes: a device that does not exist
    (hipSetDevice fails after the server object exists) and bad slot or
    thread counts; a good server is created afterwards."""
    from plqcd_amd import capi
    from plqcd_amd.serve import Server
    for kw in ({"device": 1000}, {"slots": 0}, {"slots": 33}, {"transfer_threads": 65}):
        with pytest.raises(capi.PlqcdError):
            Server(**kw)
    with Server(slots=2) as s:
        assert s is not None
