"""plqcd_hip_apply_m_noalloc (the per-RHS kappa in a static device array, no
allocation) against plqcd_hip_apply_m through the C ABI: bitwise equal
results on two calls in a row, the second on a smaller lattice with fewer
right-hand sides, M and M^+; RHS counts above the scratch are rejected."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fn():
    from plqcd_amd import capi
    return capi.lib().plqcd_hip_apply_m_noalloc


def _case(hp, L, R, kaps, dagger, seed):
    from plqcd_amd import solver as sv
    V = int(np.prod(L))
    u = hp.random_gauge_su3(V, seed)
    lat = hp.Lattice(L)
    g = hp.Gauge(lat).upload(u)
    fi, fa, fb = hp.Field(lat, nrhs=R), hp.Field(lat, nrhs=R), hp.Field(lat, nrhs=R)
    try:
        for r in range(R):
            fi.upload(hp.random_spinor(V, seed + 1 + r), r)
        sv.apply_m(g, fi, fa, kaps, dagger)
        k = np.asarray(kaps, dtype=np.float64)
        assert _fn()(g.h, fi.h, fb.h, k.ctypes.data_as(C.POINTER(C.c_double)), int(dagger)) == 0
        lat.sync()
        for r in range(R):
            assert np.array_equal(fb.download(r), fa.download(r)), (L, R, dagger, r)
    finally:
        for o in (fb, fa, fi, g, lat):
            o.close()


def test_noalloc_matches_apply_m_two_calls():
    from plqcd_amd import hopping as hp
    hp.lib()
    _case(hp, (8, 8, 8, 16), 6, [0.12, 0.121, 0.122, 0.123, 0.124, 0.1245], False, 11)
    _case(hp, (4, 4, 4, 8), 2, [0.1245, 0.12], True, 21)
    _case(hp, (4, 4, 4, 8), 1, [0.124], False, 31)


def test_noalloc_rejects_more_rhs_than_scratch():
    from plqcd_amd import capi
    from plqcd_amd import hopping as hp
    lat = hp.Lattice((4, 4, 4, 4))
    g = hp.Gauge(lat)
    fi, fo = hp.Field(lat, nrhs=65), hp.Field(lat, nrhs=65)
    try:
        k = np.full(65, 0.12)
        assert _fn()(g.h, fi.h, fo.h, k.ctypes.data_as(C.POINTER(C.c_double)), 0) == 2
        assert "scratch" in capi.last_error()
    finally:
        for o in (fo, fi, g, lat):
            o.close()
