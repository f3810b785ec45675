"""GPU tests of the host <-> device transfers of fields (plqcd_hip_field_upload /
download): every transfer goes through the lattice's two page-locked staging
buffers in 32 MiB chunks, with multi-threaded host copies, or by DMA alone
from and into page-locked caller memory (plqcd_hip_host_alloc).  Checked
bitwise on the round trip (FP64 storage holds the values exactly) at a volume
spanning several chunks with an uneven last one, into one RHS of a multi-RHS
field, from and into page-locked arrays, and from two threads uploading to
two lattices at once (one thread per lattice, as the header requires)."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# 32^3 x 34 sites x 192 B = 214 MB: six 32 MiB chunks and a partial seventh
L_CHUNKS = (32, 32, 32, 34)


def test_round_trip_across_chunks_and_rhs(hp):
    V = int(np.prod(L_CHUNKS))
    lat = hp.Lattice(L_CHUNKS)
    f = hp.Field(lat, 0, 3)
    try:
        qs = [hp.random_spinor(V, 700 + r) for r in range(3)]
        for r in (2, 0, 1):
            f.upload(qs[r], r)
        for r in range(3):
            assert np.array_equal(f.download(r), qs[r]), r
    finally:
        f.close(); lat.close()


def test_round_trip_page_locked(hp):
    from plqcd_amd.serve import pinned_empty
    V = int(np.prod(L_CHUNKS))
    lat = hp.Lattice(L_CHUNKS)
    f = hp.Field(lat)
    try:
        src = pinned_empty((V, 4, 3))
        src[...] = hp.random_spinor(V, 711)
        f.upload(src)
        dst = pinned_empty((V, 4, 3))
        dst[...] = 0
        from plqcd_amd import capi
        capi.check(capi.lib().plqcd_hip_field_download(f.h, 0, dst.ctypes.data), "field_download")
        assert np.array_equal(dst, src)
        # pageable destination of the same field
        assert np.array_equal(f.download(), src)
    finally:
        f.close(); lat.close()


def test_two_lattices_from_two_threads(hp):
    L = (16, 16, 16, 32)
    V = int(np.prod(L))
    lats = [hp.Lattice(L), hp.Lattice(L)]
    fields = [hp.Field(la) for la in lats]
    qs = [[hp.random_spinor(V, 720 + 10 * t + k) for k in range(3)] for t in range(2)]
    out = [[None] * 3 for _ in range(2)]
    errors = []

    def work(t):
        try:
            for k in range(3):
                fields[t].upload(qs[t][k])
                out[t][k] = fields[t].download()
        except Exception as e:  # reported below, on the main thread
            errors.append(e)

    try:
        th = [threading.Thread(target=work, args=(t,)) for t in range(2)]
        for x in th:
            x.start()
        for x in th:
            x.join(timeout=120)
        assert not errors, errors
        for t in range(2):
            for k in range(3):
                assert np.array_equal(out[t][k], qs[t][k]), (t, k)
    finally:
        for f in fields:
            f.close()
        for la in lats:
            la.close()
