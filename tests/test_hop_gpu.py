"""GPU parity of the hopping matrix (through the C ABI) against the oracle and
the compiled reference's golden fixtures.

Tolerances (relative max-norm per site: max_k |d_k| / max_k |ref_k| over the
24 reals of a site, maximum over sites):
  FP64                         1e-13  (the GPU contracts multiply-adds to FMA,
                                        so results are not bitwise)
  FP32, 18- or 12-real links   2e-6
  FP16 spinor + FP16 links     4e-3  (12- or 18-real)
"""
import numpy as np
import pytest

from conftest import load_golden, site_rel_err

pytestmark = pytest.mark.gpu

TOL64, TOL32, TOL16 = 1e-13, 2e-6, 4e-3


def inputs(hp, L, gseed=101, sseed=102):
    V = int(np.prod(L))
    return hp.random_spinor(V, sseed), hp.random_gauge_su3(V, gseed)


def device_hop(hp, L, q, u, mode, parity=None, prec=0, links=18):
    parity = hp.PARITY_ALL if parity is None else parity
    lat = hp.Lattice(L)
    g = hp.Gauge(lat, prec, links).upload(u)
    fi = hp.Field(lat, prec).upload(q)
    fo = hp.Field(lat, prec)
    hp.hop(g, fi, fo, parity, mode)
    lat.sync()
    out = fo.download()
    fo.close(); fi.close(); g.close(); lat.close()
    return out


# ------------------------------------------------------- drop-in vs golden

def test_dropin_harness_fill_matches_reference(hp):
    g = load_golden("harness_4x4x4x4.npz")
    raw = g["rand_ints"].astype(np.float64) / float(g["rand_max"])
    V = 256
    q = raw[: V * 24].view(np.complex128).reshape(V, 4, 3).copy()
    u = raw[V * 24:].view(np.complex128).reshape(V, 4, 3, 3).copy()
    hp.init_dims((4, 4, 4, 4))
    hp.set_mode(hp.HOP_COMPAT)
    out = hp.spinor_field(V)
    assert hp.hopping_matrix_eo(q, out, u) > 0
    assert hp.hopping_matrix_oe(q, out, u) > 0
    assert site_rel_err(out, g["out_compat"]) < TOL64
    out2 = hp.spinor_field(V)
    hp.hopping_matrix(q, out2, u)
    assert site_rel_err(out2, g["out_compat"]) < TOL64
    hp.finalize_plqcd()


@pytest.mark.parametrize("name", ["seed_4x4x4x4", "seed_8x4x6x4", "seed_8x8x8x16"])
def test_dropin_seeded_matches_reference(hp, name):
    g = load_golden(f"{name}.npz")
    L = tuple(int(x) for x in g["L"])
    q, u = inputs(hp, L, int(g["gauge_seed"]), int(g["spinor_seed"]))
    hp.init_dims(L)
    hp.set_mode(hp.HOP_COMPAT)
    out = hp.spinor_field(q.shape[0])
    for _ in range(3):  # repeated applications give the same output
        hp.hopping_matrix_eo(q, out, u)
        hp.hopping_matrix_oe(q, out, u)
    if "out_compat" in g:
        assert site_rel_err(out, g["out_compat"]) < TOL64
    idx = g["sample_idx"]
    assert site_rel_err(out[idx], g["sample_out"]) < TOL64
    norms = np.sqrt((np.abs(out) ** 2).sum(axis=(1, 2)))
    assert np.abs(norms - g["site_norm"]).max() < 1e-13 * g["site_norm"].max()
    hp.finalize_plqcd()


def test_dropin_half_calls_touch_one_parity_and_alias(hp, orc):
    L = (4, 4, 6, 4)
    V = int(np.prod(L))
    q, u = inputs(hp, L)
    hp.init_dims(L)
    hp.set_mode(hp.HOP_WILSON)
    sentinel = hp.random_spinor(V, 999)
    out = sentinel.copy()
    hp.hopping_matrix_eo(q, out, u)
    assert np.array_equal(out[V // 2:], sentinel[V // 2:])
    assert site_rel_err(out[: V // 2], orc.hopping_eo(q, u, L, orc.MODE_WILSON)[: V // 2]) < TOL64
    out = sentinel.copy()
    hp.hopping_matrix_oe(q, out, u)
    assert np.array_equal(out[: V // 2], sentinel[: V // 2])
    # in place, as eo(q,q,u); oe(q,q,u) on the host: oe sees eo's output
    qq = q.copy()
    hp.hopping_matrix(qq, qq, u)
    ref = q.copy()
    orc.hopping_eo(q, u, L, orc.MODE_WILSON, qout=ref)
    ref2 = ref.copy()
    orc.hopping_oe(ref, u, L, orc.MODE_WILSON, qout=ref2)
    assert site_rel_err(qq, ref2) < TOL64
    # gauge caching: with the cache on, a changed field at the same pointer
    # is only seen after plqcd_hip_gauge_changed()
    hp.lib().plqcd_hip_cache_gauge(1)
    a = hp.spinor_field(V)
    hp.hopping_matrix(q, a, u)
    u2 = u.copy()
    u[:] = hp.random_gauge_su3(V, 5)
    b = hp.spinor_field(V)
    hp.hopping_matrix(q, b, u)
    assert np.array_equal(a, b)
    hp.lib().plqcd_hip_gauge_changed()
    hp.hopping_matrix(q, b, u)
    assert site_rel_err(b, orc.hopping_full(q, u, L, orc.MODE_WILSON)) < TOL64
    hp.lib().plqcd_hip_cache_gauge(0)
    u[:] = u2
    hp.finalize_plqcd()


# ---------------------------------------------------- device API vs oracle

LATTICES = [(2, 2, 2, 2), (4, 4, 4, 4), (2, 4, 6, 8), (6, 4, 2, 4), (10, 6, 4, 2), (8, 8, 8, 16)]


@pytest.mark.parametrize("L", LATTICES)
@pytest.mark.parametrize("mode", [0, 1])
def test_device_f64_matches_oracle(hp, orc, L, mode):
    q, u = inputs(hp, L)
    out = device_hop(hp, L, q, u, mode)
    ref = orc.hopping_full(q, u, L, mode)
    assert site_rel_err(out, ref) < TOL64


@pytest.mark.parametrize("mode", [0, 1])
def test_device_parities(hp, orc, mode):
    L = (6, 4, 4, 4)
    V = int(np.prod(L))
    q, u = inputs(hp, L)
    lat = hp.Lattice(L)
    g = hp.Gauge(lat).upload(u)
    fi = hp.Field(lat).upload(q)
    fo = hp.Field(lat)
    sentinel = hp.random_spinor(V, 77)
    for parity, fn in ((hp.PARITY_EVEN, orc.hopping_eo), (hp.PARITY_ODD, orc.hopping_oe)):
        fo.upload(sentinel)
        hp.hop(g, fi, fo, parity, mode)
        out = fo.download()
        ref = fn(q, u, L, mode)
        half = slice(0, V // 2) if parity == hp.PARITY_EVEN else slice(V // 2, V)
        other = slice(V // 2, V) if parity == hp.PARITY_EVEN else slice(0, V // 2)
        assert site_rel_err(out[half], ref[half]) < TOL64
        assert np.array_equal(out[other], sentinel[other])


def test_multi_rhs_field_equals_single(hp, orc):
    """Each RHS of a 3-RHS field: bitwise its one-RHS run, and the oracle's
    wilson H to 1e-13 per site."""
    L = (4, 4, 4, 8)
    V = int(np.prod(L))
    _, u = inputs(hp, L)
    qs = [hp.random_spinor(V, 200 + r) for r in range(3)]
    lat = hp.Lattice(L)
    g = hp.Gauge(lat).upload(u)
    fi = hp.Field(lat, nrhs=3)
    fo = hp.Field(lat, nrhs=3)
    for r, q in enumerate(qs):
        fi.upload(q, r)
    hp.hop(g, fi, fo, hp.PARITY_ALL, hp.HOP_WILSON)
    for r, q in enumerate(qs):
        out = fo.download(r)
        assert np.array_equal(out, device_hop(hp, L, q, u, hp.HOP_WILSON))
        assert site_rel_err(out, orc.hopping_full(q, u, L, orc.MODE_WILSON)) < TOL64, r


def test_device_argument_errors(hp):
    from plqcd_amd import capi
    lat = hp.Lattice((4, 4, 4, 4))
    g = hp.Gauge(lat)
    f64 = hp.Field(lat)
    f32 = hp.Field(lat, hp.PREC_F32)
    lib = capi.lib()
    assert lib.plqcd_hip_hop(g.h, f64.h, f32.h, 2, 1) != 0
    assert lib.plqcd_hip_hop(g.h, f64.h, f64.h, 2, 1) != 0
    assert lib.plqcd_hip_hop(g.h, f64.h, hp.Field(lat).h, 5, 1) != 0
    with pytest.raises(capi.PlqcdError):
        hp.Lattice((3, 4, 4, 4))
    with pytest.raises(capi.PlqcdError):
        hp.Gauge(lat, hp.PREC_F16, 16)


# -------------------------------------------------------- reduced precision

@pytest.mark.parametrize("prec,links,tol", [(1, 18, TOL32), (1, 12, TOL32), (2, 12, TOL16), (2, 18, TOL16)])
def test_reduced_precision_error_vs_f64_oracle(hp, orc, prec, links, tol):
    L = (8, 8, 8, 16)
    q, u = inputs(hp, L)
    ref = orc.hopping_full(q, u, L, orc.MODE_WILSON)
    out = device_hop(hp, L, q, u, hp.HOP_WILSON, prec=prec, links=links)
    err = site_rel_err(out, ref)
    print(f"prec {prec} links {links}: relative max-norm error per site {err:.3e}")
    assert err < tol


def test_f64_12_real_links(hp, orc):
    L = (4, 4, 8, 4)
    q, u = inputs(hp, L)
    out = device_hop(hp, L, q, u, hp.HOP_WILSON, prec=0, links=12)
    assert site_rel_err(out, orc.hopping_full(q, u, L, orc.MODE_WILSON)) < 1e-12


# ------------------------------------------------------- benchmark shapes

@pytest.mark.parametrize("mode", [0, 1])
def test_benchmark_shape_32c64_f64(hp, orc, mode):
    """32^3 x 64 FP64 against the oracle at every site."""
    L = (32, 32, 32, 64)
    q, u = inputs(hp, L, 1, 2)
    out = device_hop(hp, L, q, u, mode)
    ref = orc.hopping_full(q, u, L, mode)
    err = site_rel_err(out, ref)
    print(f"32^3x64 FP64 mode {mode}: {err:.3e}")
    assert err < TOL64


def test_benchmark_shape_32c64_f32_error(hp, orc):
    L = (32, 32, 32, 64)
    q, u = inputs(hp, L, 1, 2)
    ref = orc.hopping_full(q, u, L, orc.MODE_WILSON)
    for links in (18, 12):
        err = site_rel_err(device_hop(hp, L, q, u, hp.HOP_WILSON, prec=1, links=links), ref)
        print(f"32^3x64 FP32 {links}-real links: {err:.3e}")
        assert err < TOL32


# ---------------------------------------------------------- launch shapes

TUNINGS = [(0, 1, 256, 0, 1, 0, 0), (1, 1, 128, 0, 1, 0, 0), (2, 4, 64, 0, 1, 0, 0), (2, 16, 512, 0, 1, 0, 0),
           (2, 16, 256, 2, 1, 0, 0), (1, 1, 128, 4, 1, 0, 0), (2, 8, 64, 8, 1, 0, 0), (0, 1, 256, 16, 1, 0, 0)]


@pytest.mark.parametrize("tuning", TUNINGS)
@pytest.mark.parametrize("prec,links,tol", [(0, 18, TOL64), (1, 12, TOL32), (2, 12, TOL16), (2, 18, TOL16)])
def test_launch_shapes_match_oracle(hp, orc, tuning, prec, links, tol):
    """Every block order, block size and t-march length gives the operator."""
    from plqcd_amd import capi
    L = (8, 4, 6, 16)
    q, u = inputs(hp, L)
    ref = orc.hopping_full(q, u, L, orc.MODE_WILSON)
    lib = capi.lib()
    assert lib.plqcd_hip_hop_tuning(prec, *tuning) == 0
    try:
        for mode in (1, 0):
            out = device_hop(hp, L, q, u, mode, prec=prec, links=links)
            r = ref if mode == 1 else orc.hopping_full(q, u, L, orc.MODE_COMPAT)
            assert site_rel_err(out, r) < tol, (tuning, mode)
    finally:
        assert lib.plqcd_hip_hop_tuning_reset(prec) == 0


# ------------------------------------------------- RHS-blocked gather kernel

def _set_tuning(prec, t):
    from plqcd_amd import capi
    assert capi.lib().plqcd_hip_hop_tuning(prec, *t) == 0


@pytest.mark.parametrize("R", [4, 12])
@pytest.mark.parametrize("prec,links", [(0, 18), (1, 12), (1, 18), (2, 12), (2, 18)])
def test_rhs_blocked_bitwise_equal_single_rhs(hp, prec, links, R):
    """Every RHS of an RHS-blocked launch (2 or 4 RHS per thread) is bitwise
    the RHS's own one-RHS-per-thread result (gather and t-march kernels), in
    both modes and on each parity."""
    L = (8, 4, 6, 8)
    V = int(np.prod(L))
    _, u = inputs(hp, L)
    qs = [hp.random_spinor(V, 300 + r) for r in range(R)]
    lat = hp.Lattice(L)
    g = hp.Gauge(lat, prec, links).upload(u)
    fi, fo = hp.Field(lat, prec, R), hp.Field(lat, prec, R)
    f1i, f1o = hp.Field(lat, prec), hp.Field(lat, prec)
    for r, q in enumerate(qs):
        fi.upload(q, r)
    try:
        for mode in (hp.HOP_WILSON, hp.HOP_COMPAT):
            for parity in (hp.PARITY_ALL, hp.PARITY_EVEN, hp.PARITY_ODD):
                singles = []
                for tm in (0, 4):
                    _set_tuning(prec, (1, 1, 256, tm, 1, 0, 0))
                    outs = []
                    for q in qs:
                        f1i.upload(q)
                        f1o.upload(np.zeros_like(q))
                        hp.hop(g, f1i, f1o, parity, mode)
                        outs.append(f1o.download())
                    singles.append(outs)
                for r in range(R):
                    assert np.array_equal(singles[0][r], singles[1][r]), ("t-march", mode, parity, r)
                for rb, blk in ((2, 128), (4, 256), (4, 64)):
                    _set_tuning(prec, (2, 4, blk, 0, rb, 0, 0))
                    for r in range(R):
                        fo.upload(np.zeros_like(qs[0]), r)
                    hp.hop(g, fi, fo, parity, mode)
                    for r in range(R):
                        assert np.array_equal(fo.download(r), singles[0][r]), (rb, blk, mode, parity, r)
    finally:
        from plqcd_amd import capi
        assert capi.lib().plqcd_hip_hop_tuning_reset(prec) == 0
        for f in (fi, fo, f1i, f1o):
            f.close()
        g.close(); lat.close()


# ---------------------------------------------------- LDS-tiled t-march kernel

TILES = [((8, 8, 8, 16), [(4, 4), (8, 2), (2, 8), (8, 4), (4, 8)]),
         ((16, 16, 8, 8), [(8, 1), (4, 2), (1, 8), (16, 1), (4, 4)])]


@pytest.mark.parametrize("L,tiles", TILES)
@pytest.mark.parametrize("prec,links,tol", [(0, 18, TOL64), (0, 12, TOL64), (1, 12, TOL32), (2, 12, TOL16),
                                            (2, 18, TOL16)])
def test_tile_kernel_bitwise_equal_tmarch(hp, orc, L, tiles, prec, links, tol):
    """The LDS-tiled t-march gives bitwise the t-march's and the gather
    kernel's output for every tile (including tiles spanning a whole extent,
    whose periodic neighbours lie inside the tile), in both modes, and matches
    the oracle."""
    from plqcd_amd import capi
    q, u = inputs(hp, L, 61, 62)
    refs = {hp.HOP_WILSON: orc.hopping_full(q, u, L, orc.MODE_WILSON),
            hp.HOP_COMPAT: orc.hopping_full(q, u, L, orc.MODE_COMPAT)}
    try:
        for mode, ref in refs.items():
            _set_tuning(prec, (1, 1, 256, 0, 1, 0, 0))
            gather = device_hop(hp, L, q, u, mode, prec=prec, links=links)
            assert site_rel_err(gather, ref) < tol, mode
            for T in (4, 8):
                _set_tuning(prec, (1, 1, 256, T, 1, 0, 0))
                assert np.array_equal(device_hop(hp, L, q, u, mode, prec=prec, links=links), gather), (mode, T)
                for ty, tz in tiles:
                    for order in (0, 1):
                        _set_tuning(prec, (order, 1, 256, T, 1, ty, tz))
                        out = device_hop(hp, L, q, u, mode, prec=prec, links=links)
                        assert np.array_equal(out, gather), (mode, T, ty, tz, order)
            # the half-lattice launches ignore the tile
            _set_tuning(prec, (1, 1, 256, 4, 1, 4, 4))
            for parity in (hp.PARITY_EVEN, hp.PARITY_ODD):
                half = device_hop(hp, L, q, u, mode, parity=parity, prec=prec, links=links)
                sl = slice(0, half.shape[0] // 2) if parity == hp.PARITY_EVEN else slice(half.shape[0] // 2, None)
                assert np.array_equal(half[sl], gather[sl]), (mode, parity)
    finally:
        assert capi.lib().plqcd_hip_hop_tuning_reset(prec) == 0


def test_tuning_rejects_bad_tiles(hp):
    from plqcd_amd import capi
    lib = capi.lib()
    assert lib.plqcd_hip_hop_tuning(0, 1, 1, 256, 4, 1, 4, 0) != 0
    assert lib.plqcd_hip_hop_tuning(0, 1, 1, 256, 4, 1, -1, 2) != 0
    assert lib.plqcd_hip_hop_tuning(0, 1, 1, 256, 4, 3, 0, 0) != 0
    assert lib.plqcd_hip_hop_tuning_reset(-1) == 0


# ------------------------------------------------------------ ragged batch

RAGGED = [[((4, 4, 4, 8), 3), ((8, 4, 6, 6), 1), ((8, 8, 8, 16), 4), ((4, 4, 4, 4), 12)],  # mixed kinds
          [((4, 4, 4, 16), 3), ((8, 8, 8, 32), 1), ((4, 4, 8, 64), 2)],                    # t-march only (F64/F32)
          [((4, 4, 4, 8), 12), ((8, 4, 4, 8), 12)],                                        # 192-lane blocks
          [((4, 4, 4, 4 * (1 + k % 2)), 1 + k % 3) for k in range(18)]]                    # two launches of <= 16


@pytest.mark.parametrize("shapes", RAGGED)
@pytest.mark.parametrize("prec,links,tol", [(0, 18, TOL64), (1, 12, TOL32), (2, 12, TOL16), (2, 18, TOL16)])
def test_ragged_batch_matches_separate_and_oracle(hp, orc, shapes, prec, links, tol):
    """One launch over lattices of different volumes and RHS counts
    (t-march items and items whose time extent takes gather blocks, mixed or
    all of one kind) gives each item bitwise its own launch's output, which
    matches the oracle."""
    from plqcd_amd import capi
    lib = capi.lib()
    lats, gs, fis, fos, fsep, inputs_ = [], [], [], [], [], []
    for k, (L, R) in enumerate(shapes):
        V = int(np.prod(L))
        lat = hp.Lattice(L)
        if lats:
            lat.set_stream(lib.plqcd_hip_lattice_stream(lats[0].h))
        u = hp.random_gauge_su3(V, 300 + k)
        qs = [hp.random_spinor(V, 400 + 13 * k + r) for r in range(R)]
        g = hp.Gauge(lat, prec, links).upload(u)
        fi, fo, fs = hp.Field(lat, prec, R), hp.Field(lat, prec, R), hp.Field(lat, prec, R)
        for r, q in enumerate(qs):
            fi.upload(q, r)
        lats.append(lat); gs.append(g); fis.append(fi); fos.append(fo); fsep.append(fs); inputs_.append((L, u, qs))
    try:
        for mode, omode in ((hp.HOP_WILSON, orc.MODE_WILSON), (hp.HOP_COMPAT, orc.MODE_COMPAT)):
            hp.hop_batch(gs, fis, fos, mode)
            for k in range(len(shapes)):
                hp.hop(gs[k], fis[k], fsep[k], hp.PARITY_ALL, mode)
            lats[0].sync()
            for k, (L, u, qs) in enumerate(inputs_):
                for r, q in enumerate(qs):
                    out = fos[k].download(r)
                    assert np.array_equal(out, fsep[k].download(r)), (mode, L, r)
                    assert site_rel_err(out, orc.hopping_full(q, u, L, omode)) < tol, (mode, L, r)
    finally:
        for objs in (fis, fos, fsep, gs):
            for o in objs:
                o.close()
        for lat in reversed(lats):
            lat.close()


def test_ragged_batch_errors(hp):
    lat1, lat2 = hp.Lattice((4, 4, 4, 4)), hp.Lattice((4, 4, 4, 8))
    g1, g2 = hp.Gauge(lat1), hp.Gauge(lat2, 1, 12)
    f1, o1 = hp.Field(lat1), hp.Field(lat1)
    f2, o2 = hp.Field(lat2, 1), hp.Field(lat2, 1)
    with pytest.raises(RuntimeError):
        hp.hop_batch([g1, g2], [f1, f2], [o1, o2])  # mixed storage formats
    with pytest.raises(RuntimeError):
        hp.hop_batch([g1], [f1], [f1])  # in == out
    for o in (f1, o1, f2, o2, g1, g2, lat2, lat1):
        o.close()
