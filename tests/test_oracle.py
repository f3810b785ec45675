"""Pins the oracle (oracle/hop_oracle.c) before it is trusted as the checker.

- compat mode against the compiled reference's golden outputs (bitwise);
- each of the two reference defects shown present (removing either one
  breaks agreement with the reference);
- wilson mode against mathematics that does not share the oracle's
  half-spinor formulation: a direct full-spinor construction from explicit
  gamma matrices, the free-field plane-wave closed form, the constant-field
  values, gamma5-hermiticity and gauge covariance.
"""
import numpy as np
import pytest

from conftest import load_golden

# ------------------------------------------------------------ gamma matrices
# Chiral basis implied by the reference's projectors (lib/dirac.c:95-114).
I = 1j
G = [
    np.array([[0, 0, -1, 0], [0, 0, 0, -1], [-1, 0, 0, 0], [0, -1, 0, 0]], complex),
    np.array([[0, 0, 0, -I], [0, 0, -I, 0], [0, I, 0, 0], [I, 0, 0, 0]], complex),
    np.array([[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]], complex),
    np.array([[0, 0, -I, 0], [0, 0, 0, I], [I, 0, 0, 0], [0, -I, 0, 0]], complex),
]
G5 = G[0] @ G[1] @ G[2] @ G[3]


def test_gamma_algebra():
    for m in range(4):
        assert np.allclose(G[m], G[m].conj().T)
        for n in range(4):
            assert np.allclose(G[m] @ G[n] + G[n] @ G[m], 2 * np.eye(4) * (m == n))
    assert np.allclose(G5, np.diag([1, 1, -1, -1]))


# --------------------------------------------------- independent formulation

def lex_coords(L):
    V = int(np.prod(L))
    lex = np.arange(V)
    x = lex % L[0]
    y = (lex // L[0]) % L[1]
    z = (lex // (L[0] * L[1])) % L[2]
    t = lex // (L[0] * L[1] * L[2])
    return np.stack([x, y, z, t], axis=1)


def eo_order(L):
    """even-odd index -> lexicographic index, from the parity of the coordinate sum."""
    c = lex_coords(L)
    par = c.sum(axis=1) % 2
    lex = np.arange(len(c))
    return np.concatenate([lex[par == 0], lex[par == 1]])


def wilson_dense(q_eo, u_eo, L):
    """Eq. 2.1 with full 4-spinors, 4x4 gamma matrices and np.roll shifts on a
    lexicographic (t,z,y,x) array; no half-spinor projection anywhere."""
    order = eo_order(L)
    V = len(order)
    shp = (L[3], L[2], L[1], L[0])
    q = np.empty((V, 4, 3), complex)
    q[order] = q_eo
    u = np.empty((V, 4, 3, 3), complex)
    u[order] = u_eo
    q = q.reshape(*shp, 4, 3)
    u = u.reshape(*shp, 4, 3, 3)
    out = np.zeros_like(q)
    axis = {0: 3, 1: 2, 2: 1, 3: 0}  # mu -> array axis
    for mu in range(4):
        ax = axis[mu]
        P_minus = np.eye(4) - G[mu]
        P_plus = np.eye(4) + G[mu]
        U = u[..., mu, :, :]
        fwd = np.roll(q, -1, axis=ax)  # psi(x + mu)
        bwd = np.roll(q, 1, axis=ax)  # psi(x - mu)
        Ub = np.roll(U, 1, axis=ax)  # U_mu(x - mu)
        out += np.einsum("...ab,st,...tb->...sa", U, P_minus, fwd)
        out += np.einsum("...ba,st,...tb->...sa", Ub.conj(), P_plus, bwd)
    return out.reshape(V, 4, 3)[order]


def random_inputs(hp, L, gseed, sseed):
    V = int(np.prod(L))
    return hp.random_spinor(V, sseed), hp.random_gauge_su3(V, gseed)


# ------------------------------------------------------------------- golden

def test_oracle_geometry_matches_reference_dump(orc):
    import gzip
    from conftest import GOLDEN
    for L in [(4, 4, 4, 4), (8, 4, 6, 4)]:
        text = gzip.decompress((GOLDEN / ("geometry_%dx%dx%dx%d.txt.gz" % L)).read_bytes()).decode()
        rows = [l.split() for l in text.splitlines() if l.strip() and l.split()[0].lstrip("-").isdigit()]
        tab = np.array([[int(v) for v in r] for r in rows if len(r) == 14])
        ipt_eo, ipt_lex, iup, idn = orc.geometry(L)
        V = int(np.prod(L))
        assert tab.shape == (V, 14)
        assert np.array_equal(tab[:, 4], np.arange(V))
        assert np.array_equal(tab[:, 5], ipt_lex)
        assert np.array_equal(tab[:, 6:10], iup)
        assert np.array_equal(tab[:, 10:14], idn)
        assert np.array_equal(ipt_eo[ipt_lex], np.arange(V))


def _harness_inputs():
    g = load_golden("harness_4x4x4x4.npz")
    raw = g["rand_ints"].astype(np.float64) / float(g["rand_max"])
    V = 256
    q = raw[: V * 24].view(np.complex128).reshape(V, 4, 3).copy()
    u = raw[V * 24:].view(np.complex128).reshape(V, 4, 3, 3).copy()
    return q, u, g["out_compat"]


def test_oracle_compat_bitwise_harness_fill(orc):
    q, u, ref = _harness_inputs()
    out = orc.hopping_full(q, u, (4, 4, 4, 4), orc.MODE_COMPAT)
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("name", ["seed_4x4x4x4", "seed_8x4x6x4"])
def test_oracle_compat_bitwise_seeded(orc, hp, name):
    g = load_golden(f"{name}.npz")
    L = tuple(int(x) for x in g["L"])
    q, u = random_inputs(hp, L, int(g["gauge_seed"]), int(g["spinor_seed"]))
    out = orc.hopping_full(q, u, L, orc.MODE_COMPAT, nthreads=2)
    assert np.array_equal(out, g["out_compat"])


def test_oracle_compat_configs0_8x8x8x16(orc, hp):
    import hashlib
    g = load_golden("seed_8x8x8x16.npz")
    L = tuple(int(x) for x in g["L"])
    q, u = random_inputs(hp, L, int(g["gauge_seed"]), int(g["spinor_seed"]))
    out = orc.hopping_full(q, u, L, orc.MODE_COMPAT)
    assert hashlib.sha256(out.tobytes()).hexdigest() == str(g["sha_out"])
    assert np.array_equal(out[g["sample_idx"]], g["sample_out"])
    assert np.array_equal(np.sqrt((np.abs(out) ** 2).sum(axis=(1, 2))), g["site_norm"])
    assert bool(g["repeat100_equal"])


@pytest.mark.parametrize("mode", [2, 3])
def test_each_reference_defect_is_present(orc, hp, mode):
    """Planted in the oracle: dropping either defect breaks agreement."""
    q, u, ref = _harness_inputs()
    out = orc.hopping_full(q, u, (4, 4, 4, 4), mode)
    assert not np.allclose(out, ref)
    g = load_golden("seed_4x4x4x4.npz")
    q, u = random_inputs(hp, (4, 4, 4, 4), int(g["gauge_seed"]), int(g["spinor_seed"]))
    out = orc.hopping_full(q, u, (4, 4, 4, 4), mode)
    V = 256
    if mode == 2:  # the link shift only touches odd output sites
        assert np.array_equal(out[: V // 2], g["out_compat"][: V // 2])
        assert not np.allclose(out[V // 2:], g["out_compat"][V // 2:])
    else:  # the mu=0 assignment only touches the lower spin components
        assert np.array_equal(out[:, :2], g["out_compat"][:, :2])
        assert not np.allclose(out[:, 2:], g["out_compat"][:, 2:])


# -------------------------------------------------------- wilson mathematics

@pytest.mark.parametrize("L", [(4, 4, 4, 4), (2, 4, 6, 4), (6, 2, 4, 8)])
def test_oracle_wilson_equals_full_spinor_construction(orc, hp, L):
    q, u = random_inputs(hp, L, 5, 6)
    out = orc.hopping_full(q, u, L, orc.MODE_WILSON)
    ref = wilson_dense(q, u, L)
    assert np.abs(out - ref).max() < 1e-12 * np.abs(ref).max()


def test_oracle_eo_oe_halves(orc, hp):
    L = (4, 4, 4, 4)
    q, u = random_inputs(hp, L, 7, 8)
    for mode in (orc.MODE_COMPAT, orc.MODE_WILSON):
        full = orc.hopping_full(q, u, L, mode)
        eo = orc.hopping_eo(q, u, L, mode)
        oe = orc.hopping_oe(q, u, L, mode)
        assert np.array_equal(eo[:128], full[:128]) and np.all(eo[128:] == 0)
        assert np.array_equal(oe[128:], full[128:]) and np.all(oe[:128] == 0)


def test_oracle_free_field_plane_wave(orc, hp):
    L = (4, 6, 4, 8)
    V = int(np.prod(L))
    order = eo_order(L)
    c = lex_coords(L)[order]
    rng = np.random.default_rng(3)
    n = np.array([1, 2, 3, 5])
    p = 2 * np.pi * n / np.array(L)
    chi = rng.normal(size=(4, 3)) + 1j * rng.normal(size=(4, 3))
    phase = np.exp(1j * (c @ p))
    q = np.ascontiguousarray(phase[:, None, None] * chi[None])
    u = hp.unit_gauge(V)
    out = orc.hopping_full(q, u, L, orc.MODE_WILSON)
    M = sum(2 * np.cos(p[m]) * np.eye(4) - 2j * np.sin(p[m]) * G[m] for m in range(4))
    ref = phase[:, None, None] * (M @ chi)[None]
    assert np.abs(out - ref).max() < 1e-12


def test_oracle_constant_field(orc, hp):
    L = (4, 4, 4, 4)
    V = 256
    rng = np.random.default_rng(4)
    cvec = rng.normal(size=(4, 3)) + 1j * rng.normal(size=(4, 3))
    q = np.ascontiguousarray(np.broadcast_to(cvec, (V, 4, 3)))
    u = hp.unit_gauge(V)
    w = orc.hopping_full(q, u, L, orc.MODE_WILSON)
    assert np.abs(w - 8 * cvec).max() < 1e-13
    cm = orc.hopping_full(q, u, L, orc.MODE_COMPAT)
    c0, c1, c2, c3 = cvec
    expect = np.stack([8 * c0, 8 * c1, (-1 + 1j) * c0 + (1 + 1j) * c1 + 4 * c2,
                       (-1 + 1j) * c0 - (1 + 1j) * c1 + 4 * c3])
    assert np.abs(cm - expect).max() < 1e-13


def test_oracle_gamma5_hermiticity(orc, hp):
    L = (4, 4, 6, 4)
    V = int(np.prod(L))
    u = hp.random_gauge_su3(V, 9)
    psi = hp.random_spinor(V, 10)
    chi = hp.random_spinor(V, 11)
    g5 = np.array([1, 1, -1, -1])[None, :, None]
    Hpsi = orc.hopping_full(psi, u, L, orc.MODE_WILSON)
    g5chi = np.ascontiguousarray(g5 * chi)
    g5Hg5chi = g5 * orc.hopping_full(g5chi, u, L, orc.MODE_WILSON)
    lhs = np.vdot(chi, Hpsi)
    rhs = np.vdot(g5Hg5chi, psi)
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)


def test_oracle_gauge_covariance(orc, hp):
    L = (4, 4, 4, 6)
    V = int(np.prod(L))
    u = hp.random_gauge_su3(V, 12)
    psi = hp.random_spinor(V, 13)
    g = hp.random_gauge_su3(V, 14)[:, 0]  # one SU(3) matrix per site
    _, _, iup, _ = orc.geometry(L)
    ug = np.empty_like(u)
    for mu in range(4):
        ug[:, mu] = g @ u[:, mu] @ g[iup[:, mu]].conj().transpose(0, 2, 1)
    psig = np.ascontiguousarray(np.einsum("xab,xsb->xsa", g, psi))
    lhs = orc.hopping_full(psig, np.ascontiguousarray(ug), L, orc.MODE_WILSON)
    rhs = np.einsum("xab,xsb->xsa", g, orc.hopping_full(psi, u, L, orc.MODE_WILSON))
    assert np.abs(lhs - rhs).max() < 1e-12 * np.abs(rhs).max()


def test_oracle_compat_eo_upper_equals_wilson(orc, hp):
    L = (4, 4, 4, 4)
    q, u = random_inputs(hp, L, 15, 16)
    w = orc.hopping_eo(q, u, L, orc.MODE_WILSON)
    c = orc.hopping_eo(q, u, L, orc.MODE_COMPAT)
    assert np.array_equal(w[:128, :2], c[:128, :2])


def su3_known_answer_inputs(L):
    """One non-zero link U_0(x0) = U of tests/test_intrinsics.c:161-173 and
    psi = (v,0,0,0) at x0 and x0+0: then out(x0).s0 = U v and
    out(x0+0).s0 = U^+ v, with the values worked out by hand."""
    V = int(np.prod(L))
    U = np.array([[0, 1j, 2j], [1, 1 + 1j, 1 + 2j], [2, 2 + 1j, 2 + 2j]])
    v = np.array([1 + 1j, 2 + 2j, 3 + 3j])
    order = eo_order(L)
    eo_of_lex = np.argsort(order)
    x0 = eo_of_lex[0]  # site (0,0,0,0), even
    x1 = eo_of_lex[1]  # site (1,0,0,0), odd
    u = np.zeros((V, 4, 3, 3), complex)
    u[x0, 0] = U
    q = np.zeros((V, 4, 3), complex)
    q[x0, 0] = v
    q[x1, 0] = v
    return q, u, x0, x1


def test_oracle_su3_known_answer(orc):
    L = (4, 4, 4, 4)
    q, u, x0, x1 = su3_known_answer_inputs(L)
    for mode in (orc.MODE_COMPAT, orc.MODE_WILSON):
        out = orc.hopping_full(q, u, L, mode)
        assert np.array_equal(out[x0, 0], [-8 + 8j, -2 + 14j, 4 + 20j])
        assert np.array_equal(out[x1, 0], [8 + 8j, 14 + 2j, 20 - 4j])
