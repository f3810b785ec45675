"""Known-answer tests of the GPU hopping matrix with a point source, through
plqcd_hip_hop, in the manner of the reference's gauge-free neighbour test
(tests/test_comms.c:120-146, :167-497) and its SU(3) known answer
(tests/test_intrinsics.c:161-173).

A source that is non-zero at one site x0 only reaches the 8 sites x0 -+ mu:
  out(x0 - mu) = U_mu(x0 - mu)  (1 - gamma_mu) psi(x0)      (forward term)
  out(x0 + mu) = U_mu(x0)^+     (1 + gamma_mu) psi(x0)      (backward term)
and every other site is exactly zero.  The expected values are built here from
the explicit gamma matrices of the reference's chiral basis (lib/dirac.c:95-114,
:163-195) as 4x4 spin matrices, independently of the kernels' projector
tables, and from hand-derived SU(3) products; all entries are small Gaussian
integers, so FP64 and FP32 results must be exact."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

I = 1j
GAMMA = [
    np.array([[0, 0, -1, 0], [0, 0, 0, -1], [-1, 0, 0, 0], [0, -1, 0, 0]], dtype=complex),
    np.array([[0, 0, 0, -I], [0, 0, -I, 0], [0, I, 0, 0], [I, 0, 0, 0]], dtype=complex),
    np.array([[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]], dtype=complex),
    np.array([[0, 0, -I, 0], [0, 0, 0, I], [I, 0, 0, 0], [0, -I, 0, 0]], dtype=complex),
]
ONE = np.eye(4, dtype=complex)
# tests/test_intrinsics.c:161-173: U and v with U v and U^+ v worked by hand
K = np.array([[0, I, 2 * I], [1, 1 + I, 1 + 2 * I], [2, 2 + I, 2 + 2 * I]])
V3 = np.array([1 + I, 2 + 2 * I, 3 + 3 * I])
KV = np.array([-8 + 8 * I, -2 + 14 * I, 4 + 20 * I])
KDV = np.array([8 + 8 * I, 14 + 2 * I, 20 - 4 * I])

L = (4, 6, 4, 8)
X0 = (1, 2, 3, 5)


def test_gamma_basis_is_consistent():
    """The spin matrices used below are Hermitian, square to one, anticommute
    and multiply to gamma5 = diag(1, 1, -1, -1)."""
    for a in range(4):
        assert np.allclose(GAMMA[a], GAMMA[a].conj().T)
        for b in range(4):
            ac = GAMMA[a] @ GAMMA[b] + GAMMA[b] @ GAMMA[a]
            assert np.allclose(ac, 2 * ONE * (a == b))
    assert np.allclose(GAMMA[0] @ GAMMA[1] @ GAMMA[2] @ GAMMA[3], np.diag([1, 1, -1, -1]))
    assert np.array_equal(K @ V3, KV) and np.array_equal(K.conj().T @ V3, KDV)


def _sites(orc):
    ipt_eo, _, iup, idn = orc.geometry(L)
    lex = X0[0] + L[0] * (X0[1] + L[1] * (X0[2] + L[2] * X0[3]))
    i0 = int(ipt_eo[lex])
    return i0, [int(idn[i0, mu]) for mu in range(4)], [int(iup[i0, mu]) for mu in range(4)]


def _run(hp, q, u, mode, prec):
    lat = hp.Lattice(L)
    g = hp.Gauge(lat, prec, 18).upload(u)
    fi, fo = hp.Field(lat, prec).upload(q), hp.Field(lat, prec)
    try:
        hp.hop(g, fi, fo, hp.PARITY_ALL, mode)
        return fo.download()
    finally:
        for o in (fo, fi, g, lat):
            o.close()


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("spin", [0, 1, 2, 3])
def test_point_source_unit_gauge(hp, orc, prec, spin):
    """Unit links, psi(x0) = e_(spin, colour 1): the 8 neighbours hold the
    columns of (1 -+ gamma_mu); in compat mode the forward terms do not reach
    the lower spin components (lib/dirac.c:373-374, :756-757)."""
    Vol = int(np.prod(L))
    i0, dn, up = _sites(orc)
    assert len(set(dn + up + [i0])) == 9
    q = np.zeros((Vol, 4, 3), complex)
    q[i0, spin, 1] = 1.0
    u = hp.unit_gauge(Vol)
    for mode in (hp.HOP_WILSON, hp.HOP_COMPAT):
        out = _run(hp, q, u, mode, prec)
        exp = np.zeros_like(out)
        for mu in range(4):
            fwd = (ONE - GAMMA[mu])[:, spin]
            if mode == hp.HOP_COMPAT:
                fwd = fwd * np.array([1, 1, 0, 0])
            exp[dn[mu], :, 1] += fwd
            exp[up[mu], :, 1] += (ONE + GAMMA[mu])[:, spin]
        assert np.array_equal(out, exp), (mode, prec, spin)


@pytest.mark.parametrize("prec", [0, 1])
def test_point_source_known_answer_links(hp, orc, prec):
    """psi(x0) = v in spin 0; the matrix K of tests/test_intrinsics.c:161-173 on
    the links U_mu(x0 - mu) (forward term at x0 - mu: K v) and U_mu(x0)
    (backward term at x0 + mu: K^+ v), unit links elsewhere; wilson mode."""
    Vol = int(np.prod(L))
    i0, dn, up = _sites(orc)
    q = np.zeros((Vol, 4, 3), complex)
    q[i0, 0, :] = V3
    u = hp.unit_gauge(Vol)
    for mu in range(4):
        u[dn[mu], mu] = K
        u[i0, mu] = K
    out = _run(hp, q, u, hp.HOP_WILSON, prec)
    exp = np.zeros_like(out)
    for mu in range(4):
        exp[dn[mu]] += np.outer((ONE - GAMMA[mu])[:, 0], KV)
        exp[up[mu]] += np.outer((ONE + GAMMA[mu])[:, 0], KDV)
    assert np.array_equal(out, exp), prec
