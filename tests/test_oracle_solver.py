"""Pins the solver oracle (oracle/solver.py) on the CPU: M^+ is the adjoint
of M, M^+M is Hermitian positive definite, and the oracle CG reaches its
tolerance as measured by an independent residual computation."""
import numpy as np


def test_m_dagger_is_adjoint(orc, hp):
    from oracle import solver as osv
    L = (4, 4, 4, 6)
    V = int(np.prod(L))
    u = hp.random_gauge_su3(V, 31)
    psi, chi = hp.random_spinor(V, 32), hp.random_spinor(V, 33)
    k = 0.124
    lhs = np.vdot(chi, osv.apply_m(psi, u, L, k))
    rhs = np.vdot(osv.apply_m(chi, u, L, k, dagger=True), psi)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    a = np.vdot(psi, osv.apply_mdagm(psi, u, L, k))
    assert abs(a.imag) < 1e-12 * abs(a) and a.real > 0
    # kappa = 0 gives the identity
    assert np.array_equal(osv.apply_m(psi, u, L, 0.0), psi)


def test_oracle_cg_converges(orc, hp):
    from oracle import solver as osv
    L = (4, 4, 4, 4)
    V = 256
    u = hp.random_gauge_su3(V, 34)
    b = hp.random_spinor(V, 35)
    x, it = osv.cg(b, u, L, 0.124, 1e-10)
    assert osv.true_residual(b, x, u, L, 0.124) < 2e-10
    assert 5 < it < 500
