"""Oracle for the solvers (test infrastructure only).

The reference has no solver (include/plqcd_global_params.h:197-201), so the
solver results are pinned by what the mathematics fixes, computed here on the
CPU with the oracle's FP64 hopping matrix (wilson mode):
  M = 1 - kappa H,  M^+ = gamma5 M gamma5,  A = M^+ M;
  true relative residual ||b - A x|| / ||b||;
  a textbook CG on A for small lattices.
"""
from __future__ import annotations

import numpy as np

from . import MODE_WILSON, hopping_full

G5 = np.array([1.0, 1.0, -1.0, -1.0])[None, :, None]


def apply_m(x, u, L, kappa, dagger=False):
    if dagger:
        return G5 * apply_m(np.ascontiguousarray(G5 * x), u, L, kappa)
    return x - kappa * hopping_full(np.ascontiguousarray(x), u, L, MODE_WILSON)


def apply_mdagm(x, u, L, kappa):
    return apply_m(apply_m(x, u, L, kappa), u, L, kappa, dagger=True)


def true_residual(b, x, u, L, kappa) -> float:
    r = b - apply_mdagm(x, u, L, kappa)
    return float(np.linalg.norm(r) / np.linalg.norm(b))


def cg(b, u, L, kappa, tol, max_iter=10000):
    """Plain FP64 CG on M^+M (Hestenes-Stiefel); returns (x, iterations)."""
    x = np.zeros_like(b)
    r = b.copy()
    p = r.copy()
    rr = np.vdot(r, r).real
    bb = rr
    for it in range(1, max_iter + 1):
        ap = apply_mdagm(p, u, L, kappa)
        alpha = rr / np.vdot(p, ap).real
        x += alpha * p
        r -= alpha * ap
        rr_new = np.vdot(r, r).real
        if np.sqrt(rr_new / bb) <= tol:
            return x, it
        p = r + (rr_new / rr) * p
        rr = rr_new
    return x, max_iter
