"""Host wrappers of the solver C ABI (plqcd_hip_cg_solve, plqcd_hip_apply_m)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import check, lib
from .hopping import Field, Gauge


def apply_m(gauge: Gauge, fin: Field, fout: Field, kappa, dagger: bool = False) -> None:
    """fout = M fin or M^+ fin with M = 1 - kappa H (per-RHS kappa allowed)."""
    k = np.broadcast_to(np.asarray(kappa, dtype=np.float64), (fin.nrhs,)).copy()
    check(lib().plqcd_hip_apply_m(gauge.h, fin.h, fout.h, k.ctypes.data_as(C.POINTER(C.c_double)), int(dagger)),
          "apply_m")


def apply_schur(gauge: Gauge, fin: Field, ftmp: Field, fout: Field, kappa, parity: int = 0) -> None:
    """fout_p = fin_p - kappa^2 H_pq H_qp fin_p on parity p (0 even, 1 odd)."""
    k = np.broadcast_to(np.asarray(kappa, dtype=np.float64), (fin.nrhs,)).copy()
    check(lib().plqcd_hip_apply_schur(gauge.h, fin.h, ftmp.h, fout.h, k.ctypes.data_as(C.POINTER(C.c_double)),
                                      int(parity)), "apply_schur")


def cg_solve(gauge: Gauge, b: Field, x: Field, kappa, tol, max_iter: int = 20000, sloppy: Gauge | None = None,
             delta: float = 0.0, use_graph: bool = True) -> list[dict]:
    """Solve M^+M x = b for every RHS of b.  kappa and tol may be scalars or
    per-RHS sequences.  Returns one stats dict per RHS."""
    R = b.nrhs
    kap = np.broadcast_to(np.asarray(kappa, dtype=np.float64), (R,)).copy()
    tl = np.broadcast_to(np.asarray(tol, dtype=np.float64), (R,)).copy()
    p = capi.CgParams()
    p.kappa = float(kap[0])
    p.tol = float(tl[0])
    p.max_iter = int(max_iter)
    p.delta = float(delta)
    p.use_graph = int(use_graph)
    p.kappa_rhs = kap.ctypes.data_as(C.POINTER(C.c_double))
    p.tol_rhs = tl.ctypes.data_as(C.POINTER(C.c_double))
    stats = (capi.CgStats * R)()
    check(lib().plqcd_hip_cg_solve(gauge.h, sloppy.h if sloppy is not None else None, b.h, x.h, C.byref(p), stats),
          "cg_solve")
    return [{k: getattr(s, k) for k, _ in capi.CgStats._fields_} for s in stats]
