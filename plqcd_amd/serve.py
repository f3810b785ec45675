"""Host wrapper of the solve server (continuous batching of CG requests).

The server runs in a worker thread inside libplqcd_hip.so; this module keeps
the host buffers of in-flight requests alive and turns results into dicts.
"""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np

from . import capi
from .capi import check, lib
from .hopping import PREC_F16, PREC_F32, PREC_F64  # noqa: F401


def pinned_empty(shape, dtype=np.complex128) -> np.ndarray:
    """An uninitialised array in page-locked host memory
    (plqcd_hip_host_alloc), freed when the last view of it is gone.  The
    server moves sources and solutions held in such arrays by DMA alone."""
    n = int(np.prod(shape))
    nbytes = max(1, n * np.dtype(dtype).itemsize)
    p = C.c_void_p()
    check(lib().plqcd_hip_host_alloc(nbytes, C.byref(p)), "host_alloc")
    buf = (C.c_char * nbytes).from_address(p.value)
    weakref.finalize(buf, lib().plqcd_hip_host_free, C.c_void_p(p.value))
    return np.frombuffer(buf, dtype=dtype, count=n).reshape(shape)


class Server:
    def __init__(self, sloppy_prec: int = PREC_F16, sloppy_links: int = 12, slots: int = 12,
                 hbm_budget: int = 0, use_graph: bool = True, device: int = 0, transfer_threads: int = 0):
        cfg = capi.ServeConfig(device, sloppy_prec, sloppy_links, slots, hbm_budget, int(use_graph),
                               transfer_threads)
        h = C.c_void_p()
        check(lib().plqcd_hip_server_create(C.byref(cfg), C.byref(h)), "server_create")
        self.h = h
        self._live = {}  # id -> (b, x, request struct)

    def submit(self, L, gauge_seed: int, b: np.ndarray, kappa: float, tol: float, max_iter: int = 20000,
               gauge: np.ndarray | None = None, out: np.ndarray | None = None) -> int:
        """b: complex128 array (nrhs, V, 4, 3).  out: optional array of the same
        shape for the solutions (for instance from pinned_empty), else a new
        one.  Returns the request id."""
        V = int(np.prod(L))
        b = np.ascontiguousarray(b, dtype=np.complex128).reshape(-1, V, 4, 3)
        if out is None:
            # the server writes every element; empty_like commits no host pages
            # until the solution arrives (zeros_like would touch them all now)
            x = np.empty_like(b)
        else:
            if out.shape != b.shape or out.dtype != np.complex128 or not out.flags.c_contiguous:
                raise ValueError(f"out must be a C-contiguous complex128 array of shape {b.shape}")
            x = out
        req = capi.ServeRequest()
        req.L = (C.c_int * 4)(*L)
        req.gauge_seed = gauge_seed
        req.gauge = gauge.ctypes.data if gauge is not None else None
        req.kappa, req.tol, req.max_iter, req.nrhs = kappa, tol, max_iter, b.shape[0]
        req.b, req.x = b.ctypes.data, x.ctypes.data
        rid = C.c_uint64()
        check(lib().plqcd_hip_server_submit(self.h, C.byref(req), C.byref(rid)), "server_submit")
        self._live[rid.value] = (b, x, req, gauge)
        return rid.value

    def poll(self) -> list[dict]:
        out = (capi.ServeResult * 256)()
        n = C.c_int()
        check(lib().plqcd_hip_server_poll(self.h, out, 256, C.byref(n)), "server_poll")
        res = []
        for r in out[: n.value]:
            d = {k: getattr(r, k) for k, _ in capi.ServeResult._fields_}
            b, x, _, _ = self._live.pop(d["id"])
            d["x"] = x
            res.append(d)
        return res

    def wait(self, count: int, timeout: float = 600.0) -> None:
        check(lib().plqcd_hip_server_wait(self.h, count, timeout), "server_wait")

    def run_all(self, timeout: float = 600.0) -> dict:
        """Waits for every submitted request; returns {id: result}."""
        done = {}
        while self._live:
            self.wait(1, timeout)
            for d in self.poll():
                done[d["id"]] = d
        return done

    def info(self) -> dict:
        v = (C.c_double * 8)()
        check(lib().plqcd_hip_server_info(self.h, v), "server_info")
        return {"budget": v[0], "used": v[1], "entries": int(v[2]), "failed": int(v[3]),
                "engines_created": int(v[4]), "evictions": int(v[5]), "transfer_bytes": v[6],
                "transfer_seconds": v[7]}

    def close(self):
        if self.h:
            lib().plqcd_hip_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
