"""Host-side mirror of PLQCD's hopping-matrix interface over libplqcd_hip.so.

Fields are numpy arrays with the reference's memory layout:
  spinor field: complex128, shape (V, 4, 3)     == spinor[V]   (include/su3.h:67-70)
  gauge field:  complex128, shape (V, 4, 3, 3)  == su3 u[4*V], u[4*i+mu] (su3.h:52-55)
with sites in even-odd order (lib/start.c:406-433).

Global-state functions mirror lib/start.c and include/dirac.h (same names and
argument meaning); the device-resident classes wrap the library's handle API.
Every call goes through the C ABI; nothing here computes the operator.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import capi
from .capi import check, lib

PREC_F64, PREC_F32, PREC_F16 = capi.PREC_F64, capi.PREC_F32, capi.PREC_F16
PARITY_EVEN, PARITY_ODD, PARITY_ALL = capi.PARITY_EVEN, capi.PARITY_ODD, capi.PARITY_ALL
HOP_COMPAT, HOP_WILSON = capi.HOP_COMPAT, capi.HOP_WILSON

FLOP_PER_SITE = 1320  # 8 x (2 x 66 SU(3) mat-vec) + 8 x 12 projection + 7 x 24 accumulation


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _check_spinor(a: np.ndarray, V: int, name: str):
    if a.dtype != np.complex128 or a.shape != (V, 4, 3) or not a.flags.c_contiguous:
        raise ValueError(f"{name} must be a C-contiguous complex128 array of shape ({V}, 4, 3)")


def _check_gauge(a: np.ndarray, V: int):
    if a.dtype != np.complex128 or a.shape != (V, 4, 3, 3) or not a.flags.c_contiguous:
        raise ValueError(f"u must be a C-contiguous complex128 array of shape ({V}, 4, 3, 3)")


# ------------------------------------------------------------ synthetic inputs

def spinor_field(V: int) -> np.ndarray:
    return np.zeros((V, 4, 3), dtype=np.complex128)


def gauge_field(V: int) -> np.ndarray:
    return np.zeros((V, 4, 3, 3), dtype=np.complex128)


def random_gauge_su3(V: int, seed: int) -> np.ndarray:
    u = gauge_field(V)
    lib().plqcd_random_gauge_su3(_ptr(u), V, seed)
    return u


def random_spinor(V: int, seed: int) -> np.ndarray:
    s = spinor_field(V)
    lib().plqcd_random_spinor_gauss(_ptr(s), V, seed)
    return s


def unit_gauge(V: int) -> np.ndarray:
    u = gauge_field(V)
    lib().plqcd_unit_gauge(_ptr(u), V)
    return u


# ------------------------------------------- global state (lib/start.c API)

class _GlobalParams(C.Structure):
    _fields_ = [
        ("nprocs", C.c_int * 4), ("nthread", C.c_int), ("cpr", C.c_int * 4), ("npr", C.c_int * 8),
        ("LX", C.c_int), ("LY", C.c_int), ("LZ", C.c_int), ("LT", C.c_int),
        ("latdims", C.c_int * 4), ("VOLUME", C.c_int), ("face", C.c_int * 4),
        ("ipt_eo", C.POINTER(C.c_int)), ("ipt_lex", C.POINTER(C.c_int)),
        ("nn_bnde", C.POINTER(C.c_int) * 8), ("nn_bndo", C.POINTER(C.c_int) * 8),
        ("iup", C.POINTER(C.POINTER(C.c_int))), ("idn", C.POINTER(C.POINTER(C.c_int))),
        ("phip", C.c_void_p * 4), ("phim", C.c_void_p * 4),
        ("ALIGN", C.c_int), ("init", C.c_int),
    ]


def plqcd_g() -> _GlobalParams:
    """The library's global `plqcd_g` struct (live view)."""
    return _GlobalParams.in_dll(lib(), "plqcd_g")


@dataclass
class Geometry:
    latdims: tuple
    VOLUME: int
    ipt_eo: np.ndarray
    ipt_lex: np.ndarray
    iup: np.ndarray  # (V, 4)
    idn: np.ndarray  # (V, 4)


def geometry() -> Geometry:
    g = plqcd_g()
    if g.init != 1:
        raise capi.PlqcdError("plqcd is not initialised")
    V = g.VOLUME
    ipt_eo = np.ctypeslib.as_array(g.ipt_eo, shape=(V,)).copy()
    ipt_lex = np.ctypeslib.as_array(g.ipt_lex, shape=(V,)).copy()
    # iup/idn rows are one contiguous block (iup[0] points at it)
    iup = np.ctypeslib.as_array(g.iup[0], shape=(V * 4,)).reshape(V, 4).copy()
    idn = np.ctypeslib.as_array(g.idn[0], shape=(V * 4,)).reshape(V, 4).copy()
    return Geometry(tuple(g.latdims), V, ipt_eo, ipt_lex, iup, idn)


def init_plqcd(input_file: str) -> int:
    """init_plqcd(argc, argv) with argv[1] = input_file (lib/start.c:19)."""
    argv = (C.c_char_p * 3)(b"plqcd", os.fsencode(input_file), None)
    return lib().init_plqcd(2, argv)


def init_dims(L, nthread: int = 1, align: int = 6) -> None:
    check(lib().plqcd_init_dims(int(L[0]), int(L[1]), int(L[2]), int(L[3]), nthread, align), "plqcd_init_dims")


def finalize_plqcd() -> None:
    lib().finalize_plqcd()


def lex2cart(ix: int) -> tuple:
    x = (C.c_int * 4)(-1, -1, -1, -1)
    lib().lex2cart(ix, x)
    return tuple(x)


def cart2lex(xc) -> int:
    return lib().cart2lex((C.c_int * 4)(*xc))


def find_site(xg) -> tuple:
    pid, xeo = C.c_int(), C.c_int()
    lib().find_site((C.c_int * 4)(*xg), C.byref(pid), C.byref(xeo))
    return pid.value, xeo.value


def print_params(path: str) -> None:
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    fp = libc.fopen(os.fsencode(path), b"w")
    if not fp:
        raise OSError(f"cannot open {path}")
    try:
        lib().print_params(fp)
    finally:
        libc.fclose(fp)


# -------------------------------------------- drop-in hopping (dirac.h API)

def set_mode(mode: int) -> None:
    if lib().plqcd_hip_set_mode(mode) != 0:
        raise ValueError(capi.last_error())


def _hop_call(fn, qin, qout, u) -> float:
    V = plqcd_g().VOLUME
    _check_spinor(qin, V, "qin")
    _check_spinor(qout, V, "qout")
    _check_gauge(u, V)
    t = fn(_ptr(qin), _ptr(qout), _ptr(u))
    if t < 0:
        raise capi.PlqcdError(capi.last_error())
    return t


def hopping_matrix_eo(qin, qout, u) -> float:
    """qout[:V/2] = H_eo qin[V/2:] (include/dirac.h:21); returns wall seconds."""
    return _hop_call(lib().plqcd_hopping_matrix_eo_hip, qin, qout, u)


def hopping_matrix_oe(qin, qout, u) -> float:
    """qout[V/2:] = H_oe qin[:V/2] (include/dirac.h:22); returns wall seconds."""
    return _hop_call(lib().plqcd_hopping_matrix_oe_hip, qin, qout, u)


def hopping_matrix(qin, qout, u) -> float:
    """eo followed by oe (tests/test_hopping.c:296-297)."""
    return _hop_call(lib().plqcd_hopping_matrix_hip, qin, qout, u)


# ------------------------------------------------------ device-resident API

class Lattice:
    def __init__(self, L):
        self.L = tuple(int(x) for x in L)
        self.V = int(np.prod(self.L))
        h = C.c_void_p()
        check(lib().plqcd_hip_lattice_create(capi.i4(*self.L), C.byref(h)), "lattice_create")
        self.h = h

    def close(self):
        if self.h:
            lib().plqcd_hip_lattice_destroy(self.h)
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

    def sync(self):
        check(lib().plqcd_hip_lattice_sync(self.h), "lattice_sync")

    def set_stream(self, stream_handle):
        check(lib().plqcd_hip_lattice_set_stream(self.h, stream_handle), "set_stream")

    def timer_start(self):
        check(lib().plqcd_hip_timer_start(self.h), "timer_start")

    def timer_stop(self) -> float:
        ms = C.c_double()
        check(lib().plqcd_hip_timer_stop(self.h, C.byref(ms)), "timer_stop")
        return ms.value


class Gauge:
    def __init__(self, lat: Lattice, prec: int = PREC_F64, links: int = 18):
        self.lat, self.prec, self.links = lat, prec, links
        h = C.c_void_p()
        check(lib().plqcd_hip_gauge_create(lat.h, prec, links, C.byref(h)), "gauge_create")
        self.h = h

    def upload(self, u: np.ndarray):
        _check_gauge(u, self.lat.V)
        check(lib().plqcd_hip_gauge_upload(self.h, _ptr(u)), "gauge_upload")
        return self

    @property
    def nbytes(self) -> int:
        return lib().plqcd_hip_gauge_bytes(self.h)

    def close(self):
        if self.h:
            lib().plqcd_hip_gauge_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Field:
    def __init__(self, lat: Lattice, prec: int = PREC_F64, nrhs: int = 1):
        self.lat, self.prec, self.nrhs = lat, prec, nrhs
        h = C.c_void_p()
        check(lib().plqcd_hip_field_create(lat.h, prec, nrhs, C.byref(h)), "field_create")
        self.h = h

    def upload(self, s: np.ndarray, rhs: int = 0):
        _check_spinor(s, self.lat.V, "spinor")
        check(lib().plqcd_hip_field_upload(self.h, rhs, _ptr(s)), "field_upload")
        return self

    def download(self, rhs: int = 0) -> np.ndarray:
        out = spinor_field(self.lat.V)
        check(lib().plqcd_hip_field_download(self.h, rhs, _ptr(out)), "field_download")
        return out

    @property
    def nbytes(self) -> int:
        return lib().plqcd_hip_field_bytes(self.h)

    def close(self):
        if self.h:
            lib().plqcd_hip_field_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def hop(gauge: Gauge, fin: Field, fout: Field, parity: int = PARITY_ALL, mode: int = HOP_WILSON) -> None:
    """fout = H fin on `parity` (asynchronous on the lattice stream)."""
    check(lib().plqcd_hip_hop(gauge.h, fin.h, fout.h, parity, mode), "hop")


def hop_batch(gauges, fins, fouts, mode: int = HOP_WILSON) -> None:
    """fouts[k] = H fins[k] (full lattice) for independent lattices in one
    launch; all items share one storage format and one stream."""
    n = len(gauges)
    if not (n == len(fins) == len(fouts)) or n == 0:
        raise ValueError("hop_batch: gauges, fins and fouts must be non-empty and of equal length")
    garr = (C.c_void_p * n)(*[g.h.value for g in gauges])
    iarr = (C.c_void_p * n)(*[f.h.value for f in fins])
    oarr = (C.c_void_p * n)(*[f.h.value for f in fouts])
    check(lib().plqcd_hip_hop_batch(n, garr, iarr, oarr, mode), "hop_batch")
