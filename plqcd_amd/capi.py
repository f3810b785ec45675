"""ctypes binding of libplqcd_hip.so (the C ABI declared in include/plqcd_hip.h).

Loading fails loudly when the library has not been built: there is no
fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libplqcd_hip.so"
HEADER_DIR = PKG.parent / "include"

PREC_F64, PREC_F32, PREC_F16 = 0, 1, 2
LINKS_18, LINKS_12 = 18, 12
PARITY_EVEN, PARITY_ODD, PARITY_ALL = 0, 1, 2
HOP_COMPAT, HOP_WILSON = 0, 1

vp = C.c_void_p
i4 = C.c_int * 4


class CgParams(C.Structure):
    _fields_ = [("kappa", C.c_double), ("tol", C.c_double), ("max_iter", C.c_int), ("delta", C.c_double),
                ("use_graph", C.c_int), ("kappa_rhs", C.POINTER(C.c_double)), ("tol_rhs", C.POINTER(C.c_double))]


class CgStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("reliable_updates", C.c_int), ("converged", C.c_int),
                ("true_residual", C.c_double), ("seconds", C.c_double), ("hop_applications", C.c_double)]


class ServeConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("sloppy_prec", C.c_int), ("sloppy_links", C.c_int), ("slots", C.c_int),
                ("hbm_budget", C.c_size_t), ("use_graph", C.c_int), ("transfer_threads", C.c_int)]


class ServeRequest(C.Structure):
    _fields_ = [("L", C.c_int * 4), ("gauge_seed", C.c_uint64), ("gauge", C.c_void_p), ("kappa", C.c_double),
                ("tol", C.c_double), ("max_iter", C.c_int), ("nrhs", C.c_int), ("b", C.c_void_p), ("x", C.c_void_p)]


class ServeResult(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int), ("submit_time", C.c_double), ("start_time", C.c_double),
                ("end_time", C.c_double), ("iterations", C.c_int), ("true_residual", C.c_double),
                ("converged", C.c_int)]


# name -> (restype, argtypes)
_PROTOS = {
    "init_plqcd": (C.c_int, [C.c_int, C.POINTER(C.c_char_p)]),
    "plqcd_init_dims": (C.c_int, [C.c_int] * 6),
    "finalize_plqcd": (None, []),
    "ipr": (C.c_int, [C.POINTER(C.c_int)]),
    "comp_cpr": (None, [C.c_int, C.POINTER(C.c_int)]),
    "set_npr": (None, []),
    "lex2cart": (None, [C.c_int, C.POINTER(C.c_int)]),
    "cart2lex": (C.c_int, [C.POINTER(C.c_int)]),
    "set_geometry": (C.c_int, []),
    "find_site": (None, [C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "print_params": (None, [vp]),
    "alloc": (vp, [C.c_size_t, C.c_size_t]),
    "amalloc": (vp, [C.c_size_t, C.c_int]),
    "afree": (None, [vp]),
    "stop_watch": (C.c_double, [C.c_double]),
    "plqcd_hip_last_error": (C.c_char_p, []),
    "plqcd_hip_set_mode": (C.c_int, [C.c_int]),
    "plqcd_hopping_matrix_eo_hip": (C.c_double, [vp, vp, vp]),
    "plqcd_hopping_matrix_oe_hip": (C.c_double, [vp, vp, vp]),
    "plqcd_hopping_matrix_hip": (C.c_double, [vp, vp, vp]),
    "plqcd_hip_cache_gauge": (None, [C.c_int]),
    "plqcd_hip_gauge_changed": (None, []),
    "plqcd_hip_lattice_create": (C.c_int, [i4, C.POINTER(vp)]),
    "plqcd_hip_max_rhs": (C.c_int, [i4]),
    "plqcd_hip_lattice_destroy": (None, [vp]),
    "plqcd_hip_lattice_set_stream": (C.c_int, [vp, vp]),
    "plqcd_hip_lattice_stream": (vp, [vp]),
    "plqcd_hip_lattice_sync": (C.c_int, [vp]),
    "plqcd_hip_timer_start": (C.c_int, [vp]),
    "plqcd_hip_timer_stop": (C.c_int, [vp, C.POINTER(C.c_double)]),
    "plqcd_hip_gauge_create": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(vp)]),
    "plqcd_hip_gauge_destroy": (None, [vp]),
    "plqcd_hip_gauge_upload": (C.c_int, [vp, vp]),
    "plqcd_hip_gauge_bytes": (C.c_size_t, [vp]),
    "plqcd_hip_field_create": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(vp)]),
    "plqcd_hip_field_destroy": (None, [vp]),
    "plqcd_hip_field_upload": (C.c_int, [vp, C.c_int, vp]),
    "plqcd_hip_field_download": (C.c_int, [vp, C.c_int, vp]),
    "plqcd_hip_field_bytes": (C.c_size_t, [vp]),
    "plqcd_hip_hop": (C.c_int, [vp, vp, vp, C.c_int, C.c_int]),
    "plqcd_hip_apply_m": (C.c_int, [vp, vp, vp, C.POINTER(C.c_double), C.c_int]),
    "plqcd_hip_apply_m_noalloc": (C.c_int, [vp, vp, vp, C.POINTER(C.c_double), C.c_int]),
    "plqcd_hip_apply_schur": (C.c_int, [vp, vp, vp, vp, C.POINTER(C.c_double), C.c_int]),
    "plqcd_hip_cg_solve": (C.c_int, [vp, vp, vp, vp, C.POINTER(CgParams), C.POINTER(CgStats)]),
    "plqcd_hip_server_create": (C.c_int, [C.POINTER(ServeConfig), C.POINTER(vp)]),
    "plqcd_hip_server_submit": (C.c_int, [vp, C.POINTER(ServeRequest), C.POINTER(C.c_uint64)]),
    "plqcd_hip_server_poll": (C.c_int, [vp, C.POINTER(ServeResult), C.c_int, C.POINTER(C.c_int)]),
    "plqcd_hip_server_wait": (C.c_int, [vp, C.c_int, C.c_double]),
    "plqcd_hip_server_info": (C.c_int, [vp, C.POINTER(C.c_double)]),
    "plqcd_hip_server_destroy": (None, [vp]),
    "plqcd_hip_autotune": (C.c_int, [vp, vp, vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "plqcd_hip_autotune_clear": (C.c_int, []),
    "plqcd_hip_autotune_lookup": (C.c_int, [vp, vp, C.c_int, C.POINTER(C.c_int)]),
    "plqcd_hip_autotune_solver": (C.c_int, [vp, vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "plqcd_hip_hop_batch": (C.c_int, [C.c_int, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.c_int]),
    "plqcd_hip_hop_tuning": (C.c_int, [C.c_int] * 8),
    "plqcd_hip_hop_get_tuning": (C.c_int, [C.c_int, C.POINTER(C.c_int)]),
    "plqcd_hip_hop_tuning_reset": (C.c_int, [C.c_int]),
    "plqcd_hip_stream_bandwidth": (C.c_int, [C.c_int, C.c_size_t, C.c_int, C.POINTER(C.c_double)]),
    "plqcd_hip_host_alloc": (C.c_int, [C.c_size_t, C.POINTER(vp)]),
    "plqcd_hip_host_free": (None, [vp]),
    "plqcd_random_gauge_su3": (None, [vp, C.c_int, C.c_uint64]),
    "plqcd_random_spinor_gauss": (None, [vp, C.c_int, C.c_uint64]),
    "plqcd_unit_gauge": (None, [vp, C.c_int]),
}

_lib = None


class PlqcdError(RuntimeError):
    pass


def lib() -> C.CDLL:
    """The loaded library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise PlqcdError(f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _PROTOS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return lib().plqcd_hip_last_error().decode()


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        raise PlqcdError(f"{what} failed ({rc}): {last_error()}")


def header_functions() -> list[str]:
    """Names of all functions declared in include/*.h (for export checks)."""
    names = []
    for h in sorted(HEADER_DIR.glob("*.h")):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text, flags=re.M):
            if m.group(0).lstrip().startswith(("typedef", "#", "return")):
                continue
            names.append(m.group(1))
    return names


def header_variables() -> list[str]:
    out = []
    for h in sorted(HEADER_DIR.glob("*.h")):
        for m in re.finditer(r"^extern\s+struct\s+\w+\s+(\w+)\s*;", h.read_text(), flags=re.M):
            out.append(m.group(1))
    return out
