"""The oracle (test infrastructure only).

ctypes wrappers of oracle/libhop_oracle.so, the plain-C restatement of the
reference's ASSYMBLY hopping matrix (lib/dirac.c:35-791) and geometry
(lib/start.c:406-526), plus a runner for oracle/_ref/hopping_ref (the
reference itself, compiled where its sources exist).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may use this package,
and only as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libhop_oracle.so"
REF_BIN = HERE / "_ref" / "hopping_ref"

MODE_COMPAT, MODE_WILSON = 0, 1

_lib = None


def build() -> None:
    """Compile the restatement (and the reference binary where its sources exist)."""
    subprocess.run(["make", "-s", "-B", "-C", str(HERE), "libhop_oracle.so"], check=True)
    if Path("/root/reference/lib/dirac.c").exists():
        subprocess.run(["make", "-s", "-B", "-C", str(HERE), "ref"], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} is not built (run `make -C oracle`)")
        L = C.CDLL(str(LIB))
        vp, ip = C.c_void_p, C.POINTER(C.c_int)
        L.oracle_geometry.argtypes = [C.c_int * 4, ip, ip, ip, ip]
        L.oracle_geometry.restype = C.c_int
        for n in ("oracle_hopping_eo", "oracle_hopping_oe", "oracle_hopping_full"):
            getattr(L, n).argtypes = [vp, vp, vp, C.c_int * 4, C.c_int, C.c_int]
            getattr(L, n).restype = None
        L.oracle_bench_setup.argtypes = [C.c_int * 4, C.c_int]
        L.oracle_bench_setup.restype = C.c_int
        L.oracle_bench_run.argtypes = [vp, vp, vp, C.c_int, C.c_int]
        L.oracle_bench_run.restype = C.c_double
        L.oracle_bench_teardown.argtypes = []
        _lib = L
    return _lib


def _args(qin, u, L):
    L = tuple(int(x) for x in L)
    V = int(np.prod(L))
    assert qin.dtype == np.complex128 and qin.shape == (V, 4, 3) and qin.flags.c_contiguous
    assert u.dtype == np.complex128 and u.shape == (V, 4, 3, 3) and u.flags.c_contiguous
    return (C.c_int * 4)(*L), V


def geometry(L):
    """(ipt_eo, ipt_lex, iup (V,4), idn (V,4)) as lib/start.c:406-526."""
    L4 = (C.c_int * 4)(*L)
    V = int(np.prod(L))
    ipt_eo = np.zeros(V, np.int32)
    ipt_lex = np.zeros(V, np.int32)
    iup = np.zeros((V, 4), np.int32)
    idn = np.zeros((V, 4), np.int32)
    ip = C.POINTER(C.c_int)
    rc = lib().oracle_geometry(L4, ipt_eo.ctypes.data_as(ip), ipt_lex.ctypes.data_as(ip),
                               iup.ctypes.data_as(ip), idn.ctypes.data_as(ip))
    if rc != 0:
        raise ValueError(f"bad lattice {L}")
    return ipt_eo, ipt_lex, iup, idn


def _run(fn, qin, u, L, mode, qout=None, nthreads=0):
    L4, V = _args(qin, u, L)
    out = np.zeros((V, 4, 3), np.complex128) if qout is None else qout
    fn(qin.ctypes.data, out.ctypes.data, u.ctypes.data, L4, mode, nthreads)
    return out


def hopping_eo(qin, u, L, mode=MODE_COMPAT, qout=None, nthreads=0):
    return _run(lib().oracle_hopping_eo, qin, u, L, mode, qout, nthreads)


def hopping_oe(qin, u, L, mode=MODE_COMPAT, qout=None, nthreads=0):
    return _run(lib().oracle_hopping_oe, qin, u, L, mode, qout, nthreads)


def hopping_full(qin, u, L, mode=MODE_COMPAT, qout=None, nthreads=0):
    return _run(lib().oracle_hopping_full, qin, u, L, mode, qout, nthreads)


def bench(qin, u, L, napply, mode=MODE_COMPAT, nthreads=0) -> float:
    """Wall seconds of `napply` eo+oe applications (scratch set up beforehand)."""
    L4, V = _args(qin, u, L)
    out = np.zeros((V, 4, 3), np.complex128)
    if lib().oracle_bench_setup(L4, nthreads) != 0:
        raise RuntimeError("oracle_bench_setup failed")
    try:
        return lib().oracle_bench_run(qin.ctypes.data, out.ctypes.data, u.ctypes.data, mode, napply)
    finally:
        lib().oracle_bench_teardown()


# ----------------------------------------------------------- the reference

def ref_available() -> bool:
    return REF_BIN.exists()


def write_input_file(path: Path, L, nthread: int = 1, align: int = 6) -> None:
    Path(path).write_text(
        f"L0 {L[0]}\nL1 {L[1]}\nL2 {L[2]}\nL3 {L[3]}\n"
        "NPROC0 1\nNPROC1 1\nNPROC2 1\nNPROC3 1\n"
        f"NTHREAD {nthread}\nALIGN {align}\n")


def ref_hop(qin, u, L, workdir: Path, nmul: int = 1, nthread: int = 1):
    """Output of the compiled reference (eo then oe, `nmul` times) on (qin, u)."""
    workdir = Path(workdir)
    inp, fin, fout = workdir / "ref.input", workdir / "ref_in.bin", workdir / "ref_out.bin"
    write_input_file(inp, L, nthread)
    with open(fin, "wb") as f:
        f.write(qin.tobytes())
        f.write(u.tobytes())
    subprocess.run([str(REF_BIN), str(inp), "hop", str(fin), str(fout), str(nmul)], check=True,
                   capture_output=True, timeout=600)
    V = int(np.prod(L))
    return np.fromfile(fout, dtype=np.complex128).reshape(V, 4, 3)
