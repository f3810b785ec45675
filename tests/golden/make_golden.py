"""Writes the golden fixtures of tests/golden/ from the compiled reference.

Runs in the build container only (it needs oracle/_ref/hopping_ref, which is
compiled from the reference sources by `make -C oracle ref`).  Inputs come
from libplqcd_hip.so's seeded generator (host code) or from the reference
harness's own rand() fill; outputs come from the reference's ASSYMBLY
eo+oe routines (lib/dirac.c:35-791).  Fixtures are data only: inputs,
checksums and outputs.

usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import gzip
import hashlib
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from plqcd_amd import hopping as hp  # noqa: E402

OUT = Path(__file__).resolve().parent
RAND_MAX = 2147483647

# (name, L, gauge seed, spinor seed, store full output)
SEEDED = [
    ("seed_4x4x4x4", (4, 4, 4, 4), 11, 12, True),
    ("seed_8x4x6x4", (8, 4, 6, 4), 21, 22, True),
    ("seed_8x8x8x16", (8, 8, 8, 16), 1, 2, False),
]
GEOMETRY = [(4, 4, 4, 4), (8, 4, 6, 4)]
N_SAMPLE = 64


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def sample_sites(V: int, n: int = N_SAMPLE) -> np.ndarray:
    return np.unique(np.linspace(0, V - 1, n).astype(np.int64))


def run_ref(args, cwd):
    subprocess.run([str(oracle.REF_BIN), *map(str, args)], cwd=cwd, check=True, capture_output=True, timeout=600)


def main():
    if not oracle.ref_available():
        sys.exit("oracle/_ref/hopping_ref is missing: run `make -C oracle ref` first")
    with tempfile.TemporaryDirectory() as td:
        td = Path(td)
        # geometry dumps (tests/test_geometry.c)
        for L in GEOMETRY:
            inp = td / "g.input"
            oracle.write_input_file(inp, L, nthread=1, align=6)
            run_ref([inp, "geometry", td / "geo.txt"], td)
            name = "geometry_%dx%dx%dx%d.txt.gz" % L
            (OUT / name).write_bytes(gzip.compress((td / "geo.txt").read_bytes(), compresslevel=9, mtime=0))
            print("wrote", name)

        # the harness's own fill (tests/test_hopping.c:178-254), 4^4
        L = (4, 4, 4, 4)
        V = 256
        inp = td / "h.input"
        oracle.write_input_file(inp, L)
        run_ref([inp, "harness", td / "h_in.bin", td / "h_out.bin"], td)
        raw = np.fromfile(td / "h_in.bin", dtype=np.float64)
        ints = np.rint(raw * RAND_MAX).astype(np.uint32)
        assert np.array_equal(ints / float(RAND_MAX), raw), "rand() integers do not round-trip"
        out = np.fromfile(td / "h_out.bin", dtype=np.complex128).reshape(V, 4, 3)
        np.savez_compressed(OUT / "harness_4x4x4x4.npz", L=np.array(L), rand_ints=ints,
                            rand_max=np.int64(RAND_MAX), out_compat=out)
        print("wrote harness_4x4x4x4.npz")

        # seeded random SU(3) links and Gaussian spinors
        for name, L, gseed, sseed, full in SEEDED:
            V = int(np.prod(L))
            u = hp.random_gauge_su3(V, gseed)
            q = hp.random_spinor(V, sseed)
            out = oracle.ref_hop(q, u, L, td, nmul=1)
            out100 = oracle.ref_hop(q, u, L, td, nmul=100)
            assert np.array_equal(out, out100), "repeated applications differ"
            rec = dict(L=np.array(L), gauge_seed=np.uint64(gseed), spinor_seed=np.uint64(sseed),
                       sha_gauge=np.array(sha(u)), sha_spinor=np.array(sha(q)),
                       site_norm=np.sqrt((np.abs(out) ** 2).sum(axis=(1, 2))),
                       sample_idx=sample_sites(V), sha_out=np.array(sha(out)),
                       repeat100_equal=np.bool_(True))
            rec["sample_out"] = out[rec["sample_idx"]]
            if full:
                rec["out_compat"] = out
            np.savez_compressed(OUT / f"{name}.npz", **rec)
            print("wrote", name)


if __name__ == "__main__":
    main()
