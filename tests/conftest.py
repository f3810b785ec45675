import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running")


def load_golden(name: str):
    return np.load(GOLDEN / name, allow_pickle=False)


def site_rel_err(a: np.ndarray, ref: np.ndarray) -> float:
    """max over sites of max_k |a_k - ref_k| / max_k |ref_k| over the 24 reals of a site."""
    V = ref.shape[0]
    ar = a.reshape(V, -1).view(np.float64)
    rr = ref.reshape(V, -1).view(np.float64)
    num = np.abs(ar - rr).max(axis=1)
    den = np.abs(rr).max(axis=1)
    den = np.where(den == 0, 1.0, den)
    return float((num / den).max())


@pytest.fixture(scope="session")
def golden():
    return load_golden


@pytest.fixture(scope="session")
def hp():
    from plqcd_amd import hopping
    hopping.lib()  # fail loudly if not built
    return hopping


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu_available():
    # counting devices does not initialise the GPU runtime in a way that matters here
    return os.path.exists("/dev/kfd")
