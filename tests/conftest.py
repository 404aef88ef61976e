import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
STEP_CASES = ["lattice512", "random512", "compressed512"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs the libmfl.so kernels")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


def load_step_case(name):
    return dict(np.load(os.path.join(GOLDEN, f"step_{name}.npz")))


def force_errors(f, ref):
    """Normwise and floored per-particle relative errors of [N,>=3] force-like fields (xyz only)."""
    f = np.asarray(f, np.float64)[:, :3]
    ref = np.asarray(ref, np.float64)[:, :3]
    d = np.linalg.norm(f - ref, axis=1)
    mag = np.linalg.norm(ref, axis=1)
    rms = np.sqrt(np.mean(mag ** 2)) if mag.size else 0.0
    normwise = np.linalg.norm(f - ref) / max(np.linalg.norm(ref), 1e-300)
    per = np.max(d / (mag + rms + 1e-300)) if d.size else 0.0
    return normwise, per
