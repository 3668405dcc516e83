import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "marineseis-fdm-2d_amd")
for p in (PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libmsfdm.so")


def rel_l2(a, b):
    """||a - b||_2 / ||b||_2 in float64."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))


@pytest.fixture(scope="session")
def gpu_available():
    import msfdm
    try:
        import ctypes
        msfdm.lib()
    except OSError as e:   # pragma: no cover
        pytest.fail(f"libmsfdm.so does not load: {e}")
    return True
