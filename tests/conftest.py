import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP step through the C ABI)")


def rel_linf(a, b):
    """Per-field relative L-infinity: max|a - b| / max|b| (b = oracle)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    out = []
    for k in range(a.shape[0]):
        den = np.abs(b[k]).max()
        out.append(float(np.abs(a[k] - b[k]).max() / den) if den > 0 else float(np.abs(a[k]).max()))
    return out


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.lib()   # raises if not built
    return pyoracle
