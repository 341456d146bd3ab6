import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def relerr(a, b):
    """max |a-b| / max |b| (0 when both vanish)."""
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / den) if den > 0 else float(np.abs(a - b).max())


@pytest.fixture(scope="session")
def gpu():
    """Fails (does not skip) when the HIP library or the GPU is missing: a GPU
    test that cannot reach the device is an error, never a silent pass."""
    import he_dft_2024_amd as H
    H.lib()
    return H
