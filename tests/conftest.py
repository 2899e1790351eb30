"""Test configuration: the `gpu` marker and shared helpers."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def rel_err(got, ref64):
    """Normwise relative error max|got-ref| / max|ref| and the max abs error."""
    got = np.asarray(got, np.float64)
    ref64 = np.asarray(ref64, np.float64)
    ab = float(np.max(np.abs(got - ref64))) if got.size else 0.0
    den = float(np.max(np.abs(ref64))) if ref64.size else 1.0
    return ab / (den if den > 0 else 1.0), ab


@pytest.fixture(scope="session")
def gpu():
    """Skip-free GPU fixture: on a machine with a GPU the library must load and see a device."""
    import mdsm_amd as M
    n = M.device_count()
    if n < 1:
        pytest.fail("no HIP device visible to libmdsm_hip.so")
    M.lib()
    return M
