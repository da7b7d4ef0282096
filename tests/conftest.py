"""Test configuration: the `gpu` marker and shared helpers."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import pyoracle
    if not os.path.exists(pyoracle.LIB_PATH):
        import subprocess
        subprocess.check_call(["bash", os.path.join(ROOT, "oracle", "build.sh")])
    return pyoracle.lib()


def rel_err(a, b):
    """max |a-b| / max |b| (normwise relative error, infinity norm)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))
