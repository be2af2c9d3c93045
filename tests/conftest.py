import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library through the C ABI)")
    # the oracle's 2D FFTs on up to 16 host threads (full-size checks at 2048^2)
    from oracle import fourierflows
    fourierflows.set_fft_workers(min(16, os.cpu_count() or 1))


@pytest.fixture(scope="session")
def gf():
    import __graft_entry__ as ge
    return ge.load_package()


def rel(a, b):
    import numpy as np
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
