import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libabhsf.so on device 0)")
    config.addinivalue_line("markers", "slow: full-size workload (seconds to minutes)")


@pytest.fixture(scope="session")
def gpu():
    from abhsf_amd import hip
    n = hip.device_count()
    if n < 1:
        pytest.fail("no GPU visible: gpu tests must run on the MI355X box")
    return hip


def read_mtx(path):
    """Coordinate Matrix Market reader for the golden fixtures (1-based -> 0-based)."""
    import numpy as np
    with open(path) as f:
        header = f.readline()
        assert header.startswith("%%MatrixMarket")
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        m, n, nnz = (int(t) for t in line.split())
        rows, cols, vals = [], [], []
        for _ in range(nnz):
            t = f.readline().split()
            rows.append(int(t[0]) - 1)
            cols.append(int(t[1]) - 1)
            vals.append(float(t[2]) if len(t) > 2 else 1.0)
    return m, n, np.array(rows), np.array(cols), np.array(vals)
