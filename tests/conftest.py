import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(TESTS, "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on the device)")


def golden_cases():
    return sorted(f[len("small_"):-len(".npz")] for f in os.listdir(GOLDEN) if f.startswith("small_"))


def load_golden(name):
    return dict(np.load(os.path.join(GOLDEN, f"small_{name}.npz"), allow_pickle=False))


def normwise_error(y, yref, rowptr, val, x, y0):
    """max|y - yref| / max(||A||_inf * ||x||_inf, ||y0||_inf, tiny)."""
    n = len(rowptr) - 1
    absrow = np.zeros(n)
    if len(val):
        np.add.at(absrow, np.repeat(np.arange(n), np.diff(rowptr)), np.abs(val.astype(np.float64)))
    anorm = absrow.max() if n else 0.0
    scale = max(anorm * float(np.abs(x).max() if len(x) else 0.0), float(np.abs(y0).max() if len(y0) else 0.0),
                1e-300)
    return float(np.abs(y.astype(np.float64) - yref.astype(np.float64)).max() / scale) if len(y) else 0.0
