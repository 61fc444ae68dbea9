import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("oracle", "pb-sam_amd"):
    path = os.path.join(ROOT, sub)
    if path not in sys.path:
        sys.path.insert(0, path)

KAPPA = 0.0737403009  # 0.05 M salt, eps 78, 298.15 K (selfp_makemat.cpp:59-60)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def orc():
    import pyoracle
    return pyoracle.Oracle(KAPPA)


@pytest.fixture(scope="session")
def golden():
    return np.load(os.path.join(ROOT, "tests", "golden", "ref_kat.npz"))


@pytest.fixture(scope="session")
def gpu():
    import pbsam_amd
    ctx = pbsam_amd.Context(KAPPA)
    yield ctx
    ctx.close()


def relerr(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / scale


def rot_close(a, b, p, tol):
    """True when block a matches block b to `tol` relative to max|b|, allowing in row n the
    rounding growth of the rotation recursion (3-4x per order): 1e-15 * 4^n."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    scale = max(np.abs(b).max(), 1e-300)
    for n in range(p):
        s = slice(n * n, (n + 1) * (n + 1))
        if np.abs(a[s] - b[s]).max() > max(tol, 1e-15 * 4.0 ** n) * scale:
            return False
    return True
