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


# Rounding error of the reference's rotation recursion (rotcoeff.cpp:398-441) in row n of a
# rotated block, relative to the block's largest value: the oracle's rotation of multipole
# blocks of charges at 0.6-0.9 of the radius against the 40-digit expansion of the rotated
# charges, largest over 8 random rotations (tests/test_oracle_math.py,
# test_rotation_rows_against_multiprecision, re-measures it).  The error does not depend on
# the pole order the rotation is built for, only on the row.
ROT_ROW_ERR = np.array([3e-16, 3e-15, 4e-15, 4e-15, 6e-15, 2e-14, 6e-14, 2e-13, 2e-13, 5e-13,
                        2e-12, 8e-12, 3e-11, 2e-10, 6e-10, 2e-9, 2e-9, 7e-9, 4e-8, 2e-7,
                        5e-7, 9e-7, 2e-6, 2e-5, 2e-5, 6e-5, 4e-4, 8e-4, 2e-3, 2e-2, 6e-2, 2e-1])


def rot_close(a, b, p, tol):
    """True when block a matches block b to `tol` relative to max|b|, allowing in row n ten
    times the measured rounding error of the rotation recursion (ROT_ROW_ERR)."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    scale = max(np.abs(b).max(), 1e-300)
    for n in range(p):
        s = slice(n * n, (n + 1) * (n + 1))
        if np.abs(a[s] - b[s]).max() > max(tol, 10.0 * ROT_ROW_ERR[n]) * scale:
            return False
    return True
