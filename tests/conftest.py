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
# rotated block, relative to that row's largest value (a rotation keeps each row's norm, so
# the figure does not depend on the block): the oracle's p = 30 rotation of multipole blocks
# of charges at 0.6-0.95 of the radius against the 40-digit expansion of the rotated charges,
# largest over 10 random rotations (test_rotation_rows_against_multiprecision re-measures
# it).  It grows ~3-4x per row; rows 26-29 of a p = 30 rotation carry no correct digits (the
# reference's own rotations included).  The error depends on the row only, not on the order
# the rotation is built for.
ROT_ROW_REL = np.array([3e-16, 6e-16, 2e-15, 3e-15, 1e-14, 7e-14, 4e-13, 1e-12, 6e-12, 8e-12,
                        5e-11, 2e-10, 6e-10, 2e-9, 1e-8, 3e-8, 2e-7, 8e-7, 3e-6, 1e-5,
                        3e-5, 9e-5, 9e-4, 2e-3, 1e-2, 3e-2, 3e-1, 6e-1, 3.0, 2e1, 1e2, 5e2])


def rot_close(a, b, p, tol):
    """True when block a matches block b to `tol` relative to max|b|, allowing in row n ten
    times the measured rounding error of the rotation recursion relative to that row
    (ROT_ROW_REL)."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    scale = max(np.abs(b).max(), 1e-300)
    for n in range(p):
        s = slice(n * n, (n + 1) * (n + 1))
        if np.abs(a[s] - b[s]).max() > max(tol * scale, 10.0 * ROT_ROW_REL[n] * np.abs(b[s]).max()):
            return False
    return True
