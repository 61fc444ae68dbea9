"""The oracle's computePotInSpace and computeTotalIntEnergy against the reference's own, compiled
from its sources (tests/golden/pot_ref.npz, written by tests/golden/make_pot_fixtures.py), and
the rotation-correct potential against what the mathematics fixes: it equals the reference's
at identity orientations and is invariant when the whole system and the points rotate
together (the reference's formula is not)."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
N = 15


def _qmul(a, b):
    w1, v1, w2, v2 = a[0], np.asarray(a[1:]), b[0], np.asarray(b[1:])
    return np.concatenate([[w1 * w2 - v1 @ v2], w1 * v2 + w2 * v1 + np.cross(v1, v2)])


def _qrot(q, r):
    w, v = q[0], np.asarray(q[1:])
    c = np.cross(v, r)
    return r + 2 * (w * c + np.cross(v, c))


@pytest.fixture(scope="module")
def setup(orc):
    import pyoracle
    fx = np.load(os.path.join(HERE, "golden", "pot_ref.npz"))
    mts = []
    for i in range(2):
        q = fx["mol%d_q" % i]
        o = pyoracle.MolType(orc, fx["mol%d_cen" % i], fx["mol%d_rad" % i], q[:, :3], q[:, 3], N, eps_in=4.0,
                             rcen=np.zeros(3))
        o.set_self(fx["mol%d_F" % i], fx["mol%d_H" % i])
        o.values()
        mts.append(o)
    return fx, mts


def _system(orc, mts, pose):
    import pyoracle
    pose = np.asarray(pose, float).reshape(2, 7)
    s = pyoracle.System(orc, mts, pose[:, :3], pose[:, 3:])
    s.set_cutoffs(10.0, 100.0)
    s.lists()
    return s


@pytest.mark.parametrize("k", [0, 1])
def test_reference_formula_matches_reference(orc, setup, k):
    fx, mts = setup
    s = _system(orc, mts, fx["pose%d" % k])
    np.testing.assert_allclose(s.pot_in_space(fx["pts"], ref_frame=True), fx["pot%d" % k], rtol=1e-13, atol=0)
    assert abs(s.total_int_energy() - fx["eint%d" % k]) <= 1e-13 * abs(fx["eint%d" % k])


def test_rotated_formula_equals_reference_at_identity(orc, setup):
    fx, mts = setup
    s = _system(orc, mts, fx["pose0"])
    np.testing.assert_allclose(s.pot_in_space(fx["pts"], ref_frame=False), fx["pot0"], rtol=1e-13, atol=0)


def test_rotated_formula_is_rotation_invariant(orc, setup):
    fx, mts = setup
    pose = fx["pose1"].reshape(2, 7).copy()
    pose[:, 3:] /= np.linalg.norm(pose[:, 3:], axis=1)[:, None]
    s = _system(orc, mts, pose)
    base = s.pot_in_space(fx["pts"], ref_frame=False)
    g = np.array([0.6, -0.2, 0.7, 0.33])
    g /= np.linalg.norm(g)
    rp = pose.copy()
    for i in range(2):
        rp[i, :3] = _qrot(g, pose[i, :3])
        rp[i, 3:] = _qmul(g, pose[i, 3:])
    s2 = _system(orc, mts, rp)
    pts2 = np.array([_qrot(g, x) for x in fx["pts"]])
    np.testing.assert_allclose(s2.pot_in_space(pts2, ref_frame=False), base, rtol=1e-11, atol=1e-18)
    # the reference's formula changes under the same rotation
    assert np.abs(s2.pot_in_space(pts2, ref_frame=True) - s.pot_in_space(fx["pts"], ref_frame=True)).max() > 1e-8
