"""The oracle's forces (restating polarize_mutual's gradient phase and
computeMol_Force_and_Torque) against central differences of its own interaction energy.

The reference multiplies the force by conj(orientation) before returning it
(moldynamics.cpp:545-546) although the sum it starts from is already lab-frame (the
gradient blocks differentiate with respect to lab-frame sphere positions), so the
lab-frame force is orientation * returned force.  That vector matches
-d<H_i, LS_i>/dr_i to about 1% of its size: the forward-mode derivative leaves out the
dependence of the truncation orders and of the surface quadrature on the positions.
The torque (c x f only, the surface term being disabled in the reference) is not a
derivative of the energy and is not checked here."""
import numpy as np
import pytest

from conftest import KAPPA


def quat_rot(q, v):
    w, u = q[0], q[1:]
    t = np.cross(u, v)
    return v + 2 * (w * t + np.cross(u, t))


@pytest.fixture(scope="module")
def pair(orc):
    import pyoracle
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pb-sam_amd"))
    from pbsam_amd import synth
    p = 10
    mols = [synth.molecule(6, seed=s) for s in (4, 5)]
    mts = []
    for m in mols:
        mt = pyoracle.MolType(orc, m["cen"], m["rad"], m["chg_xyz"], m["chg_q"], p, eps_in=4.0, rcen=m["rcen"],
                              npts=[400] * 6)
        mt.selfpol(0, 0.0, 0)
        mt.values()
        mts.append(mt)
    pos, quat = synth.place_pair(mols[0], mols[1], 2 * synth.ball_radius(6) + 6.0, seed=9)
    return mts, pos, quat


def solve(orc, mts, pos, quat):
    import pyoracle
    S = pyoracle.System(orc, mts, pos, quat)
    S.set_cutoffs(100.0, 100.0)
    rc, f, t = S.compute_forces(0, 1e-13, 5000)   # reference order, tight tolerance
    assert rc == 0
    return S, f, t


@pytest.mark.parametrize("mol", [0, 1])
def test_force_is_energy_gradient(orc, pair, mol):
    mts, pos, quat = pair
    S, f, _ = solve(orc, mts, pos, quat)
    F = quat_rot(quat[mol], f[mol])
    h = 1e-4
    fd = np.zeros(3)
    for d in range(3):
        pp, pm = pos.copy(), pos.copy()
        pp[mol, d] += h
        pm[mol, d] -= h
        fd[d] = -(solve(orc, mts, pp, quat)[0].pot(mol) - solve(orc, mts, pm, quat)[0].pot(mol)) / (2 * h)
    assert np.abs(F - fd).max() < 0.03 * np.linalg.norm(fd), (F, fd)
    # the returned vector itself is the rotated one: it does not match unless the
    # orientation is the identity
    assert np.abs(f[mol] - fd).max() > 0.1 * np.linalg.norm(fd)
