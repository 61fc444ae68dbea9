"""Potential in space and total interaction energy (pbsam_potential, pbsam_total_int_energy)
against the oracle's computePotInSpace / computeTotalIntEnergy (pinned to the reference by
tests/test_oracle_pot.py): the reference's formula at any pose, the rotation-correct one at the
identity orientation, and its invariance under a rotation of the whole system."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _qmul(a, b):
    w1, v1, w2, v2 = a[0], np.asarray(a[1:]), b[0], np.asarray(b[1:])
    return np.concatenate([[w1 * w2 - v1 @ v2], w1 * v2 + w2 * v1 + np.cross(v1, v2)])


def _qrot(q, r):
    w, v = q[0], np.asarray(q[1:])
    c = np.cross(v, r)
    return r + 2 * (w * c + np.cross(v, c))


@pytest.fixture(scope="module")
def mols(gpu, orc):
    import pbsam_amd
    import pyoracle
    from pbsam_amd import synth
    ms = [synth.molecule(6, seed=s) for s in (61, 62)]
    gs, os_ = [], []
    for m in ms:
        args = (m["cen"], m["rad"], m["chg_xyz"], m["chg_q"], 10)
        kw = dict(eps_in=4.0, rcen=m["rcen"], npts=[400] * len(m["rad"]))
        g, o = pbsam_amd.MolType(gpu, *args, **kw), pyoracle.MolType(orc, *args, **kw)
        o.selfpol(0, 0.0, 0)
        F, H = o.FH()
        g.set_self(F, H)
        o.set_self(F, H)
        gs.append(g)
        os_.append(o)
    return ms, gs, os_


def _pair(gpu, orc, mols, pos, quat, solve):
    import pbsam_amd
    import pyoracle
    ms, gs, os_ = mols
    sg = pbsam_amd.System(gpu, gs, pos, quat)
    so = pyoracle.System(orc, os_, pos, quat)
    for s in (sg, so):
        s.set_cutoffs(100.0, 100.0)
    if solve:
        sg.solve(tol=0, max_rounds=0)
        so.lists()
        so.mutual(0, 0.0, 0)
    else:
        so.lists()
    return sg, so


@pytest.mark.parametrize("solve", [False, True])
def test_potential_matches_oracle(gpu, orc, mols, solve):
    from pbsam_amd import synth
    pos, quat = synth.place_pair(mols[0][0], mols[0][1], 30.0, seed=5)
    sg, so = _pair(gpu, orc, mols, pos, quat, solve)
    rng = np.random.default_rng(8)
    pts = rng.uniform(-30, 60, (200, 3))
    for ref in (True, False):
        g, o = sg.potential(pts, reference_frame=ref), so.pot_in_space(pts, ref_frame=ref)
        np.testing.assert_allclose(g, o, rtol=1e-9, atol=1e-9 * np.abs(o).max())


def test_total_int_energy(gpu, orc, mols):
    from pbsam_amd import synth
    pos, quat = synth.place_pair(mols[0][0], mols[0][1], 30.0, seed=6)
    sg, so = _pair(gpu, orc, mols, pos, quat, True)
    e_ref = so.total_int_energy()
    assert abs(sg.total_int_energy(reference_frame=True)[0] - e_ref) <= 1e-9 * abs(e_ref)
    # the rotation-correct energy: the reference's value at identity orientations ...
    ident = np.tile([1.0, 0.0, 0.0, 0.0], (2, 1))
    sg0, so0 = _pair(gpu, orc, mols, pos, ident, True)
    e0 = so0.total_int_energy()
    assert abs(sg0.total_int_energy()[0] - e0) <= 1e-9 * abs(e0)
    # ... and unchanged when the whole system rotates
    e1 = sg.total_int_energy()[0]
    g = np.array([0.3, -0.5, 0.4, 0.7])
    g /= np.linalg.norm(g)
    rpos = np.array([_qrot(g, x) for x in pos])
    rquat = np.array([_qmul(g, q / np.linalg.norm(q)) for q in quat])
    sgr, _ = _pair(gpu, orc, mols, rpos, rquat, True)
    assert abs(sgr.total_int_energy()[0] - e1) <= 1e-8 * abs(e1)


def test_batch_energies_per_system(gpu, orc, mols):
    import pbsam_amd
    from pbsam_amd import synth
    ms, gs, _ = mols
    P, Q = [], []
    for b in range(3):
        p_, q_ = synth.place_pair(ms[0], ms[1], 28.0 + 3 * b, seed=40 + b)
        P.append(p_)
        Q.append(q_)
    batch = pbsam_amd.System(gpu, [gs[0], gs[1]] * 3, np.vstack(P), np.vstack(Q), groups=[2] * 3)
    batch.set_cutoffs(100.0, 100.0)
    batch.solve(tol=0, max_rounds=0)
    eb = batch.total_int_energy()
    for b in range(3):
        s = pbsam_amd.System(gpu, gs, P[b], Q[b])
        s.set_cutoffs(100.0, 100.0)
        s.solve(tol=0, max_rounds=0)
        e = s.total_int_energy()[0]
        assert abs(eb[b] - e) <= 1e-10 * abs(e)
