"""Brownian dynamics on batches (pbsam_bd_*) against the oracle's CBDnam restatement, which
itself matches the reference's own makeMove bit for bit (tests/test_oracle_bd.py).

* trajectories that never come within f_dist (no forces): every pose, stored rotation, time
  step and status equal to the oracle's, bit for bit, after hundreds of steps;
* trajectories near contact with forces from the device (BDnam's cutoffs 10 / 100 A): poses
  within 1e-9 A (forces differ from the oracle's at the 1e-12 level), statuses and step counts
  equal, after several steps with collision retries and docking criteria.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _moltypes(gpu, orc, n_sph, p, npts, seeds):
    import pbsam_amd
    import pyoracle
    from pbsam_amd import synth
    mols = [synth.molecule(n_sph[i], seed=s) for i, s in enumerate(seeds)]
    gs, os_ = [], []
    for m in mols:
        args = (m["cen"], m["rad"], m["chg_xyz"], m["chg_q"], p)
        kw = dict(eps_in=4.0, rcen=m["rcen"], npts=[npts] * len(m["rad"]))
        g, o = pbsam_amd.MolType(gpu, *args, **kw), pyoracle.MolType(orc, *args, **kw)
        o.selfpol(0, 0.0, 0)
        F, H = o.FH()
        g.set_self(F, H)
        o.set_self(F, H)
        gs.append(g)
        os_.append(o)
    return mols, gs, os_


def _geom(m):
    return (m["cen"] - m["rcen"], m["rad"])


def _oracle_prm(prm):
    return [prm.dtr, prm.dr[0], prm.dr[1], prm.b_dist, prm.q_dist, prm.f_dist, prm.fact_tr, prm.ikbt, prm.coul_k,
            prm.dt_fixed, prm.max_tries]


def test_bd_free_diffusion_bitwise(gpu, orc):
    """No forces (f_dist = 0: every rotation is stored, as beyond f_DIST): the library's host
    moves equal the oracle's exactly, including the random start, translation retries and
    escapes.  (With a small positive f_dist, stored rotations applied on entry can overlap the
    molecules, as in the reference; the force evaluation then fails, PBSAM_BD_FAILED.)"""
    import pbsam_amd
    import pyoracle
    mols, gs, _ = _moltypes(gpu, orc, (6, 5), 6, 200, seeds=(71, 72))
    G = 8
    seeds = [1000 + 17 * g for g in range(G)]
    pos = np.zeros((2 * G, 3))
    pos[1::2, 0] = 100.0
    quat = np.tile([1.0, 0.0, 0.0, 0.0], (2 * G, 1))
    batch = pbsam_amd.System(gpu, [gs[0], gs[1]] * G, pos, quat, groups=[2] * G)
    batch.set_cutoffs(10.0, 100.0)
    prm = pbsam_amd.System.bd_params(dtr=6.0, dr=(4e-3, 3e-3), f_dist=0.0)
    batch.bd_start(seeds, place=True, params=prm)
    nsteps = 300
    rc, st, steps = batch.bd_run(nsteps)
    assert rc == pbsam_amd.OK
    for g in range(G):
        b = pyoracle.BDTrajectory([_geom(mols[0]), _geom(mols[1])], seeds[g], True, prm=_oracle_prm(prm))
        for _ in range(nsteps):
            if b.step(np.zeros(9)) != 2:
                break
        s, got = b.state(), batch.bd_get(g)
        assert steps[g] == s["steps"]
        np.testing.assert_array_equal(got["pos"], s["pos"])
        np.testing.assert_array_equal(got["quat"], s["quat"])
        np.testing.assert_array_equal(got["rot"], s["rot"])
        assert got["dt"] == s["dt"]
    assert np.any(st == pbsam_amd.BD_ESCAPED) or np.all(steps == nsteps)


def test_bd_with_forces_matches_oracle(gpu, orc):
    """Near contact, forces and torques from the device each step (BDnam's cutoffs)."""
    import pbsam_amd
    import pyoracle
    from pbsam_amd import synth
    p = 6
    mols, gs, os_ = _moltypes(gpu, orc, (6, 5), p, 200, seeds=(81, 82))
    G = 4
    d0 = synth.ball_radius(6) + synth.ball_radius(5) + 3.0
    poses = []
    for g in range(G):
        pp, qq = synth.place_pair(mols[0], mols[1], d0 + 1.5 * g, seed=300 + g)
        poses.append((pp, qq))
    pos = np.vstack([pp for pp, _ in poses])
    quat = np.vstack([qq for _, qq in poses])
    batch = pbsam_amd.System(gpu, [gs[0], gs[1]] * G, pos, quat, groups=[2] * G)
    batch.set_cutoffs(10.0, 100.0)
    prm = pbsam_amd.System.bd_params(dtr=0.2, dr=(2e-3, 2e-3))
    seeds = [77 + g for g in range(G)]
    batch.bd_start(seeds, place=False, params=prm)
    defs = [(2, [(0, 0), (1, 1), (2, 2)], [1.5, 1.5, 1.5])]
    batch.bd_contacts(defs)
    nsteps = 25
    rc, st, steps = batch.bd_run(nsteps)
    assert rc in (pbsam_amd.OK, pbsam_amd.ERR_NOCONV)
    for g in range(G):
        osys = pyoracle.System(orc, [os_[0], os_[1]], poses[g][0], poses[g][1])
        osys.set_cutoffs(10.0, 100.0)
        b = pyoracle.BDTrajectory([_geom(mols[0]), _geom(mols[1])], seeds[g], False, poses[g][0], poses[g][1],
                                  prm=_oracle_prm(prm))
        b.attach(osys)
        for nc, pairs, dist in defs:
            b.add_contact(nc, pairs, dist)
        last = 2
        for _ in range(nsteps):
            last = b.step()
            if last != 2:
                break
        s, got = b.state(), batch.bd_get(g)
        assert int(st[g]) == last and steps[g] == s["steps"], (g, st[g], last, steps[g], s["steps"])
        np.testing.assert_allclose(got["pos"], s["pos"], rtol=0, atol=1e-9)
        np.testing.assert_allclose(got["quat"], s["quat"], rtol=0, atol=1e-10)
        np.testing.assert_allclose(got["rot"], s["rot"], rtol=0, atol=1e-10)
