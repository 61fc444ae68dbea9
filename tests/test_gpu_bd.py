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


def _qmul(a, b):
    w1, v1, w2, v2 = a[0], a[1:], b[0], b[1:]
    return np.r_[w1 * w2 - v1 @ v2, w1 * v2 + w2 * v1 + np.cross(v1, v2)]


def _qconj(q):
    return np.r_[q[0], -q[1:]]


def _qaxis(o):
    a = np.linalg.norm(o)
    return np.r_[np.cos(0.5 * a), np.sin(0.5 * a) * o / a]


def _rotvec(Q):
    """axis * angle of a unit quaternion (cos a/2, sin a/2 u)"""
    Q = Q / np.linalg.norm(Q)
    s = np.linalg.norm(Q[1:])
    return 2.0 * np.arctan2(s, Q[0]) * Q[1:] / s


def test_bd_step_moves_by_the_adjoint_forces(gpu):
    """PBSAM_BD_ADJOINT: one step of one trajectory moves molecule 1 by a F_1 and turns each
    molecule by r_i tau_i plus the random parts, with F and tau the adjoint's lab-frame -dE/dr
    and -dE/domega at the start poses (a = D_t dt FACT2, r_i = D_r,i dt IKbT COUL_K).  The same
    step without the flag (forces of pbsam_compute_forces, conj(q) F and the reference's torque)
    draws the same random numbers; those are recovered from it and the adjoint step predicted
    from them."""
    import pbsam_amd
    from pbsam_amd import synth
    mols = [synth.molecule(6, seed=s) for s in (61, 62)]
    mts = []
    for m in mols:
        g = pbsam_amd.MolType(gpu, m["cen"], m["rad"], m["chg_xyz"], m["chg_q"], 10, eps_in=4.0, rcen=m["rcen"],
                              npts=[400] * 6)
        g.selfpol(tol=0, max_rounds=0)
        g.values()
        mts.append(g)
    pos, quat = synth.place_pair(mols[0], mols[1], 2 * synth.ball_radius(6) + 6.0, seed=63)
    prm = pbsam_amd.System.bd_params(dtr=0.01, dr=(1e-4, 1e-4), dt_fixed=2.0, f_dist=1e9)

    def step(flags):
        B = pbsam_amd.System(gpu, mts, pos, quat, groups=[2])
        B.set_cutoffs(100.0, 100.0)
        B.set_flags(flags)
        B.bd_start(np.array([77], dtype=np.int64), place=False, params=prm)
        rc, st, n = B.bd_run(1)
        assert rc == 0 and st[0] == pbsam_amd.BD_RUNNING and n[0] == 1
        out = B.bd_get(0)
        B.close()
        return out

    A = step(pbsam_amd.EXACT_DERIVATIVE | pbsam_amd.BD_ADJOINT)
    Bm = step(pbsam_amd.EXACT_DERIVATIVE)
    S = pbsam_amd.System(gpu, mts, pos, quat, groups=[2])
    S.set_cutoffs(100.0, 100.0)
    S.set_flags(pbsam_amd.EXACT_DERIVATIVE)
    _, rc, fB, tB = S.compute_forces(tol=0.0, max_rounds=0)
    assert rc == 0
    S.solve_exact(tol=1e-16, max_rounds=2000)
    fA, tA, _ = S.adjoint_forces(tol=1e-16, max_rounds=2000)
    S.close()
    a = prm.dtr * 2.0 * prm.fact_tr
    r = [prm.dr[i] * 2.0 * prm.ikbt * prm.coul_k for i in range(2)]
    # translation of molecule 1: the same random vector in both runs
    v = Bm["pos"][1] - pos[1] - a * fB[1]
    pred = pos[1] + a * fA[1] + v
    print("translation: adjoint step vs prediction %.3e A, vs the plain step %.3e A"
          % (np.abs(A["pos"][1] - pred).max(), np.abs(A["pos"][1] - Bm["pos"][1]).max()), flush=True)
    assert np.abs(A["pos"][1] - pred).max() < 1e-10
    assert np.abs(A["pos"][1] - Bm["pos"][1]).max() > 1e-6
    assert np.all(A["pos"][0] == pos[0])
    # turns: Q = q_new conj(q_old) = qaxis(r tau + w), w the same random vector in both runs
    for i in range(2):
        w = _rotvec(_qmul(Bm["quat"][i], _qconj(quat[i]))) - r[i] * tB[i]
        qp = _qmul(_qaxis(r[i] * tA[i] + w), quat[i])
        qa = A["quat"][i]
        err = min(np.abs(qa - qp).max(), np.abs(qa + qp).max())
        print("turn of molecule %d: adjoint step vs prediction %.3e, vs the plain step %.3e"
              % (i, err, np.abs(qa - Bm["quat"][i]).max()), flush=True)
        assert err < 1e-12
        assert np.abs(qa - Bm["quat"][i]).max() > 1e-10
    for g in mts:
        g.close()


def test_bd_adjoint_batch_runs(gpu):
    """PBSAM_BD_ADJOINT on a batch of three trajectories for a few steps: every trajectory
    running, finite poses; without the exact-derivative flags the mode is refused"""
    import pbsam_amd
    from pbsam_amd import synth
    mols = [synth.molecule(5, seed=s) for s in (64, 65)]
    mts = []
    for m in mols:
        g = pbsam_amd.MolType(gpu, m["cen"], m["rad"], m["chg_xyz"], m["chg_q"], 10, eps_in=4.0, rcen=m["rcen"],
                              npts=[400] * 5)
        g.selfpol(tol=0, max_rounds=0)
        g.values()
        mts.append(g)
    d = 2 * synth.ball_radius(5) + 6.0
    poses = [synth.place_pair(mols[0], mols[1], d + k, seed=66 + k) for k in range(3)]
    B = pbsam_amd.System(gpu, mts * 3, np.vstack([p for p, _ in poses]), np.vstack([q for _, q in poses]),
                         groups=[2] * 3)
    B.set_cutoffs(100.0, 100.0)
    prm = pbsam_amd.System.bd_params(dtr=0.01, dr=(1e-4, 1e-4), dt_fixed=2.0, f_dist=1e9)
    B.set_flags(pbsam_amd.BD_ADJOINT)
    B.bd_start(np.arange(3, dtype=np.int64) + 5, place=False, params=prm)
    with pytest.raises(pbsam_amd.PbsamError) as e:
        B.bd_run(1)
    assert e.value.code == pbsam_amd.ERR_STATE
    B.set_flags(pbsam_amd.EXACT_DERIVATIVE | pbsam_amd.BD_ADJOINT)
    B.bd_start(np.arange(3, dtype=np.int64) + 5, place=False, params=prm)
    rc, st, n = B.bd_run(3)
    assert rc == 0 and np.all(st == pbsam_amd.BD_RUNNING) and np.all(n == 3)
    for g in range(3):
        o = B.bd_get(g)
        assert np.isfinite(o["pos"]).all() and np.isfinite(o["quat"]).all()
    B.close()
    for g in mts:
        g.close()
