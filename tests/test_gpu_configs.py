"""The benchmark's workloads on the device, at the sizes they are measured at.

* Two molecules of 60 spheres at p = 30 with 1800 points per sphere, centres 50 A apart,
  random orientations, inter-polarization cutoff 100 A (the benchmark's force evaluation):
  forces, torques and energies against the oracle in the reference's order and stock
  tolerances.  (Central differences of the exact-derivative mode need the mutual solve
  converged to ~1e-16, which the reference's Gauss-Seidel schedule does not reach at this
  size within 1500 rounds; they are checked at 2 x 6 spheres, p = 10, and 2 x 3 spheres,
  p = 30, in tests/test_gpu_exact.py.)
* Brownian dynamics of two 100-sphere molecules at p = 10 with 400 points per sphere
  (D_t = 0.01 A^2/ps, D_r = 1e-4 rad^2/ps, 2 ps steps, cutoffs 10 / 100 A, forces every
  step): a few trajectories against the oracle's CBDnam restatement.
* Eight molecules in one system (the large-system workload, reduced): 8 x 4 spheres at
  p = 10 against the oracle; 8 x 60 spheres at p = 30 with 2000 points per sphere, checked
  by what a rigid motion of the whole system must leave unchanged (energies, returned
  forces and torques are invariant under a global rotation and translation).
* The large-system workload at full size: eight distinct 500-sphere molecules at p = 30
  with 2000 points per sphere on a grid with 10 A gaps (about 200 GB of device memory;
  operators past the store cap are made per use), by the same invariances; there also the
  adjoint gradient of all 48 rigid-body coordinates (exact mode), by its covariance under a
  rigid motion and central differences of one force and one torque component.
"""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu


def quat_mul(a, b):
    w1, x1, y1, z1 = a
    w2, x2, y2, z2 = b
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


def quat_rot(q, v):
    w, u = q[0], q[1:]
    t = np.cross(u, v)
    return v + 2 * (w * t + np.cross(u, t))


def self_polarized(gpu, orc, mols, p, npts, oracle=True):
    import pbsam_amd
    import pyoracle
    gs, os_ = [], []
    for m in mols:
        n = len(m["rad"])
        args = (m["cen"], m["rad"], m["chg_xyz"], m["chg_q"], p)
        kw = dict(eps_in=4.0, rcen=m["rcen"], npts=[npts] * n)
        g = pbsam_amd.MolType(gpu, *args, **kw)
        g.selfpol(tol=0, max_rounds=0)
        g.values()
        gs.append(g)
        if oracle:
            o = pyoracle.MolType(orc, *args, **kw)
            o.set_self(*g.FH())
            o.values()
            os_.append(o)
    return gs, os_


@pytest.fixture(scope="module")
def two_by_60(gpu, orc):
    from pbsam_amd import synth
    mols = [synth.molecule(60, seed=s) for s in (101, 102)]
    gs, os_ = self_polarized(gpu, orc, mols, 30, 1800)
    pos, quat = synth.place_pair(mols[0], mols[1], 50.0, seed=7)
    return gs, os_, pos, quat


def test_two_by_60_p30_against_oracle(gpu, orc, two_by_60):
    import pbsam_amd
    import pyoracle
    gs, os_, pos, quat = two_by_60
    S = pbsam_amd.System(gpu, gs, pos, quat)
    S.set_cutoffs(100.0, 100.0)
    rounds, rc, fg, tg = S.compute_forces()
    assert rc == 0
    eg = S.energy()
    So = pyoracle.System(orc, os_, pos, quat)
    So.set_cutoffs(100.0, 100.0)
    rco, fo, to = So.compute_forces(0, 0.0, 0)
    assert rco == 0
    # measured 4.7e-9 (forces) and 3.0e-9 (torques); the bar is 1e-6
    assert relerr(fg, fo) < 1e-7, (fg, fo)
    assert relerr(tg, to) < 1e-7, (tg, to)
    assert relerr(eg, [So.pot(0), So.pot(1)]) < 1e-9


def test_bd_two_by_100_p10_against_oracle(gpu, orc):
    import pbsam_amd
    import pyoracle
    from pbsam_amd import synth
    mols = [synth.molecule(100, seed=s) for s in (301, 302)]
    gs, os_ = self_polarized(gpu, orc, mols, 10, 400)
    G, nsteps = 3, 3
    rng = np.random.default_rng(5007)
    d0 = 2 * synth.ball_radius(100) + 6.0
    poses = [synth.place_pair(mols[0], mols[1], d0 + rng.uniform(0.0, 10.0), seed=5100 + g) for g in range(G)]
    batch = pbsam_amd.System(gpu, [gs[0], gs[1]] * G, np.vstack([a for a, _ in poses]),
                             np.vstack([b for _, b in poses]), groups=[2] * G)
    batch.set_cutoffs(10.0, 100.0)
    prm = pbsam_amd.System.bd_params(dtr=0.01, dr=(1e-4, 1e-4), dt_fixed=2.0, f_dist=1e9)
    seeds = np.arange(G, dtype=np.int64) + 424242
    batch.bd_start(seeds, place=False, params=prm)
    rc, st, steps = batch.bd_run(nsteps)
    assert rc in (pbsam_amd.OK, pbsam_amd.ERR_NOCONV)
    oprm = [prm.dtr, prm.dr[0], prm.dr[1], prm.b_dist, prm.q_dist, prm.f_dist, prm.fact_tr, prm.ikbt, prm.coul_k,
            prm.dt_fixed, prm.max_tries]
    for g in range(G):
        osys = pyoracle.System(orc, os_, poses[g][0], poses[g][1])
        osys.set_cutoffs(10.0, 100.0)
        b = pyoracle.BDTrajectory([(m["cen"] - m["rcen"], m["rad"]) for m in mols], int(seeds[g]), False,
                                  poses[g][0], poses[g][1], prm=oprm)
        b.attach(osys)
        last = 2
        for _ in range(nsteps):
            last = b.step()
            if last != 2:
                break
        s, got = b.state(), batch.bd_get(g)
        assert int(st[g]) == last and steps[g] == s["steps"]
        np.testing.assert_allclose(got["pos"], s["pos"], rtol=0, atol=1e-9)
        np.testing.assert_allclose(got["quat"], s["quat"], rtol=0, atol=1e-10)


def eight(mols, gap, seed):
    """eight molecules on the corners of a cube, random orientations, no overlaps"""
    from pbsam_amd import synth
    R = max(synth.ball_radius(len(m["rad"])) for m in mols)
    side = 2 * R + gap
    pos = np.array([[x, y, z] for x in (0, side) for y in (0, side) for z in (0, side)], float)
    rng = np.random.default_rng(seed)
    quat = np.array([synth.random_quat(rng) for _ in range(8)])
    return pos, quat


def test_eight_molecules_p10_against_oracle(gpu, orc):
    import pbsam_amd
    import pyoracle
    from pbsam_amd import synth
    mols = [synth.molecule(4, seed=900 + i) for i in range(8)]
    gs, os_ = self_polarized(gpu, orc, mols, 10, 300)
    pos, quat = eight(mols, 10.0, 3)
    S = pbsam_amd.System(gpu, gs, pos, quat)
    S.set_cutoffs(100.0, 100.0)
    _, rc, fg, tg = S.compute_forces()
    assert rc == 0
    So = pyoracle.System(orc, os_, pos, quat)
    So.set_cutoffs(100.0, 100.0)
    rco, fo, to = So.compute_forces(0, 0.0, 0)
    assert rco == 0
    assert relerr(fg, fo) < 1e-8 and relerr(tg, to) < 1e-8
    assert relerr(S.energy(), [So.pot(i) for i in range(8)]) < 1e-9


def test_eight_by_60_p30_rigid_motion_invariance(gpu, orc):
    """Eight 60-sphere molecules (two types) at p = 30 with 2000 points per sphere, >= 10 A
    gaps: rotating and translating the whole system leaves every molecule's energy and
    returned force unchanged (the returned force is in the molecule frame), translating it
    leaves the returned torque unchanged too.
    After a fixed number of simultaneous rounds (no stopping decisions) the energies agree to
    rounding.  The forces come from the reference's schedule with its stock stopping rule
    (mean squared relative change <= 1e-4), whose selective recomputes can go differently on
    inputs that differ only by rounding: measured 3.7e-5, held to 1e-3."""
    import pbsam_amd
    from pbsam_amd import synth
    types = [synth.molecule(60, seed=s) for s in (101, 102)]
    gs, _ = self_polarized(gpu, orc, types, 30, 2000, oracle=False)
    mols = [types[i % 2] for i in range(8)]
    pos, quat = eight(mols, 10.0, 11)
    S = pbsam_amd.System(gpu, [gs[i % 2] for i in range(8)], pos, quat)
    S.set_cutoffs(100.0, 100.0)
    S.rounds(2)
    j0 = S.energy()
    _, rc, f0, t0 = S.compute_forces()
    assert rc == 0
    e0 = S.energy()
    Q = synth.random_quat(np.random.default_rng(12))
    shift = np.array([13.0, -7.0, 21.0])
    pos2 = np.array([quat_rot(Q, x) for x in pos]) + shift
    quat2 = np.array([quat_mul(Q, q) for q in quat])
    S2 = pbsam_amd.System(gpu, [gs[i % 2] for i in range(8)], pos2, quat2)
    S2.set_cutoffs(100.0, 100.0)
    S2.rounds(2)
    assert relerr(S2.energy(), j0) < 1e-9
    _, rc2, f1, t1 = S2.compute_forces()
    assert rc2 == 0
    assert relerr(S2.energy(), e0) < 1e-4
    assert relerr(f1, f0) < 1e-3, (f1, f0)
    # the reference's torque sums c x f with c the molecule-frame sphere centre and f the
    # lab-frame force (moldynamics.cpp:540), so it is not invariant under a rotation of the
    # whole system; under a translation it is
    assert relerr(t1, t0) > 1e-2
    S3 = pbsam_amd.System(gpu, [gs[i % 2] for i in range(8)], pos + shift, quat)
    S3.set_cutoffs(100.0, 100.0)
    _, rc3, f3, t3 = S3.compute_forces()
    assert rc3 == 0 and relerr(f3, f0) < 1e-3 and relerr(t3, t0) < 1e-3, (t3, t0)
    assert np.abs(f0).max() > 0


def test_eight_by_500_p30_full_size():
    """Eight distinct 500-sphere molecules at p = 30, 2000 points per sphere, packing balls
    10 A apart, spheres of different molecules within 20 A polarized, interactions to 100 A:
    the whole set-up and the force evaluation fit one device with part of the far-pair
    translation matrices and all numerical maps made per use (the store cap of a quarter of
    the device's memory is reached), and a rigid motion of the whole system leaves the
    energies after two simultaneous rounds unchanged to 1e-9 and the returned forces (stock
    schedule) to 1e-3; the adjoint forces and torques of all eight molecules turn with it
    and match central differences of the solved energy.  Most spheres here interact without being polarized; their forces are
    rotation covariant with PBSAM_SELF_FRAME (the reference's pairing of the unrotated self
    H with a lab-frame gradient is not, tests/test_oracle_forces.py), which is set."""
    import pbsam_amd
    from pbsam_amd import synth
    from conftest import KAPPA
    ctx = pbsam_amd.Context(KAPPA)
    gs = []
    for i in range(8):
        m = synth.molecule(500, seed=7001 + i)
        g = pbsam_amd.MolType(ctx, m["cen"], m["rad"], m["chg_xyz"], m["chg_q"], 30, eps_in=4.0, rcen=m["rcen"],
                              npts=[2000] * 500)
        g.selfpol(tol=0, max_rounds=0)
        g.values()
        gs.append(g)
    stored, limit = ctx.stored()
    assert 0 < stored <= limit
    a = 2 * synth.ball_radius(500) + 10.0
    pos = np.array([[a * (i & 1), a * ((i >> 1) & 1), a * ((i >> 2) & 1)] for i in range(8)], dtype=float)
    rng = np.random.default_rng(7000)
    quat = np.array([synth.random_quat(rng) for _ in range(8)])
    Q = synth.random_quat(np.random.default_rng(12))
    shift = np.array([13.0, -7.0, 21.0])
    pos2 = np.array([quat_rot(Q, x) for x in pos]) + shift
    quat2 = np.array([quat_mul(Q, q) for q in quat])
    res = []
    for P_, Q_ in ((pos, quat), (pos2, quat2)):
        S = pbsam_amd.System(ctx, gs, P_, Q_)
        S.set_cutoffs(20.0, 100.0)
        S.set_flags(pbsam_amd.SELF_FRAME)
        S.rounds(2)
        n_inter, n_pol = S.counts()
        assert n_inter > 0 and n_pol > 0
        j = S.energy()
        _, rc, f, t = S.compute_forces()
        assert rc == 0
        res.append((j, f))
        S.close()
    assert relerr(res[1][0], res[0][0]) < 1e-9
    assert np.abs(res[0][1]).max() > 0 and np.isfinite(res[0][1]).all()
    assert relerr(res[1][1], res[0][1]) < 1e-3
    # The adjoint gradient of every molecule's energy with respect to all 48 rigid-body
    # coordinates (one transposed solve per molecule), interactions cut at the polarization
    # cutoff: the exact-mode fixed point by GMRES, then the adjoint.  It needs the solution
    # only (no iterates kept).  The rigid motion turns the lab-frame forces and torques with
    # it, and central differences of molecule 0's solved energy give one force and one torque
    # component.
    from test_gpu_torque import turned

    def exact(P_, Q_):
        S = pbsam_amd.System(ctx, gs, P_, Q_)
        S.set_cutoffs(20.0, 20.0)
        S.set_flags(pbsam_amd.EXACT_DERIVATIVE)
        n, rc = S.solve_exact(tol=1e-24, max_rounds=3000)
        assert rc == 0, n
        return S

    S = exact(pos, quat)
    f0, t0, napp = S.adjoint_forces(tol=1e-20, max_rounds=5000)
    S.close()
    print("8 x 500 adjoint: applications over eight solves", napp, flush=True)
    assert np.abs(f0).max() > 0 and np.isfinite(f0).all() and np.isfinite(t0).all()
    S = exact(pos2, quat2)
    f1, t1, _ = S.adjoint_forces(tol=1e-20, max_rounds=5000)
    S.close()
    assert relerr(f1, np.array([quat_rot(Q, x) for x in f0])) < 1e-6, (f1, f0)
    assert relerr(t1, np.array([quat_rot(Q, x) for x in t0])) < 1e-6, (t1, t0)
    h = 1e-4

    def e0(P_, Q_):
        S = exact(P_, Q_)
        e = S.energy()[0]
        S.close()
        return e

    pp, pm = pos.copy(), pos.copy()
    pp[0, 0] += h
    pm[0, 0] -= h
    fd = -(e0(pp, quat) - e0(pm, quat)) / (2 * h)
    qp, qm = quat.copy(), quat.copy()
    qp[0] = turned(quat[0], 2, h)
    qm[0] = turned(quat[0], 2, -h)
    td = -(e0(pos, qp) - e0(pos, qm)) / (2 * h)
    print("8 x 500 molecule 0: force x adjoint %.9e fd %.9e, torque z adjoint %.9e fd %.9e"
          % (f0[0, 0], fd, t0[0, 2], td), flush=True)
    assert abs(f0[0, 0] - fd) < 1e-5 * np.linalg.norm(f0[0])
    assert abs(t0[0, 2] - td) < 1e-5 * np.linalg.norm(t0[0])
    for g in gs:
        g.close()
    ctx.close()
