"""State rules of the system handle and a batch whose groups' gradient lists end at
different steps.

* Changing a pose or the cutoffs discards the last solution: the gradient and the forces
  are refused (PBSAM_ERR_STATE) until the next mutual solve, instead of returning the old
  pose's values.
* A batch of groups made of molecule types with different sphere counts: each group's
  gradient list (ascending in its own molecule flag, molecule and sphere) runs out at a
  different key of the batch's union of keys, so the sweep's first pass meets items past
  their last entry (the pass once read one key beyond such an item's list).  Every group
  must equal the same molecules solved as a system of their own."""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu


def moltypes(gpu, orc, sizes, p, npts, seed0=20):
    import pbsam_amd
    import pyoracle
    from pbsam_amd import synth
    mols, gs = [], []
    for t, n in enumerate(sizes):
        m = synth.molecule(n, seed=seed0 + t)
        args = (m["cen"], m["rad"], m["chg_xyz"], m["chg_q"], p)
        kw = dict(eps_in=4.0, rcen=m["rcen"], npts=[npts] * n)
        o = pyoracle.MolType(orc, *args, **kw)
        o.selfpol(0, 0.0, 0)
        g = pbsam_amd.MolType(gpu, *args, **kw)
        g.set_self(*o.FH())
        mols.append(m)
        gs.append(g)
    return mols, gs


def test_pose_change_discards_solution(gpu, orc):
    import pbsam_amd
    from pbsam_amd import synth
    mols, gs = moltypes(gpu, orc, [4, 5], 6, 200)
    d = synth.ball_radius(4) + synth.ball_radius(5) + 6.0
    pos, quat = synth.place_pair(mols[0], mols[1], d, seed=3)
    S = pbsam_amd.System(gpu, gs, pos, quat)
    S.set_cutoffs(100.0, 100.0)
    _, rc, f0, _ = S.compute_forces()
    assert rc == 0
    S.set_pose(1, pos[1] + np.array([1.0, 0.0, 0.0]), quat[1])
    for call in (S.forces, S.gradient, lambda: S.get_grad(0, 0)):
        with pytest.raises(pbsam_amd.PbsamError) as e:
            call()
        assert e.value.code == pbsam_amd.ERR_STATE
    _, rc, f1, _ = S.compute_forces()
    assert rc == 0 and relerr(f1, f0) > 1e-4  # a new solve gives the new pose's forces
    S.set_cutoffs(100.0, 90.0)
    with pytest.raises(pbsam_amd.PbsamError) as e:
        S.forces()
    assert e.value.code == pbsam_amd.ERR_STATE


def test_batch_groups_with_uneven_lists(gpu, orc):
    import pbsam_amd
    from pbsam_amd import synth
    sizes = [3, 7, 5, 2]
    mols, gs = moltypes(gpu, orc, sizes, 8, 300)
    pairs = [(0, 1), (1, 2), (3, 1), (2, 0), (3, 3)]
    poses = []
    for b, (a, c) in enumerate(pairs):
        d = synth.ball_radius(sizes[a]) + synth.ball_radius(sizes[c]) + 4.0 + b
        poses.append(synth.place_pair(mols[a], mols[c], d, seed=60 + b))
    mts = [gs[t] for pr in pairs for t in pr]
    pos = np.vstack([pp for pp, _ in poses])
    quat = np.vstack([qq for _, qq in poses])
    batch = pbsam_amd.System(gpu, mts, pos, quat, groups=[2] * len(pairs))
    batch.set_cutoffs(100.0, 100.0)
    _, rc, fb, tb = batch.compute_forces()
    assert rc == 0
    st, br, bp = batch.status()
    assert len(set(bp)) > 1 or len(set(sizes)) > 1
    for b, (a, c) in enumerate(pairs):
        single = pbsam_amd.System(gpu, [gs[a], gs[c]], *poses[b])
        single.set_cutoffs(100.0, 100.0)
        _, rc1, f1, t1 = single.compute_forces()
        _, sr, sp = single.status()
        assert rc1 == 0 and (br[b], bp[b]) == (sr[0], sp[0]), b
        assert relerr(fb[2 * b:2 * b + 2], f1) < 1e-10, (b, fb[2 * b:2 * b + 2], f1)
        assert relerr(tb[2 * b:2 * b + 2], t1) < 1e-10, b


def test_forces_with_interaction_partners_only(gpu, orc):
    """Molecules closer than the interaction cutoff but farther than the polarization
    cutoff (the Brownian-dynamics cutoffs, 10 / 100 A): no pair is polarized, the spheres
    with interaction partners use their rotated self values, and the forces are not zero.
    Single systems and a batch of them against the oracle."""
    import pbsam_amd
    import pyoracle
    from conftest import relerr
    from pbsam_amd import synth
    mols, gs = moltypes(gpu, orc, [6, 5], 8, 300, seed0=40)
    os_ = []
    for m, g in zip(mols, gs):
        o = pyoracle.MolType(orc, m["cen"], m["rad"], m["chg_xyz"], m["chg_q"], 8, eps_in=4.0, rcen=m["rcen"],
                             npts=[300] * len(m["rad"]))
        o.set_self(*g.FH())
        o.values()
        os_.append(o)
    d = synth.ball_radius(6) + synth.ball_radius(5) + 14.0
    poses = [synth.place_pair(mols[0], mols[1], d + b, seed=70 + b) for b in range(3)]
    batch = pbsam_amd.System(gpu, gs * 3, np.vstack([p for p, _ in poses]), np.vstack([q for _, q in poses]),
                             groups=[2] * 3)
    batch.set_cutoffs(10.0, 100.0)
    _, rcb, fb, tb = batch.compute_forces()
    assert rcb == 0
    for b, (pos, quat) in enumerate(poses):
        S = pbsam_amd.System(gpu, gs, pos, quat)
        S.set_cutoffs(10.0, 100.0)
        _, rc, f, t = S.compute_forces()
        assert rc == 0 and S.counts()[1] == 0
        So = pyoracle.System(orc, os_, pos, quat)
        So.set_cutoffs(10.0, 100.0)
        rco, fo, to = So.compute_forces(0, 0.0, 0)
        assert rco == 0 and np.abs(fo).max() > 0
        assert relerr(f, fo) < 1e-8 and relerr(t, to) < 1e-8, (b, f, fo)
        assert relerr(fb[2 * b:2 * b + 2], fo) < 1e-8, b
