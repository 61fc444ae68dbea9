"""Batches of independent systems (pbsam_batch_create): every group's forces, torques,
energies, mutual rounds and gradient passes against the same molecules solved as a system
of their own, groups of one, two and three molecules, an overlapping group (no pairs, the
others still solved), and one group against the oracle."""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu


def moltypes(gpu, orc, n_sph, p, npts, seeds):
    import pbsam_amd
    import pyoracle
    from pbsam_amd import synth
    mols = [synth.molecule(n_sph, seed=s) for s in seeds]
    gs, os_ = [], []
    for m in mols:
        args = (m["cen"], m["rad"], m["chg_xyz"], m["chg_q"], p)
        kw = dict(eps_in=4.0, rcen=m["rcen"], npts=[npts] * n_sph)
        g, o = pbsam_amd.MolType(gpu, *args, **kw), pyoracle.MolType(orc, *args, **kw)
        o.selfpol(0, 0.0, 0)
        F, H = o.FH()
        g.set_self(F, H)
        o.set_self(F, H)
        gs.append(g)
        os_.append(o)
    return mols, gs, os_


def centres(mol, pos, q):
    from pbsam_amd import synth
    return np.array([pos + synth.quat_rot(q, c - mol["rcen"]) for c in mol["cen"]])


def place_third(mols, pos, quat, mol_c, at, seed):
    """a random orientation for mol_c at `at` that overlaps none of the placed molecules"""
    from pbsam_amd import synth
    rng = np.random.default_rng(seed)
    for _ in range(1000):
        q = synth.random_quat(rng)
        cc = centres(mol_c, at, q)
        ok = True
        for m, p_, q_ in zip(mols, pos, quat):
            c = centres(m, p_, q_)
            gap = np.linalg.norm(cc[:, None, :] - c[None, :, :], axis=2) - mol_c["rad"][:, None] - m["rad"][None, :]
            ok = ok and gap.min() > 0.5
        if ok:
            return q
    raise RuntimeError("no overlap-free orientation")


def build_batch(mols, gs, n_sph):
    """groups: (A,B) (A,C) (B) (A,B,C) (A,B overlapping) (C,A); returns the per-group
    molecule-type indices, positions and orientations"""
    from pbsam_amd import synth
    d = 2 * synth.ball_radius(n_sph) + 6.0
    groups = []
    p, q = synth.place_pair(mols[0], mols[1], d, seed=9)
    groups.append(([0, 1], p, q))
    p, q = synth.place_pair(mols[0], mols[2], d + 3.0, seed=11)
    groups.append(([0, 2], p, q))
    groups.append(([1], np.zeros((1, 3)), np.array([[1.0, 0.0, 0.0, 0.0]])))
    p, q = synth.place_pair(mols[0], mols[1], d, seed=12)
    at = np.array([0.5 * d, 0.9 * d, 0.0])
    q3 = place_third([mols[0], mols[1]], p, q, mols[2], at, seed=13)
    groups.append(([0, 1, 2], np.vstack([p, at]), np.vstack([q, q3])))
    groups.append(([0, 1], np.array([[0.0, 0.0, 0.0], [2.0, 0.0, 0.0]]), np.tile([1.0, 0.0, 0.0, 0.0], (2, 1))))
    p, q = synth.place_pair(mols[2], mols[0], d + 6.0, seed=14)
    groups.append(([2, 0], p, q))
    return groups


@pytest.mark.parametrize("n_sph,p,npts", [(6, 10, 400), (5, 30, 1800)])
def test_batch_equals_single_systems(gpu, orc, n_sph, p, npts):
    import pbsam_amd
    import pyoracle
    mols, gs, os_ = moltypes(gpu, orc, n_sph, p, npts, seeds=(4, 5, 6))
    groups = build_batch(mols, gs, n_sph)
    mts = [gs[t] for g in groups for t in g[0]]
    pos = np.vstack([g[1] for g in groups])
    quat = np.vstack([g[2] for g in groups])
    sizes = [len(g[0]) for g in groups]
    batch = pbsam_amd.System(gpu, mts, pos, quat, groups=sizes)
    batch.set_cutoffs(100.0, 100.0)
    rounds, rc, fb, tb = batch.compute_forces(tol=0, max_rounds=0)
    eb = batch.energy()
    st, br, bp = batch.status()
    assert rc == pbsam_amd.ERR_OVERLAP  # the first non-OK group
    assert list(st) == [0, 0, 0, 0, pbsam_amd.ERR_OVERLAP, 0]
    off = np.cumsum([0] + sizes)
    for gi, (types, gp, gq) in enumerate(groups):
        sl = slice(off[gi], off[gi + 1])
        if gi == 4 or len(types) == 1:
            assert np.all(fb[sl] == 0.0) and np.all(tb[sl] == 0.0), gi
            continue
        single = pbsam_amd.System(gpu, [gs[t] for t in types], gp, gq)
        single.set_cutoffs(100.0, 100.0)
        r1, rc1, f1, t1 = single.compute_forces(tol=0, max_rounds=0)
        _, sr, sp = single.status()
        assert rc1 == 0 and (br[gi], bp[gi]) == (sr[0], sp[0]), (gi, br[gi], bp[gi], sr[0], sp[0])
        # the same kernels on the same data; only launch shapes (order-class split,
        # point splits) change the rounding
        assert relerr(fb[sl], f1) < 1e-10, (gi, fb[sl], f1)
        assert relerr(tb[sl], t1) < 1e-10, (gi, tb[sl], t1)
        assert relerr(eb[sl], single.energy()) < 1e-12, gi
    # group 0 against the oracle (stock tolerances, reference order)
    so = pyoracle.System(orc, [os_[0], os_[1]], groups[0][1], groups[0][2])
    so.set_cutoffs(100.0, 100.0)
    rco, fo, to = so.compute_forces(0, 0.0, 0)
    assert rco == 0
    assert relerr(fb[0:2], fo) < 1e-8
    assert relerr(tb[0:2], to) < 1e-8


@pytest.mark.parametrize("n_sph,p,npts", [(6, 10, 400), (5, 30, 1800)])
def test_batch_of_snapshots(gpu, orc, n_sph, p, npts):
    """Six poses of the same two molecule types: every step's stored maps and surface
    harmonics are shared by all groups (the bundled matrix-core products)."""
    import pbsam_amd
    import pyoracle
    from pbsam_amd import synth
    mols, gs, os_ = moltypes(gpu, orc, n_sph, p, npts, seeds=(4, 5))
    d0 = 2 * synth.ball_radius(n_sph) + 4.0
    poses = [synth.place_pair(mols[0], mols[1], d0 + 1.5 * b, seed=40 + b) for b in range(6)]
    pos = np.vstack([pp for pp, _ in poses])
    quat = np.vstack([qq for _, qq in poses])
    batch = pbsam_amd.System(gpu, [gs[0], gs[1]] * 6, pos, quat, groups=[2] * 6)
    batch.set_cutoffs(100.0, 100.0)
    _, rc, fb, tb = batch.compute_forces(tol=0, max_rounds=0)
    assert rc == 0
    st, br, bp = batch.status()
    for b, (pp, qq) in enumerate(poses):
        single = pbsam_amd.System(gpu, [gs[0], gs[1]], pp, qq)
        single.set_cutoffs(100.0, 100.0)
        _, rc1, f1, t1 = single.compute_forces(tol=0, max_rounds=0)
        _, sr, sp = single.status()
        assert rc1 == 0 and (br[b], bp[b]) == (sr[0], sp[0]), b
        assert relerr(fb[2 * b:2 * b + 2], f1) < 1e-10, (b, fb[2 * b:2 * b + 2], f1)
        assert relerr(tb[2 * b:2 * b + 2], t1) < 1e-10, b
    so = pyoracle.System(orc, [os_[0], os_[1]], poses[3][0], poses[3][1])
    so.set_cutoffs(100.0, 100.0)
    rco, fo, to = so.compute_forces(0, 0.0, 0)
    assert rco == 0
    assert relerr(fb[6:8], fo) < 1e-8
    assert relerr(tb[6:8], to) < 1e-8


def test_batch_of_mixed_pole_orders(gpu, orc):
    """Systems at p = 10 and p = 30 with uneven sphere counts in one batch handle (the
    mixed-order batch workload, reduced): each group equals its molecules solved alone,
    indices and statuses come back in the caller's order, and one group of each order
    matches the oracle."""
    import pbsam_amd
    import pyoracle
    from pbsam_amd import synth
    spec = [(10, 5, 400), (30, 3, 1800), (10, 8, 400), (30, 4, 1800), (10, 3, 400)]
    types = {}
    for t, (p, n, npts) in enumerate(spec):
        m = synth.molecule(n, seed=700 + t)
        args = (m["cen"], m["rad"], m["chg_xyz"], m["chg_q"], p)
        kw = dict(eps_in=4.0, rcen=m["rcen"], npts=[npts] * n)
        o = pyoracle.MolType(orc, *args, **kw)
        o.selfpol(0, 0.0, 0)
        g = pbsam_amd.MolType(gpu, *args, **kw)
        g.set_self(*o.FH())
        o.values()
        types[t] = (m, g, o)
    groups = [(0, 2), (1, 3), (2, 4), (3, 1), (4, 0), (1, 1)]
    rng = np.random.default_rng(31)
    poses = []
    for a, b in groups:
        d = synth.ball_radius(len(types[a][0]["rad"])) + synth.ball_radius(len(types[b][0]["rad"])) + \
            rng.uniform(4.0, 12.0)
        poses.append(synth.place_pair(types[a][0], types[b][0], d, seed=800 + len(poses)))
    batch = pbsam_amd.System(gpu, [types[t][1] for g in groups for t in g], np.vstack([p for p, _ in poses]),
                             np.vstack([q for _, q in poses]), groups=[2] * len(groups))
    batch.set_cutoffs(100.0, 100.0)
    _, rc, fb, tb = batch.compute_forces()
    assert rc == 0
    eb = batch.energy()
    st, br, bp = batch.status()
    assert list(st) == [0] * len(groups)
    checked = set()
    for gi, (a, b) in enumerate(groups):
        single = pbsam_amd.System(gpu, [types[a][1], types[b][1]], *poses[gi])
        single.set_cutoffs(100.0, 100.0)
        _, rc1, f1, t1 = single.compute_forces()
        _, sr, sp = single.status()
        assert rc1 == 0 and (br[gi], bp[gi]) == (sr[0], sp[0]), gi
        sl = slice(2 * gi, 2 * gi + 2)
        assert relerr(fb[sl], f1) < 1e-10, (gi, fb[sl], f1)
        assert relerr(tb[sl], t1) < 1e-10, gi
        assert relerr(eb[sl], single.energy()) < 1e-12, gi
        p = spec[a][0]
        if p not in checked:
            checked.add(p)
            so = pyoracle.System(orc, [types[a][2], types[b][2]], *poses[gi])
            so.set_cutoffs(100.0, 100.0)
            rco, fo, to = so.compute_forces(0, 0.0, 0)
            assert rco == 0
            assert relerr(fb[sl], fo) < 1e-8 and relerr(tb[sl], to) < 1e-8, (gi, p)
            # per-sphere blocks come back at the molecule's own order
            assert len(batch.get(2 * gi, 0)["H"]) == p * p
    assert checked == {10, 30}


def test_autotune_keeps_results(gpu, orc):
    """pbsam_autotune times complete evaluations over launch-shape settings and keeps the
    fastest; the forces stay those of the default settings to rounding, and a second batch
    of the same shape picks the tuned settings up."""
    import pbsam_amd
    mols, gs, _ = moltypes(gpu, orc, 5, 10, 400, seeds=(4, 5))
    from pbsam_amd import synth
    poses = [synth.place_pair(mols[0], mols[1], 2 * synth.ball_radius(5) + 4.0 + b, seed=90 + b) for b in range(6)]
    pos = np.vstack([pp for pp, _ in poses])
    quat = np.vstack([qq for _, qq in poses])

    def batch():
        b = pbsam_amd.System(gpu, [gs[0], gs[1]] * 6, pos, quat, groups=[2] * 6)
        b.set_cutoffs(100.0, 100.0)
        return b

    b0 = batch()
    _, rc, f0, t0 = b0.compute_forces()
    assert rc == 0
    b1 = batch()
    ms, settings = b1.autotune(reps=1)
    assert ms > 0 and set(settings) == {"grad_window", "xsplit_max", "xbin_min", "mgemv_wg", "pgemm_min",
                                        "p2l_wg", "surf_wg"}
    _, rc, f1, t1 = b1.compute_forces()
    assert rc == 0 and relerr(f1, f0) < 1e-12 and relerr(t1, t0) < 1e-12
    _, rc, f2, _ = batch().compute_forces()   # same shape: the tuned settings
    assert rc == 0 and relerr(f2, f1) < 1e-15
