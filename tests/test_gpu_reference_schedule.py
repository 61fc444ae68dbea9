"""Device solves in the reference's Gauss-Seidel order against the oracle's restatement
of polarize_self / polarize_mutual (same order, same selective-recompute and stopping
rules, stock tolerances)."""
import numpy as np
import pytest

from conftest import relerr, rot_close

pytestmark = pytest.mark.gpu


def moltypes(gpu, orc, n_sph, p, npts, seed):
    import pbsam_amd
    import pyoracle
    from pbsam_amd import synth
    mol = synth.molecule(n_sph, seed=seed)
    args = (mol["cen"], mol["rad"], mol["chg_xyz"], mol["chg_q"], p)
    kw = dict(eps_in=4.0, rcen=mol["rcen"], npts=[npts] * n_sph)
    return mol, pbsam_amd.MolType(gpu, *args, **kw), pyoracle.MolType(orc, *args, **kw)


@pytest.mark.parametrize("n_sph,p,npts", [(20, 10, 400), (8, 30, 1800)])
def test_selfpol_reference_order(gpu, orc, n_sph, p, npts):
    _, g, o = moltypes(gpu, orc, n_sph, p, npts, seed=1)
    rg, _ = g.selfpol(tol=0, max_rounds=0)
    ro = o.selfpol(0, 0.0, 0)
    assert rg == ro
    for k in range(g.ns):
        a, b = g.get(k), o.get(k)
        assert relerr(a["F"], b["F"]) < 1e-9, k
        assert relerr(a["H"], b["H"]) < 1e-9, k


@pytest.mark.parametrize("n_sph,p,npts", [(10, 10, 400), (6, 30, 1800), (30, 10, 400), (24, 30, 1800)])
def test_mutual_reference_order(gpu, orc, n_sph, p, npts):
    import pbsam_amd
    import pyoracle
    from pbsam_amd import synth
    molA, gA, oA = moltypes(gpu, orc, n_sph, p, npts, seed=4)
    molB, gB, oB = moltypes(gpu, orc, n_sph, p, npts, seed=5)
    for m in (gA, gB):
        m.selfpol(tol=0, max_rounds=0)
    for m in (oA, oB):
        m.selfpol(0, 0.0, 0)
    # identical self values on both sides, then the mutual solve
    for g, o in ((gA, oA), (gB, oB)):
        F, H = o.FH()
        g.set_self(F, H)
        o.set_self(F, H)
    dist = 2 * synth.ball_radius(n_sph) + 6.0
    pos, quat = synth.place_pair(molA, molB, dist, seed=9)
    sg = pbsam_amd.System(gpu, [gA, gB], pos, quat)
    so = pyoracle.System(orc, [oA, oB], pos, quat)
    for s in (sg, so):
        s.set_cutoffs(100.0, 100.0)
    rg, rc = sg.solve(tol=0, max_rounds=0)
    assert so.lists() == 0
    ro = so.mutual(0, 0.0, 0)
    assert rg == ro
    assert sg.counts() == so.counts()
    for i in range(2):
        for k in range(n_sph):
            a, b = sg.get(i, k), so.get(i, k)
            for key in ("F", "H", "LS"):
                assert relerr(a[key], b[key]) < 1e-9, (i, k, key)
            # H rotated to the lab frame: the rotation's top rows carry its rounding growth
            assert rot_close(a["Hrot"], b["Hrot"], p, 1e-9), (i, k)
    e = sg.energy()
    for i in range(2):
        assert abs(e[i] - so.pot(i)) < 1e-9 * abs(so.pot(i))
