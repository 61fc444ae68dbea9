"""Device gradient phase, forces and torques against the oracle's restatement of
polarize_mutual's gradient part and computeMol_Force_and_Torque, in the reference's
order with its stock tolerances (pair lists, mutual solve, gradient, forces)."""
import numpy as np
import pytest

from conftest import relerr, rot_close

pytestmark = pytest.mark.gpu


def systems(gpu, orc, n_sph, p, npts, seeds=(4, 5), pose_seed=9, gap=6.0):
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
    dist = 2 * synth.ball_radius(n_sph) + gap
    pos, quat = synth.place_pair(mols[0], mols[1], dist, seed=pose_seed)
    sg = pbsam_amd.System(gpu, gs, pos, quat)
    so = pyoracle.System(orc, os_, pos, quat)
    for s in (sg, so):
        s.set_cutoffs(100.0, 100.0)
    return sg, so, (gs, os_)


@pytest.mark.parametrize("n_sph,p,npts", [(6, 10, 400), (12, 10, 400), (6, 30, 1800)])
def test_forces_reference_order(gpu, orc, n_sph, p, npts):
    sg, so, keep = systems(gpu, orc, n_sph, p, npts)
    rounds, rc, fg, tg = sg.compute_forces(tol=0, max_rounds=0)
    rco, fo, to = so.compute_forces(0, 0.0, 0)
    assert rco == 0
    for i in range(2):
        for k in range(n_sph):
            gH, gL = sg.get_grad(i, k)
            oH, oL = so.get_grad(i, k)
            for d in range(3):
                s = slice(d * p * p, (d + 1) * p * p)
                assert rot_close(gH[s], oH[s], p, 1e-8), (i, k, d, "gH")
                assert rot_close(gL[s], oL[s], p, 1e-8), (i, k, d, "gLHN")
    # forces and torques: 1e-6 relative is required, far tighter is expected
    assert relerr(fg, fo) < 1e-8, (fg, fo)
    assert relerr(tg, to) < 1e-8, (tg, to)


@pytest.mark.parametrize("self_frame", [False, True])
def test_forces_with_interaction_only_spheres(gpu, orc, self_frame):
    """Spheres whose partners interact but do not polarize them (interRcutoff 5 A,
    interactRcutoff 100 A) take their forces from self values; the device matches the
    oracle with the reference's pairing of the unrotated self H and with PBSAM_SELF_FRAME
    (the rotated self H)."""
    import pbsam_amd
    import pyoracle
    sg, so, keep = systems(gpu, orc, 6, 10, 400, gap=8.0)
    for s in (sg, so):
        s.set_cutoffs(5.0, 100.0)
    if self_frame:
        sg.set_flags(pbsam_amd.SELF_FRAME)
    pyoracle.self_frame(self_frame)
    try:
        _, rc, fg, tg = sg.compute_forces(tol=0, max_rounds=0)
        rco, fo, to = so.compute_forces(0, 0.0, 0)
    finally:
        pyoracle.self_frame(False)
    assert rc == 0 and rco == 0
    n_inter, n_pol = sg.counts()
    assert n_pol < 12  # some spheres with partners are not polarized
    assert np.abs(fo).max() > 0
    assert relerr(fg, fo) < 1e-8, (fg, fo)
    assert relerr(tg, to) < 1e-8, (tg, to)
