"""HIP graphs of the gradient sweep: windows of steps whose launches recur are captured the
second time they come up and replayed after that.  Every replayed evaluation must equal
the same evaluation with graphs off, bit for bit, including after the poses change (same
buffers, new contents) and with eager work and synchronizations between the replays."""
import numpy as np
import pytest

from conftest import KAPPA

pytestmark = pytest.mark.gpu


def build(ctx, mols, p, npts, pos, quat):
    import pbsam_amd
    gs = []
    for m in mols:
        n = len(m["rad"])
        g = pbsam_amd.MolType(ctx, m["cen"], m["rad"], m["chg_xyz"], m["chg_q"], p, eps_in=4.0, rcen=m["rcen"],
                              npts=[npts] * n)
        g.selfpol(tol=0, max_rounds=0)
        g.values()
        gs.append(g)
    S = pbsam_amd.System(ctx, gs, pos, quat)
    S.set_cutoffs(100.0, 100.0)
    return gs, S


@pytest.mark.parametrize("nsph,p,npts", [(12, 10, 400), (8, 30, 1200)])
def test_graph_replays_equal_direct_launches(nsph, p, npts):
    import pbsam_amd
    from pbsam_amd import synth
    mols = [synth.molecule(nsph, seed=s) for s in (61, 62)]
    d = 2 * synth.ball_radius(nsph) + 6.0
    poses = [synth.place_pair(mols[0], mols[1], d + dd, seed=63 + k) for k, dd in enumerate((0.0, 0.0, 2.0))]
    poses[1] = poses[0]  # the same evaluation twice, then a new pose
    out = {}
    for on in (False, True):
        ctx = pbsam_amd.Context(KAPPA)
        ctx.set_graphs(on)
        gs, S = build(ctx, mols, p, npts, *poses[0])
        res = []
        for rep in range(3):
            for pos, quat in poses:
                for i in range(2):
                    S.set_pose(i, pos[i], quat[i])
                _, rc, f, t = S.compute_forces()
                assert rc == 0
                res.append((f.copy(), t.copy(), S.energy()))
                ctx.sync()
                S.get(0, 0)  # eager work between replays
        out[on] = (res, S.graph_stats())
        S.close()
        for g in gs:
            g.close()
        ctx.close()
    (r0, st0), (r1, st1) = out[False], out[True]
    assert st0 == (0, 0)
    assert st1[0] > 0 and st1[1] > 0, st1  # windows were replayed from graphs
    for a, b in zip(r0, r1):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
