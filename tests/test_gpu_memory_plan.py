"""The memory plan: molecule types whose far-pair translation matrices and numerical
re-expansion maps are not stored (the context's store cap at zero) make them again in every
gradient step that needs them (the reference's LowMemory recomputation,
src/molecule.cpp:1286-1293).  The per-step operators come from the same kernels and the same
inputs as the stored ones, so forces, torques, energies and the exact-derivative torques
equal those of types that store everything, bit for bit."""
import numpy as np
import pytest

from conftest import KAPPA

pytestmark = pytest.mark.gpu


def build(ctx, mols, p, npts, pos, quat, intra_rc):
    import pbsam_amd
    mts = []
    for m in mols:
        n = len(m["rad"])
        g = pbsam_amd.MolType(ctx, m["cen"], m["rad"], m["chg_xyz"], m["chg_q"], p, eps_in=4.0, rcen=m["rcen"],
                              npts=[npts] * n, intra_rcutoff=intra_rc)
        g.selfpol(tol=0, max_rounds=0)
        g.values()
        mts.append(g)
    S = pbsam_amd.System(ctx, mts, pos, quat)
    S.set_cutoffs(100.0, 100.0)
    return mts, S


@pytest.mark.parametrize("p,nsph,npts,exact", [(10, 16, 400, False), (10, 12, 400, True), (30, 6, 1200, False)])
def test_recomputed_operators_equal_stored(p, nsph, npts, exact):
    import pbsam_amd
    from pbsam_amd import synth
    mols = [synth.molecule(nsph, seed=s) for s in (71, 72)]
    # an intra cutoff of 6 A leaves far pairs (frozen values, analytical gradient pairs)
    # beside near numerical pairs (surfaces closer than 5 A)
    intra_rc = 6.0
    c = mols[0]["cen"]
    d = np.linalg.norm(c[:, None] - c[None], axis=2) - mols[0]["rad"][:, None] - mols[0]["rad"][None]
    np.fill_diagonal(d, 1e9)
    assert (d > intra_rc).any() and (d < 5.0).any(), "the molecule must have far and numerical pairs"
    pos, quat = synth.place_pair(mols[0], mols[1], 2 * synth.ball_radius(nsph) + 6.0, seed=73)
    res = []
    for limit in (None, 0):
        ctx = pbsam_amd.Context(KAPPA)
        if limit is not None:
            ctx.set_store_limit(limit)
        mts, S = build(ctx, mols, p, npts, pos, quat, intra_rc)
        if exact:
            S.set_flags(pbsam_amd.EXACT_DERIVATIVE)
            S.solve(tol=1e-14, max_rounds=400)
            S.gradient(tol=1e-14, max_rounds=400)
            f, t = S.forces()
            tx = S.torques_exact(tol=1e-14, max_rounds=400)
        else:
            _, _, f, t = S.compute_forces()
            tx = None
        e = S.energy()
        res.append((f, t, e, tx, ctx.stored()[0]))
        S.close()
        for g in mts:
            g.close()
        ctx.close()
    (f0, t0, e0, tx0, st0), (f1, t1, e1, tx1, st1) = res
    assert st0 > 0 and st1 == 0
    assert np.abs(f0).max() > 0
    assert np.array_equal(f0, f1) and np.array_equal(t0, t1) and np.array_equal(e0, e1)
    if exact:
        assert np.array_equal(tx0, tx1)


def test_store_limit_accounting():
    """Bytes stored grow with each molecule type that stores, never pass the cap, and are
    returned when a type is destroyed."""
    import pbsam_amd
    from pbsam_amd import synth
    ctx = pbsam_amd.Context(KAPPA)
    m = synth.molecule(12, seed=81)
    args = (m["cen"], m["rad"], m["chg_xyz"], m["chg_q"], 10)
    kw = dict(eps_in=4.0, rcen=m["rcen"], npts=[400] * 12, intra_rcutoff=6.0)
    a = pbsam_amd.MolType(ctx, *args, **kw)
    a.selfpol(tol=0, max_rounds=0)
    a.values()
    one = ctx.stored()[0]
    assert one > 0
    ctx.set_store_limit(2 * one + one // 2)  # room for one more type's far matrices, not two
    b = pbsam_amd.MolType(ctx, *args, **kw)
    b.selfpol(tol=0, max_rounds=0)
    b.values()
    two = ctx.stored()[0]
    assert two == 2 * one
    c = pbsam_amd.MolType(ctx, *args, **kw)
    c.selfpol(tol=0, max_rounds=0)
    c.values()
    assert ctx.stored()[0] == two
    b.close()
    assert ctx.stored()[0] == one
    for g in (a, c):
        g.close()
    ctx.close()
