"""Forces and torques from the adjoint of the converged fixed point (pbsam_adjoint_forces).

Under PBSAM_EXACT_DERIVATIVE the adjoint solve gives -dE_i/dr_i and -dE_i/domega_i of every
molecule's energy sum_k <H_k, LS_k>.  Checked against: the forward-mode gradient phase in the
same mode (orientation * pbsam_forces, and pbsam_torques_exact), which the oracle pins
(tests/test_gpu_exact.py, tests/test_gpu_torque.py); central differences of the device energy
under 1e-4 A moves and 1e-4 rad turns; the Debye-Hueckel closed forms of two charged spheres;
the same systems solved alone against a batch; and torch.autograd.gradcheck through
pbsam_amd.autograd.RigidEnergy."""
import numpy as np
import pytest

from conftest import relerr
from test_gpu_exact import pair, quat_rot

pytestmark = pytest.mark.gpu

ADJ_TOL = 1e-22


def system(gpu, gs, pos, quat, tol=1e-18):
    import pbsam_amd
    S = pbsam_amd.System(gpu, gs, pos, quat)
    S.set_cutoffs(100.0, 100.0)
    S.set_flags(pbsam_amd.EXACT_DERIVATIVE)
    rounds, rc = S.solve(tol=tol, max_rounds=4000)
    assert rc == 0, rounds
    return S


def turned(q, a, h):
    from test_gpu_torque import turned as t
    return t(q, a, h)


@pytest.mark.parametrize("ns,p,npts", [(4, 10, 400), (6, 10, 400), (3, 30, 1200)])
def test_adjoint_equals_forward_mode(gpu, orc, ns, p, npts):
    """every polarized pair of these molecules is within the intra cutoff, so the forward
    sweeps and the adjoint linearize the same map"""
    gs, _, pos, quat = pair(gpu, orc, ns, p, npts)
    S = system(gpu, gs, pos, quat)
    fa, ta, sweeps = S.adjoint_forces(tol=ADJ_TOL, max_rounds=20000)
    print("adjoint sweeps", sweeps, flush=True)
    S.gradient(tol=1e-20, max_rounds=4000)
    f, _ = S.forces()
    ff = np.array([quat_rot(quat[i], f[i]) for i in range(2)])
    tf = S.torques_exact(tol=1e-20, max_rounds=4000)
    assert relerr(fa, ff) < 1e-8, (fa, ff)
    assert relerr(ta, tf) < 1e-8, (ta, tf)
    # Newton's third law holds only approximately for the per-molecule energies: not tested


@pytest.mark.parametrize("ns,p,npts", [(5, 10, 400), (3, 30, 1200)])
def test_adjoint_is_the_energy_gradient(gpu, orc, ns, p, npts):
    gs, _, pos, quat = pair(gpu, orc, ns, p, npts, seeds=(7, 8))
    S = system(gpu, gs, pos, quat)
    fa, ta, _ = S.adjoint_forces(tol=ADJ_TOL, max_rounds=20000)
    h = 1e-4
    for mol in range(2):
        fd, td = np.zeros(3), np.zeros(3)
        for d in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[mol, d] += h
            pm[mol, d] -= h
            ep = system(gpu, gs, pp, quat).energy()[mol]
            em = system(gpu, gs, pm, quat).energy()[mol]
            fd[d] = -(ep - em) / (2 * h)
            qp, qm = quat.copy(), quat.copy()
            qp[mol] = turned(quat[mol], d, h)
            qm[mol] = turned(quat[mol], d, -h)
            ep = system(gpu, gs, pos, qp).energy()[mol]
            em = system(gpu, gs, pos, qm).energy()[mol]
            td[d] = -(ep - em) / (2 * h)
        assert np.abs(fa[mol] - fd).max() < 1e-6 * np.linalg.norm(fd), (mol, fa[mol], fd)
        assert np.abs(ta[mol] - td).max() < 1e-6 * np.linalg.norm(td), (mol, ta[mol], td)


@pytest.mark.parametrize("d,tol", [(40.0, 1e-3), (80.0, 1e-4)])
def test_adjoint_debye_hueckel(gpu, d, tol):
    """two one-sphere molecules (radii 3 and 2 A, +1 and -1 at the centres, the first sphere
    off its molecule centre): force -f1 on molecule 0, torque (R off) x (-f1)"""
    import pbsam_amd
    from test_gpu_exact import dh_pair
    q0 = np.array([0.9, 0.3, -0.2, 0.1])
    q0 /= np.linalg.norm(q0)
    off = (1.0, -0.5, 0.7)
    _, _, f1, offs, quat, mts = dh_pair(gpu, d, pbsam_amd.EXACT_DERIVATIVE, q0, (1.0, 0.0, 0.0, 0.0), off)
    c0, c1 = np.zeros(3), np.array([d, 0.0, 0.0])
    pos = np.array([c0 - quat_rot(quat[0], offs[0]), c1])
    S = pbsam_amd.System(gpu, mts, pos, quat)
    S.set_cutoffs(1e4, 1e4)
    S.set_flags(pbsam_amd.EXACT_DERIVATIVE)
    S.solve(tol=1e-24, max_rounds=400)
    fa, ta, _ = S.adjoint_forces(tol=ADJ_TOL, max_rounds=2000)
    assert np.abs(fa[0] + f1).max() < tol * np.linalg.norm(f1)
    assert np.abs(fa[1] - f1).max() < tol * np.linalg.norm(f1)
    tau = np.cross(quat_rot(quat[0], offs[0]), -f1)
    assert np.abs(ta[0] - tau).max() < tol * np.linalg.norm(f1) * np.linalg.norm(offs[0])


def test_adjoint_batch_equals_single_systems(gpu, orc):
    import pbsam_amd
    from pbsam_amd import synth
    gs, _, _, _ = pair(gpu, orc, 4, 10, 400)
    d = 2 * synth.ball_radius(4) + 6.0
    mols = [synth.molecule(4, seed=s) for s in (4, 5)]
    poses = [synth.place_pair(mols[0], mols[1], d + b, seed=30 + b) for b in range(3)]
    alone = []
    for pos, quat in poses:
        S = system(gpu, gs, pos, quat)
        fa, ta, _ = S.adjoint_forces(tol=ADJ_TOL, max_rounds=20000)
        alone.append((fa, ta))
    pos = np.vstack([pp for pp, _ in poses])
    quat = np.vstack([qq for _, qq in poses])
    B = pbsam_amd.System(gpu, gs * 3, pos, quat, groups=[2] * 3)
    B.set_cutoffs(100.0, 100.0)
    B.set_flags(pbsam_amd.EXACT_DERIVATIVE)
    B.solve(tol=1e-18, max_rounds=4000)
    fb, tb, _ = B.adjoint_forces(tol=ADJ_TOL, max_rounds=20000)
    for g, (fa, ta) in enumerate(alone):
        assert relerr(fb[2 * g:2 * g + 2], fa) < 1e-9
        assert relerr(tb[2 * g:2 * g + 2], ta) < 1e-9


def test_adjoint_needs_the_exact_mode(gpu, orc):
    import pbsam_amd
    gs, _, pos, quat = pair(gpu, orc, 3, 10, 400)
    S = pbsam_amd.System(gpu, gs, pos, quat)
    S.set_cutoffs(100.0, 100.0)
    S.solve(tol=1e-12, max_rounds=400)
    with pytest.raises(pbsam_amd.PbsamError) as e:
        S.adjoint_forces()
    assert e.value.code == pbsam_amd.ERR_STATE
    Si = pbsam_amd.System(gpu, gs, pos, quat)
    Si.set_cutoffs(10.0, 100.0)  # interaction partners beyond the polarization cutoff
    Si.set_flags(pbsam_amd.EXACT_DERIVATIVE)
    Si.solve(tol=1e-12, max_rounds=400)
    with pytest.raises(pbsam_amd.PbsamError) as e:
        Si.adjoint_forces()
    assert e.value.code == pbsam_amd.ERR_UNSUPPORTED


def triangle(gpu, orc, ns=4, p=10, npts=400, seed=21):
    """three molecules (the two types of pair(), the first twice) at the corners of a triangle
    of side 2 R + 6, R the radius of the ball holding each molecule's spheres (so no spheres
    of different molecules overlap), random orientations"""
    from pbsam_amd import synth
    gs, _, _, _ = pair(gpu, orc, ns, p, npts)
    d = 2 * synth.ball_radius(ns) + 6.0
    rng = np.random.default_rng(seed)
    pos = np.array([[0.0, 0.0, 0.0], [d, 0.0, 0.0], [0.5 * d, 0.5 * np.sqrt(3.0) * d, 0.4]])
    quat = np.array([synth.random_quat(rng) for _ in range(3)])
    return [gs[0], gs[1], gs[0]], pos, quat


def exact_forward(S, quat):
    """forces (lab frame) and torques of the forward gradient sweeps in the exact-derivative mode"""
    S.gradient(tol=1e-20, max_rounds=4000)
    f, _ = S.forces()
    ff = np.array([quat_rot(quat[i], f[i]) for i in range(len(quat))])
    return ff, S.torques_exact(tol=1e-20, max_rounds=4000)


def test_forward_sweeps_drop_the_coupling_through_a_third_molecule(gpu, orc):
    """quirk: when the gradient planes of molecule j's move are re-expanded into a sphere of
    another molecule, the reference takes only the partners in molecule j
    (reexpandGrad_LowMemory, molecule.cpp:1779-1788), so with three molecules the forward
    sweeps leave out the change of the third molecule's H; the adjoint, the derivative of the
    solved energy (central differences in the next test), keeps it: in a triangle the two
    differ at 1e-5.  In a chain 0 - 1 - 2 with the cutoff between the neighbours' and the
    ends' separations, the middle molecule's partners do not polarize each other and its
    derivatives agree (the ends' differ only through the loop 1 - 2 - 1, at 1e-10 here)."""
    from pbsam_amd import synth
    gs, pos, quat = triangle(gpu, orc)
    S = system(gpu, gs, pos, quat)
    fa, ta, sweeps = S.adjoint_forces(tol=ADJ_TOL, max_rounds=20000)
    print("triangle: adjoint applications over three solves", sweeps, flush=True)
    ff, tf = exact_forward(S, quat)
    gap = max(relerr(fa, ff), relerr(ta, tf))
    print("triangle: forward sweeps vs adjoint", relerr(fa, ff), relerr(ta, tf), flush=True)
    assert 1e-6 < gap < 1e-3, gap
    # chain along x: neighbours' surface separations <= d + 2 R - 4, the ends' >= 2 d - 2 R
    R = synth.ball_radius(4)
    d = 50.0
    cut = 0.5 * ((d + 2 * R - 4) + (2 * d - 2 * R))
    pos = np.array([[0.0, 0.0, 0.0], [d, 0.0, 0.0], [2 * d, 0.0, 0.0]])
    S = cut_system(gpu, gs, pos, quat, cut)
    fa, ta, _ = S.adjoint_forces(tol=ADJ_TOL, max_rounds=20000)
    ff, tf = exact_forward(S, quat)
    print("chain: forward sweeps vs adjoint per molecule",
          [(relerr(fa[m], ff[m]), relerr(ta[m], tf[m])) for m in range(3)], flush=True)
    assert relerr(fa[1], ff[1]) < 1e-9 and relerr(ta[1], tf[1]) < 1e-9


def cut_system(gpu, gs, pos, quat, cut, tol=1e-18):
    """system() with both cutoffs at cut (no interaction-only partners)"""
    import pbsam_amd
    S = pbsam_amd.System(gpu, gs, pos, quat)
    S.set_cutoffs(cut, cut)
    S.set_flags(pbsam_amd.EXACT_DERIVATIVE)
    rounds, rc = S.solve(tol=tol, max_rounds=4000)
    assert rc == 0, rounds
    return S


def test_adjoint_with_a_distant_third_molecule_equals_the_pair(gpu, orc):
    """a third molecule beyond the cutoff of both others: per-molecule solves whose right-hand
    sides reduce to the pair's shared one; its own force and torque are zero"""
    gs, _, pos, quat = pair(gpu, orc, 4, 10, 400)
    S = system(gpu, gs, pos, quat)
    fa, ta, _ = S.adjoint_forces(tol=ADJ_TOL, max_rounds=20000)
    pos3 = np.vstack([pos, pos[:1] + np.array([0.0, 400.0, 0.0])])
    quat3 = np.vstack([quat, quat[:1]])
    S3 = system(gpu, gs + gs[:1], pos3, quat3)
    f3, t3, _ = S3.adjoint_forces(tol=ADJ_TOL, max_rounds=20000)
    assert relerr(f3[:2], fa) < 1e-10
    assert relerr(t3[:2], ta) < 1e-10
    assert np.all(f3[2] == 0.0) and np.all(t3[2] == 0.0)


def test_adjoint_three_molecules_is_the_energy_gradient(gpu, orc):
    gs, pos, quat = triangle(gpu, orc, seed=23)
    S = system(gpu, gs, pos, quat)
    fa, ta, _ = S.adjoint_forces(tol=ADJ_TOL, max_rounds=20000)
    h = 1e-4
    for mol in range(3):
        fd, td = np.zeros(3), np.zeros(3)
        for d in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[mol, d] += h
            pm[mol, d] -= h
            ep = system(gpu, gs, pp, quat).energy()[mol]
            em = system(gpu, gs, pm, quat).energy()[mol]
            fd[d] = -(ep - em) / (2 * h)
            qp, qm = quat.copy(), quat.copy()
            qp[mol] = turned(quat[mol], d, h)
            qm[mol] = turned(quat[mol], d, -h)
            ep = system(gpu, gs, pos, qp).energy()[mol]
            em = system(gpu, gs, pos, qm).energy()[mol]
            td[d] = -(ep - em) / (2 * h)
        assert np.abs(fa[mol] - fd).max() < 1e-6 * np.linalg.norm(fd), (mol, fa[mol], fd)
        assert np.abs(ta[mol] - td).max() < 1e-6 * np.linalg.norm(td), (mol, ta[mol], td)


def test_adjoint_eight_molecules_all_48_coordinates(gpu, orc):
    """eight molecules on the corners of a cube (10 A gaps between packing balls): the adjoint's
    force and torque on every molecule (eight transposed solves) against central differences
    of that molecule's solved energy, all 48 rigid-body coordinates"""
    from pbsam_amd import synth
    from test_gpu_configs import eight, self_polarized
    mols = [synth.molecule(4, seed=900 + i) for i in range(8)]
    gs, _ = self_polarized(gpu, orc, mols, 10, 300, oracle=False)
    pos, quat = eight(mols, 10.0, 3)
    S = system(gpu, gs, pos, quat)
    fa, ta, sweeps = S.adjoint_forces(tol=ADJ_TOL, max_rounds=20000)
    print("eight molecules: adjoint applications over eight solves", sweeps, flush=True)
    h = 1e-4
    for mol in range(8):
        fd, td = np.zeros(3), np.zeros(3)
        for d in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[mol, d] += h
            pm[mol, d] -= h
            fd[d] = -(system(gpu, gs, pp, quat).energy()[mol] - system(gpu, gs, pm, quat).energy()[mol]) / (2 * h)
            qp, qm = quat.copy(), quat.copy()
            qp[mol] = turned(quat[mol], d, h)
            qm[mol] = turned(quat[mol], d, -h)
            td[d] = -(system(gpu, gs, pos, qp).energy()[mol] - system(gpu, gs, pos, qm).energy()[mol]) / (2 * h)
        assert np.abs(fa[mol] - fd).max() < 1e-6 * np.linalg.norm(fd), (mol, fa[mol], fd)
        assert np.abs(ta[mol] - td).max() < 1e-6 * np.linalg.norm(td), (mol, ta[mol], td)


def test_adjoint_mixed_group_sizes_equal_single_systems(gpu, orc):
    """a group of three molecules and a group of two in one handle: the three-molecule group's
    per-molecule solves run beside the pair's (zero right-hand side after the first pass)"""
    import pbsam_amd
    gs3, pos3, quat3 = triangle(gpu, orc, seed=25)
    gs2, _, pos2, quat2 = pair(gpu, orc, 4, 10, 400)
    alone = []
    for gs, pos, quat in ((gs3, pos3, quat3), (gs2, pos2, quat2)):
        S = system(gpu, gs, pos, quat)
        fa, ta, _ = S.adjoint_forces(tol=ADJ_TOL, max_rounds=20000)
        alone.append((fa, ta))
    B = pbsam_amd.System(gpu, gs3 + gs2, np.vstack([pos3, pos2]), np.vstack([quat3, quat2]), groups=[3, 2])
    B.set_cutoffs(100.0, 100.0)
    B.set_flags(pbsam_amd.EXACT_DERIVATIVE)
    B.solve(tol=1e-18, max_rounds=4000)
    fb, tb, _ = B.adjoint_forces(tol=ADJ_TOL, max_rounds=20000)
    assert relerr(fb[:3], alone[0][0]) < 1e-9
    assert relerr(tb[:3], alone[0][1]) < 1e-9
    assert relerr(fb[3:], alone[1][0]) < 1e-9
    assert relerr(tb[3:], alone[1][1]) < 1e-9


@pytest.mark.parametrize("ns,p,npts", [(4, 10, 400), (3, 30, 1200)])
def test_gradcheck_rigid_energy(gpu, ns, p, npts):
    import pbsam_amd
    pbsam_amd.load()   # the library's HIP runtime first, then torch
    import torch
    from pbsam_amd import autograd, synth
    mols = [synth.molecule(ns, seed=s) for s in (24, 25)]
    mts = []
    for m in mols:
        mt = pbsam_amd.MolType(gpu, m["cen"], m["rad"], m["chg_xyz"], m["chg_q"], p, eps_in=4.0, rcen=m["rcen"],
                               npts=[npts] * ns)
        mt.selfpol(tol=0, max_rounds=0)
        mt.values()
        mts.append(mt)
    pos, quat = synth.place_pair(mols[0], mols[1], 2 * synth.ball_radius(ns) + 5.0, seed=22)
    for j in range(2):
        ev = autograd.Evaluator(gpu, mts, pos, quat, j, tol=1e-18, max_rounds=4000, adj_tol=ADJ_TOL,
                                adj_sweeps=20000)
        x = torch.zeros(6, dtype=torch.float64, requires_grad=True)
        _, g = autograd.solve_rigid(ev, np.zeros(6), True)
        scale = float(np.linalg.norm(g))
        assert scale > 0
        assert torch.autograd.gradcheck(lambda y: autograd.RigidEnergy.apply(y, ev), (x,), eps=1e-4,
                                        atol=1e-6 * scale, rtol=1e-6)


# ---- the exact-mode fixed point by GMRES (pbsam_solve_exact) ----

def test_solve_exact_equals_converged_rounds(gpu, orc):
    """where the reference's Gauss-Seidel rounds converge, GMRES lands on the same state"""
    import pbsam_amd
    gs, _, pos, quat = pair(gpu, orc, 5, 30, 1200)
    S = system(gpu, gs, pos, quat, tol=1e-24)
    e0 = S.energy()
    H0 = [S.get(i, k)["H"] for i in range(2) for k in range(5)]
    f0, t0, _ = S.adjoint_forces(tol=ADJ_TOL, max_rounds=2000)
    n, rc = S.solve_exact(tol=1e-26, max_rounds=500)
    assert rc == pbsam_amd.OK and n > 0
    assert relerr(S.energy(), e0) < 1e-10
    H1 = [S.get(i, k)["H"] for i in range(2) for k in range(5)]
    assert max(relerr(a, b) for a, b in zip(H1, H0)) < 1e-9
    f1, t1, _ = S.adjoint_forces(tol=ADJ_TOL, max_rounds=2000)
    assert relerr(f1, f0) < 1e-9 and relerr(t1, t0) < 1e-9


def exact_state(gpu, mts, pos, quat, tol=1e-24):
    import pbsam_amd
    S = pbsam_amd.System(gpu, mts, pos, quat)
    S.set_cutoffs(100.0, 100.0)
    S.set_flags(pbsam_amd.EXACT_DERIVATIVE)
    n, rc = S.solve_exact(tol=tol, max_rounds=1000)
    assert rc == pbsam_amd.OK, n
    return S


def test_full_size_forces_and_torques_by_central_differences(gpu):
    """The benchmark's two 60-sphere molecules (p = 30, 1800 points per sphere, 50 A apart):
    nearly touching spheres (gaps down to 0.004 A) make the reference's mutual rounds wander
    (Jacobi rounds diverge), so no converged fixed point comes from them; GMRES converges the
    exact-mode system, and the adjoint's forces and torques are central differences (1e-4 A,
    1e-4 rad) of that solved energy to 1e-6."""
    import pbsam_amd
    from pbsam_amd import synth
    mols = [synth.molecule(60, seed=s) for s in (101, 102)]
    mts = []
    for m in mols:
        mt = pbsam_amd.MolType(gpu, m["cen"], m["rad"], m["chg_xyz"], m["chg_q"], 30, eps_in=synth.EPS_IN,
                               rcen=m["rcen"], npts=[1800] * 60)
        mt.selfpol(tol=0, max_rounds=0)
        mt.values()
        mts.append(mt)
    pos, quat = synth.place_pair(mols[0], mols[1], 50.0, seed=7)
    S = exact_state(gpu, mts, pos, quat)
    e = S.energy()
    assert relerr(e[0], e[1]) < 1e-9  # the two molecules' energies agree at the fixed point
    fa, ta, n = S.adjoint_forces(tol=ADJ_TOL, max_rounds=1000)
    print("adjoint applications", n, flush=True)
    assert relerr(fa[0], -fa[1]) < 1e-9
    h = 1e-4
    for mol in range(2):
        fd, td = np.zeros(3), np.zeros(3)
        for d in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[mol, d] += h
            pm[mol, d] -= h
            ep, em = (exact_state(gpu, mts, x, quat).energy()[mol] for x in (pp, pm))
            fd[d] = -(ep - em) / (2 * h)
            qp, qm = quat.copy(), quat.copy()
            qp[mol] = turned(quat[mol], d, h)
            qm[mol] = turned(quat[mol], d, -h)
            ep, em = (exact_state(gpu, mts, pos, x).energy()[mol] for x in (qp, qm))
            td[d] = -(ep - em) / (2 * h)
        print("molecule", mol, fa[mol], fd, ta[mol], td, flush=True)
        assert np.abs(fa[mol] - fd).max() < 1e-6 * np.linalg.norm(fd)
        assert np.abs(ta[mol] - td).max() < 1e-6 * np.linalg.norm(td)


def test_solve_exact_batch_equals_single_systems(gpu, orc):
    import pbsam_amd
    from pbsam_amd import synth
    gs, _, _, _ = pair(gpu, orc, 4, 10, 400)
    d = 2 * synth.ball_radius(4) + 6.0
    mols = [synth.molecule(4, seed=s) for s in (4, 5)]
    poses = [synth.place_pair(mols[0], mols[1], d + b, seed=40 + b) for b in range(3)]
    alone = []
    for pos, quat in poses:
        S = exact_state(gpu, gs, pos, quat, tol=1e-26)
        alone.append((S.energy(), S.adjoint_forces(tol=ADJ_TOL, max_rounds=1000)[:2]))
    B = pbsam_amd.System(gpu, gs * 3, np.vstack([pp for pp, _ in poses]), np.vstack([qq for _, qq in poses]),
                         groups=[2] * 3)
    B.set_cutoffs(100.0, 100.0)
    B.set_flags(pbsam_amd.EXACT_DERIVATIVE)
    n, rc = B.solve_exact(tol=1e-26, max_rounds=1000)
    assert rc == pbsam_amd.OK
    eb = B.energy()
    fb, tb, _ = B.adjoint_forces(tol=ADJ_TOL, max_rounds=1000)
    for g, (e, (fa, ta)) in enumerate(alone):
        assert relerr(eb[2 * g:2 * g + 2], e) < 1e-10
        assert relerr(fb[2 * g:2 * g + 2], fa) < 1e-9
        assert relerr(tb[2 * g:2 * g + 2], ta) < 1e-9


def far_pair_types(gpu, store_limit=None, p=10, nsph=12, npts=400):
    """two 12-sphere molecules with an intra cutoff of 6 A: far pairs (frozen self values in the
    mutual solve) beside near numerical ones"""
    import pbsam_amd
    from pbsam_amd import synth
    ctx = pbsam_amd.Context(gpu.kappa) if store_limit is not None else gpu
    if store_limit is not None:
        ctx.set_store_limit(store_limit)
    mols = [synth.molecule(nsph, seed=s) for s in (71, 72)]
    mts = []
    for m in mols:
        g = pbsam_amd.MolType(ctx, m["cen"], m["rad"], m["chg_xyz"], m["chg_q"], p, eps_in=4.0, rcen=m["rcen"],
                              npts=[npts] * nsph, intra_rcutoff=6.0)
        g.selfpol(tol=0, max_rounds=0)
        g.values()
        mts.append(g)
    pos, quat = synth.place_pair(mols[0], mols[1], 2 * synth.ball_radius(nsph) + 6.0, seed=73)
    return ctx, mts, pos, quat


def test_adjoint_with_far_pairs_is_the_energy_gradient(gpu):
    """the adjoint linearizes the mutual solve itself (intra coupling through near pairs only), so
    it is the derivative of the solved energy also where far pairs exist; the reference's forward
    sweeps couple every polarized pair and differ there"""
    import pbsam_amd
    _, mts, pos, quat = far_pair_types(gpu)
    S = exact_state(gpu, mts, pos, quat, tol=1e-26)
    fa, ta, _ = S.adjoint_forces(tol=ADJ_TOL, max_rounds=1000)
    h = 1e-4
    for mol in range(2):
        fd, td = np.zeros(3), np.zeros(3)
        for d in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[mol, d] += h
            pm[mol, d] -= h
            ep, em = (exact_state(gpu, mts, x, quat, tol=1e-26).energy()[mol] for x in (pp, pm))
            fd[d] = -(ep - em) / (2 * h)
            qp, qm = quat.copy(), quat.copy()
            qp[mol] = turned(quat[mol], d, h)
            qm[mol] = turned(quat[mol], d, -h)
            ep, em = (exact_state(gpu, mts, pos, x, tol=1e-26).energy()[mol] for x in (qp, qm))
            td[d] = -(ep - em) / (2 * h)
        assert np.abs(fa[mol] - fd).max() < 1e-6 * np.linalg.norm(fd), (mol, fa[mol], fd)
        assert np.abs(ta[mol] - td).max() < 1e-6 * np.linalg.norm(td), (mol, ta[mol], td)
    S.gradient(tol=1e-22, max_rounds=4000)
    f, _ = S.forces()
    ff = np.array([quat_rot(quat[i], f[i]) for i in range(2)])
    gap = relerr(ff, fa)
    print("forward-mode vs adjoint with far pairs:", gap, flush=True)
    assert gap > 1e-6  # measured 2.2e-5: the forward sweeps' far-pair coupling is not the solve's


def test_exact_path_without_stored_maps(gpu):
    """with the store cap at zero the exact solve makes the numerical re-expansions in the kernel
    (k_p2l) and the intra re-expansions generate their rotations; the results equal the stored
    path's to rounding"""
    import pbsam_amd
    res = []
    for limit in (None, 0):
        ctx, mts, pos, quat = far_pair_types(gpu, store_limit=limit)
        S = exact_state(ctx, mts, pos, quat, tol=1e-26)
        f, t, _ = S.adjoint_forces(tol=ADJ_TOL, max_rounds=1000)
        res.append((S.energy(), f, t, ctx.stored()[0]))
        S.close()
        for m in mts:
            m.close()
        if limit is not None:
            ctx.close()
    (e0, f0, t0, st0), (e1, f1, t1, st1) = res
    assert st0 > 0 and st1 == 0
    assert relerr(e1, e0) < 1e-11 and relerr(f1, f0) < 1e-10 and relerr(t1, t0) < 1e-10


def test_adjoint_mixed_pole_orders_equal_single_systems(gpu):
    """a batch handle with systems at p = 10 and p = 30 (split into one part per order): a pair
    and a group of three at p = 10, a pair at p = 30; the exact solve and the adjoint of every
    group equal the same systems solved alone"""
    import pbsam_amd
    from pbsam_amd import synth
    types = {}
    for t, (p, n, npts, seed) in enumerate([(10, 4, 400, 951), (10, 5, 400, 952), (30, 3, 1200, 953),
                                            (30, 3, 1200, 954)]):
        m = synth.molecule(n, seed=seed)
        g = pbsam_amd.MolType(gpu, m["cen"], m["rad"], m["chg_xyz"], m["chg_q"], p, eps_in=4.0, rcen=m["rcen"],
                              npts=[npts] * n)
        g.selfpol(tol=0, max_rounds=0)
        g.values()
        types[t] = (m, g)
    rng = np.random.default_rng(955)
    systems = []
    for a, b in ((0, 1), (2, 3)):
        d = synth.ball_radius(len(types[a][0]["rad"])) + synth.ball_radius(len(types[b][0]["rad"])) + 6.0
        pos, quat = synth.place_pair(types[a][0], types[b][0], d, seed=956 + a)
        systems.append(([a, b], pos, quat))
    d = 2 * synth.ball_radius(5) + 6.0
    pos = np.array([[0.0, 0.0, 0.0], [d, 0.0, 0.0], [0.5 * d, 0.5 * np.sqrt(3.0) * d, -0.3]])
    systems.insert(1, ([0, 1, 1], pos, np.array([synth.random_quat(rng) for _ in range(3)])))

    def solved(mts, pos, quat, groups=None):
        S = pbsam_amd.System(gpu, mts, pos, quat, groups=groups)
        S.set_cutoffs(100.0, 100.0)
        S.set_flags(pbsam_amd.EXACT_DERIVATIVE)
        n, rc = S.solve_exact(tol=1e-24, max_rounds=3000)
        assert rc == 0, n
        f, t, _ = S.adjoint_forces(tol=ADJ_TOL, max_rounds=20000)
        e = S.energy()
        S.close()
        return f, t, e

    alone = [solved([types[k][1] for k in ks], pos, quat) for ks, pos, quat in systems]
    fb, tb, eb = solved([types[k][1] for ks, _, _ in systems for k in ks], np.vstack([p for _, p, _ in systems]),
                        np.vstack([q for _, _, q in systems]), groups=[len(ks) for ks, _, _ in systems])
    o = 0
    for (ks, _, _), (f, t, e) in zip(systems, alone):
        n = len(ks)
        assert relerr(fb[o:o + n], f) < 1e-9, (ks, fb[o:o + n], f)
        assert relerr(tb[o:o + n], t) < 1e-9
        assert relerr(eb[o:o + n], e) < 1e-10
        o += n


def test_adjoint_after_a_move_equals_a_fresh_system(gpu, orc):
    """the inter rotations and full-order translation matrices the exact solve leaves for the
    adjoint are made again after a move (set_pose) and after mutual rounds that write the
    matrices: the results equal a fresh handle's at the same poses"""
    import pbsam_amd
    gs, _, pos, quat = pair(gpu, orc, 4, 10, 400)

    def exact(S):
        n, rc = S.solve_exact(tol=1e-24, max_rounds=3000)
        assert rc == 0, n
        return S.adjoint_forces(tol=ADJ_TOL, max_rounds=20000)[:2]

    def fresh(pos, quat):
        S = pbsam_amd.System(gpu, gs, pos, quat)
        S.set_cutoffs(100.0, 100.0)
        S.set_flags(pbsam_amd.EXACT_DERIVATIVE)
        return S

    S = fresh(pos, quat)
    exact(S)
    pos2, quat2 = pos.copy(), quat.copy()
    pos2[1] += np.array([0.7, -0.4, 0.3])
    quat2[1] = turned(quat[1], 2, 0.2)
    S.set_pose(1, pos2[1], quat2[1])
    f2, t2 = exact(S)
    f2r, t2r = exact(fresh(pos2, quat2))
    assert relerr(f2, f2r) < 1e-10 and relerr(t2, t2r) < 1e-10
    assert relerr(f2, exact(fresh(pos, quat))[0]) > 1e-3  # the move changes the forces
    # mutual rounds between the exact solve and the adjoint write the translation matrices
    S.solve_exact(tol=1e-24, max_rounds=3000)
    S.solve(tol=1e-18, max_rounds=4000)
    f3, t3 = S.adjoint_forces(tol=ADJ_TOL, max_rounds=20000)[:2]
    assert relerr(f3, f2r) < 1e-8 and relerr(t3, t2r) < 1e-8


def test_gradcheck_rigid_energy_three_molecules(gpu, orc):
    """torch.autograd.gradcheck of every molecule's energy in the triangle of three molecules
    over its six rigid-body coordinates (the adjoint's per-molecule solves in backward)"""
    import pbsam_amd
    pbsam_amd.load()   # the library's HIP runtime first, then torch
    import torch
    from pbsam_amd import autograd
    gs, pos, quat = triangle(gpu, orc, seed=27)
    for j in range(3):
        ev = autograd.Evaluator(gpu, gs, pos, quat, j, tol=1e-18, max_rounds=4000, adj_tol=ADJ_TOL,
                                adj_sweeps=20000)
        x = torch.zeros(6, dtype=torch.float64, requires_grad=True)
        _, g = autograd.solve_rigid(ev, np.zeros(6), True)
        scale = float(np.linalg.norm(g))
        assert scale > 0
        assert torch.autograd.gradcheck(lambda y: autograd.RigidEnergy.apply(y, ev), (x,), eps=1e-4,
                                        atol=1e-6 * scale, rtol=1e-6)
