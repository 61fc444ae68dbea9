"""torch.autograd.gradcheck of the device's exact-derivative forces (pbsam_amd.autograd):
the analytic derivative of each molecule's energy with respect to its centre, from the
gradient phase with PBSAM_EXACT_DERIVATIVE, against torch's central differences of the
device energy (step 1e-4 A), both molecules of a two-molecule system at p = 10 and p = 30."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ns,p,npts", [(5, 10, 400), (3, 30, 1200)])
def test_gradcheck_molecule_energy(gpu, ns, p, npts):
    import pbsam_amd
    pbsam_amd.load()   # the library's HIP runtime first, then torch
    import torch
    from pbsam_amd import autograd, synth
    mols = [synth.molecule(ns, seed=s) for s in (14, 15)]
    mts = []
    for m in mols:
        mt = pbsam_amd.MolType(gpu, m["cen"], m["rad"], m["chg_xyz"], m["chg_q"], p, eps_in=4.0, rcen=m["rcen"],
                               npts=[npts] * ns)
        mt.selfpol(tol=0, max_rounds=0)
        mt.values()
        mts.append(mt)
    pos, quat = synth.place_pair(mols[0], mols[1], 2 * synth.ball_radius(ns) + 5.0, seed=21)
    for j in range(2):
        ev = autograd.Evaluator(gpu, mts, pos, quat, j, tol=1e-16, max_rounds=2000)
        r = torch.tensor(pos[j], dtype=torch.float64, requires_grad=True)
        e, f = ev.solve(pos[j], True)
        scale = float(np.linalg.norm(f))
        assert scale > 0
        assert torch.autograd.gradcheck(lambda x: autograd.MoleculeEnergy.apply(x, ev), (r,), eps=1e-4,
                                        atol=1e-6 * scale, rtol=1e-6)
