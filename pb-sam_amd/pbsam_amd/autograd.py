"""torch.autograd wrapper of the device energy and forces, for gradient checks only.

`MoleculeEnergy.apply(r, ev)` is the energy of molecule j, sum over its spheres of <H, LS>
(CMolecule::computePot, src/moldynamics.cpp:509-521), as a function of its centre r (a
CPU float64 tensor of 3) with every other pose fixed.  Its backward is -(orientation * the
force pbsam_forces returns), computed on the device with PBSAM_EXACT_DERIVATIVE, under
which the reference's forward-mode gradient phase is the derivative of that energy.  It
exists so that torch.autograd.gradcheck can compare the device's analytic derivative with
central differences of the device's own energy; nothing in the library depends on torch.

Load the library (pbsam_amd.load()) before importing torch: both ship a HIP runtime with
the same soname, and the process keeps the first one loaded.
"""
import numpy as np
import torch

import pbsam_amd


def _quat_rot(q, v):
    w, u = q[0], q[1:]
    t = np.cross(u, v)
    return v + 2 * (w * t + np.cross(u, t))


class Evaluator:
    """Molecule types, poses and solve settings of one system; molecule j's centre varies."""

    def __init__(self, ctx, moltypes, pos, quat, j, cutoffs=(100.0, 100.0), tol=1e-16, grad_tol=1e-16,
                 max_rounds=2000):
        self.ctx, self.mts = ctx, list(moltypes)
        self.pos, self.quat = np.array(pos, float), np.array(quat, float)
        self.j, self.cutoffs = j, cutoffs
        self.tol, self.grad_tol, self.max_rounds = tol, grad_tol, max_rounds

    def solve(self, r, forces):
        pos = self.pos.copy()
        pos[self.j] = r
        S = pbsam_amd.System(self.ctx, self.mts, pos, self.quat)
        S.set_cutoffs(*self.cutoffs)
        S.set_flags(pbsam_amd.EXACT_DERIVATIVE)
        rounds, rc = S.solve(tol=self.tol, max_rounds=self.max_rounds)
        if rc != pbsam_amd.OK:
            raise RuntimeError("mutual solve did not converge (%d rounds)" % rounds)
        e = float(S.energy()[self.j])
        if not forces:
            return e, None
        S.gradient(tol=self.grad_tol, max_rounds=self.max_rounds)
        f, _ = S.forces()
        return e, _quat_rot(self.quat[self.j], f[self.j])   # lab-frame force on molecule j


class MoleculeEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, r, ev):
        e, f = ev.solve(r.detach().cpu().numpy().astype(np.float64), True)
        ctx.save_for_backward(torch.as_tensor(-f, dtype=torch.float64))
        return torch.tensor(e, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_e):
        (de_dr,) = ctx.saved_tensors
        return grad_e * de_dr, None
