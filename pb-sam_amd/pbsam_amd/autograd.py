"""torch.autograd wrapper of the device energy and forces, for gradient checks only.

`MoleculeEnergy.apply(r, ev)` is the energy of molecule j, sum over its spheres of <H, LS>
(CMolecule::computePot, src/moldynamics.cpp:509-521), as a function of its centre r (a
CPU float64 tensor of 3) with every other pose fixed.  Its backward is -(orientation * the
force pbsam_forces returns), computed on the device with PBSAM_EXACT_DERIVATIVE, under
which the reference's forward-mode gradient phase is the derivative of that energy.

`RigidEnergy.apply(x, ev)` is the same energy as a function of x = (displacement of molecule
j's centre, rotation vector of a lab-frame turn of j about its centre), six float64 values;
its backward at x is -(force, torque) of pbsam_adjoint_forces at the displaced pose (exact at
x = 0, where the turn's derivative is the torque), from the adjoint solve (any number of
molecules: the adjoint solves per molecule when there are more than two).

Both exist so that torch.autograd.gradcheck can compare the device's analytic derivatives
with central differences of the device's own energy; nothing in the library depends on torch.

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
                 max_rounds=2000, adj_tol=1e-20, adj_sweeps=5000):
        self.ctx, self.mts = ctx, list(moltypes)
        self.pos, self.quat = np.array(pos, float), np.array(quat, float)
        self.j, self.cutoffs = j, cutoffs
        self.tol, self.grad_tol, self.max_rounds = tol, grad_tol, max_rounds
        self.adj_tol, self.adj_sweeps = adj_tol, adj_sweeps

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


def _turn(x, q):
    """quaternion of a lab-frame turn by the rotation vector x, composed with q"""
    a = float(np.linalg.norm(x))
    t = np.r_[1.0, 0.0, 0.0, 0.0] if a == 0.0 else np.r_[np.cos(a / 2), np.sin(a / 2) * np.asarray(x) / a]
    w1, x1, y1, z1 = t
    w2, x2, y2, z2 = q
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


def solve_rigid(ev, x, grads):
    """energy of molecule ev.j after moving it by x[:3] and turning it by x[3:], and (grads)
    its adjoint force and torque (lab frame)"""
    pos, quat = ev.pos.copy(), ev.quat.copy()
    pos[ev.j] = pos[ev.j] + x[:3]
    quat[ev.j] = _turn(x[3:], quat[ev.j])
    S = pbsam_amd.System(ev.ctx, ev.mts, pos, quat)
    S.set_cutoffs(*ev.cutoffs)
    S.set_flags(pbsam_amd.EXACT_DERIVATIVE)
    rounds, rc = S.solve(tol=ev.tol, max_rounds=ev.max_rounds)
    if rc != pbsam_amd.OK:
        raise RuntimeError("mutual solve did not converge (%d rounds)" % rounds)
    e = float(S.energy()[ev.j])
    if not grads:
        return e, None
    f, t, _ = S.adjoint_forces(tol=ev.adj_tol, max_rounds=ev.adj_sweeps)
    return e, np.r_[f[ev.j], t[ev.j]]


class RigidEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, ev):
        e, g = solve_rigid(ev, x.detach().cpu().numpy().astype(np.float64), True)
        ctx.save_for_backward(torch.as_tensor(-g, dtype=torch.float64))
        return torch.tensor(e, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_e):
        (de_dx,) = ctx.saved_tensors
        return grad_e * de_dx, None


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
