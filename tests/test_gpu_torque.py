"""Exact torques (pbsam_torques_exact): the gradient sweeps of polarize_mutual run for a
rotation of each molecule about its centre, under PBSAM_EXACT_DERIVATIVE.  -dE/domega of each
molecule's energy against central differences of the device energy under rotations of
1e-4 rad (lab axes), against the oracle's restatement of the same sweeps, and the
reference's c x f torque for comparison (it is not the derivative: its surface term is
disabled and its arm is the molecule-frame centre)."""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu


def qmul(a, b):
    w1, x1, y1, z1 = a
    w2, x2, y2, z2 = b
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


def turned(q, a, h):
    e = np.zeros(3)
    e[a] = 1.0
    return qmul(np.r_[np.cos(h / 2), np.sin(h / 2) * e], q)


@pytest.mark.parametrize("ns,p,npts", [(4, 10, 400), (3, 30, 1200)])
def test_exact_torque_is_the_energy_derivative(gpu, orc, ns, p, npts):
    import pbsam_amd
    import pyoracle
    from test_gpu_exact import pair
    gs, os_, pos, quat = pair(gpu, orc, ns, p, npts)

    def system(q):
        S = pbsam_amd.System(gpu, gs, pos, q)
        S.set_cutoffs(100.0, 100.0)
        S.set_flags(pbsam_amd.EXACT_DERIVATIVE)
        r, rc = S.solve(tol=1e-16, max_rounds=2000)
        assert rc == 0, r
        return S

    S = system(quat)
    tq = S.torques_exact()
    h = 1e-4
    for mol in range(2):
        fd = np.zeros(3)
        for a in range(3):
            qp, qm = quat.copy(), quat.copy()
            qp[mol] = turned(quat[mol], a, h)
            qm[mol] = turned(quat[mol], a, -h)
            fd[a] = -(system(qp).energy()[mol] - system(qm).energy()[mol]) / (2 * h)
        assert np.abs(tq[mol] - fd).max() < 1e-6 * np.linalg.norm(fd), (mol, tq[mol], fd)
    # the oracle's restatement of the same sweeps
    pyoracle.exact_derivative(True)
    try:
        So = pyoracle.System(orc, os_, pos, quat)
        So.set_cutoffs(100.0, 100.0)
        So.lists()
        So.mutual(0, 1e-16, 2000, hrot_refresh=True)
        So.gradient_rot(1e-16, 2000)
        to = np.array([So.grad_energy(i) for i in range(2)])
    finally:
        pyoracle.exact_derivative(False)
    assert relerr(tq, to) < 1e-7, (tq, to)
    # forces after the torques need the translational sweeps again
    with pytest.raises(pbsam_amd.PbsamError):
        S.forces()
    S.gradient(tol=1e-16, max_rounds=2000)
    f, t_ref = S.forces()
    assert relerr(t_ref, tq) > 1e-3  # the reference's torque is another quantity


def test_exact_torque_needs_the_settings(gpu, orc):
    import pbsam_amd
    from test_gpu_exact import pair
    gs, _, pos, quat = pair(gpu, orc, 3, 10, 300)
    S = pbsam_amd.System(gpu, gs, pos, quat)
    S.set_cutoffs(100.0, 100.0)
    S.solve()
    with pytest.raises(pbsam_amd.PbsamError) as e:
        S.torques_exact()
    assert e.value.code == pbsam_amd.ERR_STATE
    S.set_flags(pbsam_amd.EXACT_DERIVATIVE)
    S.set_cutoffs(5.0, 100.0)   # interaction partners without polarization partners
    S.solve()
    with pytest.raises(pbsam_amd.PbsamError) as e:
        S.torques_exact()
    assert e.value.code == pbsam_amd.ERR_UNSUPPORTED
