"""HIP kernels against the CPU oracle, one batch of independent jobs per call."""
import numpy as np
import pytest

from conftest import relerr, rot_close

pytestmark = pytest.mark.gpu
TOL = 1e-11


def pairs(rng, n, near=False):
    P, a = [], []
    for i in range(n):
        a_s, a_t = rng.uniform(2, 6), rng.uniform(2, 6)
        gap = rng.uniform(0.2, 4.0) if near else rng.uniform(5.0, 40.0)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        if i == 0:
            d = np.array([0.0, 0.0, 1.0])   # theta = 0 (singular rotation)
        if i == 1:
            d = np.array([0.0, 0.0, -1.0])  # theta = pi
        P.append(d * (a_s + a_t + gap))
        a.append((a_s, a_t))
    return np.array(P), np.array(a)


@pytest.mark.parametrize("p", [6, 10, 30])
def test_xform_h(gpu, orc, p):
    rng = np.random.default_rng(p)
    n = 12
    P, a = pairs(rng, n)
    orders = rng.integers(1, p + 1, n)
    orders[2] = p
    blocks = rng.normal(size=(n, p * p)) * 0.1
    got = gpu.xform_h(P.ravel(), a.ravel(), orders, blocks, p)
    for i in range(n):
        q = int(orders[i])
        ref = np.zeros(p * p)
        ref[:q * q] = orc.xform_h(P[i], a[i, 0], a[i, 1], q, blocks[i, :q * q])
        assert relerr(got[i], ref) < TOL, (i, q)


@pytest.mark.parametrize("p", [10, 30])
def test_xform_intra(gpu, orc, p):
    rng = np.random.default_rng(100 + p)
    n = 8
    P, a = pairs(rng, n)
    orders = rng.integers(1, p + 1, 2 * n)
    H = rng.normal(size=(n, p * p)) * 0.1
    F = rng.normal(size=(n, p * p)) * 0.1
    gH, gF = gpu.xform_intra(P.ravel(), a.ravel(), orders, H, F, p)
    for i in range(n):
        pH, pF = int(orders[2 * i]), int(orders[2 * i + 1])
        q = max(pH, pF)
        rH, rF = orc.xform_intra(P[i], a[i, 0], a[i, 1], pH, pF, H[i, :q * q], F[i, :q * q])
        eH, eF = np.zeros(p * p), np.zeros(p * p)
        eH[:q * q], eF[:q * q] = rH, rF
        assert relerr(gH[i], eH) < TOL and relerr(gF[i], eF) < TOL, (i, pH, pF)


@pytest.mark.parametrize("p,npts", [(10, 400), (30, 1800)])
def test_xform_numerical(gpu, orc, p, npts):
    rng = np.random.default_rng(7 + p)
    a_s = 3.5
    pts = orc.sphere_points(npts, a_s)
    qF, qH = rng.normal(size=npts) * 1e-3, rng.normal(size=npts) * 1e-3
    qF[5] = 0.0
    tQF, tQH = np.abs(qF).sum(), np.abs(qH).sum()
    n = 6
    P, a = pairs(rng, n, near=True)
    a[:, 0] = a_s
    P = np.array([d / np.linalg.norm(d) * (a_s + a[i, 1] + 0.5 + i) for i, d in enumerate(P)])
    gF, gH = gpu.xform_num(P.ravel(), a.ravel(), pts.ravel(), qF, qH, tQF, tQH, p)
    for i in range(n):
        rF, rH = orc.xform_num(P[i], pts.ravel(), qF, qH, tQF, tQH, a[i, 0], a[i, 1], p)
        assert relerr(gF[i], rF) < TOL and relerr(gH[i], rH) < TOL, i


@pytest.mark.parametrize("p", [10, 30])
def test_rotate_quaternion(gpu, orc, p):
    """Row n agrees to TOL, or to 1e-15 * 4^n where that is larger: the rotation
    recursion (the reference's, rotcoeff.cpp:398-441, restated by both sides) amplifies
    rounding by 3-4x per level m <= n, so two correct evaluations of it in different
    operation orders differ by that much."""
    rng = np.random.default_rng(55 + p)
    n = 6
    qs = rng.normal(size=(n, 4))
    qs[0] = [1.0, 0.0, 0.0, 0.0]
    qs[1] = [0.6, 0.0, 0.0, 0.8]       # theta = 0 with xi != 0
    qs[2] = [0.0, 1.0, 0.0, 0.0]       # theta = pi
    qs /= np.linalg.norm(qs, axis=1)[:, None]
    blocks = rng.normal(size=(n, p * p))
    for fwd in (1, 0):
        got = gpu.rotate_quat(qs.ravel(), fwd, blocks, p)
        for i in range(n):
            ref = orc.rotate_quat(qs[i], p, fwd, blocks[i])
            scale = np.abs(ref).max()
            for k in range(p):
                err = np.abs(got[i, k * k:(k + 1) * (k + 1)] - ref[k * k:(k + 1) * (k + 1)]).max()
                assert err <= max(TOL, 1e-15 * 4.0 ** k) * scale, (i, fwd, k)


@pytest.mark.parametrize("p", [10, 30])
def test_xform_h_grad(gpu, orc, p):
    """Gradient re-expansion (k_xgrad + k_trans grad pass) against the oracle, including
    the singular directions theta = 0 and pi (closed forms with the Q table)."""
    rng = np.random.default_rng(300 + p)
    n = 8
    P, a = pairs(rng, n)
    orders = rng.integers(1, p + 1, n)
    orders[:3] = p
    scale = 0.6 ** np.repeat(np.arange(p), 2 * np.arange(p) + 1)   # decaying like physical blocks
    blocks = rng.normal(size=(n, p * p)) * scale
    got = gpu.xform_h_grad(P.ravel(), a.ravel(), orders, blocks, p)
    for i in range(n):
        q = int(orders[i])
        r = orc.xform_h_grad(P[i], a[i, 0], a[i, 1], q, blocks[i, :q * q]).reshape(3, q * q)
        for d in range(3):
            ref = np.zeros(p * p)
            ref[:q * q] = r[d]
            assert rot_close(got[i, d * p * p:(d + 1) * p * p], ref, p, 1e-10), (i, q, d)
