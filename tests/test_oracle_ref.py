"""The CPU oracle against known-answer vectors of the reference itself.

tests/golden/ref_kat.npz holds inputs and outputs of the reference's own code (built
from its sources by oracle/Makefile.ref, driven by oracle/ref_harness.cpp and written by
tests/golden/make_ref_fixtures.py).  The reference is compiled with N_POLES = 15.
"""
import numpy as np
import pytest

from conftest import relerr

TOL = 2e-12


def test_spherical_harmonics(orc, golden):
    for i, (P, th, ph) in enumerate(golden["sh_in"]):
        got = orc.sh(int(P), th, ph)
        assert relerr(got, golden["out_sh_%d" % i]) < TOL, i


def test_bessel(orc, golden):
    for i, (L, v) in enumerate(golden["bes_in"]):
        assert relerr(orc.bessel_k(int(L), v), golden["out_besk_%d" % i]) < TOL
        assert relerr(orc.bessel_i(int(L), v), golden["out_besi_%d" % i]) < TOL


@pytest.mark.parametrize("i,kind", [(0, "lexp"), (1, "lexp"), (2, "mexp"), (3, "mexp")])
def test_point_expansions(orc, golden, i, kind):
    a = golden["%s_%d_in" % (kind, i)]
    p, bk, sc = int(a[0]), int(a[1]), a[2]
    n = (len(a) - 3) // 4
    q, x = a[3:3 + n], a[3 + n:].reshape(n, 3)
    f = orc.local_expan if kind == "lexp" else orc.mul_expan
    assert relerr(f(q, x, p, bk, sc), golden["out_%s_%d" % (kind, i)]) < TOL


def test_gradient_local_expansion(orc, golden):
    a = golden["gradlocal_in"]
    p, bk, sc = int(a[0]), int(a[1]), a[2]
    n = (len(a) - 3) // 6
    g, x = a[3:3 + 3 * n].reshape(n, 3), a[3 + 3 * n:].reshape(n, 3)
    assert relerr(orc.grad_local(g, x.ravel(), p, bk, sc), golden["out_gradlocal"]) < TOL


def test_rotate(orc, golden):
    for i in range(int(golden["rotate_n"])):
        a = golden["rotate_%d_in" % i]
        th, ph, xi, p, p2, fwd = a[:6]
        got = orc.rotate(th, ph, xi, int(p), int(p2), int(fwd), a[6:])
        ref = golden["out_rotate_%d" % i]
        assert relerr(got, ref[:-1]) < TOL, i


def test_rotate_derivatives(orc, golden):
    for i in range(6):
        a = golden["drot_%d_in" % i]
        th, ph, xi, p, which, fwd = a[:6]
        got = orc.rotate_grad(th, ph, xi, int(p), int(which), int(fwd), a[6:])
        assert relerr(got, golden["out_drot_%d" % i]) < TOL, i


def test_rotate_quaternion(orc, golden):
    for i in range(4):
        a = golden["rotquat_%d_in" % i]
        q, p, fwd = a[:4], int(a[4]), int(a[5])
        assert relerr(orc.rotate_quat(q, p, fwd, a[6:]), golden["out_rotquat_%d" % i]) < TOL, i


def test_rotate_quaternion_singular(orc, golden):
    """orientations whose Euler split has theta = 0 or pi, both directions"""
    for i in range(4):
        for fwd in (0, 1):
            a = golden["rotqs_%d_%d_in" % (i, fwd)]
            got = orc.rotate_quat(a[:4], int(a[4]), int(a[5]), a[6:])
            assert relerr(got, golden["out_rotqs_%d_%d" % (i, fwd)]) < TOL, (i, fwd)


def test_translate(orc, golden):
    for i in range(4):
        a = golden["trans_%d_in" % i]
        rho, kap, p, si, so, grad = a[:6]
        got = orc.translate(rho, kap, int(p), si, so, int(grad), a[6:])
        assert relerr(got, golden["out_trans_%d" % i]) < TOL, i


def test_xform_h(orc, golden):
    for i in range(4):
        a = golden["xformh_%d_in" % i]
        P, a_s, a_t, p = a[:3], a[3], a[4], int(a[5])
        assert relerr(orc.xform_h(P, a_s, a_t, p, a[6:]), golden["out_xformh_%d" % i]) < TOL, i


def test_xform_h_gradient(orc, golden):
    for i in range(2):
        a = golden["xformhg_%d_in" % i]
        P, a_s, a_t, p = a[:3], a[3], a[4], int(a[5])
        assert relerr(orc.xform_h_grad(P, a_s, a_t, p, a[6:]), golden["out_xformhg_%d" % i]) < TOL, i


def test_xform_intra(orc, golden):
    for i in range(2):
        a = golden["xformi_%d_in" % i]
        P, a_s, a_t, pH, pF = a[:3], a[3], a[4], int(a[5]), int(a[6])
        p = max(pH, pF)
        h, f = a[7:7 + p * p], a[7 + p * p:]
        oH, oF = orc.xform_intra(P, a_s, a_t, pH, pF, h, f)
        ref = golden["out_xformi_%d" % i]
        assert relerr(oH, ref[:p * p]) < TOL and relerr(oF, ref[p * p:]) < TOL, i


def test_xform_numerical(orc, golden):
    for i in range(2):
        a = golden["xformn_%d_in" % i]
        P, a_s, a_t, tqf, tqh = a[:3], a[3], a[4], a[5], a[6]
        n = (len(a) - 7) // 5
        pts = a[7:7 + 3 * n]
        qF, qH = a[7 + 3 * n:7 + 4 * n], a[7 + 4 * n:]
        oF, oH = orc.xform_num(P, pts, qF, qH, tqf, tqh, a_s, a_t, 15)
        ref = golden["out_xformn_%d" % i]
        assert relerr(oH, ref[:225]) < TOL and relerr(oF, ref[225:]) < TOL, i


def test_compute_order(orc, golden):
    for i, (rho, tq, sc) in enumerate(golden["order_in"]):
        assert orc.compute_order(rho, tq, sc, 15) == int(golden["out_order_%d" % i][0])


def test_sphere_points(orc, golden):
    for i, (N, r) in enumerate(golden["spx_in"]):
        got = orc.sphere_points(int(N), r)
        assert relerr(got.ravel(), golden["out_spx_%d" % i]) < 1e-15


def test_imat_buried(orc, golden):
    got = orc.imat_build(15, golden["imat_spe"], golden["imat_spb"])
    assert relerr(got, golden["out_imat"]) < 1e-12


def _moltype(orc, golden, i, rc=30.0):
    import pyoracle
    return pyoracle.MolType(orc, golden["mol_%d_cen" % i], golden["mol_%d_rad" % i], golden["mol_%d_xyz" % i],
                            golden["mol_%d_q" % i], 15, eps_in=4.0, rcen=golden["mol_%d_rcen" % i], npts=None,
                            npts_cap=10000, intra_rcutoff=rc)


@pytest.mark.parametrize("i", [0, 1])
def test_moltype_setup(orc, golden, i):
    """points, constants, fixed-charge expansions and initial surface charges"""
    mt = _moltype(orc, golden, i)
    out = golden["out_moltype_%d" % i]
    o = 0
    for k in range(mt.ns):
        nexp, nspx, nneigh, nnear = (int(x) for x in out[o:o + 4])
        o += 4
        n_got, tot = mt.npts(k)
        assert (n_got, tot) == (nexp, nspx)
        near, far = mt.lists(k)
        assert len(near) == nnear
        pts = out[o:o + 3 * nexp]
        o += 3 * nexp
        assert relerr(mt.points(k).ravel(), pts) < 1e-15
        consts, dfix, vfix, e0 = mt.consts(k)
        ref_c = out[o:o + 150].reshape(10, 15)
        o += 150
        assert relerr(consts, ref_c) < TOL
        assert relerr(dfix, out[o:o + 225]) < TOL
        o += 225
        assert relerr(vfix, out[o:o + 225]) < TOL
        o += 225
        assert relerr(e0, out[o:o + 225]) < TOL
        o += 225
        qH = mt.get(k)["qH"]
        assert relerr(qH, out[o:o + nexp]) < TOL
        o += nexp
    assert o == len(out)


@pytest.mark.parametrize("i", [0, 1])
def test_moltype_values(orc, golden, i):
    """exposed self charges and frozen intra-molecular re-expansions (near numerical,
    near analytical and far analytical pairs)"""
    rc = float(golden["molvals_%d_rc" % i])
    mt = _moltype(orc, golden, i, rc)
    mt.set_self(golden["molvals_%d_F" % i], golden["molvals_%d_H" % i])
    mt.values()
    out = golden["out_molvals_%d" % i]
    o = 0
    import pyoracle
    lib = pyoracle.lib()
    for k in range(mt.ns):
        n, _ = mt.npts(k)
        qF = out[o:o + n]
        qH = out[o + n:o + 2 * n]
        o += 2 * n + 2
    # the oracle stores the self charges internally; compare the re-expansions
    for k in range(mt.ns):
        LF, LH, LFf, LHf = mt.intra(k)
        ref = out[o:o + 4 * 225].reshape(4, 225)
        o += 4 * 225
        for got, r in zip((LF, LH, LFf, LHf), ref):
            if np.abs(r).max() == 0:
                assert np.abs(got).max() == 0
            else:
                assert relerr(got, r) < 1e-11, k
    assert o == len(out)
    assert lib is not None
