"""The transposition rules the adjoint kernels rest on, checked on the oracle's operators as
explicit matrices (columns = the operator applied to unit blocks).  With W = diag(1 for the
m = 0 coefficient, 2 for the real and imaginary parts of m > 0) -- the weights of the reference's
block inner product (expansion.h:706-741):

    X(b <- a)^T = W X(a <- b) W^-1     re-expansion of a's block onto b, screened (H) and
                                       unscreened (F), and its reverse (-P, radii swapped)
    R_fwd^T     = W R_back W^-1        rotation of a block by an orientation and its inverse

A transposed operand, a dropped weight or the wrong direction fails them by O(1)."""
import numpy as np
import pytest

import pyoracle

P_ORD = 6
L = P_ORD * P_ORD


def weights(p):
    w = np.full(p * p, 2.0)
    for n in range(p):
        w[n * n] = 1.0
    return w


def matrix(f):
    cols = []
    for c in range(L):
        e = np.zeros(L)
        e[c] = 1.0
        cols.append(f(e))
    return np.array(cols).T


@pytest.fixture(scope="module")
def orc():
    return pyoracle.Oracle(0.073)


@pytest.mark.parametrize("P,a,b", [((3.0, -7.0, 12.0), 2.5, 4.0), ((-9.0, 2.0, -1.5), 5.0, 2.0)])
def test_reexpansion_transpose_is_the_reverse(orc, P, a, b):
    P = np.array(P)
    W = weights(P_ORD)
    X = matrix(lambda e: orc.xform_h(P, a, b, P_ORD, e))
    Xr = matrix(lambda e: orc.xform_h(-P, b, a, P_ORD, e))
    assert np.abs(X.T - W[:, None] * Xr / W[None, :]).max() < 1e-14 * np.abs(X).max()
    # and the plain transpose is not the reverse
    assert np.abs(X.T - Xr).max() > 0.1 * np.abs(X).max()
    z = np.zeros(L)
    for which in (0, 1):  # the intra re-expansion's H (screened) and F (unscreened) parts
        def fwd(e, P=P, a=a, b=b):
            oH, oF = orc.xform_intra(P, a, b, P_ORD, P_ORD, e if which == 0 else z, e if which == 1 else z)
            return oH if which == 0 else oF

        def rev(e, P=P, a=a, b=b):
            oH, oF = orc.xform_intra(-P, b, a, P_ORD, P_ORD, e if which == 0 else z, e if which == 1 else z)
            return oH if which == 0 else oF
        Xi, Xir = matrix(fwd), matrix(rev)
        assert np.abs(Xi.T - W[:, None] * Xir / W[None, :]).max() < 1e-14 * np.abs(Xi).max()


def test_rotation_transpose_is_the_inverse(orc):
    q = np.array([0.3, -0.5, 0.7, 0.2])
    q /= np.linalg.norm(q)
    W = weights(P_ORD)
    Rf = matrix(lambda e: orc.rotate_quat(q, P_ORD, 1, e))
    Rb = matrix(lambda e: orc.rotate_quat(q, P_ORD, 0, e))
    assert np.abs(Rb @ Rf - np.eye(L)).max() < 1e-12
    assert np.abs(Rf.T - W[:, None] * Rb / W[None, :]).max() < 1e-14
