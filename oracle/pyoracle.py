"""ctypes bindings of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, the smoke check and bench.py's CPU-baseline leg import this module.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

D = ctypes.c_double
I = ctypes.c_int
P = ctypes.c_void_p


def _p(a):
    return None if a is None else a.ctypes.data_as(P)


class MolDesc(ctypes.Structure):
    _fields_ = [("n_sph", I), ("cen", P), ("rad", P), ("n_chg", I), ("chg_xyz", P),
                ("chg_q", P), ("rcen", P), ("p", I), ("eps_in", D), ("npts", P),
                ("npts_cap", I), ("intra_rcutoff", D), ("imats", P), ("imat_npts", I)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.orc_create.restype = P
        L.orc_create.argtypes = [D, D, I]
        L.orc_destroy.argtypes = [P]
        for name, args in [
            ("orc_sh", [P, I, D, D, P]),
            ("orc_bessel_k", [P, I, D, P]),
            ("orc_bessel_i", [P, I, D, P]),
            ("orc_local_expan", [P, I, P, P, I, I, D, P]),
            ("orc_mul_expan", [P, I, P, P, I, I, D, P]),
            ("orc_rotate", [P, D, D, D, I, I, I, P, P]),
            ("orc_rotate_quat", [P, P, I, I, P, P]),
            ("orc_rotate_grad", [P, D, D, D, I, I, I, P, P]),
            ("orc_trans_coeff", [P, D, D, I, D, D, P, P]),
            ("orc_translate", [P, D, D, I, D, D, I, P, P]),
            ("orc_xform_h", [P, P, D, D, I, P, P]),
            ("orc_xform_intra", [P, P, D, D, I, I, P, P, P, P]),
            ("orc_xform_h_grad", [P, P, D, D, I, P, P]),
            ("orc_xform_num", [P, P, I, P, P, P, D, D, D, D, I, P, P]),
            ("orc_grad_local", [P, I, P, P, I, I, D, P]),
            ("orc_sphere_points", [I, D, P]),
            ("orc_imat_build", [P, I, I, P, I, P, P]),
            ("orc_moltype_destroy", [P]),
            ("orc_moltype_points", [P, I, P]),
            ("orc_moltype_consts", [P, I, P, P, P, P]),
            ("orc_selfpol_round_jacobi", [P, P]),
            ("orc_moltype_get", [P, I, P, P, P, P, P]),
            ("orc_moltype_set_self", [P, P, P]),
            ("orc_moltype_get_intra", [P, I, P, P, P, P]),
            ("orc_system_destroy", [P]),
            ("orc_system_set_cutoffs", [P, D, D]),
            ("orc_system_set_pose", [P, I, P, P]),
            ("orc_forces", [P, P, P]),
            ("orc_system_get", [P, I, I, P, P, P, P]),
            ("orc_system_get_grad", [P, I, I, P, P]),
        ]:
            f = getattr(L, name)
            f.argtypes = args
            f.restype = None
        for name, args in [
            ("orc_compute_order", [D, D, D, I]),
            ("orc_moltype_values", [P]),
            ("orc_rot_coeff", [P, D, D, D, I, I, P, P]),
            ("orc_npts_rule", [D, I]),
            ("orc_moltype_nsph", [P]),
            ("orc_moltype_npts", [P, I, P]),
            ("orc_moltype_lists", [P, I, P, P, P]),
            ("orc_selfpol", [P, I, D, I]),
            ("orc_system_lists", [P]),
            ("orc_mutual", [P, I, D, I, I]),
            ("orc_gradient", [P, I, D, I]),
            ("orc_compute_forces", [P, I, D, I, P, P]),
            ("orc_system_counts", [P, P, P]),
        ]:
            f = getattr(L, name)
            f.argtypes = args
            f.restype = I
        L.orc_debug_full_order.argtypes = [I]
        L.orc_debug_full_order.restype = None
        L.orc_set_grad_skip.argtypes = [D]
        L.orc_set_grad_skip.restype = None
        L.orc_set_ext_sign.argtypes = [D]
        L.orc_set_ext_sign.restype = None
        L.orc_set_self_frame.argtypes = [ctypes.c_int]
        L.orc_set_self_frame.restype = None
        L.orc_rot_generator.argtypes = [I, P, P, I]
        L.orc_rot_generator.restype = None
        L.orc_gradient_rot.argtypes = [P, D, I]
        L.orc_gradient_rot.restype = I
        L.orc_grad_energy.argtypes = [P, I, P]
        L.orc_grad_energy.restype = None
        L.orc_moltype_create.restype = P
        L.orc_moltype_create.argtypes = [P, ctypes.POINTER(MolDesc)]
        L.orc_system_create.restype = P
        L.orc_system_create.argtypes = [P, I, P, P, P]
        for name in ("orc_moltype_energy_self",):
            getattr(L, name).restype = D
            getattr(L, name).argtypes = [P]
        L.orc_bd_create.restype = P
        L.orc_bd_create.argtypes = [I, P, P, I, P, P, P, ctypes.c_longlong, I, P, P]
        for name, args in [("orc_bd_destroy", [P]), ("orc_bd_set_system", [P, P]),
                           ("orc_bd_add_contact", [P, I, I, P, P]),
                           ("orc_bd_get", [P, P, P, P, P, P]),
                           ("orc_bd_rng", [ctypes.c_longlong, I, P])]:
            f = getattr(L, name)
            f.argtypes = args
            f.restype = None
        L.orc_bd_step.restype = I
        L.orc_bd_step.argtypes = [P, P]
        L.orc_pot_in_space.restype = None
        L.orc_pot_in_space.argtypes = [P, I, P, P, I]
        L.orc_system_pot.restype = D
        L.orc_system_pot.argtypes = [P, I]
        L.orc_system_total_int_energy.restype = D
        L.orc_system_total_int_energy.argtypes = [P]
        _LIB = L
    return _LIB


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """Context with the physical constants (kappa in 1/A, solvent dielectric)."""

    def __init__(self, kappa, eps_out=78.0, pmax=32):
        self.L = lib()
        self.kappa = kappa
        self.h = self.L.orc_create(kappa, eps_out, pmax)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    # ---- primitives ----
    def sh(self, P, theta, phi):
        out = np.zeros(P * P)
        self.L.orc_sh(self.h, P, theta, phi, _p(out))
        return out

    def bessel_k(self, L, val):
        out = np.zeros(L)
        self.L.orc_bessel_k(self.h, L, val, _p(out))
        return out

    def bessel_i(self, L, val):
        out = np.zeros(L)
        self.L.orc_bessel_i(self.h, L, val, _p(out))
        return out

    def local_expan(self, q, xyz, p, bkappa, scale):
        q, xyz = f64(q), f64(xyz)
        out = np.zeros(p * p)
        self.L.orc_local_expan(self.h, len(q), _p(q), _p(xyz), p, int(bkappa), scale, _p(out))
        return out

    def mul_expan(self, q, xyz, p, bkappa, scale):
        q, xyz = f64(q), f64(xyz)
        out = np.zeros(p * p)
        self.L.orc_mul_expan(self.h, len(q), _p(q), _p(xyz), p, int(bkappa), scale, _p(out))
        return out

    def grad_local(self, g, xyz, p, bkappa, scale):
        g, xyz = f64(g), f64(xyz)
        out = np.zeros(3 * p * p)
        self.L.orc_grad_local(self.h, len(xyz) // 3 if xyz.ndim == 1 else len(xyz), _p(g), _p(xyz), p,
                              int(bkappa), scale, _p(out))
        return out

    def rotate(self, theta, phi, xi, p, p2, fwd, block):
        block = f64(block)
        out = np.zeros(p2 * p2)
        self.L.orc_rotate(self.h, theta, phi, xi, p, p2, int(fwd), _p(block), _p(out))
        return out

    def rotate_quat(self, q, p, fwd, block):
        q, block = f64(q), f64(block)
        out = np.zeros(p * p)
        self.L.orc_rotate_quat(self.h, _p(q), p, int(fwd), _p(block), _p(out))
        return out

    def rotate_grad(self, theta, phi, xi, p, which, fwd, block):
        block = f64(block)
        out = np.zeros(p * p)
        self.L.orc_rotate_grad(self.h, theta, phi, xi, p, which, int(fwd), _p(block), _p(out))
        return out

    def rot_coeff(self, theta, phi, xi, p, nmax=None, grad=False):
        nmax = p if nmax is None else nmax
        n_el = sum((2 * n + 1) * (n + 1) for n in range(nmax))
        out = np.zeros(2 * n_el)
        dout = np.zeros(2 * n_el) if grad else None
        sing = self.L.orc_rot_coeff(self.h, theta, phi, xi, p, nmax, _p(out), _p(dout))
        R = out[0::2] + 1j * out[1::2]
        if grad:
            return R, dout[0::2] + 1j * dout[1::2], sing
        return R, sing

    def trans_coeff(self, rho, kappa_seed, p, scalei, scaleo, grad=False):
        n = sum(1 for m in range(p) for nn in range(m, p) for l in range(nn, p))
        T = np.zeros(n)
        dT = np.zeros(n) if grad else None
        self.L.orc_trans_coeff(self.h, rho, kappa_seed, p, scalei, scaleo, _p(T), _p(dT))
        return (T, dT) if grad else T

    def translate(self, rho, kappa_seed, p, scalei, scaleo, grad, block):
        block = f64(block)
        out = np.zeros(p * p)
        self.L.orc_translate(self.h, rho, kappa_seed, p, scalei, scaleo, int(grad), _p(block), _p(out))
        return out

    def xform_h(self, P, a_src, a_tgt, p, block):
        P, block = f64(P), f64(block)
        out = np.zeros(p * p)
        self.L.orc_xform_h(self.h, _p(P), a_src, a_tgt, p, _p(block), _p(out))
        return out

    def xform_intra(self, P, a_src, a_tgt, pH, pF, H, F):
        p = max(pH, pF)
        P, H, F = f64(P), f64(H), f64(F)
        oH, oF = np.zeros(p * p), np.zeros(p * p)
        self.L.orc_xform_intra(self.h, _p(P), a_src, a_tgt, pH, pF, _p(H), _p(F), _p(oH), _p(oF))
        return oH, oF

    def xform_h_grad(self, P, a_src, a_tgt, p, block):
        P, block = f64(P), f64(block)
        out = np.zeros(3 * p * p)
        self.L.orc_xform_h_grad(self.h, _p(P), a_src, a_tgt, p, _p(block), _p(out))
        return out

    def xform_num(self, P, xyz, qF, qH, tQF, tQH, a_src, a_tgt, p):
        P, xyz, qF, qH = f64(P), f64(xyz), f64(qF), f64(qH)
        oF, oH = np.zeros(p * p), np.zeros(p * p)
        self.L.orc_xform_num(self.h, _p(P), len(qF), _p(xyz), _p(qF), _p(qH), tQF, tQH, a_src, a_tgt, p,
                             _p(oF), _p(oH))
        return oF, oH

    def imat_build(self, p, spe, spb):
        spe, spb = f64(spe).reshape(-1, 3), f64(spb).reshape(-1, 3)
        out = np.zeros(p ** 4)
        self.L.orc_imat_build(self.h, p, len(spe), _p(spe), len(spb), _p(spb), _p(out))
        return out

    @staticmethod
    def sphere_points(N, r):
        out = np.zeros(3 * N)
        lib().orc_sphere_points(N, r, _p(out))
        return out.reshape(N, 3)

    @staticmethod
    def compute_order(rho, tq, scale, p):
        return lib().orc_compute_order(rho, tq, scale, p)

    @staticmethod
    def npts_rule(r, nin=10000):
        return lib().orc_npts_rule(r, nin)


def exact_derivative(on):
    """Switch the oracle between the reference's gradient phase (on=False, the default) and
    the exact derivative of its energy (on=True): full truncation orders everywhere
    (computeOrder, xforms.cpp:402-414, returns p), no skipping of small gradient charges
    (triexpan.cpp:136,146), and the external spheres' geometric derivative with the sign
    of d/d(position of molecule j) (prepareDTA_LowMemory, molecule.cpp:1495-1521, adds
    d/dP with P = c_i - c_m for a sphere m of molecule j, whose motion changes P by minus
    its displacement).  The mutual solve must refresh Hrot for an exact fixed point
    (System.mutual(hrot_refresh=True)).  Process-wide: restore with on=False."""
    L = lib()
    L.orc_debug_full_order(1 if on else 0)
    L.orc_set_grad_skip(0.0 if on else 1e-15)
    L.orc_set_ext_sign(-1.0 if on else 1.0)
    L.orc_set_self_frame(1 if on else 0)


def self_frame(on):
    """Forces of spheres with interaction partners only from their rotated self H (on=True)
    or, as the reference does (expcenter.cpp:140), the unrotated one.  Process-wide."""
    lib().orc_set_self_frame(1 if on else 0)


class MolType:
    """A molecule type: spheres, charges, surface points, constants, self-pol state."""

    def __init__(self, orc, cen, rad, chg_xyz, chg_q, p, eps_in=4.0, rcen=None, npts=None,
                 npts_cap=10000, intra_rcutoff=30.0, imats=None, imat_npts=0):
        self.orc = orc
        L = orc.L
        self._keep = [f64(cen).reshape(-1), f64(rad), f64(chg_xyz).reshape(-1), f64(chg_q),
                      f64(np.zeros(3) if rcen is None else rcen)]
        cen_a, rad_a, cx, cq, rc = self._keep
        npts_a = None if npts is None else np.ascontiguousarray(npts, dtype=np.int32)
        im = None if imats is None else f64(imats)
        self._keep += [npts_a, im]
        d = MolDesc(len(rad_a), _p(cen_a), _p(rad_a), len(cq), _p(cx), _p(cq), _p(rc), p, eps_in,
                    _p(npts_a), npts_cap, intra_rcutoff, _p(im), imat_npts)
        self.h = L.orc_moltype_create(orc.h, ctypes.byref(d))
        if not self.h:
            raise ValueError("charge outside every sphere")
        self.p = p
        self.ns = len(rad_a)

    def __del__(self):
        if getattr(self, "h", None):
            self.orc.L.orc_moltype_destroy(self.h)
            self.h = None

    def npts(self, k):
        tot = ctypes.c_int()
        n = self.orc.L.orc_moltype_npts(self.h, k, ctypes.byref(tot))
        return n, tot.value

    def points(self, k):
        n, _ = self.npts(k)
        out = np.zeros(3 * n)
        self.orc.L.orc_moltype_points(self.h, k, _p(out))
        return out.reshape(n, 3)

    def lists(self, k):
        nfar = ctypes.c_int()
        nn = self.orc.L.orc_moltype_lists(self.h, k, None, None, ctypes.byref(nfar))
        near = np.zeros(max(nn, 1), dtype=np.int32)
        far = np.zeros(max(nfar.value, 1), dtype=np.int32)
        self.orc.L.orc_moltype_lists(self.h, k, _p(near), _p(far), None)
        return near[:nn].tolist(), far[:nfar.value].tolist()

    def consts(self, k):
        p = self.p
        c = np.zeros(10 * p)
        dfix, vfix, e0 = np.zeros(p * p), np.zeros(p * p), np.zeros(p * p)
        self.orc.L.orc_moltype_consts(self.h, k, _p(c), _p(dfix), _p(vfix), _p(e0))
        return c.reshape(10, p), dfix, vfix, e0

    def selfpol(self, schedule=1, tol=1e-24, max_rounds=400):
        return self.orc.L.orc_selfpol(self.h, schedule, tol, max_rounds)

    def round_jacobi(self):
        md = np.zeros(1)
        self.orc.L.orc_selfpol_round_jacobi(self.h, _p(md))
        return md[0]

    def get(self, k):
        p = self.p
        n, _ = self.npts(k)
        F, H = np.zeros(p * p), np.zeros(p * p)
        qF, qH, tq = np.zeros(max(n, 1)), np.zeros(max(n, 1)), np.zeros(3)
        self.orc.L.orc_moltype_get(self.h, k, _p(F), _p(H), _p(qF), _p(qH), _p(tq))
        return dict(F=F, H=H, qF=qF[:n], qH=qH[:n], TQF=tq[0], TQH=tq[1], dev=tq[2])

    def FH(self):
        Fs = np.stack([self.get(k)["F"] for k in range(self.ns)])
        Hs = np.stack([self.get(k)["H"] for k in range(self.ns)])
        return Fs, Hs

    def set_self(self, F, H):
        F, H = f64(F), f64(H)
        self.orc.L.orc_moltype_set_self(self.h, _p(F), _p(H))

    def values(self):
        if self.orc.L.orc_moltype_values(self.h) != 0:
            raise RuntimeError("self values missing: run selfpol or set_self first")

    def intra(self, k):
        p = self.p
        a = [np.zeros(p * p) for _ in range(4)]
        self.orc.L.orc_moltype_get_intra(self.h, k, *[_p(x) for x in a])
        return a


class System:
    def __init__(self, orc, moltypes, pos, quat):
        self.orc = orc
        self.mts = moltypes
        arr = (ctypes.c_void_p * len(moltypes))(*[m.h for m in moltypes])
        self._arr = arr
        pos, quat = f64(pos).reshape(-1), f64(quat).reshape(-1)
        self.h = orc.L.orc_system_create(orc.h, len(moltypes), ctypes.cast(arr, P), _p(pos), _p(quat))
        if not self.h:
            raise RuntimeError("self values missing: run selfpol or set_self first")
        self.n = len(moltypes)
        self.p = moltypes[0].p

    def __del__(self):
        if getattr(self, "h", None):
            self.orc.L.orc_system_destroy(self.h)
            self.h = None

    def set_cutoffs(self, inter, interact):
        self.orc.L.orc_system_set_cutoffs(self.h, inter, interact)

    def set_pose(self, i, pos, quat):
        pos, quat = f64(pos), f64(quat)
        self.orc.L.orc_system_set_pose(self.h, i, _p(pos), _p(quat))

    def lists(self):
        return self.orc.L.orc_system_lists(self.h)

    def mutual(self, schedule=1, tol=1e-24, max_rounds=400, hrot_refresh=False):
        return self.orc.L.orc_mutual(self.h, schedule, tol, max_rounds, int(hrot_refresh))

    def gradient(self, schedule=1, tol=1e-24, max_rounds=400):
        return self.orc.L.orc_gradient(self.h, schedule, tol, max_rounds)

    def gradient_rot(self, tol=1e-24, max_rounds=400):
        """the gradient sweeps for a rotation of each molecule about its centre"""
        return self.orc.L.orc_gradient_rot(self.h, tol, max_rounds)

    def grad_energy(self, i):
        """-sum <H, gLHN> + <gH, LS> of molecule i per plane of the last gradient sweeps"""
        out = np.zeros(3)
        self.orc.L.orc_grad_energy(self.h, i, _p(out))
        return out

    def forces(self):
        f, t = np.zeros(3 * self.n), np.zeros(3 * self.n)
        self.orc.L.orc_forces(self.h, _p(f), _p(t))
        return f.reshape(-1, 3), t.reshape(-1, 3)

    def compute_forces(self, schedule=1, tol=1e-24, max_rounds=400):
        f, t = np.zeros(3 * self.n), np.zeros(3 * self.n)
        rc = self.orc.L.orc_compute_forces(self.h, schedule, tol, max_rounds, _p(f), _p(t))
        return rc, f.reshape(-1, 3), t.reshape(-1, 3)

    def pot_in_space(self, xyz, ref_frame=False):
        xyz = f64(xyz).reshape(-1)
        out = np.zeros(len(xyz) // 3)
        self.orc.L.orc_pot_in_space(self.h, len(out), _p(xyz), _p(out), 1 if ref_frame else 0)
        return out

    def get(self, i, k):
        p = self.p
        a = [np.zeros(p * p) for _ in range(4)]
        self.orc.L.orc_system_get(self.h, i, k, *[_p(x) for x in a])
        return dict(F=a[0], H=a[1], LS=a[2], Hrot=a[3])

    def get_grad(self, i, k):
        p = self.p
        gH, gL = np.zeros(3 * p * p), np.zeros(3 * p * p)
        self.orc.L.orc_system_get_grad(self.h, i, k, _p(gH), _p(gL))
        return gH, gL

    def pot(self, i):
        return self.orc.L.orc_system_pot(self.h, i)

    def total_int_energy(self):
        return self.orc.L.orc_system_total_int_energy(self.h)

    def counts(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        self.orc.L.orc_system_counts(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value


# CBDnam's constants: m_Dtr, m_Dr1, m_Dr2 (BDnam.cpp:143), b_DIST, q_DIST, f_DIST (BDnam.h:23-25),
# FACT2, IKbT, COUL_K (BDnam.h:10,18-19), compute_dt schedule, 500 tries (BDnam.cpp:268)
BD_DEFAULTS = [0.03, 4.5e-5, 4e-5, 100.0, 500.0, 100.0, 5.604586e2, 1.0 / (1.3806503e-23 * 298.15),
               2.30708e-18, 0.0, 500]


class BDTrajectory:
    """One two-molecule Brownian-dynamics trajectory (CBDnam) on the oracle.

    geom: [(cen (ns,3) relative to the molecule centre, rad (ns,))] * 2.  With place=True the
    start is CBDnam::run's random placement from srand48(seed); else pos (2,3), quat (2,4).
    """

    def __init__(self, geom, seed, place=True, pos=None, quat=None, prm=None):
        self.L = lib()
        prm = f64(BD_DEFAULTS if prm is None else prm)
        (c0, r0), (c1, r1) = [(f64(c), f64(r)) for c, r in geom]
        self._keep = (c0, r0, c1, r1, prm)
        pos = None if pos is None else f64(pos)
        quat = None if quat is None else f64(quat)
        self.h = self.L.orc_bd_create(len(r0), _p(c0), _p(r0), len(r1), _p(c1), _p(r1), _p(prm),
                                      int(seed), 1 if place else 0, _p(pos), _p(quat))
        self.system = None

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_bd_destroy(self.h)
            self.h = None

    def attach(self, system):
        """Take forces from an oracle System of the same two molecule types."""
        self.system = system
        self.L.orc_bd_set_system(self.h, system.h)

    def add_contact(self, ncontact, pairs, dist):
        pairs = np.ascontiguousarray(pairs, dtype=np.int32).ravel()
        dist = f64(dist)
        self.L.orc_bd_add_contact(self.h, int(ncontact), len(dist), _p(pairs), _p(dist))

    def step(self, ft=None):
        ft = None if ft is None else f64(ft)
        return self.L.orc_bd_step(self.h, _p(ft))

    def state(self):
        pos, quat, rot, dt = np.zeros(6), np.zeros(8), np.zeros(8), np.zeros(1)
        steps = ctypes.c_longlong(0)
        self.L.orc_bd_get(self.h, _p(pos), _p(quat), _p(rot), _p(dt), ctypes.byref(steps))
        return dict(pos=pos.reshape(2, 3), quat=quat.reshape(2, 4), rot=rot.reshape(2, 4), dt=dt[0],
                    steps=steps.value)


def bd_rng(seed, n):
    """(first n drand48 values after srand48(seed), first n normRand values)."""
    out = np.zeros(2 * n)
    lib().orc_bd_rng(int(seed), n, _p(out))
    return out[:n], out[n:]
