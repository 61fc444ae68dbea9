"""ctypes mirror of the pbsam_amd C ABI (include/pbsam_amd.h).

The classes follow the reference's objects on this path: MolType ~ a molecule type
set up by `pbsam spol` (CMolecule self-pol constructor + polarize_self,
src/selfp_makemat.cpp:37-118), System ~ the BD drivers' vector<CMolecule*>
(CMolecule::computeForces / polarize_mutual, src/moldynamics.cpp:554-587).
All compute runs in libpbsam_amd.so; importing this module fails loudly if the
library is missing.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(os.path.dirname(HERE), "libpbsam_amd.so")

OK, ERR_BADARG, ERR_OVERLAP, ERR_NOCONV, ERR_HIP, ERR_OOM, ERR_UNSUPPORTED, ERR_CHARGE, ERR_STATE = range(9)
HROT_REFRESH = 1
FULL_ORDERS, NO_GRAD_SKIP, EXT_SIGN, SELF_FRAME = 2, 4, 8, 16
EXACT_DERIVATIVE = HROT_REFRESH | FULL_ORDERS | NO_GRAD_SKIP | EXT_SIGN | SELF_FRAME
BD_ADJOINT = 32  # with EXACT_DERIVATIVE: Brownian-dynamics steps move by the adjoint's forces
PMAX = 32

D = ctypes.c_double
I = ctypes.c_int
P = ctypes.c_void_p

SYMBOLS = [
    "pbsam_ctx_create", "pbsam_ctx_destroy", "pbsam_last_error", "pbsam_kappa_from_salt",
    "pbsam_moltype_create", "pbsam_moltype_destroy", "pbsam_moltype_info", "pbsam_moltype_points",
    "pbsam_selfpol", "pbsam_selfpol_rounds", "pbsam_moltype_set_self", "pbsam_moltype_values",
    "pbsam_moltype_get", "pbsam_moltype_get_intra", "pbsam_system_create", "pbsam_system_destroy",
    "pbsam_system_set_pose", "pbsam_system_set_cutoffs", "pbsam_solve_mutual", "pbsam_mutual_rounds",
    "pbsam_system_get", "pbsam_energy", "pbsam_system_counts", "pbsam_xform_h_batch",
    "pbsam_xform_intra_batch", "pbsam_xform_num_batch", "pbsam_rotate_quat_batch",
    "pbsam_profile", "pbsam_profile_read", "pbsam_sync", "pbsam_gradient", "pbsam_forces",
    "pbsam_compute_forces", "pbsam_system_get_grad", "pbsam_xform_h_grad_batch",
    "pbsam_batch_create", "pbsam_batch_status", "pbsam_imat_build",
    "pbsam_bd_defaults", "pbsam_bd_start", "pbsam_bd_contacts", "pbsam_bd_run", "pbsam_bd_get",
    "pbsam_read_pqr", "pbsam_write_imat", "pbsam_read_imat", "pbsam_write_expansion", "pbsam_read_expansion",
    "pbsam_read_config", "pbsam_write_config", "pbsam_read_imats", "pbsam_read_expansions",
    "pbsam_potential", "pbsam_total_int_energy", "pbsam_system_set_flags", "pbsam_autotune",
    "pbsam_torques_exact", "pbsam_adjoint_forces", "pbsam_solve_exact",
    "pbsam_xform_h_batch_noalloc", "pbsam_xform_h_batch_noalloc_bytes",
    "pbsam_ctx_set_store_limit", "pbsam_ctx_stored", "pbsam_ctx_set_graphs", "pbsam_system_graph_stats",
]
POT_REFERENCE_FRAME = 1
BD_ESCAPED, BD_DOCKED, BD_RUNNING, BD_STUCK, BD_FAILED = range(5)

KERNELS = {"xform": 0, "p2l": 1, "solve": 2, "surface": 3, "xgrad": 4, "surfgrad": 5}


class PbsamError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pbsam error {code}: {msg}")
        self.code = code


class MolDesc(ctypes.Structure):
    _fields_ = [("n_sph", I), ("cen", P), ("rad", P), ("n_chg", I), ("chg_xyz", P),
                ("chg_q", P), ("rcen", P), ("p", I), ("eps_in", D), ("npts", P),
                ("npts_cap", I), ("intra_rcutoff", D), ("imats", P), ("imat_npts", I)]


class BDParams(ctypes.Structure):
    _fields_ = [("dtr", D), ("dr", D * 2), ("b_dist", D), ("q_dist", D), ("f_dist", D), ("fact_tr", D),
                ("ikbt", D), ("coul_k", D), ("dt_fixed", D), ("max_tries", I)]


_LIB = None


def _p(a):
    return None if a is None else a.ctypes.data_as(P)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def load():
    """Load libpbsam_amd.so (built in-tree by `make -C pb-sam_amd`)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIBPATH):
        raise RuntimeError(f"{LIBPATH} is missing: build it with `make -C pb-sam_amd`")
    L = ctypes.CDLL(LIBPATH)
    sig = {
        "pbsam_ctx_create": ([D, D, I, P], I),
        "pbsam_ctx_destroy": ([P], None),
        "pbsam_last_error": ([P], ctypes.c_char_p),
        "pbsam_ctx_set_store_limit": ([P, ctypes.c_longlong], I),
        "pbsam_ctx_stored": ([P, P, P], I),
        "pbsam_ctx_set_graphs": ([P, I], I),
        "pbsam_system_graph_stats": ([P, P, P], I),
        "pbsam_kappa_from_salt": ([D, D, D], D),
        "pbsam_moltype_create": ([P, ctypes.POINTER(MolDesc), P], I),
        "pbsam_moltype_destroy": ([P], None),
        "pbsam_moltype_info": ([P, P, P, P], I),
        "pbsam_moltype_points": ([P, I, P, P, P], I),
        "pbsam_selfpol": ([P, D, I, P], I),
        "pbsam_selfpol_rounds": ([P, I, P], I),
        "pbsam_moltype_set_self": ([P, P, P], I),
        "pbsam_moltype_values": ([P], I),
        "pbsam_moltype_get": ([P, I, P, P, P, P, P], I),
        "pbsam_moltype_get_intra": ([P, I, P, P, P, P], I),
        "pbsam_system_create": ([P, I, P, P, P, P], I),
        "pbsam_batch_create": ([P, I, P, P, P, P, P], I),
        "pbsam_batch_status": ([P, P, P, P], I),
        "pbsam_imat_build": ([P, I, I, P, I, P, P], I),
        "pbsam_bd_defaults": ([ctypes.POINTER(BDParams)], None),
        "pbsam_bd_start": ([P, ctypes.POINTER(BDParams), P, I], I),
        "pbsam_bd_contacts": ([P, I, P, P, P, P], I),
        "pbsam_bd_run": ([P, I, P, P], I),
        "pbsam_bd_get": ([P, I, P, P, P, P], I),
        "pbsam_read_pqr": ([ctypes.c_char_p, P, P, P, P, P, P, ctypes.c_char_p, I], I),
        "pbsam_write_imat": ([ctypes.c_char_p, I, P], I),
        "pbsam_read_imat": ([ctypes.c_char_p, I, P, P], I),
        "pbsam_write_expansion": ([ctypes.c_char_p, I, D, D, P], I),
        "pbsam_read_expansion": ([ctypes.c_char_p, I, P, P, P, P], I),
        "pbsam_read_config": ([ctypes.c_char_p, P, P, P, P], I),
        "pbsam_write_config": ([ctypes.c_char_p, I, P, P, P], I),
        "pbsam_read_imats": ([ctypes.c_char_p, I, I, P], I),
        "pbsam_read_expansions": ([ctypes.c_char_p, ctypes.c_char_p, I, I, P, P, P], I),
        "pbsam_potential": ([P, I, P, P, ctypes.c_uint], I),
        "pbsam_total_int_energy": ([P, ctypes.c_uint, P], I),
        "pbsam_system_destroy": ([P], None),
        "pbsam_system_set_pose": ([P, I, P, P], I),
        "pbsam_system_set_cutoffs": ([P, D, D], I),
        "pbsam_system_set_flags": ([P, ctypes.c_uint], I),
        "pbsam_autotune": ([P, I, P, P], I),
        "pbsam_torques_exact": ([P, D, I, P], I),
        "pbsam_adjoint_forces": ([P, D, I, P, P, P], I),
        "pbsam_solve_exact": ([P, D, I, P], I),
        "pbsam_solve_mutual": ([P, D, I, ctypes.c_uint, P], I),
        "pbsam_mutual_rounds": ([P, I, ctypes.c_uint, P], I),
        "pbsam_system_get": ([P, I, I, P, P, P, P], I),
        "pbsam_energy": ([P, P], I),
        "pbsam_system_counts": ([P, P, P], I),
        "pbsam_xform_h_batch": ([P, I, I, P, P, P, P, P], I),
        "pbsam_xform_intra_batch": ([P, I, I, P, P, P, P, P, P, P], I),
        "pbsam_xform_num_batch": ([P, I, I, P, P, I, P, P, P, D, D, P, P], I),
        "pbsam_rotate_quat_batch": ([P, I, I, P, I, P, P], I),
        "pbsam_gradient": ([P, D, I, P], I),
        "pbsam_forces": ([P, P, P], I),
        "pbsam_compute_forces": ([P, D, I, P, P, P], I),
        "pbsam_system_get_grad": ([P, I, I, P, P], I),
        "pbsam_xform_h_grad_batch": ([P, I, I, P, P, P, P, P], I),
        "pbsam_profile": ([P, I], I),
        "pbsam_profile_read": ([P, I, P, P, P, P], I),
        "pbsam_sync": ([P], I),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _LIB = L
    return L


# ---- the reference's file formats (host only) ----
def _io_check(rc, what):
    if rc != OK:
        raise PbsamError(rc, what)


def read_pqr(path):
    """readpqr: (sphere centres (n,3), radii (n,), charge positions (m,3), charges (m,))"""
    L = load()
    ns, nq = ctypes.c_int(0), ctypes.c_int(0)
    err = ctypes.create_string_buffer(256)
    _io_check(L.pbsam_read_pqr(os.fsencode(path), ctypes.byref(ns), None, None, ctypes.byref(nq), None, None,
                               err, 256), err.value.decode() or str(path))
    cen, rad = np.zeros(3 * ns.value), np.zeros(ns.value)
    xyz, q = np.zeros(3 * nq.value), np.zeros(nq.value)
    _io_check(L.pbsam_read_pqr(os.fsencode(path), ctypes.byref(ns), _p(cen), _p(rad), ctypes.byref(nq), _p(xyz),
                               _p(q), err, 256), err.value.decode() or str(path))
    return cen.reshape(-1, 3), rad, xyz.reshape(-1, 3), q


def write_imat(path, imat, p):
    m = f64(imat).reshape(-1)
    if m.size != p ** 4:
        raise ValueError("an IMat of order p holds p^4 values")
    _io_check(load().pbsam_write_imat(os.fsencode(path), p, _p(m)), str(path))


def read_imat(path, p):
    """(p^2 x p^2 column-major matrix as p^4 values, the file's pole order)"""
    m, pf = np.zeros(p ** 4), ctypes.c_int(0)
    _io_check(load().pbsam_read_imat(os.fsencode(path), p, _p(m), ctypes.byref(pf)), str(path))
    return m, pf.value


def write_expansion(path, coef, p, scale, rcut):
    c = f64(coef).reshape(-1)
    if c.size < p * p:
        raise ValueError("an expansion of order p holds p^2 values")
    _io_check(load().pbsam_write_expansion(os.fsencode(path), p, scale, rcut, _p(c)), str(path))


def read_expansion(path, p):
    """(coefficients (p^2,), scale, rcut, the file's pole order)"""
    c, sc, rc, pf = np.zeros(p * p), ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
    _io_check(load().pbsam_read_expansion(os.fsencode(path), p, ctypes.byref(sc), ctypes.byref(rc), _p(c),
                                          ctypes.byref(pf)), str(path))
    return c, sc.value, rc.value, pf.value


def read_config(path):
    """(positions (n,3), quaternions (n,4) w x y z, molecule types (n,))"""
    L = load()
    n = ctypes.c_int(0)
    _io_check(L.pbsam_read_config(os.fsencode(path), ctypes.byref(n), None, None, None), str(path))
    pos, quat, t = np.zeros(3 * n.value), np.zeros(4 * n.value), np.zeros(n.value, dtype=np.int32)
    _io_check(L.pbsam_read_config(os.fsencode(path), ctypes.byref(n), _p(pos), _p(quat), _p(t)), str(path))
    return pos.reshape(-1, 3), quat.reshape(-1, 4), t


def write_config(path, pos, quat, moltype):
    pos, quat = f64(pos).reshape(-1), f64(quat).reshape(-1)
    t = np.ascontiguousarray(moltype, dtype=np.int32)
    _io_check(load().pbsam_write_config(os.fsencode(path), len(t), _p(pos), _p(quat), _p(t)), str(path))


def read_imats(directory, ns, p):
    m = np.zeros(ns * p ** 4)
    _io_check(load().pbsam_read_imats(os.fsencode(directory), ns, p, _p(m)), str(directory))
    return m.reshape(ns, p ** 4)


def read_expansions(directory, run, ns, p):
    """(F (ns,p^2), H (ns,p^2), intra cutoff) from <dir>/<run>.<k>.{F,H}.exp"""
    F, H, rc = np.zeros(ns * p * p), np.zeros(ns * p * p), ctypes.c_double()
    _io_check(load().pbsam_read_expansions(os.fsencode(directory), run.encode(), ns, p, _p(F), _p(H),
                                           ctypes.byref(rc)), str(directory))
    return F.reshape(ns, -1), H.reshape(ns, -1), rc.value


def kappa_from_salt(salt_molar, eps_out=78.0, temp=298.15):
    return load().pbsam_kappa_from_salt(salt_molar, eps_out, temp)


class Context:
    """One HIP device + stream with the physical constants (CSystem::initConstants)."""

    def __init__(self, kappa, eps_out=78.0, device=0):
        self.L = load()
        h = ctypes.c_void_p()
        rc = self.L.pbsam_ctx_create(kappa, eps_out, device, ctypes.byref(h))
        if rc != OK:
            raise PbsamError(rc, "context creation failed (no HIP device?)")
        self.h = h.value
        self.kappa = kappa

    def check(self, rc):
        if rc not in (OK,):
            raise PbsamError(rc, self.L.pbsam_last_error(self.h).decode())
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.L.pbsam_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def sync(self):
        self.check(self.L.pbsam_sync(self.h))

    def set_store_limit(self, nbytes):
        """Cap on the optional operator stores (far-pair translation matrices, numerical
        maps) of molecule types set up after the call; 0 makes them all per use."""
        self.check(self.L.pbsam_ctx_set_store_limit(self.h, int(nbytes)))

    def set_graphs(self, on):
        """Replay recurring gradient sweep windows as captured HIP graphs (default on)."""
        self.check(self.L.pbsam_ctx_set_graphs(self.h, int(bool(on))))

    def stored(self):
        """(bytes of optional operator stores, their cap)"""
        b, lim = ctypes.c_longlong(), ctypes.c_longlong()
        self.check(self.L.pbsam_ctx_stored(self.h, ctypes.byref(b), ctypes.byref(lim)))
        return b.value, lim.value

    def profile(self, on):
        """Start (reset) or stop the per-kernel HIP-event timing and work counters."""
        self.check(self.L.pbsam_profile(self.h, int(bool(on))))

    def profile_read(self, kernel):
        """{'launches', 'ms', 'flops', 'bytes'} of one kernel since profiling started."""
        n, ms, fl, by = ctypes.c_longlong(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self.check(self.L.pbsam_profile_read(self.h, KERNELS[kernel], ctypes.byref(n), ctypes.byref(ms),
                                             ctypes.byref(fl), ctypes.byref(by)))
        return dict(launches=n.value, ms=ms.value, flops=fl.value, bytes=by.value)

    # kernel-level batches
    def xform_h(self, P3, a, order, blocks, p):
        P3, a, blocks = f64(P3), f64(a), f64(blocks)
        order = np.ascontiguousarray(order, dtype=np.int32)
        n = len(order)
        out = np.zeros((n, p * p))
        self.check(self.L.pbsam_xform_h_batch(self.h, n, p, _p(P3), _p(a), _p(order), _p(blocks), _p(out)))
        return out

    def imat_build(self, p, spe, spb):
        """build_IMat from exposed and buried points (n x 3 each): p^2 x p^2 column-major"""
        spe, spb = f64(spe).reshape(-1), f64(spb).reshape(-1)
        out = np.zeros(p ** 4)
        self.check(self.L.pbsam_imat_build(self.h, p, len(spe) // 3, _p(spe), len(spb) // 3, _p(spb), _p(out)))
        return out

    def xform_intra(self, P3, a, orders, H, F, p):
        P3, a, H, F = f64(P3), f64(a), f64(H), f64(F)
        orders = np.ascontiguousarray(orders, dtype=np.int32)
        n = len(orders) // 2
        oH, oF = np.zeros((n, p * p)), np.zeros((n, p * p))
        self.check(self.L.pbsam_xform_intra_batch(self.h, n, p, _p(P3), _p(a), _p(orders), _p(H), _p(F),
                                                  _p(oH), _p(oF)))
        return oH, oF

    def xform_num(self, P3, a, xyz, qF, qH, tQF, tQH, p):
        P3, a, xyz, qF, qH = f64(P3), f64(a), f64(xyz), f64(qF), f64(qH)
        n = len(P3) // 3
        oF, oH = np.zeros((n, p * p)), np.zeros((n, p * p))
        self.check(self.L.pbsam_xform_num_batch(self.h, n, p, _p(P3), _p(a), len(qF), _p(xyz), _p(qF), _p(qH),
                                                tQF, tQH, _p(oF), _p(oH)))
        return oF, oH

    def xform_h_grad(self, P3, a, order, blocks, p):
        P3, a, blocks = f64(P3), f64(a), f64(blocks)
        order = np.ascontiguousarray(order, dtype=np.int32)
        n = len(order)
        out = np.zeros((n, 3 * p * p))
        self.check(self.L.pbsam_xform_h_grad_batch(self.h, n, p, _p(P3), _p(a), _p(order), _p(blocks), _p(out)))
        return out

    def rotate_quat(self, quats, fwd, blocks, p):
        quats, blocks = f64(quats), f64(blocks)
        n = len(quats) // 4
        out = np.zeros((n, p * p))
        self.check(self.L.pbsam_rotate_quat_batch(self.h, n, p, _p(quats), int(fwd), _p(blocks), _p(out)))
        return out


class MolType:
    """Molecule type on the device: spheres, charges, surface points, self-pol state."""

    def __init__(self, ctx, cen, rad, chg_xyz, chg_q, p, eps_in=4.0, rcen=None, npts=None, npts_cap=10000,
                 intra_rcutoff=30.0, imats=None, imat_npts=0):
        self.ctx = ctx
        L = ctx.L
        keep = [f64(cen).reshape(-1), f64(rad), f64(chg_xyz).reshape(-1), f64(chg_q),
                f64(np.zeros(3) if rcen is None else rcen)]
        cen_a, rad_a, cx, cq, rc = keep
        npts_a = None if npts is None else np.ascontiguousarray(npts, dtype=np.int32)
        im = None if imats is None else f64(imats)
        self._keep = keep + [npts_a, im]
        d = MolDesc(len(rad_a), _p(cen_a), _p(rad_a), len(cq), _p(cx), _p(cq), _p(rc), p, eps_in, _p(npts_a),
                    npts_cap, intra_rcutoff, _p(im), imat_npts)
        h = ctypes.c_void_p()
        ctx.check(L.pbsam_moltype_create(ctx.h, ctypes.byref(d), ctypes.byref(h)))
        self.h = h.value
        self.p = p
        self.ns = len(rad_a)

    def close(self):
        if getattr(self, "h", None):
            self.ctx.L.pbsam_moltype_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def points(self, k):
        ne, nt = ctypes.c_int(), ctypes.c_int()
        self.ctx.check(self.ctx.L.pbsam_moltype_points(self.h, k, None, ctypes.byref(ne), ctypes.byref(nt)))
        out = np.zeros(3 * ne.value)
        self.ctx.check(self.ctx.L.pbsam_moltype_points(self.h, k, _p(out), None, None))
        return out.reshape(-1, 3), nt.value

    def selfpol(self, tol=0.0, max_rounds=0):
        """polarize_self in the reference's order; tol/max_rounds <= 0: its stock values.
        Returns (rounds, status) with status OK or ERR_NOCONV."""
        r = ctypes.c_int()
        rc = self.ctx.L.pbsam_selfpol(self.h, tol, max_rounds, ctypes.byref(r))
        if rc not in (OK, ERR_NOCONV):
            self.ctx.check(rc)
        return r.value, rc

    def rounds(self, n):
        md = ctypes.c_double()
        self.ctx.check(self.ctx.L.pbsam_selfpol_rounds(self.h, n, ctypes.byref(md)))
        return md.value

    def set_self(self, F, H):
        F, H = f64(F), f64(H)
        self.ctx.check(self.ctx.L.pbsam_moltype_set_self(self.h, _p(F), _p(H)))

    def values(self):
        self.ctx.check(self.ctx.L.pbsam_moltype_values(self.h))

    def get(self, k):
        p = self.p
        pts, _ = self.points(k)
        n = len(pts)
        F, H = np.zeros(p * p), np.zeros(p * p)
        qF, qH, tq = np.zeros(max(n, 1)), np.zeros(max(n, 1)), np.zeros(3)
        self.ctx.check(self.ctx.L.pbsam_moltype_get(self.h, k, _p(F), _p(H), _p(qF), _p(qH), _p(tq)))
        return dict(F=F, H=H, qF=qF[:n], qH=qH[:n], TQF=tq[0], TQH=tq[1], dev=tq[2])

    def FH(self):
        g = [self.get(k) for k in range(self.ns)]
        return np.stack([x["F"] for x in g]), np.stack([x["H"] for x in g])

    def intra(self, k):
        a = [np.zeros(self.p * self.p) for _ in range(4)]
        self.ctx.check(self.ctx.L.pbsam_moltype_get_intra(self.h, k, *[_p(x) for x in a]))
        return a


class System:
    """Rigid molecules with poses; mutual polarization and energies on the device."""

    def __init__(self, ctx, moltypes, pos, quat, groups=None):
        """groups: molecules per independent system (pbsam_batch_create), or None for one
        system of all molecules (pbsam_system_create)."""
        self.ctx = ctx
        self.mts = list(moltypes)
        arr = (ctypes.c_void_p * len(moltypes))(*[m.h for m in moltypes])
        self._arr = arr
        pos, quat = f64(pos).reshape(-1), f64(quat).reshape(-1)
        h = ctypes.c_void_p()
        if groups is None:
            ctx.check(ctx.L.pbsam_system_create(ctx.h, len(moltypes), ctypes.cast(arr, P), _p(pos), _p(quat),
                                                ctypes.byref(h)))
            self.groups = [len(moltypes)]
        else:
            g = np.ascontiguousarray(groups, dtype=np.int32)
            if int(g.sum()) != len(moltypes):
                raise ValueError("groups must add up to the number of molecules")
            ctx.check(ctx.L.pbsam_batch_create(ctx.h, len(g), g.ctypes.data_as(P), ctypes.cast(arr, P), _p(pos),
                                               _p(quat), ctypes.byref(h)))
            self.groups = [int(x) for x in g]
        self.h = h.value
        self.n = len(moltypes)
        self.p = moltypes[0].p

    def close(self):
        if getattr(self, "h", None):
            self.ctx.L.pbsam_system_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_cutoffs(self, inter, interact):
        self.ctx.check(self.ctx.L.pbsam_system_set_cutoffs(self.h, inter, interact))

    def set_flags(self, flags):
        """pbsam_system_set_flags: EXACT_DERIVATIVE (or its parts) for forces that are the
        exact derivative of the energy; 0 for the reference's behaviour."""
        self.ctx.check(self.ctx.L.pbsam_system_set_flags(self.h, int(flags)))

    def autotune(self, reps=1):
        """pbsam_autotune: (best ms per evaluation, dict of the chosen settings)"""
        ms, st = ctypes.c_double(), np.zeros(7, dtype=np.int32)
        self.ctx.check(self.ctx.L.pbsam_autotune(self.h, reps, ctypes.byref(ms), st.ctypes.data_as(P)))
        keys = ("grad_window", "xsplit_max", "xbin_min", "mgemv_wg", "pgemm_min", "p2l_wg", "surf_wg")
        return ms.value, dict(zip(keys, (int(v) for v in st)))

    def set_pose(self, i, pos, quat):
        pos, quat = f64(pos), f64(quat)
        self.ctx.check(self.ctx.L.pbsam_system_set_pose(self.h, i, _p(pos), _p(quat)))

    def solve(self, tol=0.0, max_rounds=0, refresh=False):
        """polarize_mutual in the reference's order; tol/max_rounds <= 0: its stock values.
        Returns (rounds, status) with status OK or ERR_NOCONV."""
        r = ctypes.c_int()
        rc = self.ctx.L.pbsam_solve_mutual(self.h, tol, max_rounds, HROT_REFRESH if refresh else 0,
                                           ctypes.byref(r))
        if rc not in (OK, ERR_NOCONV) and not (len(self.groups) > 1 and rc == ERR_OVERLAP):
            self.ctx.check(rc)
        return r.value, rc

    def rounds(self, n, refresh=False):
        md = ctypes.c_double()
        rc = self.ctx.L.pbsam_mutual_rounds(self.h, n, HROT_REFRESH if refresh else 0, ctypes.byref(md))
        self.ctx.check(rc)
        return md.value

    def get(self, i, k):
        p = self.mts[i].p
        a = [np.zeros(p * p) for _ in range(4)]
        self.ctx.check(self.ctx.L.pbsam_system_get(self.h, i, k, *[_p(x) for x in a]))
        return dict(F=a[0], H=a[1], LS=a[2], Hrot=a[3])

    def gradient(self, tol=0.0, max_rounds=0):
        """Gradient phase in the reference's order; returns the number of passes."""
        r = ctypes.c_int()
        self.ctx.check(self.ctx.L.pbsam_gradient(self.h, tol, max_rounds, ctypes.byref(r)))
        return r.value

    def torques_exact(self, tol=1e-16, max_rounds=2000):
        """pbsam_torques_exact: -dE/domega of each molecule (lab frame), exact-derivative mode"""
        t = np.zeros(3 * self.n)
        self.ctx.check(self.ctx.L.pbsam_torques_exact(self.h, tol, max_rounds, _p(t)))
        return t.reshape(-1, 3)

    def solve_exact(self, tol=1e-24, max_rounds=2000):
        """pbsam_solve_exact: the exact-mode fixed point by GMRES; returns (applications, status)
        with status OK or ERR_NOCONV"""
        n = ctypes.c_int()
        rc = self.ctx.L.pbsam_solve_exact(self.h, tol, max_rounds, ctypes.byref(n))
        if rc not in (OK, ERR_NOCONV) and not (len(self.groups) > 1 and rc == ERR_OVERLAP):
            self.ctx.check(rc)
        return n.value, rc

    def adjoint_forces(self, tol=1e-20, max_rounds=5000):
        """pbsam_adjoint_forces: (force, torque, operator applications) of each molecule from the
        transposed solves (one for two molecules, one per molecule for more), -dE/dr and
        -dE/domega in the lab frame (exact-derivative mode, after a solve)"""
        f, t, n = np.zeros(3 * self.n), np.zeros(3 * self.n), ctypes.c_int()
        self.ctx.check(self.ctx.L.pbsam_adjoint_forces(self.h, tol, max_rounds, _p(f), _p(t), ctypes.byref(n)))
        return f.reshape(-1, 3), t.reshape(-1, 3), n.value

    def forces(self):
        f, t = np.zeros(3 * self.n), np.zeros(3 * self.n)
        self.ctx.check(self.ctx.L.pbsam_forces(self.h, _p(f), _p(t)))
        return f.reshape(-1, 3), t.reshape(-1, 3)

    def compute_forces(self, tol=0.0, max_rounds=0):
        """computeForces: returns (rounds, status, force[n,3], torque[n,3])."""
        f, t, r = np.zeros(3 * self.n), np.zeros(3 * self.n), ctypes.c_int()
        rc = self.ctx.L.pbsam_compute_forces(self.h, tol, max_rounds, _p(f), _p(t), ctypes.byref(r))
        if rc not in (OK, ERR_NOCONV) and not (len(self.groups) > 1 and rc == ERR_OVERLAP):
            self.ctx.check(rc)
        return r.value, rc, f.reshape(-1, 3), t.reshape(-1, 3)

    def get_grad(self, i, k):
        p = self.mts[i].p
        gH, gL = np.zeros(3 * p * p), np.zeros(3 * p * p)
        self.ctx.check(self.ctx.L.pbsam_system_get_grad(self.h, i, k, _p(gH), _p(gL)))
        return gH, gL

    def energy(self):
        e = np.zeros(self.n)
        self.ctx.check(self.ctx.L.pbsam_energy(self.h, _p(e)))
        return e

    def potential(self, xyz, reference_frame=False):
        """computePotInSpace at points (n,3) from the current state"""
        xyz = f64(xyz).reshape(-1)
        out = np.zeros(len(xyz) // 3)
        self.ctx.check(self.ctx.L.pbsam_potential(self.h, len(out), _p(xyz), _p(out),
                                                  POT_REFERENCE_FRAME if reference_frame else 0))
        return out

    def total_int_energy(self, reference_frame=False):
        """computeTotalIntEnergy per system of the handle"""
        e = np.zeros(len(self.groups))
        self.ctx.check(self.ctx.L.pbsam_total_int_energy(self.h, POT_REFERENCE_FRAME if reference_frame else 0,
                                                         _p(e)))
        return e

    def status(self):
        """per group: (status, mutual rounds, gradient passes) of the last solve / gradient"""
        n = len(self.groups)
        st, r, g = (np.zeros(n, dtype=np.int32) for _ in range(3))
        self.ctx.check(self.ctx.L.pbsam_batch_status(self.h, st.ctypes.data_as(P), r.ctypes.data_as(P),
                                                     g.ctypes.data_as(P)))
        return st, r, g

    def graph_stats(self):
        """(graph launches, graph captures) of this system's gradient phases so far"""
        a, b = ctypes.c_longlong(), ctypes.c_longlong()
        self.ctx.check(self.ctx.L.pbsam_system_graph_stats(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def counts(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        self.ctx.check(self.ctx.L.pbsam_system_counts(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    # ---- Brownian dynamics: one two-molecule trajectory per group (CBDnam) ----
    @staticmethod
    def bd_params(**kw):
        """CBDnam's parameters (pbsam_bd_defaults) with keyword overrides (dr as a pair)."""
        prm = BDParams()
        load().pbsam_bd_defaults(ctypes.byref(prm))
        for k, v in kw.items():
            if k == "dr":
                prm.dr[0], prm.dr[1] = v
            else:
                setattr(prm, k, v)
        return prm

    def bd_start(self, seeds, place=True, params=None):
        seeds = np.ascontiguousarray(seeds, dtype=np.int64)
        if len(seeds) != len(self.groups):
            raise ValueError("one seed per trajectory")
        prm = params if params is not None else self.bd_params()
        self._bd_prm = prm
        self.ctx.check(self.ctx.L.pbsam_bd_start(self.h, ctypes.byref(prm), seeds.ctypes.data_as(P),
                                                 1 if place else 0))

    def bd_contacts(self, defs):
        """defs: [(n_contact, [(k0, k1), ...], [dist, ...]), ...]"""
        nc = np.array([d[0] for d in defs], dtype=np.int32)
        npair = np.array([len(d[1]) for d in defs], dtype=np.int32)
        pairs = np.ascontiguousarray(np.concatenate([np.asarray(d[1], np.int32).ravel() for d in defs])
                                     if defs else np.zeros(0, np.int32), dtype=np.int32)
        dist = f64(np.concatenate([np.asarray(d[2], float) for d in defs]) if defs else np.zeros(0))
        self.ctx.check(self.ctx.L.pbsam_bd_contacts(self.h, len(defs), nc.ctypes.data_as(P),
                                                    npair.ctypes.data_as(P), pairs.ctypes.data_as(P), _p(dist)))

    def bd_run(self, n_steps):
        """Advance every running trajectory by up to n_steps; returns (rc, status[g], steps[g])."""
        n = len(self.groups)
        st = np.zeros(n, dtype=np.int32)
        steps = np.zeros(n, dtype=np.int64)
        rc = self.ctx.L.pbsam_bd_run(self.h, int(n_steps), st.ctypes.data_as(P), steps.ctypes.data_as(P))
        if rc not in (OK, ERR_NOCONV):
            self.ctx.check(rc)
        return rc, st, steps

    def bd_get(self, g):
        pos, quat, rot, dt = np.zeros(6), np.zeros(8), np.zeros(8), np.zeros(1)
        self.ctx.check(self.ctx.L.pbsam_bd_get(self.h, int(g), _p(pos), _p(quat), _p(rot), _p(dt)))
        return dict(pos=pos.reshape(2, 3), quat=quat.reshape(2, 4), rot=rot.reshape(2, 4), dt=dt[0])
