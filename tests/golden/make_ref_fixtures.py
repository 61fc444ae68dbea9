"""Writes reference known-answer vectors to tests/golden/ref_kat.npz.

Runs oracle/_ref/pbsam_ref_kat (built by `make -C oracle ref` from the reference's own
sources, present only in the build container) on fixed, seeded inputs and stores the
inputs and the reference's outputs.  The reference is compiled with N_POLES = 15, so
every vector here has p <= 15.

    python tests/golden/make_ref_fixtures.py
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
KAT = os.path.join(ROOT, "oracle", "_ref", "pbsam_ref_kat")
KAPPA = 0.0737403009
SDIEL = 78.0


def fmt(v):
    return " ".join("%.17g" % x for x in np.ravel(v))


def run(cmds):
    txt = "init %.17g %.17g\n" % (KAPPA, SDIEL) + "\n".join(cmds) + "\n"
    out = subprocess.run([KAT], input=txt.encode(), capture_output=True, check=True).stdout.decode()
    lines = out.strip().split("\n")
    assert lines[0].startswith("ok"), lines[0]
    return [np.array([float(x) for x in ln.split()]) for ln in lines[1:]]


def rand_block(rng, p, scale=1.0):
    return rng.normal(size=p * p) * scale


def small_molecule(seed, ns=6, overlap=True):
    rng = np.random.default_rng(seed)
    cen, rad = [], []
    while len(rad) < ns:
        r = rng.uniform(2.0, 4.0)
        c = rng.uniform(-7, 7, 3)
        ok = True
        for c2, r2 in zip(cen, rad):
            d = np.linalg.norm(c - c2)
            lim = 0.6 * (r + r2) if overlap else (r + r2 + 0.2)
            if d < lim:
                ok = False
        if ok:
            cen.append(c)
            rad.append(r)
    cen, rad = np.array(cen), np.array(rad)
    xyz, q = [], []
    for c, r in zip(cen, rad):
        for _ in range(4):
            v = rng.uniform(-1, 1, 3)
            v = v / max(1.0, np.linalg.norm(v))
            xyz.append(c + 0.7 * r * v)
            q.append(rng.choice([-1.0, 1.0]))
    return cen, rad, np.array(xyz), np.array(q)


def main():
    if not os.path.exists(KAT):
        sys.exit("build the reference harness first: make -C oracle ref")
    rng = np.random.default_rng(20260101)
    fx = {"kappa": KAPPA, "sdiel": SDIEL}
    cmds, keys = [], []

    # spherical harmonics and Bessel functions
    sh_in = []
    for i in range(6):
        P = [5, 15, 29, 29, 10, 20][i]
        th, ph = rng.uniform(0, np.pi), rng.uniform(-np.pi, np.pi)
        sh_in.append((P, th, ph))
        cmds.append("sh %d %.17g %.17g" % (P, th, ph))
        keys.append("sh_%d" % i)
    fx["sh_in"] = np.array(sh_in)
    bes_in = [(16, 0.25), (16, 1.7), (30, 4.4), (30, 0.0)]
    for i, (L, v) in enumerate(bes_in):
        cmds.append("besk %d %.17g" % (L, v))
        keys.append("besk_%d" % i)
        cmds.append("besi %d %.17g" % (L, v))
        keys.append("besi_%d" % i)
    fx["bes_in"] = np.array(bes_in)

    # single-point-set expansions
    for i, (kind, p, bk) in enumerate([("lexp", 10, 1), ("lexp", 15, 0), ("mexp", 10, 0), ("mexp", 15, 1)]):
        n = 7
        q = rng.normal(size=n)
        if kind == "lexp":
            x = rng.normal(size=(n, 3)) * 3 + np.array([9.0, -2.0, 4.0])
        else:
            x = rng.normal(size=(n, 3)) * 1.2
        sc = 3.3
        fx["%s_%d_in" % (kind, i)] = np.concatenate([[p, bk, sc], q, x.ravel()])
        cmds.append("%s %d %d %.17g %d %s" % (kind, p, bk, sc, n, " ".join(
            "%.17g %.17g %.17g %.17g" % (q[j], *x[j]) for j in range(n))))
        keys.append("%s_%d" % (kind, i))
    # gradient local expansion
    n = 6
    g = rng.normal(size=(n, 3))
    x = rng.normal(size=(n, 3)) * 2 + np.array([0.0, 8.0, -3.0])
    fx["gradlocal_in"] = np.concatenate([[10, 1, 2.7], g.ravel(), x.ravel()])
    cmds.append("gradlocal 10 1 2.7 %d %s" % (n, " ".join("%.17g %.17g %.17g %.17g %.17g %.17g" % (*g[j], *x[j])
                                                           for j in range(n))))
    keys.append("gradlocal")

    # rotations (regular and the two singular directions)
    rot_in = []
    for i, (th, ph, xi, p, p2, fwd) in enumerate([
            (0.7, 1.3, np.pi, 10, 10, 1), (0.7, 1.3, np.pi, 10, 7, 0), (2.2, -2.6, 0.4, 15, 15, 1),
            (2.2, -2.6, 0.4, 15, 15, 0), (0.0, 0.0, np.pi, 10, 10, 1), (np.pi, 0.0, np.pi, 10, 10, 1),
            (1e-13, 0.5, np.pi, 8, 8, 0)]):
        b = rand_block(rng, p2)
        rot_in.append(np.concatenate([[th, ph, xi, p, p2, fwd], b]))
        cmds.append("rotate %.17g %.17g %.17g %d %d %d %s" % (th, ph, xi, p, p2, fwd, fmt(b)))
        keys.append("rotate_%d" % i)
    fx["rotate_n"] = len(rot_in)
    for i, r in enumerate(rot_in):
        fx["rotate_%d_in" % i] = r
    for i, (th, ph, xi, p, which, fwd) in enumerate([(0.9, 0.4, np.pi, 10, 0, 1), (0.9, 0.4, np.pi, 10, 1, 0),
                                                       (2.5, -1.0, np.pi, 12, 0, 0), (2.5, -1.0, np.pi, 12, 1, 1),
                                                       (0.0, 0.0, np.pi, 8, 0, 1), (np.pi, 0.0, np.pi, 8, 1, 1)]):
        b = rand_block(rng, p)
        fx["drot_%d_in" % i] = np.concatenate([[th, ph, xi, p, which, fwd], b])
        cmds.append("drot %.17g %.17g %.17g %d %d %d %s" % (th, ph, xi, p, which, fwd, fmt(b)))
        keys.append("drot_%d" % i)
    for i in range(4):
        qq = rng.normal(size=4)
        qq /= np.linalg.norm(qq)
        if i == 3:
            qq = np.array([0.8, 0.0, 0.0, 0.6])  # theta = 0
        p = [10, 15, 12, 9][i]
        fwd = i % 2
        b = rand_block(rng, p)
        fx["rotquat_%d_in" % i] = np.concatenate([qq, [p, fwd], b])
        cmds.append("rotquat %s %d %d %s" % (fmt(qq), p, fwd, fmt(b)))
        keys.append("rotquat_%d" % i)
    # singular orientations (theta = 0 or pi after the Euler split), both directions
    rs = np.random.default_rng(77)
    for i, qq in enumerate([(0.0, 1.0, 0.0, 0.0), (0.0, 0.6, 0.8, 0.0), (0.6, 0.0, 0.0, 0.8),
                            (0.0, 0.0, 0.0, 1.0)]):
        for fwd in (0, 1):
            p = 9
            b = rand_block(rs, p)
            fx["rotqs_%d_%d_in" % (i, fwd)] = np.concatenate([qq, [p, fwd], b])
            cmds.append("rotquat %s %d %d %s" % (fmt(qq), p, fwd, fmt(b)))
            keys.append("rotqs_%d_%d" % (i, fwd))

    # translations
    for i, (rho, kap, p, si, so, grad) in enumerate([(11.0, KAPPA, 10, 3.0, 4.5, 0), (11.0, 0.0, 10, 3.0, 4.5, 0),
                                                     (17.3, KAPPA, 15, 5.5, 2.1, 0), (9.4, KAPPA, 12, 2.5, 3.5, 1)]):
        b = rand_block(rng, p)
        fx["trans_%d_in" % i] = np.concatenate([[rho, kap, p, si, so, grad], b])
        cmds.append("trans %.17g %.17g %d %.17g %.17g %d %s" % (rho, kap, p, si, so, grad, fmt(b)))
        keys.append("trans_%d" % i)

    # re-expansions
    for i, (P, a_s, a_t, p) in enumerate([((6.0, 7.0, -3.0), 3.0, 4.0, 10), ((-12.0, 1.0, 9.0), 5.5, 2.2, 15),
                                          ((0.0, 0.0, 13.0), 4.0, 4.0, 10), ((0.0, 0.0, -13.0), 4.0, 3.0, 10)]):
        b = rand_block(rng, p, 0.1)
        fx["xformh_%d_in" % i] = np.concatenate([P, [a_s, a_t, p], b])
        cmds.append("xformh %s %.17g %.17g %d %s" % (fmt(P), a_s, a_t, p, fmt(b)))
        keys.append("xformh_%d" % i)
    for i, (P, a_s, a_t, p) in enumerate([((6.0, 7.0, -3.0), 3.0, 4.0, 10), ((0.0, 0.0, 12.0), 4.0, 3.0, 8)]):
        b = rand_block(rng, p, 0.1)
        fx["xformhg_%d_in" % i] = np.concatenate([P, [a_s, a_t, p], b])
        cmds.append("xformhg %s %.17g %.17g %d %s" % (fmt(P), a_s, a_t, p, fmt(b)))
        keys.append("xformhg_%d" % i)
    for i, (P, a_s, a_t, pH, pF) in enumerate([((6.0, 7.0, -3.0), 3.0, 4.0, 10, 7), ((-9.0, -8.0, 2.0), 2.5, 5.0,
                                                                                   6, 12)]):
        p = max(pH, pF)
        h, f = rand_block(rng, p, 0.1), rand_block(rng, p, 0.1)
        fx["xformi_%d_in" % i] = np.concatenate([P, [a_s, a_t, pH, pF], h, f])
        cmds.append("xformi %s %.17g %.17g %d %d %s %s" % (fmt(P), a_s, a_t, pH, pF, fmt(h), fmt(f)))
        keys.append("xformi_%d" % i)
    for i, (P, a_s, a_t, tqf, tqh) in enumerate([((7.2, 0.5, -0.4), 3.0, 4.0, 0.4, 0.7),
                                                 ((0.3, 5.9, 0.2), 3.0, 2.5, 1e-12, 0.2)]):
        n = 60
        pts = np.array([[np.sin(t) * np.cos(5 * t), np.sin(t) * np.sin(5 * t), np.cos(t)]
                        for t in np.linspace(0.05, np.pi - 0.05, n)]) * a_s
        qF, qH = rng.normal(size=n) * 0.01, rng.normal(size=n) * 0.01
        qF[3] = 0.0
        fx["xformn_%d_in" % i] = np.concatenate([P, [a_s, a_t, tqf, tqh], pts.ravel(), qF, qH])
        cmds.append("xformn %s %.17g %.17g %.17g %.17g %d %s" % (fmt(P), a_s, a_t, tqf, tqh, n, " ".join(
            "%.17g %.17g %.17g %.17g %.17g" % (*pts[j], qF[j], qH[j]) for j in range(n))))
        keys.append("xformn_%d" % i)
    order_in = [(12.0, 0.5, 4.0), (8.0, 0.5, 4.0), (30.0, 2.0, 3.0), (30.0, 1e-11, 3.0), (45.0, 0.03, 5.0)]
    fx["order_in"] = np.array(order_in)
    for i, o in enumerate(order_in):
        cmds.append("order %.17g %.17g %.17g" % o)
        keys.append("order_%d" % i)
    for i, (N, r) in enumerate([(400, 3.0), (1801, 5.5)]):
        cmds.append("spx %d %.17g" % (N, r))
        keys.append("spx_%d" % i)
    fx["spx_in"] = np.array([(400, 3.0), (1801, 5.5)])

    # IMat with buried points (p = N_POLES = 15)
    N = 3000
    pts = []
    t = np.linspace(-1, 1, N)
    ph = np.cumsum(np.r_[0, 3.6 / np.sqrt(N * (1 - t[1:-1] ** 2)), 0])
    for h, f in zip(t, ph):
        s = np.sqrt(max(0.0, 1 - h * h))
        pts.append([s * np.cos(f), s * np.sin(f), h])
    pts = np.array(pts) * 3.0
    bur = pts[:, 0] > 2.2
    fx["imat_spe"] = pts[~bur]
    fx["imat_spb"] = pts[bur]
    cmds.append("imat %d %d %s %s" % ((~bur).sum(), bur.sum(), fmt(pts[~bur]), fmt(pts[bur])))
    keys.append("imat")

    # molecule-type set-up: points, constants, fixed charges, initial surface charges
    for i, (seed, overlap) in enumerate([(5, True), (9, False)]):
        cen, rad, xyz, q = small_molecule(seed, ns=6, overlap=overlap)
        rc = cen.mean(0)
        fx["mol_%d_cen" % i] = cen
        fx["mol_%d_rad" % i] = rad
        fx["mol_%d_xyz" % i] = xyz
        fx["mol_%d_q" % i] = q
        fx["mol_%d_rcen" % i] = rc
        desc = "%d %s %d %s %s 4.0 {RC}" % (
            len(rad), " ".join("%.17g %.17g %.17g %.17g" % (*cen[k], rad[k]) for k in range(len(rad))), len(q),
            " ".join("%.17g %.17g %.17g %.17g" % (*xyz[k], q[k]) for k in range(len(q))), fmt(rc))
        cmds.append("moltype " + desc.replace("{RC}", "30.0"))
        keys.append("moltype_%d" % i)
        L = 15 * 15
        Fs = rng.normal(size=(len(rad), L)) * 0.01
        Hs = rng.normal(size=(len(rad), L)) * 0.01
        fx["molvals_%d_F" % i] = Fs
        fx["molvals_%d_H" % i] = Hs
        cmds.append("molvals " + desc.replace("{RC}", "%.17g" % (30.0 if i == 0 else 6.0)) + " " + fmt(Fs) + " " +
                    fmt(Hs))
        keys.append("molvals_%d" % i)
        fx["molvals_%d_rc" % i] = 30.0 if i == 0 else 6.0

    res = run(cmds)
    assert len(res) == len(keys), (len(res), len(keys))
    for k, v in zip(keys, res):
        fx["out_" + k] = v
    np.savez_compressed(os.path.join(HERE, "ref_kat.npz"), **fx)
    print("wrote", len(keys), "reference vectors")


if __name__ == "__main__":
    main()
