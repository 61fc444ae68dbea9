"""Seeded synthetic molecules for the benchmark and tests.

Spheres: radii uniform in [2, 6] A, rejection-packed without overlap inside a ball of
radius 20*(n/20)^(1/3) A (20 A for 20 spheres; a 20 A ball cannot hold 60+ spheres).
Charges: 10-40 per sphere, uniform in the sphere's interior (radius 0.9 r), values +-1 e
with the net charge brought into [-3, 3].  Orientations: uniform random unit quaternions
(the method of CQuat::chooseRandom, src/util.h:488-499).
"""
import numpy as np

SALT_M = 0.05       # mol/L
EPS_IN = 4.0
EPS_OUT = 78.0
TEMP = 298.15


def ball_radius(n):
    return 20.0 * (n / 20.0) ** (1.0 / 3.0)


def molecule(n_sph, seed, rmin=2.0, rmax=6.0, qmin=10, qmax=40, ball=None, max_tries=200000, gap=0.0):
    rng = np.random.default_rng(seed)
    R = ball_radius(n_sph) if ball is None else ball
    cen, rad = [], []
    tries = 0
    while len(rad) < n_sph:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("sphere packing failed")
        r = rng.uniform(rmin, rmax)
        if r >= R:
            continue
        c = rng.uniform(-1, 1, 3) * (R - r)
        if np.linalg.norm(c) + r > R:
            continue
        ok = True
        for c2, r2 in zip(cen, rad):
            if np.linalg.norm(c - c2) < r + r2 + gap:
                ok = False
                break
        if ok:
            cen.append(c)
            rad.append(r)
    cen = np.array(cen)
    rad = np.array(rad)
    xyz, q = [], []
    for c, r in zip(cen, rad):
        nq = int(rng.integers(qmin, qmax + 1))
        for _ in range(nq):
            while True:
                v = rng.uniform(-1, 1, 3)
                if v @ v <= 1.0:
                    break
            xyz.append(c + 0.9 * r * v)
            q.append(1.0 if rng.random() < 0.5 else -1.0)
    q = np.array(q)
    xyz = np.array(xyz)
    while abs(q.sum()) > 3:
        idx = np.flatnonzero(q == np.sign(q.sum()))
        q[rng.choice(idx)] *= -1
    return dict(cen=cen, rad=rad, chg_xyz=xyz, chg_q=q, rcen=np.zeros(3))


def random_quat(rng):
    s = rng.random()
    sig1, sig2 = np.sqrt(1.0 - s), np.sqrt(s)
    t1, t2 = 2.0 * np.pi * rng.random(), 2.0 * np.pi * rng.random()
    q = np.array([np.cos(t2) * sig2, np.sin(t1) * sig1, np.cos(t1) * sig1, np.sin(t2) * sig2])
    return q / np.linalg.norm(q)


def quat_rot(q, v):
    w, x = q[0], q[1:]
    qr = np.cross(x, v)
    return v + 2 * (w * qr + np.cross(x, qr))


def place_pair(mol_a, mol_b, dist, seed, max_tries=1000):
    """Poses for two molecules `dist` apart (along x) with random orientations chosen so
    that no two spheres of different molecules overlap."""
    rng = np.random.default_rng(seed)
    for _ in range(max_tries):
        qa, qb = random_quat(rng), random_quat(rng)
        pa, pb = np.zeros(3), np.array([dist, 0.0, 0.0])
        ca = np.array([pa + quat_rot(qa, c - mol_a["rcen"]) for c in mol_a["cen"]])
        cb = np.array([pb + quat_rot(qb, c - mol_b["rcen"]) for c in mol_b["cen"]])
        d = np.linalg.norm(ca[:, None, :] - cb[None, :, :], axis=2)
        gap = d - mol_a["rad"][:, None] - mol_b["rad"][None, :]
        if gap.min() > 0.5:
            return np.stack([pa, pb]), np.stack([qa, qb])
    raise RuntimeError("no overlap-free orientation found")
