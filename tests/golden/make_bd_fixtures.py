"""Writes reference Brownian-dynamics trajectories to tests/golden/bd_ref.npz.

Runs the "bd" command of oracle/_ref/pbsam_ref_kat (built by `make -C oracle ref` from the
reference's own sources, present only in the build container): the reference's
CBDnam::compute_dt and CBDnam::makeMove, driven as CBDnam::run drives them (BDnam.cpp:196-246),
with a fixed force and torques in place of CMolecule::computeForces (whose solver needs MKL,
absent here).  One process per trajectory, because normRand keeps a static spare value.
Each row of a trajectory: status, dt, pos0(3), quat0(4), pos1(3), quat1(4), m_rot1(4), m_rot2(4).

    python tests/golden/make_bd_fixtures.py
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
KAT = os.path.join(ROOT, "oracle", "_ref", "pbsam_ref_kat")
DEF_D = [0.03, 4.5e-5, 4e-5]  # m_Dtr, m_Dr1, m_Dr2 (BDnam.cpp:143)


def fmt(v):
    return " ".join("%.17g" % x for x in np.ravel(v))


def molecule(ns, R, seed):
    """Non-overlapping spheres (radii 2-5 A) inside a ball of radius R, one charge each."""
    rng = np.random.default_rng(seed)
    cen, rad = [], []
    while len(rad) < ns:
        r = rng.uniform(2, 5)
        c = rng.uniform(-R, R, 3)
        if np.linalg.norm(c) + r > R:
            continue
        if all(np.linalg.norm(c - c2) > r + r2 + 0.5 for c2, r2 in zip(cen, rad)):
            cen.append(c)
            rad.append(r)
    cen, rad = np.array(cen), np.array(rad)
    q = np.concatenate([cen, rng.choice([-1.0, 1.0], (ns, 1))], axis=1)
    return cen, rad, q


# name: seed, place, steps, (Dtr, Dr1, Dr2), force1+torque0+torque1, start pose, contacts
CASES = {
    # random start at b_DIST, reference diffusion: stored rotations, crossing f_DIST
    "far": dict(seed=11, place=1, nsteps=400, D=[3.0, 4.5e-3, 4e-3], ft=np.zeros(9)),
    # start just outside f_DIST moving in and out: the "apply the stored rotations" branch
    "cross": dict(seed=21, place=0, nsteps=300, D=[2.0, 1e-3, 1e-3], ft=[-0.01, 0.002, 0, 0.01, 0, 0.02, 0, 0.01, 0],
                  pose=[0, 0, 0, 1, 0, 0, 0, 100.5, 1, 0, 0.8, 0.2, -0.1, 0.3]),
    # near contact, attractive force: collision retries in translation and rotation, docking
    "contact": dict(seed=32, place=0, nsteps=300, D=[0.05, 2e-4, 2e-4], ft=[-0.02, 0.001, 0, 0.02, 0.01, 0, 0, 0.02, 0.01],
                    pose=[0, 0, 0, 1, 0, 0, 0, 26.5, 0.5, 0.2, 0.9, 0.1, 0.3, 0.2],
                    defs=[(3, [(0, 0), (1, 1), (2, 3), (3, 2)], [0.5, 0.5, 0.5, 0.5]),
                          (2, [(4, 1), (0, 1), (3, 1)], [1.0, 1.0, 1.0])]),
    "stuck_translate": dict(seed=33, place=0, nsteps=50, D=[0.002, 2e-4, 2e-4], ft=[-2.0, 0, 0, 0, 0, 0, 0, 0, 0],
                            pose=[0, 0, 0, 1, 0, 0, 0, 26.5, 0.5, 0.2, 0.9, 0.1, 0.3, 0.2]),
    "contact_nodock": dict(seed=32, place=0, nsteps=250, D=[0.05, 2e-4, 2e-4],
                           ft=[-0.02, 0.001, 0, 0.02, 0.01, 0, 0, 0.02, 0.01],
                           pose=[0, 0, 0, 1, 0, 0, 0, 26.5, 0.5, 0.2, 0.9, 0.1, 0.3, 0.2]),
    # strong torque at contact: stuck in rotation
    "stuck": dict(seed=12, place=0, nsteps=50, D=[0.3, 0.01, 0.01], ft=[-0.02, 0, 0, 0.1, 0.2, 0.3, -0.1, 0, 0.2],
                  pose=[0, 0, 0, 1, 0, 0, 0, 27.5, 0, 0, 0.9, 0.1, 0.3, 0.2]),
    # fast diffusion from b_DIST: escape past q_DIST
    "escape": dict(seed=41, place=1, nsteps=400, D=[40.0, 1e-3, 1e-3], ft=np.zeros(9)),
}


def run_case(m0, m1, seed, place, nsteps, D, ft, pose=None, defs=()):
    txt = "init 0.0737403009 78\nbd %d %d %d %s %s " % (seed, place, nsteps, fmt(D), fmt(ft))
    for cen, rad, q in (m0, m1):
        txt += "%d %s %d %s " % (len(rad), fmt(np.concatenate([cen, rad[:, None]], 1)), len(q), fmt(q))
    if not place:
        txt += fmt(pose) + " "
    txt += "%d " % len(defs)
    for nc, pairs, dist in defs:
        txt += "%d %d " % (nc, len(pairs))
        for (a, b), d in zip(pairs, dist):
            txt += "%d %d %.17g " % (a, b, d)
    out = subprocess.run([KAT], input=(txt + "\n").encode(), capture_output=True, check=True).stdout.decode()
    lines = out.strip().split("\n")
    assert lines[0].startswith("ok"), lines[0]
    return np.array([float(x) for x in lines[1].split()]).reshape(-1, 24)


def main():
    if not os.path.exists(KAT):
        sys.exit("build the reference harness first: make -C oracle ref")
    m0, m1 = molecule(8, 14, 1), molecule(6, 12, 2)
    fx = {"mol0_cen": m0[0], "mol0_rad": m0[1], "mol0_q": m0[2], "mol1_cen": m1[0], "mol1_rad": m1[1],
          "mol1_q": m1[2], "cases": np.array(sorted(CASES))}
    for name, c in sorted(CASES.items()):
        rows = run_case(m0, m1, **c)
        fx[name + "_rows"] = rows
        fx[name + "_seed"] = c["seed"]
        fx[name + "_place"] = c["place"]
        fx[name + "_nsteps"] = c["nsteps"]
        fx[name + "_D"] = np.array(c["D"], float)
        fx[name + "_ft"] = np.array(c["ft"], float)
        fx[name + "_pose"] = np.array(c.get("pose", np.zeros(14)), float)
        defs = c.get("defs", [])
        fx[name + "_ndef"] = len(defs)
        for i, (nc, pairs, dist) in enumerate(defs):
            fx["%s_def%d_n" % (name, i)] = nc
            fx["%s_def%d_pairs" % (name, i)] = np.array(pairs, np.int32)
            fx["%s_def%d_dist" % (name, i)] = np.array(dist, float)
        print("%-15s steps %4d final status %d" % (name, len(rows), int(rows[-1, 0])))
    np.savez_compressed(os.path.join(HERE, "bd_ref.npz"), **fx)


if __name__ == "__main__":
    main()
