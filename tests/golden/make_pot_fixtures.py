"""Writes the reference's potential and interaction-energy vectors to tests/golden/pot_ref.npz.

Runs the "potspace" command of oracle/_ref/pbsam_ref_kat (built by `make -C oracle ref` from the
reference's own sources, present only in the build container): two molecules made by the
reference's BD constructor from given self expansions F, H (order N_POLES = 15), posed,
then CMolecule::computePotInSpace at seeded points and CMolecule::computeTotalIntEnergy.
Two poses: identity orientations, and rotated molecules (where the reference's formula uses
the unrotated centres).

    python tests/golden/make_pot_fixtures.py
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from make_bd_fixtures import KAT, fmt, molecule  # noqa: E402

N = 15


def main():
    if not os.path.exists(KAT):
        sys.exit("build the reference harness first: make -C oracle ref")
    rng = np.random.default_rng(4)
    mols = [molecule(5, 12, 11), molecule(4, 10, 12)]
    fx = {}
    for i, (cen, rad, q) in enumerate(mols):
        dec = np.array([0.5 ** n for n in range(N) for _ in range(2 * n + 1)])
        fx["mol%d_cen" % i], fx["mol%d_rad" % i], fx["mol%d_q" % i] = cen, rad, q
        fx["mol%d_F" % i] = rng.normal(size=(len(rad), N * N)) * dec * 0.05
        fx["mol%d_H" % i] = rng.normal(size=(len(rad), N * N)) * dec * 0.05
    pts = rng.uniform(-40, 40, (12, 3))
    fx["pts"] = pts
    poses = [np.array([0, 0, 0, 1, 0, 0, 0, 32.0, 1, -2, 1, 0, 0, 0]),
             np.array([0, 0, 0, 0.9, 0.1, -0.3, 0.2, 32.0, 1, -2, 0.5, 0.5, 0.5, -0.5])]
    for k, pose in enumerate(poses):
        txt = "init 0.0737403009 78\npotspace %d %s " % (len(pts), fmt(pts))
        for i, (cen, rad, q) in enumerate(mols):
            txt += "%d %s %d %s %s %s " % (len(rad), fmt(np.concatenate([cen, rad[:, None]], 1)), len(q), fmt(q),
                                          fmt(fx["mol%d_F" % i]), fmt(fx["mol%d_H" % i]))
        txt += fmt(pose)
        out = subprocess.run([KAT], input=(txt + "\n").encode(), capture_output=True, check=True).stdout.decode()
        res = np.array([float(x) for x in out.strip().split("\n")[1].split()])
        fx["pose%d" % k] = pose
        fx["pot%d" % k] = res[:-1]
        fx["eint%d" % k] = res[-1]
    np.savez_compressed(os.path.join(HERE, "pot_ref.npz"), **fx)
    print("wrote pot_ref.npz", fx["pot0"][:3], fx["eint0"], fx["eint1"])


if __name__ == "__main__":
    main()
