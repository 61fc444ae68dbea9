"""Writes the reference's file-format vectors to tests/golden/io_ref.npz.

Runs oracle/_ref/pbsam_ref_kat (built by `make -C oracle ref` from the reference's own sources,
present only in the build container) on small seeded inputs: readpqr on a PQR text (spheres
marked CEN at column 18, charges, a record at the last line without a newline),
writeMat_Unformatted / readMat_Unformatted (same order, and a truncating read),
writeExpansion / readExpansion (same order, truncating, padding), readConfigWithMolType.
The files the reference wrote are stored as bytes beside what its readers returned.

    python tests/golden/make_io_fixtures.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
KAT = os.path.join(ROOT, "oracle", "_ref", "pbsam_ref_kat")


def fmt(v):
    return " ".join("%.17g" % x for x in np.ravel(v))


def run(cmds):
    txt = "init 0.0737403009 78\n" + "\n".join(cmds) + "\n"
    out = subprocess.run([KAT], input=txt.encode(), capture_output=True, check=True).stdout.decode()
    lines = out.strip().split("\n")
    assert lines[0].startswith("ok"), lines[0]
    return [np.array([float(x) for x in ln.split()]) for ln in lines[1:]]


def pqr_text(rng, with_tail):
    """fixed columns as the reference reads them: residue name at column 18 (0-based; CEN for
    spheres), coordinates from column 30, then charge and radius"""
    lines = ["REMARK synthetic molecule"]
    k = 1
    for i in range(5):
        x, y, z = rng.uniform(-10, 10, 3)
        lines.append("ATOM  %5d  C    CEN X%4d   %8.3f%8.3f%8.3f %7.4f %6.4f" % (k, i, x, y, z, 0.0,
                                                                                   rng.uniform(2, 5)))
        k += 1
    for i in range(12):
        x, y, z = rng.uniform(-10, 10, 3)
        lines.append("ATOM  %5d  N    LYS A%4d   %8.3f%8.3f%8.3f %7.4f %6.4f" % (k, i, x, y, z,
                                                                                   rng.choice([-1.0, 0.5, 1.0]),
                                                                                   1.5))
        k += 1
    lines.append("TER")
    x, y, z = rng.uniform(-10, 10, 3)
    last = "ATOM  %5d  O    GLU A%4d   %8.3f%8.3f%8.3f %7.4f %6.4f" % (k, 99, x, y, z, -1.0, 1.4)
    return "\n".join(lines + [last]) + ("" if with_tail else "\n")


def main():
    if not os.path.exists(KAT):
        sys.exit("build the reference harness first: make -C oracle ref")
    rng = np.random.default_rng(20261203)
    fx = {}
    with tempfile.TemporaryDirectory() as d:
        cmds = []
        for tail in (0, 1):
            t = pqr_text(rng, tail == 0)
            fx["pqr_text_%d" % tail] = np.frombuffer(t.encode(), dtype=np.uint8)
            with open(os.path.join(d, "m%d.pqr" % tail), "w") as f:
                f.write(t)
            cmds.append("pqr %s" % os.path.join(d, "m%d.pqr" % tail))
        # surface-integral matrix of order 5, read back at 5 and (truncating) at 3
        p = 5
        mat = rng.normal(size=p ** 4)
        fx["imat_in"] = mat
        cmds.append("wrmat %s %d %s" % (os.path.join(d, "m.bin"), p, fmt(mat)))
        cmds.append("rdmat %s 5" % os.path.join(d, "m.bin"))
        cmds.append("rdmat %s 3" % os.path.join(d, "m.bin"))
        # expansion of order 4 (small values become 0 on writing), read at 4, 3, 6
        pe = 4
        coef = rng.normal(size=pe * pe) * np.array([10.0 ** -rng.integers(0, 18) for _ in range(pe * pe)])
        coef[3] = 1e-16
        coef[5] = 123456789.123
        fx["exp_in"] = coef
        fx["exp_scale_arg"] = 0.4
        fx["exp_rcut"] = 30.0
        cmds.append("wrexp %s %d 0.4 30 %s" % (os.path.join(d, "e.exp"), pe, fmt(coef)))
        cmds.append("rdexp %s 4" % os.path.join(d, "e.exp"))
        cmds.append("rdexp %s 3" % os.path.join(d, "e.exp"))
        cmds.append("rdexp %s 6" % os.path.join(d, "e.exp"))
        # configuration file (the last record has no newline after it: dropped)
        cfg = []
        for i in range(4):
            x, y, z = rng.uniform(-50, 50, 3)
            q = rng.normal(size=4)
            cfg.append("%.6f %.6f %.6f %.6f %.6f %.6f %.6f %d" % (x, y, z, *q, i % 2))
        cfg_text = "\n".join(cfg) + "\n" + "1 2 3 1 0 0 0 1"
        fx["cfg_text"] = np.frombuffer(cfg_text.encode(), dtype=np.uint8)
        with open(os.path.join(d, "c.cfg"), "w") as f:
            f.write(cfg_text)
        cmds.append("rdcfg %s" % os.path.join(d, "c.cfg"))
        res = run(cmds)
        fx["pqr_out_0"], fx["pqr_out_1"] = res[0], res[1]
        fx["imat_file"] = np.frombuffer(open(os.path.join(d, "m.bin"), "rb").read(), dtype=np.uint8)
        fx["imat_read5"], fx["imat_read3"] = res[3], res[4]
        fx["exp_file"] = np.frombuffer(open(os.path.join(d, "e.exp"), "rb").read(), dtype=np.uint8)
        fx["exp_read4"], fx["exp_read3"], fx["exp_read6"] = res[6], res[7], res[8]
        fx["cfg_out"] = res[9]
    np.savez_compressed(os.path.join(HERE, "io_ref.npz"), **fx)
    print("wrote io_ref.npz:", {k: np.shape(v) for k, v in fx.items()})


if __name__ == "__main__":
    main()
