"""The oracle's Brownian-dynamics restatement (CBDnam) against the reference.

* drand48: against the C library's own generator (the reference calls it through srand48 /
  drand48, BDnam_main.cpp:40-41, util.h:491-525).
* whole trajectories: against the reference's own CBDnam::compute_dt / CBDnam::makeMove compiled
  from its sources (tests/golden/bd_ref.npz, written by tests/golden/make_bd_fixtures.py), bit
  for bit, every step: status, time step, both poses and the stored rotations.
"""
import ctypes
import os

import numpy as np
import pytest

import pyoracle

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "bd_ref.npz")


def test_drand48_matches_libc():
    libc = ctypes.CDLL("libc.so.6")
    libc.drand48.restype = ctypes.c_double
    libc.srand48.argtypes = [ctypes.c_long]
    for seed in (0, 1, 12345, -7, 2 ** 33 + 5, 1733270400):
        libc.srand48(seed)
        ref = np.array([libc.drand48() for _ in range(2000)])
        u, _ = pyoracle.bd_rng(seed, 2000)
        np.testing.assert_array_equal(u, ref)


def test_normrand_polar_method():
    # normRand (util.h:514-537): pairs (v1, v2) uniform in the unit disc, returned as fac*v2 then
    # fac*v1 with fac = sqrt(-2 ln(rsq)/rsq) -- rebuilt here from the drand48 stream
    u, nrm = pyoracle.bd_rng(99, 4000)
    out, i = [], 0
    while len(out) < 4000 and i + 1 < len(u):
        v1, v2 = 2.0 * u[i] - 1.0, 2.0 * u[i + 1] - 1.0
        i += 2
        rsq = v1 * v1 + v2 * v2
        if rsq >= 1.0 or rsq == 0.0:
            continue
        fac = np.sqrt(-2.0 * np.log(rsq) / rsq)
        out += [fac * v2, v1 * fac]
    n = min(len(out), 1500)  # u holds enough uniforms for at least this many
    np.testing.assert_allclose(nrm[:n], out[:n], rtol=1e-15, atol=0)
    assert abs(np.mean(nrm)) < 0.06 and abs(np.var(nrm) - 1.0) < 0.08


def _case(fx, name):
    c = dict(seed=int(fx[name + "_seed"]), place=bool(fx[name + "_place"]), nsteps=int(fx[name + "_nsteps"]),
             D=fx[name + "_D"], ft=fx[name + "_ft"], pose=fx[name + "_pose"].reshape(2, 7))
    c["defs"] = [(int(fx["%s_def%d_n" % (name, i)]), fx["%s_def%d_pairs" % (name, i)],
                  fx["%s_def%d_dist" % (name, i)]) for i in range(int(fx[name + "_ndef"]))]
    return c


def _run_oracle(fx, c):
    prm = list(pyoracle.BD_DEFAULTS)
    prm[0:3] = list(c["D"])
    geom = [(fx["mol0_cen"], fx["mol0_rad"]), (fx["mol1_cen"], fx["mol1_rad"])]
    b = pyoracle.BDTrajectory(geom, c["seed"], c["place"], c["pose"][:, :3], c["pose"][:, 3:], prm)
    for nc, pairs, dist in c["defs"]:
        b.add_contact(nc, pairs, dist)
    rows = []
    for _ in range(c["nsteps"]):
        st = b.step(c["ft"])
        s = b.state()
        rows.append(np.concatenate([[st, s["dt"]], s["pos"][0], s["quat"][0], s["pos"][1], s["quat"][1],
                                    s["rot"].ravel()]))
        if st != 2:
            break
    return np.array(rows)


@pytest.mark.parametrize("name", ["far", "cross", "contact", "contact_nodock", "stuck", "stuck_translate",
                                  "escape"])
def test_trajectory_matches_reference(name):
    fx = np.load(FIX)
    ref = fx[name + "_rows"]
    got = _run_oracle(fx, _case(fx, name))
    assert got.shape == ref.shape, (got.shape, ref.shape)
    np.testing.assert_array_equal(got, ref)


def test_fixtures_cover_every_branch():
    """The fixtures exercise every exit and rotation branch of makeMove."""
    fx = np.load(FIX)
    final = {str(n): int(fx[str(n) + "_rows"][-1, 0]) for n in fx["cases"]}
    assert final["escape"] == 0 and final["contact"] == 1 and final["stuck"] == 3
    assert final["stuck_translate"] == 3 and final["far"] == 2
    rows = fx["cross_rows"]
    c2c = np.linalg.norm(rows[:, 9:12] - rows[:, 2:5], axis=1)
    prev = np.concatenate([[np.linalg.norm(fx["cross_pose"][7:10] - fx["cross_pose"][0:3])], c2c[:-1]])
    stored = np.abs(rows[:, 16] - 1.0) + np.abs(rows[:, 20] - 1.0) > 0
    # stored rotations while beyond f_DIST, applied and reset on the step that comes inside
    assert np.any((c2c > 100) & stored)
    enter = (prev > 100) & (c2c <= 100)
    assert np.any(enter) and np.all(~stored[enter])
    assert np.any((prev <= 100) & (c2c <= 100))
    # compute_dt's distance-dependent branch and its floor of 2
    assert fx["far_rows"][:, 1].max() > 2.0 and np.any(fx["contact_rows"][:, 1] == 2.0)
