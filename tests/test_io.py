"""The reference's file formats (pbsam_read_* / pbsam_write_*) against the reference's own
readers and writers (readpqr, writeMat_Unformatted / readMat_Unformatted, writeExpansion /
readExpansion, readConfigWithMolType compiled from its sources; tests/golden/io_ref.npz,
written by tests/golden/make_io_fixtures.py).  Host code only: no device is touched.

Every test runs twice, once in the CPU set and once marked `gpu`, so that the run on the
MI355X also checks the readers and writers of the library it loads there."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "io_ref.npz")


@pytest.fixture(scope="module")
def fx():
    return np.load(FIX)


@pytest.fixture(scope="module", params=[pytest.param("host"), pytest.param("gpu-set", marks=pytest.mark.gpu)])
def lib(request):
    import pbsam_amd
    pbsam_amd.load()
    return pbsam_amd


@pytest.mark.parametrize("k", [0, 1])
def test_read_pqr_matches_reference(lib, fx, tmp_path, k):
    path = tmp_path / "m.pqr"
    path.write_bytes(bytes(fx["pqr_text_%d" % k]))
    cen, rad, xyz, q = lib.read_pqr(path)
    ref = fx["pqr_out_%d" % k]
    ns = int(ref[0])
    rcen, rrad = ref[1:1 + 3 * ns].reshape(-1, 3), ref[1 + 3 * ns:1 + 4 * ns]
    o = 1 + 4 * ns
    nq = int(ref[o])
    rxyz, rq = ref[o + 1:o + 1 + 3 * nq].reshape(-1, 3), ref[o + 1 + 3 * nq:]
    np.testing.assert_array_equal(cen, rcen)
    np.testing.assert_array_equal(rad, rrad)
    np.testing.assert_array_equal(xyz, rxyz)
    np.testing.assert_array_equal(q, rq)


def test_pqr_last_line_rule(lib, fx):
    # the reference processes a last line only when a newline ends it
    assert len(fx["pqr_out_1"]) == len(fx["pqr_out_0"]) + 4


def test_pqr_long_line_is_an_error(lib, tmp_path):
    path = tmp_path / "long.pqr"
    path.write_text("REMARK\nATOM " + "x" * 120 + "\nATOM\n")
    with pytest.raises(lib.PbsamError):
        lib.read_pqr(path)


def test_imat_write_is_the_reference_file(lib, fx, tmp_path):
    path = tmp_path / "m.bin"
    lib.write_imat(path, fx["imat_in"], 5)
    assert path.read_bytes() == bytes(fx["imat_file"])


def test_imat_read_matches_reference(lib, fx, tmp_path):
    path = tmp_path / "ref.bin"
    path.write_bytes(bytes(fx["imat_file"]))
    m5, pf = lib.read_imat(path, 5)
    assert pf == 5
    np.testing.assert_array_equal(m5, fx["imat_read5"])
    m3, pf = lib.read_imat(path, 3)
    assert pf == 5
    np.testing.assert_array_equal(m3, fx["imat_read3"])
    # the truncation is the leading 9 x 9 block of the column-major 25 x 25 matrix
    np.testing.assert_array_equal(m3.reshape(9, 9), fx["imat_in"].reshape(25, 25)[:9, :9])
    with pytest.raises(lib.PbsamError):
        lib.read_imat(path, 6)


def test_expansion_write_is_the_reference_file(lib, fx, tmp_path):
    path = tmp_path / "e.exp"
    # the reference writes CMulExpan's scale, the reciprocal of the radius it was given
    lib.write_expansion(path, fx["exp_in"], 4, 1.0 / float(fx["exp_scale_arg"]), float(fx["exp_rcut"]))
    assert path.read_bytes() == bytes(fx["exp_file"])


@pytest.mark.parametrize("p", [4, 3, 6])
def test_expansion_read_matches_reference(lib, fx, tmp_path, p):
    path = tmp_path / "ref.exp"
    path.write_bytes(bytes(fx["exp_file"]))
    c, sc, rc, pf = lib.read_expansion(path, p)
    ref = fx["exp_read%d" % p]
    assert pf == 4
    assert sc == ref[0] and rc == ref[1]
    np.testing.assert_array_equal(c, ref[2:])


def test_config_read_matches_reference(lib, fx, tmp_path):
    path = tmp_path / "c.cfg"
    path.write_bytes(bytes(fx["cfg_text"]))
    pos, quat, t = lib.read_config(path)
    ref = fx["cfg_out"]
    n = int(ref[0])
    rows = ref[1:].reshape(n, 8)
    assert len(t) == n
    np.testing.assert_array_equal(pos, rows[:, 0:3])
    np.testing.assert_array_equal(quat, rows[:, 3:7])
    np.testing.assert_array_equal(t, rows[:, 7].astype(int))


def test_config_round_trip(lib, tmp_path):
    rng = np.random.default_rng(3)
    pos = rng.uniform(-50, 50, (5, 3))
    quat = rng.normal(size=(5, 4))
    quat /= np.linalg.norm(quat, axis=1)[:, None]
    t = np.array([0, 1, 1, 0, 2])
    path = tmp_path / "w.cfg"
    lib.write_config(path, pos, quat, t)
    p2, q2, t2 = lib.read_config(path)
    # saveConfig's default stream precision: 6 significant digits
    np.testing.assert_allclose(p2, pos, rtol=1e-5, atol=1e-4)
    np.testing.assert_allclose(q2, quat, atol=2e-6)
    np.testing.assert_array_equal(t2, t)


def test_directory_layouts(lib, tmp_path):
    rng = np.random.default_rng(5)
    ns, p = 3, 4
    mats = rng.normal(size=(ns, p ** 4))
    F, H = rng.normal(size=(ns, p * p)), rng.normal(size=(ns, p * p))
    for k in range(ns):
        lib.write_imat(tmp_path / ("imat.sp%d.out.bin" % k), mats[k], p)
        lib.write_expansion(tmp_path / ("run.%d.F.exp" % k), F[k], p, 0.25, 30.0)
        lib.write_expansion(tmp_path / ("run.%d.H.exp" % k), H[k], p, 0.25, 30.0)
    np.testing.assert_array_equal(lib.read_imats(tmp_path, ns, p), mats)
    F2, H2, rc = lib.read_expansions(tmp_path, "run", ns, p)
    assert rc == 30.0
    # text at precision 8
    np.testing.assert_allclose(F2, F, rtol=1e-7, atol=1e-15)
    np.testing.assert_allclose(H2, H, rtol=1e-7, atol=1e-15)


@pytest.mark.parametrize("head", ["-6 0.25 30.0\n", "0 0.25 30.0\n", "100000 0.25 30.0\n", "x 1 2\n"])
def test_expansion_bad_order_is_an_error(lib, tmp_path, head):
    # a malformed order (negative, zero, absurd or not a number) is rejected rather than read
    # into a buffer sized for the caller's order
    path = tmp_path / "bad.exp"
    path.write_text(head + " ".join(["1.0"] * 50) + "\n")
    with pytest.raises(lib.PbsamError):
        lib.read_expansion(path, 4)
