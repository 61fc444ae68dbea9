"""Self-polarization rounds on the device against the oracle (same round structure)."""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu


def make(gpu, orc, n_sph, p, npts, seed):
    import pbsam_amd
    import pyoracle
    from pbsam_amd import synth
    mol = synth.molecule(n_sph, seed=seed)
    args = (mol["cen"], mol["rad"], mol["chg_xyz"], mol["chg_q"], p)
    kw = dict(eps_in=4.0, rcen=mol["rcen"], npts=[npts] * n_sph)
    return pbsam_amd.MolType(gpu, *args, **kw), pyoracle.MolType(orc, *args, **kw)


@pytest.mark.parametrize("p,npts", [(10, 400), (30, 1800)])
def test_initial_state(gpu, orc, p, npts):
    g, o = make(gpu, orc, 8, p, npts, seed=3)
    for k in range(g.ns):
        a, b = g.get(k), o.get(k)
        assert relerr(a["H"], b["H"]) < 1e-12
        assert relerr(a["qH"], b["qH"]) < 1e-12
        assert abs(a["TQH"] - b["TQH"]) < 1e-12 * b["TQH"]
        pg, ng = g.points(k)
        assert ng == npts and relerr(pg.ravel(), o.points(k).ravel()) < 1e-15


@pytest.mark.parametrize("p,npts,rounds", [(10, 400, 3), (30, 1800, 2)])
def test_jacobi_rounds(gpu, orc, p, npts, rounds):
    g, o = make(gpu, orc, 12, p, npts, seed=11)
    for r in range(rounds):
        mg = g.rounds(1)
        mo = o.round_jacobi()
        assert abs(mg - mo) <= 1e-9 * abs(mo) + 1e-300, r
        for k in range(g.ns):
            a, b = g.get(k), o.get(k)
            assert relerr(a["F"], b["F"]) < 1e-10, (r, k)
            assert relerr(a["H"], b["H"]) < 1e-10, (r, k)
            assert relerr(a["qH"], b["qH"]) < 1e-10, (r, k)
