"""The C-ABI library loads without a GPU and exports every declared entry point."""
import os
import re

import pytest

from conftest import ROOT


def declared():
    txt = open(os.path.join(ROOT, "include", "pbsam_amd.h")).read()
    return sorted(set(re.findall(r"\b(pbsam_[a-z0-9_]+)\s*\(", txt)))


def test_header_lists_entry_points():
    names = declared()
    assert "pbsam_solve_mutual" in names and "pbsam_selfpol" in names
    import pbsam_amd
    assert set(names) == set(pbsam_amd.SYMBOLS)


def test_library_exports_all_symbols():
    import pbsam_amd
    lib = pbsam_amd.load()
    for name in declared():
        assert hasattr(lib, name), name


def test_kappa_from_salt():
    import pbsam_amd
    k = pbsam_amd.kappa_from_salt(0.05)
    assert abs(k - 0.0737403009) < 1e-9


def test_bad_arguments_rejected_without_device():
    import ctypes
    import pbsam_amd
    lib = pbsam_amd.load()
    h = ctypes.c_void_p()
    assert lib.pbsam_ctx_create(-1.0, 78.0, 0, ctypes.byref(h)) == pbsam_amd.ERR_BADARG
    assert lib.pbsam_moltype_create(None, None, None) == pbsam_amd.ERR_BADARG
