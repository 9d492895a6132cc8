"""The C ABI library loads without a GPU and exports every entry point include/fair_hip.h
declares; the Python binding covers all of them; the Fortran drop-in is built."""
import os
import re

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "fair_hip.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(fair_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert len(names) >= 30
    for must in ("fair_fanlineartomo", "fair_paralleltomo", "fair_matvecmul", "fair_transvecmul",
                 "fair_row_norms", "fair_transpose", "fair_sirt", "fair_gradient_descent", "fair_art",
                 "fair_normest", "fair_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from fortran_airtools_amd import _lib
    L = _lib.load()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from fortran_airtools_amd import _lib
    assert set(declared()) <= set(_lib.SIGNATURES)


def test_fortran_dropin_built():
    pkg = os.path.join(ROOT, "fortran_airtools_amd")
    assert os.path.exists(os.path.join(pkg, "libfair_fortran.a"))
    assert os.path.exists(os.path.join(pkg, "bin", "dropin_check"))


def test_missing_library_fails_loudly(tmp_path):
    from fortran_airtools_amd import _lib
    import pytest
    with pytest.raises(_lib.FairError, match="not built"):
        _lib.load.__wrapped__(str(tmp_path / "nope.so")) if hasattr(_lib.load, "__wrapped__") else \
            _check_missing(_lib, str(tmp_path / "nope.so"))


def _check_missing(_lib, path):
    saved = _lib._lib
    _lib._lib = None
    try:
        _lib.load(path)
    finally:
        _lib._lib = saved
