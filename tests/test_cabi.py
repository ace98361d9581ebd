"""The C-ABI library loads (no GPU needed) and exports every function the
public headers declare; the ctypes table covers the solver header."""
import ctypes
import os
import re

from jsolver_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(jsv_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for header in ("jsolver_amd.h", "jsolver_amd_unet.h"):
        names = _declared(header)
        assert len(names) > 5, header
        missing = [n for n in names if not hasattr(lib, n)]
        assert not missing, (header, missing)


def test_ctypes_table_matches_header():
    assert set(_declared("jsolver_amd.h")) == set(_lib.SIGNATURES)


def test_abi_version_and_error_path():
    lib = _lib.load()
    assert lib.jsv_abi_version() == 1
    plan = ctypes.c_void_p()
    # a non-power-of-2 resolution is rejected before any device work
    st = lib.jsv_plan_create(2, 100, 128, 1, 1.0, 1.0, 1.0, 0, ctypes.byref(plan))
    assert st != 0
    assert b"not a power of 2" in lib.jsv_last_error()
