"""The C-ABI library loads and exports every symbol include/pot3d_hip.h
declares (no GPU work is launched)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "pot3d_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", src)
    skip = {"if", "defined", "struct", "sizeof"}
    return sorted({n for n in names if n not in skip})


def test_header_declares_legacy_and_whole_solve_symbols():
    d = _declared()
    for n in ("load_external_lu_solver", "external_lu_solver", "unload_external_lu_solver",
              "p3d_create", "p3d_solve", "p3d_destroy"):
        assert n in d


def test_library_exports_all_declared_symbols():
    from pot3d_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libpot3d_hip.so not built (run __graft_entry__.build())")
    L = _lib.load()
    for n in _declared():
        assert hasattr(L, n), n
    assert set(_lib.SIGNATURES) == set(_declared())


def test_error_path_without_gpu_work():
    """p3d_create rejects bad shapes before touching the device."""
    import ctypes as C
    from pot3d_amd import _lib
    L = _lib.load()
    h = C.c_void_p()
    rc = L.p3d_create(C.byref(h), 2, 2, 2, *([None] * 10), -1.0, 1)
    assert rc == _lib.P3D_EINVAL
    assert "too small" in _lib.last_error()
