"""The C-ABI library loads and exports everything include/am.h declares (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "am.h")
LIB = os.path.join(ROOT, "analyticmesh_amd", "libam.so")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:am_status|const char\*)\s+(am_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_reference_calls():
    syms = declared_symbols()
    for s in ("am_create", "am_run", "am_combine", "am_export_ply", "am_destroy", "am_last_error",
              "am_get_counts", "am_get_mesh", "am_get_regions", "am_get_stats", "am_set_option"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    for s in declared_symbols():
        assert hasattr(lib, s), f"{s} missing from libam.so"


def test_create_reports_errors_without_exit():
    lib = ctypes.CDLL(LIB)
    P = ctypes.c_void_p
    lib.am_create.restype = ctypes.c_int
    lib.am_create.argtypes = [P, ctypes.c_int, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.am_last_error.restype = ctypes.c_char_p
    lib.am_last_error.argtypes = [P]
    lib.am_destroy.argtypes = [P]
    ctx = ctypes.c_void_p()
    nodes = np.array([3, 14, 1], dtype=np.int32)
    arc = np.zeros((1, 1), dtype=np.int32)
    # bad nodes -> AM_ERR_ARG
    bad = np.array([2, 14, 1], dtype=np.int32)
    rc = lib.am_create(ctypes.byref(ctx), 3, bad.ctypes.data, arc.ctypes.data, 1, 1, 0, 2)
    assert rc == 1 and b"nodes" in lib.am_last_error(None)
    # bad arc table -> AM_ERR_ARG
    rc = lib.am_create(ctypes.byref(ctx), 3, nodes.ctypes.data, arc.ctypes.data, 2, 1, 0, 2)
    assert rc == 1
    # valid arguments: OK on a GPU box, AM_ERR_NOGPU here
    rc = lib.am_create(ctypes.byref(ctx), 3, nodes.ctypes.data, arc.ctypes.data, 1, 1, 0, 2)
    assert rc in (0, 5)
    if rc == 0:
        lib.am_destroy(ctx)


def test_pybind_module_imports():
    from analyticmesh_amd.engine import native
    m = native()
    assert m.PREC_FP32 == 0 and m.PREC_SPLIT_BF16X3 == 1 and m.PREC_FP64 == 2


def test_calls_out_of_order_fail_cleanly():
    lib = ctypes.CDLL(LIB)
    lib.am_combine.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p]
    lib.am_destroy.argtypes = [ctypes.c_void_p]
    assert lib.am_combine(None, 1.0, None) == 1
    assert lib.am_destroy(None) == 1


def test_ply_roundtrip(tmp_path):
    from analyticmesh_amd.ply import read_ply, write_ply
    V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0.5]], dtype=np.float64)
    F = [np.array([0, 1, 2, 3])]
    write_ply(str(tmp_path / "a.ply"), V, F, is_float32=False)
    V2, F2 = read_ply(str(tmp_path / "a.ply"))
    np.testing.assert_array_equal(V, V2)
    assert [list(f) for f in F2] == [[0, 1, 2, 3]]
    write_ply(str(tmp_path / "b.ply"), V, F, is_float32=True, is_polymesh=False)
    V3, F3 = read_ply(str(tmp_path / "b.ply"))
    assert [list(f) for f in F3] == [[0, 1, 2], [0, 2, 3]]
