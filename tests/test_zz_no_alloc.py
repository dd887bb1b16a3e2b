"""am_combine_noalloc (the combine that allocates nothing) against am_combine, through the C-ABI.

Two calls in a row on one context's GPU, the second on a smaller mesh: the shared static
scratch keeps the first call's contents, and the second result must still equal the
allocating entry point's."""
import ctypes
import os
import sys

import numpy as np
import pytest

from _common import workload

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "analyticmesh_amd", "libam.so")


def _lib():
    lib = ctypes.CDLL(LIB)
    P, LL = ctypes.c_void_p, ctypes.c_longlong
    lib.am_combine.argtypes = [P, ctypes.c_double, P]
    lib.am_combine.restype = ctypes.c_int
    lib.am_get_counts.argtypes = [P, P, P, P, P]
    lib.am_get_counts.restype = ctypes.c_int
    lib.am_get_mesh.argtypes = [P, P, P, P]
    lib.am_get_mesh.restype = ctypes.c_int
    lib.am_combine_noalloc.argtypes = [P, ctypes.c_double, P, P, LL, P, LL, P, LL, P]
    lib.am_combine_noalloc.restype = ctypes.c_int
    lib.am_last_error.argtypes = [P]
    lib.am_last_error.restype = ctypes.c_char_p
    return lib


def _mesh_allocating(lib, ctx, scale, center):
    c3 = np.asarray(center, dtype=np.float64)
    assert lib.am_combine(ctx, scale, c3.ctypes.data) == 0
    n = [ctypes.c_longlong() for _ in range(4)]
    assert lib.am_get_counts(ctx, *[ctypes.byref(x) for x in n]) == 0
    V, F, NI = n[2].value, n[1].value, n[3].value
    v = np.zeros((V, 3)); off = np.zeros(F + 1, np.int32); idx = np.zeros(NI, np.int32)
    assert lib.am_get_mesh(ctx, v.ctypes.data, off.ctypes.data, idx.ctypes.data) == 0
    return v, off, idx


def _mesh_noalloc(lib, ctx, scale, center, caps):
    import torch
    c3 = np.asarray(center, dtype=np.float64)
    dv = torch.full((caps[0], 3), np.nan, dtype=torch.float64, device="cuda")
    doff = torch.full((caps[1],), -7, dtype=torch.int32, device="cuda")
    didx = torch.full((caps[2],), -7, dtype=torch.int32, device="cuda")
    cnt = np.zeros(3, dtype=np.int64)
    torch.cuda.synchronize()
    rc = lib.am_combine_noalloc(ctx, scale, c3.ctypes.data, dv.data_ptr(), caps[0], doff.data_ptr(), caps[1],
                                didx.data_ptr(), caps[2], cnt.ctypes.data)
    assert rc == 0, lib.am_last_error(ctx)
    V, F, NI = (int(x) for x in cnt)
    return dv[:V].cpu().numpy(), doff[:F + 1].cpu().numpy(), didx[:NI].cpu().numpy()


def test_noalloc_combine_matches_allocating_combine():
    from analyticmesh_amd.engine import Engine
    lib = _lib()
    caps = (1 << 19, (1 << 18) + 1, 1 << 20)
    cases = [("sphere", 16, 2.0, (0.5, -1.0, 0.25)), ("sphere", 1, 1.0, (0.0, 0.0, 0.0))]
    for name, n_seeds, scale, center in cases:
        spec, pts, st = workload(name, n_seeds=n_seeds)
        print(f"[noalloc] {name} x{n_seeds}", flush=True)
        eng = Engine(spec.nodes, precision="fp64")
        res = eng.run([w.astype(np.float64) for w in spec.weights], [b.astype(np.float64) for b in spec.biases],
                      pts, st)
        ctx = ctypes.c_void_p(eng._eng.ctx_address())
        a = _mesh_noalloc(lib, ctx, scale, center, caps)
        b = _mesh_allocating(lib, ctx, scale, center)
        assert len(a[0]) > 100
        for x, y in zip(a, b):
            assert x.shape == y.shape
            assert np.array_equal(x, y)
        del res


def test_noalloc_combine_rejects_small_outputs():
    from analyticmesh_amd.engine import Engine
    lib = _lib()
    spec, pts, st = workload("sphere", n_seeds=1)
    eng = Engine(spec.nodes, precision="fp64")
    eng.run([w.astype(np.float64) for w in spec.weights], [b.astype(np.float64) for b in spec.biases], pts, st)
    ctx = ctypes.c_void_p(eng._eng.ctx_address())
    import torch
    dv = torch.zeros((8, 3), dtype=torch.float64, device="cuda")
    doff = torch.zeros(8, dtype=torch.int32, device="cuda")
    didx = torch.zeros(8, dtype=torch.int32, device="cuda")
    cnt = np.zeros(3, dtype=np.int64)
    rc = lib.am_combine_noalloc(ctx, 1.0, None, dv.data_ptr(), 8, doff.data_ptr(), 8, didx.data_ptr(), 8,
                                cnt.ctypes.data)
    assert rc == 4 and b"too small" in lib.am_last_error(ctx)
    host = np.zeros(64)
    rc = lib.am_combine_noalloc(ctx, 1.0, None, host.ctypes.data, 8, doff.data_ptr(), 8, didx.data_ptr(), 8,
                                cnt.ctypes.data)
    assert rc == 1 and b"device memory" in lib.am_last_error(ctx)
