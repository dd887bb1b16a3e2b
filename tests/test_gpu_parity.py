"""GPU engine vs the CPU restatement (oracle) on the same seeded inputs."""
import os
import sys

import numpy as np
import pytest

from _common import assert_same_topology, max_vertex_error, workload

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def oracle_mod():
    import oracle
    return oracle


@pytest.mark.parametrize("name", ["polytope", "sphere"])
def test_fp64_bitexact_vs_oracle_bfs(name, oracle_mod):
    """FP64 path = the oracle's breadth-first order, record for record, bit for bit."""
    from analyticmesh_amd.engine import march
    spec, pts, st = workload(name)
    _say(f"[fp64 bit-exact] {name} L={spec.L}")
    g = march(spec, pts, st, precision="fp64")
    o = oracle_mod.run_oracle(spec, pts, st, traversal="bfs")
    assert len(g.masks) == len(o.masks)
    np.testing.assert_array_equal(g.masks, o.masks)
    np.testing.assert_array_equal(g.info[:, :6], o.info[:, :6])
    np.testing.assert_array_equal(g.tabs, o.tabs)
    np.testing.assert_array_equal(g.vpairs, o.vpairs)
    np.testing.assert_array_equal(g.points, o.points)
    assert np.array_equal(g.verts.view(np.uint64), o.verts.view(np.uint64)), "vertex bits differ"


@pytest.mark.parametrize("name", ["polytope", "sphere"])
def test_fp64_topology_vs_oracle_lifo(name, oracle_mod):
    """Same regions, faces and vertex keys as the reference's own LIFO order."""
    from analyticmesh_amd.engine import march
    spec, pts, st = workload(name)
    _say(f"[fp64 vs lifo] {name}")
    g = march(spec, pts, st, precision="fp64").records()
    o = oracle_mod.run_oracle(spec, pts, st, traversal="lifo").records()
    assert_same_topology(g, o)
    # the same vertex solved from different regions / plane orders: rounding
    # amplified by the condition number (~1e7 for the polytope's +- plane pairs)
    assert max_vertex_error(g, o) < (1e-7 if name == "polytope" else 1e-10)


def test_fp64_mesh_and_ply(tmp_path, oracle_mod):
    from analyticmesh_amd.engine import march
    from analyticmesh_amd.ply import read_ply
    spec, pts, st = workload("sphere")
    _say("[mesh+ply] sphere")
    res = march(spec, pts, st, precision="fp64")
    V, F = res.combine(scale=2.0, center=(0.5, -1.0, 0.25))
    Vr, Fr = res.records().keyed_mesh()
    assert len(V) == len(Vr) and len(F) == len(Fr)
    np.testing.assert_array_equal(V, Vr * 2.0 + np.array([0.5, -1.0, 0.25]))
    for a, b in zip(F, Fr):
        assert list(a) == list(b)
    p = tmp_path / "m.ply"
    res.export_ply(p, is_polymesh=True, is_float32=False)
    V2, F2 = read_ply(str(p))
    np.testing.assert_array_equal(V2, V)
    assert [list(f) for f in F2] == [list(f) for f in F]
    p3 = tmp_path / "t.ply"
    res.export_ply(p3, is_polymesh=False, is_float32=True)
    V3, F3 = read_ply(str(p3))
    assert len(F3) == sum(len(f) - 2 for f in F)
    np.testing.assert_array_equal(V3, V.astype(np.float32))


def test_multi_seed_dedup(oracle_mod):
    from analyticmesh_amd.engine import march
    spec, pts, st = workload("sphere", n_seeds=64)
    _say("[multi-seed] sphere x64")
    g = march(spec, pts, st, precision="fp64")
    o = oracle_mod.run_oracle(spec, pts, st, traversal="bfs")
    np.testing.assert_array_equal(g.masks, o.masks)
    np.testing.assert_array_equal(g.tabs, o.tabs)


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_device_inputs_match_host_inputs(prec, dtype):
    """Weights, seeds and states given as device tensors (no host staging) give the same
    records and mesh as the same arrays given in host memory."""
    import torch
    from analyticmesh_amd.engine import Engine
    spec, pts, st = workload("sphere", n_seeds=64)
    _say(f"[device inputs] {prec} {dtype}")
    W = [np.ascontiguousarray(w, dtype) for w in spec.weights]
    B = [np.ascontiguousarray(b, dtype) for b in spec.biases]
    P = np.ascontiguousarray(pts, dtype)
    out = []
    for on_device in (False, True):
        eng = Engine(spec.nodes, precision=prec)
        if on_device:
            dev = torch.device("cuda:0")
            args = ([torch.from_numpy(w).to(dev) for w in W], [torch.from_numpy(b).to(dev) for b in B],
                    torch.from_numpy(P).to(dev), torch.from_numpy(np.ascontiguousarray(st, np.uint8)).to(dev))
        else:
            args = (W, B, P, st)
        res = eng.run(*args)
        V, faces = res.combine()
        out.append((res.masks.copy(), res.info.copy(), res.verts.copy(), V.copy(), faces.offsets.copy(),
                    faces.indices.copy()))
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("prec", ["fp64", "fp64ray", "fp32"])
def test_iso_flip_vm_vs_oracle(prec, oracle_mod):
    """A non-zero level (iso), inside-out orientation (flip) and a raised polygon cap (vm): every
    form gives the oracle's topology; the FP64 form its records bit for bit."""
    from analyticmesh_amd import configs
    from analyticmesh_amd.engine import march
    spec = configs.sphere(0)
    iso, vm = 0.05, 32
    pts = configs.seeds_fibonacci(spec, 8, iso=iso)
    st = configs.mask_of(spec, pts)
    _say(f"[iso/flip/vm] {prec}")
    g = march(spec, pts, st, precision=prec, vm=vm, iso=iso, flip=True)
    traversal = "bfs" if prec == "fp64" else "lifo"
    o = oracle_mod.run_oracle(spec, pts, st, traversal=traversal, vm=vm, iso=iso, flip=True)
    assert len(g.masks) > 100
    if prec == "fp64":
        np.testing.assert_array_equal(g.masks, o.masks)
        np.testing.assert_array_equal(g.info[:, :6], o.info[:, :6])
        np.testing.assert_array_equal(g.tabs, o.tabs)
        assert np.array_equal(g.verts.view(np.uint64), o.verts.view(np.uint64))
    else:
        a, b = g.records(), o.records()
        assert_same_topology(a, b)
        assert max_vertex_error(a, b) < (1e-10 if prec == "fp64ray" else 1e-5)
