"""Forms with a checked walk (FP32 MFMA, split-BF16 MFMA, FP64 planes with the
ray walk) against FP64.

Every form must give the FP64 topology exactly (same visited regions, faces
and vertex keys).  Vertices (VERT_TOL, relative to |v|): the FP64 ray form
solves them from the FP64 planes, so only the order of the 3x3 solve's rows
can differ from the scan's.  The reduced forms solve each vertex in FP64 from
their own planes and keep it where the face kernel's error estimate puts it
within 1e-5; the other regions' vertices are re-solved from FP64 planes."""
import sys

import numpy as np
import pytest

from _common import assert_same_topology, max_vertex_error, workload

pytestmark = pytest.mark.gpu

VERT_TOL = {"fp64ray": 1e-12, "fp32": 1e-5, "bf16x3": 1e-5}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec,walk", [("fp64ray", "ray"), ("fp32", "ray"), ("bf16x3", "ray"), ("fp32", "scan")])
@pytest.mark.parametrize("name", ["polytope", "sphere"])
def test_fast_topology_vs_oracle(name, prec, walk):
    import oracle
    from analyticmesh_amd.engine import march
    spec, pts, st = workload(name)
    _say(f"[{prec}/{walk}] {name}")
    g = march(spec, pts, st, precision=prec, walk=walk)
    o = oracle.run_oracle(spec, pts, st, traversal="lifo")
    a, b = g.records(), o.records()
    assert_same_topology(a, b)
    err = max_vertex_error(a, b)
    _say(f"  redo {int(g.stats[9])}/{len(g.masks)} max vertex rel err {err:.2e}")
    # the polytope's planes come in near-parallel pairs (1e-7 noise): FP64 itself is only good to 1e-7 there
    assert err < (VERT_TOL[prec] if name != "polytope" else 1e-4)


@pytest.mark.parametrize("prec", ["fp64ray", "fp32", "bf16x3"])
def test_fast_deep_vs_fp64(prec):
    from analyticmesh_amd.engine import march
    spec, pts, st = workload("deep", n_seeds=64)
    _say(f"[{prec}] deep x64 seeds")
    ref = march(spec, pts, st, precision="fp64", max_regions=60000)
    g = march(spec, pts, st, precision=prec, max_regions=60000)
    a, b = g.records(), ref.records()
    # both stop after the same number of BFS levels; compare what both finished
    assert_same_topology(a, b)
    err = max_vertex_error(a, b)
    _say(f"  regions {len(a)} redo {int(g.stats[9])} re-solved {int(g.stats[33])} Newton {int(g.stats[38])} "
         f"max vertex rel err {err:.2e}")
    if prec != "fp64ray":
        assert int(g.stats[9]) + int(g.stats[33]) <= len(a)
    assert err < VERT_TOL[prec]


@pytest.mark.parametrize("mode", ["newton_where_needed", "newton", "estimate", "fp64"])
def test_fp32_vertex_modes(mode):
    """Each way the FP32 form gets its vertices to 1e-5: the same topology as the FP64 form and
    vertices within 1e-5 of its (relative to |v|); the Newton step runs on at most every vertex."""
    from analyticmesh_amd.engine import march
    spec, pts, st = workload("deep", n_seeds=64)
    _say(f"[fp32 reduced_vertices={mode}] deep x64 seeds")
    ref = march(spec, pts, st, precision="fp64", max_regions=30000)
    g = march(spec, pts, st, precision="fp32", max_regions=30000, reduced_vertices=mode)
    a, b = g.records(), ref.records()
    assert_same_topology(a, b)
    err = max_vertex_error(a, b)
    newton = int(g.stats[38])
    _say(f"  regions {len(a)} redo {int(g.stats[9])} re-solved {int(g.stats[33])} Newton {newton} "
         f"max vertex rel err {err:.2e}")
    assert err < 1e-5
    if mode in ("estimate", "fp64"):
        assert newton == 0
    else:
        assert 0 < newton <= int(np.sum(a.info[:, 2]))
    if mode == "fp64":
        assert 0 < int(g.stats[9]) + int(g.stats[33]) <= len(a)
