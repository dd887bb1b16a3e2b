"""Reduced-precision forms (FP32 MFMA, split-BF16 MFMA) against FP64.

The reduced forms must give the FP64 topology exactly (same visited regions,
faces and vertex keys) and vertices within 1e-5 relative of FP64."""
import sys

import numpy as np
import pytest

from _common import assert_same_topology, max_vertex_error, workload

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec,walk", [("fp32", "ray"), ("bf16x3", "ray"), ("fp32", "scan")])
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
    assert err < (1e-5 if name != "polytope" else 1e-4)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
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
    _say(f"  regions {len(a)} redo {int(g.stats[9])} max vertex rel err {err:.2e}")
    assert err < 1e-5
