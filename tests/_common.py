"""Shared helpers for the tests: workloads with explicit seeds."""
import numpy as np

from analyticmesh_amd import configs


def workload(name, seed=0, n_seeds=1):
    spec = getattr(configs, name)(seed)
    if n_seeds == 1:
        pts = configs.bisect_seed(spec, (1.0, 0.0, 0.0))[None]
    else:
        pts = configs.seeds_fibonacci(spec, n_seeds)
    states = configs.mask_of(spec, pts)
    return spec, pts, states


def assert_same_topology(a, b):
    """a, b: RegionRecords.  Visited set, faces and vertex keys must agree exactly."""
    va, vb = a.visited(), b.visited()
    assert va == vb, f"visited sets differ: {len(va - vb)} only in first, {len(vb - va)} only in second"
    fa, fb = a.faces(), b.faces()
    assert fa.keys() == fb.keys()
    diff = [k for k in fa if fa[k] != fb[k]]
    assert not diff, f"{len(diff)} faces differ"
    assert a.vertex_keys() == b.vertex_keys()


def max_vertex_error(a, b):
    """Largest |coordinate difference| between the same vertex key in a and b (relative to |v|)."""
    ca, cb = a.vertex_coords(), b.vertex_coords()
    worst = 0.0
    for k, lst in ca.items():
        ref = cb[k][0]
        for v in lst:
            worst = max(worst, float(np.max(np.abs(np.asarray(v) - ref)) / max(1e-12, np.max(np.abs(ref)))))
    return worst
