"""64 independent MLPs in one job: each network's topology equals its own FP64 march."""
import sys

import numpy as np
import pytest

from analyticmesh_amd import configs

pytestmark = pytest.mark.gpu


def _seeds(spec, n):
    pts = configs.seeds_fibonacci(spec, n)
    return pts, configs.mask_of(spec, pts)


def test_batch64_topology():
    from analyticmesh_amd.engine import march_batch
    specs = configs.batch64(0)
    # keep the test short: the smallest 16 networks of the batch
    specs = sorted(specs, key=lambda s: s.L)[:16]
    seeds = [_seeds(s, 8) for s in specs]
    print(f"[batch] {len(specs)} nets", flush=True)
    fast, t_fast = march_batch(specs, seeds, precision="fp32", workers=8)
    ref, t_ref = march_batch(specs, seeds, precision="fp64", workers=8)
    for a, b, s in zip(fast, ref, specs):
        ra, rb = a.records(), b.records()
        assert ra.visited() == rb.visited(), s.name
        assert ra.faces() == rb.faces(), s.name
    nf = sum(r.records().face_count() for r in fast)
    print(f"  faces {nf}, fp32 {t_fast:.2f}s, fp64 {t_ref:.2f}s", flush=True)
