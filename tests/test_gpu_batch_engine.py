"""Many networks in one job (am_batch_*): one set of kernel launches per breadth-first level for
all networks.  Per network the records must equal the oracle's breadth-first records bit for bit
(FP64 form), and every network's topology the oracle's at full size (committed fingerprints)."""
import json
import os
import sys

import numpy as np
import pytest

from analyticmesh_amd import configs

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "topology_fingerprints.json")
KEYS = ("regions", "faces", "vertex_uses", "visited", "faces_hash", "vertex_keys")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _seeds(specs, n):
    out = []
    for s in specs:
        p = configs.seeds_fibonacci(s, n)
        out.append((p, configs.mask_of(s, p)))
    return out


def test_batch_engine_records_vs_oracle():
    """The 24 narrowest nets of the batch workload (widths 32 and 64), complete marches: FP64
    records of every network bit for bit equal to its own oracle march; fp64ray the same topology."""
    import oracle
    from analyticmesh_amd.canonical import same_topology
    from analyticmesh_amd.engine import BatchEngine
    specs = sorted(configs.batch64(0), key=lambda s: (s.nodes[1], s.L))[:24]
    seeds = _seeds(specs, 16)
    _say(f"[batch engine vs oracle] {len(specs)} nets")
    res = BatchEngine(specs, precision="fp64").run(seeds)
    ray = BatchEngine(specs, precision="fp64ray").run(seeds)
    for i, (s, (p, st)) in enumerate(zip(specs, seeds)):
        o = oracle.run_oracle(s, p, st, traversal="bfs")
        g = res.regions(i)
        assert len(g["masks"]) == len(o.masks), s.name
        np.testing.assert_array_equal(g["masks"], o.masks, err_msg=s.name)
        np.testing.assert_array_equal(g["info"][:, :6], o.info[:, :6], err_msg=s.name)
        np.testing.assert_array_equal(g["tabs"], o.tabs, err_msg=s.name)
        np.testing.assert_array_equal(g["vpairs"], o.vpairs, err_msg=s.name)
        np.testing.assert_array_equal(g["points"], o.points, err_msg=s.name)
        assert np.array_equal(g["verts"].view(np.uint64), o.verts.view(np.uint64)), s.name
        ok, why = same_topology(ray.records(i), o.records())
        assert ok, f"{s.name} fp64ray: {why}"
    nf = sum(res.fingerprint(i)["faces"] for i in range(len(specs)))
    _say(f"  {nf} faces; levels {int(res.stats[6])}, host waits {int(res.stats[35])}")


@pytest.mark.parametrize("prec", ["fp64ray", "fp64"])
def test_batch_engine_full_size_vs_oracle(prec):
    """All 64 nets of the batch workload (16 seeds each, 25.3M regions) in one job: every network's
    topology fingerprint equals the oracle's; then one combine meshes them all."""
    from analyticmesh_amd.engine import BatchEngine
    with open(GOLDEN) as f:
        ref = json.load(f)["batch64"]
    specs = configs.batch64(0)
    seeds = _seeds(specs, 16)
    _say(f"[batch engine full size] {prec}")
    res = BatchEngine(specs, precision=prec).run(seeds)
    bad = [s.name for i, (s, g) in enumerate(zip(specs, ref))
           if {k: res.fingerprint(i)[k] for k in KEYS} != {k: g[k] for k in KEYS}]
    _say(f"  run {res.run_ms:.1f} ms, levels {int(res.stats[6])}, {len(bad)} nets differ {bad[:4]}")
    assert not bad
    V, F, NI = res.combine()
    assert F == sum(g["faces"] for g in ref) and NI == sum(g["vertex_uses"] for g in ref)
