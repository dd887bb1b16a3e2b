"""Full-size workloads against the oracle through topology fingerprints.

The oracle's fingerprints of the complete marches (the 4x256 net with 1024 seeds, 985,814
regions; each of the 64 batch nets with 16 seeds, 25.3M regions in all) are committed in
tests/golden/topology_fingerprints.json (written by make_topology_fingerprints.py beside it).
The engine computes the same fingerprint on the device (am_topology_fingerprint); its rows and
hash are checked here against the numpy restatement in analyticmesh_amd.canonical first."""
import json
import os
import sys

import numpy as np
import pytest

from _common import workload
from analyticmesh_amd import configs
from analyticmesh_amd.canonical import fingerprint

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "topology_fingerprints.json")
KEYS = ("regions", "faces", "vertex_uses", "visited", "faces_hash", "vertex_keys")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _golden():
    with open(GOLDEN) as f:
        return json.load(f)


def _same(fp, ref):
    return {k: fp[k] for k in KEYS} == {k: ref[k] for k in KEYS}


def test_device_fingerprint_equals_numpy_fingerprint():
    """The device kernel hashes the same rows the same way as canonical.fingerprint (sphere with
    extra planes, so that pairs past L occur, and a capped 4x256 run with 32 mask words)."""
    from analyticmesh_amd.engine import march
    spec, pts, st = workload("sphere", n_seeds=8)
    ew = np.array([[1.0 + 1e-8, 1e-8, 1e-8]])
    eb = np.array([-0.2])
    _say("[fingerprint kernel vs numpy]")
    for prec in ("fp64", "fp32"):
        r = march(spec, pts, st, precision=prec, extra_w=ew, extra_b=eb)
        assert _same(r.fingerprint(), fingerprint(r.records())), prec
    spec, pts, st = workload("deep", n_seeds=16)
    r = march(spec, pts, st, precision="fp64", max_regions=20000)
    assert _same(r.fingerprint(), fingerprint(r.records()))
    # a changed record changes the fingerprint (the numpy side, planted on the CPU)
    rec = r.records()
    rec.vpairs = rec.vpairs.copy()
    rec.vpairs[7, 0, 0] += 1
    assert not _same(r.fingerprint(), fingerprint(rec))


@pytest.mark.parametrize("prec", ["fp64", "fp64ray", "fp32", "bf16x3"])
def test_deep_full_size_vs_oracle(prec):
    """The benchmark workload in full (985,814 regions): every form has the oracle's topology."""
    from analyticmesh_amd.engine import march
    ref = _golden()["deep_1024"]
    spec = configs.deep(0)
    pts = configs.seeds_fibonacci(spec, ref["seeds"])
    st = configs.mask_of(spec, pts)
    _say(f"[deep full size] {prec}")
    r = march(spec, pts, st, precision=prec)
    fp = r.fingerprint()
    _say(f"  {fp}")
    assert _same(fp, ref)


def test_batch64_full_size_vs_oracle():
    """All 64 nets of the batch workload marched completely in one job, FP64 and FP32: every
    net's topology equals the oracle's (25.3M regions in all)."""
    from analyticmesh_amd.engine import march_batch
    ref = _golden()["batch64"]
    specs = configs.batch64(0)
    assert len(ref) == len(specs)
    seeds = []
    for s in specs:
        p = configs.seeds_fibonacci(s, 16)
        seeds.append((p, configs.mask_of(s, p)))
    for prec in ("fp64", "fp32"):
        _say(f"[batch64 full size] {prec}")
        res, dt = march_batch(specs, seeds, precision=prec, combine=False)
        bad = [s.name for s, r, g in zip(specs, res, ref) if not _same(r.fingerprint(), g)]
        nf = sum(int(g["faces"]) for g in ref)
        _say(f"  {nf} faces in {dt:.2f} s; {len(bad)} nets differ {bad[:4]}")
        assert not bad
        del res
