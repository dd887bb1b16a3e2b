"""Launch shapes chosen by the autotuner: every candidate gives the same records, and the tuner
picks one of them and stores it for the engine to apply."""
import json
import sys

import numpy as np
import pytest

from _common import workload

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name,n_seeds,cap", [("sphere", 8, None), ("deep", 16, 30000)])
def test_ray_waves_do_not_change_records(name, n_seeds, cap):
    """The FP64 ray walk with 2, 4 and 8 waves per region writes the same records bit for bit."""
    from analyticmesh_amd.engine import march
    spec, pts, st = workload(name, n_seeds=n_seeds)
    _say(f"[ray_waves] {name}")
    runs = []
    for w in (2, 4, 8):
        from analyticmesh_amd.engine import Engine
        eng = Engine(spec.nodes, precision="fp64ray", ray_waves=w, scan_blocks=512 if w == 8 else None)
        if cap:
            eng.set_option("max_regions", cap)
        r = eng.run([x.astype(np.float64) for x in spec.weights], [x.astype(np.float64) for x in spec.biases],
                    pts, st)
        runs.append((r.masks.copy(), r.info.copy(), r.tabs.copy(), r.vpairs.copy(), r.verts.copy()))
    for other in runs[1:]:
        for a, b in zip(runs[0], other):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8))


def test_autotuner_picks_and_stores(tmp_path):
    from analyticmesh_amd import autotune
    spec, pts, st = workload("sphere", n_seeds=8)
    path = str(tmp_path / "tuning.json")
    _say("[autotune] sphere")
    e = autotune.tune(spec, pts, st, repeats=1, path=path, log=_say)
    assert e["best"]["ray_waves"] in autotune.CANDIDATES["ray_waves"]
    assert e["best"]["scan_blocks"] in autotune.CANDIDATES["scan_blocks"]
    table = json.load(open(path))
    assert table[autotune.shape_key(spec.nodes)]["best"] == e["best"]
    assert autotune.lookup(spec.nodes, path=path) == e["best"]


@pytest.mark.parametrize("prec", ["fp64", "fp64ray", "fp32", "bf16x3"])
def test_graph_levels_give_the_same_records(prec):
    """Each level captured as a hipGraph (updated in place from level to level) and launched as one:
    the same records as the launch-by-launch run, bit for bit."""
    from analyticmesh_amd.engine import Engine
    spec, pts, st = workload("sphere", n_seeds=8)
    W = [x.astype(np.float64) for x in spec.weights]
    B = [x.astype(np.float64) for x in spec.biases]
    _say(f"[graph levels] {prec}")
    out = []
    for g in (False, True):
        r = Engine(spec.nodes, precision=prec, graph_levels=g).run(W, B, pts, st)
        out.append((r.masks.copy(), r.info.copy(), r.tabs.copy(), r.vpairs.copy(), r.verts.copy()))
        if g:
            assert int(r.stats[36]) == int(r.stats[6]) > 10
    for a, b in zip(*out):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))


def test_graph_levels_batch():
    """The batch engine with graph-captured levels gives every network the same records."""
    from analyticmesh_amd import configs
    from analyticmesh_amd.engine import BatchEngine
    specs = sorted(configs.batch64(0), key=lambda s: s.L)[:12]
    seeds = []
    for s in specs:
        p = configs.seeds_fibonacci(s, 8)
        seeds.append((p, configs.mask_of(s, p)))
    _say("[graph levels] batch of 12")
    a = BatchEngine(specs, precision="fp64ray").run(seeds)
    b = BatchEngine(specs, precision="fp64ray", graph_levels=True).run(seeds)
    assert int(b.stats[36]) == int(b.stats[6])
    for i in range(len(specs)):
        ra, rb = a.regions(i), b.regions(i)
        for k in ("masks", "info", "tabs", "vpairs", "verts"):
            assert np.array_equal(ra[k].view(np.uint8), rb[k].view(np.uint8)), (specs[i].name, k)
