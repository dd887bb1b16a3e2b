"""Autotuner for the launch shapes of the marching kernels.

Two shapes are chosen per network shape (widest hidden layer, depth):

* ``ray_waves``: waves per region of the FP64 ray walk (2, 4 or 8).  More waves
  split each region's constraint sweeps further but add block-wide
  reductions; the best count depends on the number of constraints.
* ``scan_blocks``: the grid of the candidate scan over the regions the ray
  walk leaves undecided (a block-stride loop over a count the device holds).

``tune()`` marches the given workload once per candidate, reads the kernel time
of the stage each shape affects from the engine's HIP-event counters, checks
that every candidate gives the same topology fingerprint (the shapes must not
change the result), and stores the fastest in ``tuning.json`` beside this file.
``Engine`` and ``BatchEngine`` apply the stored shape of the nearest network
shape unless the caller passes one.

    python -m analyticmesh_amd.autotune          # tunes the benchmark workloads
"""
import json
import os
import time

import numpy as np

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
CANDIDATES = {"ray_waves": (2, 4, 8), "scan_blocks": (512, 2048, 8192)}
DEFAULTS = {"ray_waves": 4, "scan_blocks": 2048}
STAT_RAY64_MS, STAT_WALK64_MS, STAT_RUN_MS = 28, 14, 0


def shape_key(nodes):
    hidden = list(nodes[1:-1])
    return f"w{max(hidden)}d{len(hidden)}"


def load_table(path=TABLE):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def lookup(nodes, path=TABLE):
    """Stored shapes for the network shape nearest to ``nodes`` (widest layer first, then depth);
    the defaults when the table is empty."""
    table = load_table(path)
    entries = {k: v for k, v in table.items() if k.startswith("w") and "best" in v}
    if not entries:
        return dict(DEFAULTS)
    hidden = list(nodes[1:-1])
    w, d = max(hidden), len(hidden)

    def dist(k):
        kw, kd = k[1:].split("d")
        return (abs(np.log2(int(kw) / w)), abs(int(kd) - d))

    best = min(entries, key=dist)
    out = dict(DEFAULTS)
    out.update(entries[best]["best"])
    return out


def _run(spec, pts, st, precision, opts, repeats):
    from .engine import Engine
    eng = Engine(spec.nodes, None, 0, precision=precision, **opts)
    W = [np.ascontiguousarray(w, dtype=np.float64) for w in spec.weights]
    B = [np.ascontiguousarray(b, dtype=np.float64) for b in spec.biases]
    best, fp = None, None
    for _ in range(repeats):
        res = eng.run(W, B, pts, st)
        s = np.array(res.stats)
        fp = res.fingerprint()
        best = s if best is None or s[STAT_RUN_MS] < best[STAT_RUN_MS] else best
    return best, fp


def tune(spec, pts, st, precision="fp64ray", repeats=3, write=True, path=TABLE, log=print):
    """Tune ray_waves, then scan_blocks, on one workload; returns the table entry."""
    key = shape_key(spec.nodes)
    times = {"ray_waves": {}, "scan_blocks": {}}
    ref_fp = None
    t0 = time.time()
    for w in CANDIDATES["ray_waves"]:
        s, fp = _run(spec, pts, st, precision, {"ray_waves": w, "scan_blocks": DEFAULTS["scan_blocks"]}, repeats)
        if ref_fp is None:
            ref_fp = fp
        if fp != ref_fp:
            raise RuntimeError(f"ray_waves={w} changed the topology: {fp} vs {ref_fp}")
        times["ray_waves"][w] = float(s[STAT_RAY64_MS])
        log(f"[autotune] {key} ray_waves={w}: ray walk {s[STAT_RAY64_MS]:.3f} ms, run {s[STAT_RUN_MS]:.3f} ms")
    rw = min(times["ray_waves"], key=times["ray_waves"].get)
    for sb in CANDIDATES["scan_blocks"]:
        s, fp = _run(spec, pts, st, precision, {"ray_waves": rw, "scan_blocks": sb}, repeats)
        if fp != ref_fp:
            raise RuntimeError(f"scan_blocks={sb} changed the topology")
        times["scan_blocks"][sb] = float(s[STAT_WALK64_MS])
        log(f"[autotune] {key} scan_blocks={sb}: scan {s[STAT_WALK64_MS]:.3f} ms, run {s[STAT_RUN_MS]:.3f} ms")
    sb = min(times["scan_blocks"], key=times["scan_blocks"].get)
    entry = {"best": {"ray_waves": rw, "scan_blocks": sb}, "kernel_ms": times, "workload": spec.name,
             "seeds": int(len(pts)), "regions": int(ref_fp["regions"]), "precision": precision,
             "seconds": round(time.time() - t0, 2)}
    if write:
        table = load_table(path)
        table[key] = entry
        with open(path, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
    return entry


def main():
    import argparse
    from . import configs
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", type=int, default=256)
    ap.add_argument("--repeats", type=int, default=2)
    ap.add_argument("--out", default=TABLE, help="table to update (default: the package's tuning.json)")
    args = ap.parse_args()
    # the benchmark net and one net of each width of the batch workload
    work = [configs.deep(0)]
    seen = set()
    for s in configs.batch64(0):
        k = shape_key(s.nodes)
        if k not in seen and s.nodes[1] in (32, 64, 128, 256):
            seen.add(k)
            work.append(s)
    for spec in work:
        pts = configs.seeds_fibonacci(spec, args.seeds if spec.L > 512 else 16)
        st = configs.mask_of(spec, pts)
        e = tune(spec, pts, st, repeats=args.repeats, path=args.out)
        print(json.dumps({shape_key(spec.nodes): e["best"], "workload": spec.name}), flush=True)


if __name__ == "__main__":
    main()
