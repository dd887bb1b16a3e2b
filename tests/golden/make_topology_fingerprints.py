"""Writes topology_fingerprints.json: the oracle's (FP64, breadth-first) topology fingerprints
of the full-size workloads, for the GPU tests and the benchmark to compare against.

    python tests/golden/make_topology_fingerprints.py [--threads N]

Workloads: the 4x256 net (seed 0) with 1024 Fibonacci bisection seeds, and each of the 64
nets of the batch workload (seed 0) with 16 seeds.  A fingerprint holds the region, face and
vertex-use counts and order-independent hashes of the visited masks, the canonical faces and
the vertex keys (analyticmesh_amd.canonical.fingerprint).  The full batch takes about 20
minutes on 8 cores."""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from analyticmesh_amd import configs  # noqa: E402
from analyticmesh_amd.canonical import fingerprint  # noqa: E402
import oracle  # noqa: E402


def one(spec, n_seeds, threads):
    pts = configs.seeds_fibonacci(spec, n_seeds)
    st = configs.mask_of(spec, pts)
    t = time.time()
    r = oracle.run_oracle(spec, pts, st, precision="fp64", traversal="bfs", threads=threads)
    fp = fingerprint(r.records())
    fp["seeds"] = n_seeds
    fp["oracle_seconds"] = round(time.time() - t, 2)
    print(spec.name, fp, flush=True)
    return fp


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--out", default=os.path.join(HERE, "topology_fingerprints.json"))
    args = ap.parse_args()
    out = {"generator": "oracle FP64 breadth-first; analyticmesh_amd.canonical.fingerprint"}
    out["deep_1024"] = one(configs.deep(0), 1024, args.threads)
    out["batch64"] = [one(s, 16, args.threads) for s in configs.batch64(0)]
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
