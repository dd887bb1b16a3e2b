#!/usr/bin/env python3
"""Mesh faces/s of the marching engine on the 4x256 ReLU MLP (one MI355X).

Workload: the deep network 3->256->256->256->256->1 (geometric init, r=0.6,
N(0, 0.05^2) hidden-weight perturbation, PCG64 seed 0, float32 parameters)
with 1024 seeds found by FP64 bisection along Fibonacci-sphere directions.
One step = one complete marching run: seeds and weights in, every region's
face walked, vertices deduplicated and the indexed mesh copied to the host
(``am_run`` + ``am_combine``).  Seed search and PLY writing are outside the
step, as in the reference's ``am_time``.

Prints one JSON line.  ``value`` is the plain FP32 form's faces/s (the
drop-in); the split-BF16 form and the FP64 form are measured in the same
process and reported beside it, with a topology comparison.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"fp32": 157.3, "bf16x3": 2516.6 / 3.0, "fp64": 78.6, "fp64ray": 78.6}
GEMM_KERNEL = {"fp32": "k_gemm<0>", "bf16x3": "k_gemm<1>", "fp64ray": "k_chain_tile_mfma", "fp64": "k_chain_tile_fp64"}
# the FP64 chain of the certified forms runs on the matrix cores; the bit-exact FP64 form's on VALU
CHAIN_KERNEL = {"fp32": "k_chain_tile_mfma", "bf16x3": "k_chain_tile_mfma", "fp64ray": "k_chain_tile_mfma",
                "fp64": "k_chain_tile_fp64"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(n_seeds):
    from analyticmesh_amd import configs
    spec = configs.deep(0)
    pts = configs.seeds_fibonacci(spec, n_seeds)
    states = configs.mask_of(spec, pts)
    return spec, pts, states


def run_form(spec, pts, states, form, steps, warmup):
    """form: a precision name, with "+graph" for graph-captured levels."""
    import torch
    from analyticmesh_amd.engine import Engine
    precision, _, opt = form.partition("+")
    eng = Engine(spec.nodes, None, 0, precision=precision, vm=20, graph_levels=(opt == "graph"))
    dev = torch.device("cuda:0")
    W = [torch.from_numpy(w).to(dev) for w in spec.weights]
    B = [torch.from_numpy(b).to(dev) for b in spec.biases]
    P = torch.from_numpy(pts.astype(np.float32)).to(dev)
    S = torch.from_numpy(states.astype(np.uint8)).to(dev)
    res = None
    for _ in range(warmup):
        res = eng.run(W, B, P, S)
        res.combine()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    for _ in range(steps):
        res = eng.run(W, B, P, S)
        res.combine()
        stats.append(np.array(res.stats))
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    st = np.mean(stats, axis=0)
    V, _ = res.combine()                      # outside the timed region: the mesh for the checks below
    fp = res.fingerprint()
    return dict(sec_per_step=dt, faces=int(fp["faces"]), regions=int(st[7]), stats=st, res=res, eng=eng, V=V,
                fingerprint=fp)


FP_KEYS = ("regions", "faces", "vertex_uses", "visited", "faces_hash", "vertex_keys")


def same_fp(a, b):
    return a is not None and b is not None and all(a[k] == b[k] for k in FP_KEYS)


def golden_fingerprints():
    path = os.path.join(ROOT, "tests", "golden", "topology_fingerprints.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def residual_fp64(spec, V):
    """max |f(v)| over the mesh vertices, f evaluated in FP64 (torch on the GPU)."""
    import torch
    x = torch.from_numpy(np.ascontiguousarray(V, dtype=np.float64)).cuda()
    h = x
    for i, (w, b) in enumerate(zip(spec.weights, spec.biases)):
        h = h @ torch.from_numpy(w.astype(np.float64)).cuda().T + torch.from_numpy(b.astype(np.float64)).cuda()
        if i + 1 < len(spec.weights):
            h = torch.relu(h)
    return float(h.abs().max().item())


def vertex_error_vs(res, ref, L):
    """Largest vertex difference (relative to |v|) between two runs with the same regions, matched by
    region mask and, within a region, by the vertex's pair of tight constraints."""
    idx = {m.tobytes(): i for i, m in enumerate(ref.masks)}
    perm = np.fromiter((idx[m.tobytes()] for m in res.masks), dtype=np.int64, count=len(res.masks))
    big = np.int64(1) << 40

    def keyed(verts, pairs, nv):
        p = np.sort(pairs.astype(np.int64), axis=2)
        key = p[:, :, 0] * big + p[:, :, 1]
        slot = np.arange(key.shape[1])[None, :]
        key = np.where(slot < nv[:, None], key, np.iinfo(np.int64).max)
        o = np.argsort(key, axis=1)
        return np.take_along_axis(verts, o[:, :, None], axis=1), np.take_along_axis(key, o, axis=1)

    nv = res.info[:, 2].astype(np.int64)
    va, ka = keyed(res.verts, res.vpairs, nv)
    vb, kb = keyed(ref.verts[perm], ref.vpairs[perm], ref.info[perm, 2].astype(np.int64))
    valid = (np.arange(va.shape[1])[None, :] < nv[:, None]) & (nv[:, None] >= 3)
    if not np.array_equal(np.where(valid, ka, 0), np.where(valid, kb, 0)):
        return None
    d = np.abs(va - vb).max(axis=2) / np.maximum(np.abs(vb).max(axis=2), 1e-12)
    return float(np.where(valid, d, 0.0).max())


def traversal_bytes(res, Wd, L):
    """Algorithmic HBM bytes of the visited-set stage: R*4W (pop mask) + E*(4W+4) (one probe per emitted
    hidden-plane edge) + R*(8W+16) (insert key + frontier record)."""
    R = len(res.masks)
    ntab = res.info[:, 3].astype(np.int64)
    pos = np.arange(res.tabs.shape[1])[None, :]
    t = res.tabs
    E = int(((pos < (ntab - 1)[:, None]) & (t >= 0) & (t < L)).sum())
    return R * 4 * Wd + E * (4 * Wd + 4) + R * (8 * Wd + 16), E


def copy_calibration(mb):
    """A plain device copy of known size, for scaling FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = mb * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    return {"bytes_read": n * 4, "bytes_written": n * 4, "launches": 3}


def profile_value(name, kernel):
    """Per-region figure of `kernel` from a committed counter summary under profiles/ (its
    machine-read lines are exactly "kernel value"), or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 2 and parts[0] == kernel:
            try:
                return float(parts[1])
            except ValueError:
                return None
    return None


def traffic_per_launch(kernel, regions, launches):
    """HBM bytes per launch: the counter summary's scaled bytes per region times this run's
    regions per launch."""
    b = profile_value("hbm_traffic.txt", kernel)
    return None if b is None else round(b * regions / max(1.0, launches), 1)


VALU_PEAK_G = 256 * 4 * 2.4 / 4   # G wave64 VALU instructions/s: one per SIMD every 4 cycles at 2.4 GHz


def walk_roofline(kernel, ms, regions, launches):
    """VALU-issue roofline of a walk kernel: instructions per region from the committed counter
    summary, times the regions this run walked, over this run's kernel time."""
    per = profile_value("walk_counters.txt", kernel)
    if per is None or ms <= 0 or regions <= 0:
        return None
    ach = per * regions / (ms / 1e3) / 1e9
    return {"bound": "valu issue", "kernel": kernel, "achieved": round(ach, 1), "peak": VALU_PEAK_G,
            "unit": "G VALU wave-instructions/s", "frac": round(ach / VALU_PEAK_G, 4),
            "traffic": traffic_per_launch(kernel, regions, launches), "kernel_ms": round(ms, 3),
            "launches": int(launches), "regions": int(regions), "valu_per_region": per}


def batch64_extra(n_nets, seeds_per_net, golden, steps=2):
    """The 64-net batch workload: the batch engine (every network's levels through one launch of
    each kernel; FP64 ray walk, without and with graph-captured levels) and, for comparison, one
    context per network driven by a pool of host threads (FP32 form).  Per-net topology against
    the oracle's committed fingerprints."""
    from analyticmesh_amd import configs
    from analyticmesh_amd.engine import BatchEngine, march_batch
    specs = configs.batch64(0)[:n_nets]
    seeds = []
    for s in specs:
        p = configs.seeds_fibonacci(s, seeds_per_net)
        seeds.append((p, configs.mask_of(s, p)))
    ref = (golden or {}).get("batch64")
    out = {"workload": f"batch64(seed 0): {len(specs)} nets, depth U{{2..6}}, width U{{32,64,128,256}}, "
                       f"{seeds_per_net} seeds each", "nets": len(specs)}

    def check(fps):
        return sum(1 for a, b in zip(fps, ref) if same_fp(a, b)) if ref else None

    for name, graph in (("batch_engine_fp64ray", False), ("batch_engine_fp64ray_graph", True)):
        eng = BatchEngine(specs, precision="fp64ray", graph_levels=graph)
        eng.run(seeds[:]).combine()                       # warm up
        ts = []
        for _ in range(steps):
            t0 = time.perf_counter()
            res = eng.run(seeds)
            res.combine()
            ts.append(time.perf_counter() - t0)
        fps = [res.fingerprint(i) for i in range(len(specs))]
        faces = sum(int(x["faces"]) for x in fps)
        dt = min(ts)
        st = res.stats
        out[name] = {"faces": faces, "seconds": round(dt, 4), "faces_per_s": round(faces / dt, 1),
                     "levels": int(st[6]), "host_waits": int(st[35]), "graph_launches": int(st[36]),
                     "chain_ms": round(float(st[13]), 3), "ray_walk_ms": round(float(st[28]), 3),
                     "insert_ms": round(float(st[2]), 3), "combine_ms": round(float(res.combine_ms), 3),
                     "nets_identical_to_oracle": check(fps)}
        del res, eng
    march_batch(specs[:2], seeds[:2], precision="fp32", workers=2, combine=False)   # warm up
    res, dt = march_batch(specs, seeds, precision="fp32", combine=True)
    fps = [r.fingerprint() for r in res]
    faces = sum(int(x["faces"]) for x in fps)
    out["thread_pool_fp32"] = {"faces": faces, "seconds": round(dt, 4), "faces_per_s": round(faces / dt, 1),
                               "contexts": len(specs), "nets_identical_to_oracle": check(fps)}
    return out


def adjoint_extra(reps=3):
    from analyticmesh_amd import configs
    from analyticmesh_amd.engine import march
    spec = configs.diff(0)
    pts = configs.seeds_fibonacci(spec, 64)
    st = configs.mask_of(spec, pts)
    res = march(spec, pts, st, precision="fp64")
    V, _ = res.combine()
    g = np.random.Generator(np.random.PCG64(0)).normal(size=V.shape)
    shapes = [w.shape for w in spec.weights]
    res.adjoint(g, shapes)                      # warm up
    ts, ks = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        res.adjoint(g, shapes)
        ts.append(time.perf_counter() - t0)
        ks.append(float(res._eng.stats()[12]))   # device part: kernels from first launch to the stream's end
    return {"workload": "diff 3->128x4->1, r=0.5, sigma=0.03, 64 seeds", "vertices": int(len(V)),
            "regions": int(len(res.masks)), "backward_ms": round(min(ts) * 1e3, 3),
            "backward_device_ms": round(min(ks), 3), "forward_run_ms": round(float(res.stats[0]), 3)}


def cpu_baseline(spec, pts, states, max_regions):
    """The oracle (FP64, the reference's LIFO order) timed on the host cores; returns the baseline
    entry and the run's topology fingerprint (None when the run was cut short)."""
    import oracle
    from analyticmesh_amd.canonical import fingerprint
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    r = oracle.run_oracle(spec, pts, states, precision="fp64", traversal="lifo", max_regions=max_regions,
                          threads=threads)
    faces = int((r.info[:, 2] >= 3).sum())
    fp = None if r.truncated else fingerprint(r.records())
    return dict(value=faces / r.seconds, unit="faces/s", cores=threads, kind="port",
                sample=(f"{'first ' if r.truncated else 'all '}{len(r.masks)} regions of the LIFO march of the "
                        f"same 4x256 net and seeds, FP64 (the reference default dtype), {r.seconds:.1f} s")), fp


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--seeds", type=int, default=1024)
    ap.add_argument("--cpu-regions", type=int, default=-1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="fp32,fp32+graph,bf16x3,fp64ray,fp64ray+graph,fp64")
    ap.add_argument("--calib-copy-mb", type=int, default=0, help="plain copy of this size for counter scaling")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-MLP batch")
    ap.add_argument("--adjoint-only", action="store_true", help="measure only the adjoint (one JSON line)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding is designed but not built",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)

    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_extra()}))
        return

    spec, pts, states = make_workload(args.seeds)
    flops_region = spec.gemm_flops_per_region()
    forms = [f for f in args.forms.split(",") if f]
    out = {}
    for f in forms:
        log(f"[bench] form {f}: warmup {args.warmup}, steps {args.steps}")
        try:
            out[f] = run_form(spec, pts, states, f, args.steps, args.warmup)
            log(f"[bench]   {out[f]['faces']} faces, {out[f]['sec_per_step']*1e3:.1f} ms/step")
        except Exception as e:  # a form that is not built yet is reported, not fatal
            log(f"[bench]   form {f} failed: {e}")
            out[f] = None

    head = "fp32" if out.get("fp32") else next(f for f in forms if out.get(f))
    h = out[head]
    st = h["stats"]
    # plane kernels of the head form: its own GEMM (reduced forms) and the FP64
    # chain (every region in every form: the reduced forms re-solve the
    # vertices of the regions they decide from FP64 planes and re-decide the
    # others on them); the roofline line is the one that took longer
    kernels = []
    if head in ("fp32", "bf16x3"):
        kernels.append(dict(kernel=GEMM_KERNEL[head], ms=float(st[3]), regions=h["regions"], peak=PEAK_TFLOPS[head],
                            launches=max(1.0, float(st[11]))))
        if float(st[13]) > 0:
            # the FP64 chain runs on the regions the reduced walk left to FP64 (undecided, or
            # decided with vertices re-solved from FP64 planes)
            kernels.append(dict(kernel=CHAIN_KERNEL[head], ms=float(st[13]), regions=int(st[9]) + int(st[33]),
                                peak=PEAK_TFLOPS["fp64"], launches=max(1.0, float(st[31]))))
    else:
        kernels.append(dict(kernel=CHAIN_KERNEL[head.split("+")[0]], ms=float(st[13]), regions=h["regions"],
                            peak=PEAK_TFLOPS["fp64"], launches=max(1.0, float(st[31]))))
    lines = []
    for k in kernels:
        ach = flops_region * k["regions"] / (k["ms"] / 1e3) / 1e12 if k["ms"] > 0 else 0.0
        lines.append({"bound": "fp64 valu" if k["kernel"] == "k_chain_tile_fp64" else "mfma", "kernel": k["kernel"],
                      "achieved": round(ach, 3), "peak": k["peak"], "unit": "TFLOP/s",
                      "frac": round(ach / k["peak"], 5),
                      "traffic": traffic_per_launch(k["kernel"], k["regions"], k["launches"]),
                      "kernel_ms": round(k["ms"], 3), "launches": int(k["launches"]),
                      "avg_launch_ms": round(k["ms"] / k["launches"], 4),
                      "regions": k["regions"], "flops_per_region": flops_region})
    lines.sort(key=lambda r: -r["kernel_ms"])
    roofline = lines[0]
    # walk kernels of the head form: the reduced-precision walk over every region, the FP64 ray
    # walk over the regions it left undecided (every region in the fp64ray form)
    walks = []
    nlev = max(1.0, float(st[6]))
    if head in ("fp32", "bf16x3"):
        walks.append(walk_roofline("k_walk_ray_f32", float(st[4]), h["regions"], nlev))
        walks.append(walk_roofline("k_walk_ray64<4>", float(st[28]), int(st[9]), float(st[31])))
    elif head == "fp64ray":
        walks.append(walk_roofline("k_walk_ray64<4>", float(st[28]), h["regions"], float(st[31])))
    walks = [w for w in walks if w]

    topo = {}
    vert_err = {}
    if out.get("fp64"):
        log("[bench] topology and vertex error against FP64")
        ref_res = out["fp64"]["res"]
        for f in forms:
            if f != "fp64" and out.get(f):
                same = same_fp(out[f]["fingerprint"], out["fp64"]["fingerprint"])
                topo[f] = "identical" if same else "DIFFERENT"
                if same and "+" not in f:
                    vert_err[f] = vertex_error_vs(out[f]["res"], ref_res, spec.L)
    golden = golden_fingerprints()
    gold = (golden or {}).get("deep_1024") if args.seeds == 1024 else None
    topo_oracle = {f: ("identical" if same_fp(out[f]["fingerprint"], gold) else "DIFFERENT")
                   for f in forms if out.get(f) and gold}
    resid = {f: residual_fp64(spec, out[f]["V"]) for f in forms if out.get(f)}
    # traversal (visited set) stage against HBM bandwidth
    tbytes, n_edges = traversal_bytes(h["res"], (spec.L + 31) // 32, spec.L)
    ins_ms = float(st[2])
    traversal = {"bound": "hbm", "stage": "visited-set insert + compaction", "bytes": int(tbytes),
                 "emitted_edges": n_edges, "ms": round(ins_ms, 3),
                 "achieved": round(tbytes / (ins_ms / 1e3) / 1e9, 1) if ins_ms > 0 else None,
                 "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": round(tbytes / (ins_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4) if ins_ms > 0 else None,
                 "traffic": (None if profile_value("hbm_traffic.txt", "k_insert") is None or
                             profile_value("hbm_traffic.txt", "k_scan_commit") is None else
                             round((profile_value("hbm_traffic.txt", "k_insert") +
                                    profile_value("hbm_traffic.txt", "k_scan_commit")) * h["regions"], 1))}

    line = {
        "metric": "mesh faces/s, 4x256 ReLU MLP: plain FP32 vs split-BF16 MFMA, same topology",
        "value": round(h["faces"] / h["sec_per_step"], 1),
        "unit": "faces/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(h["sec_per_step"] * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32", "bf16x3": "bf16x3", "fp64": "f64", "fp64ray": "f64"}[head.split("+")[0]],
        "data": "synthetic (seeded geometric-init weights, bisection seeds)",
        "config": {"workload": "deep 3->256x4->1 ReLU MLP, r=0.6, sigma=0.05, seed 0, 1024 seeds",
                   "faces": h["faces"], "regions": h["regions"], "form": head},
        "forms": {f: (None if not out.get(f) else {
            "faces_per_s": round(out[f]["faces"] / out[f]["sec_per_step"], 1),
            "vertices": int(len(out[f]["V"])),
            "vertices_per_s": round(len(out[f]["V"]) / out[f]["sec_per_step"], 1),
            "max_abs_f_at_vertices_fp64": resid.get(f),
            "max_vertex_err_vs_fp64": vert_err.get(f),
            "ms_per_step": round(out[f]["sec_per_step"] * 1e3, 3),
            "faces": out[f]["faces"],
            "run_ms": round(float(out[f]["stats"][0]), 3),
            "region_kernels_ms": round(float(out[f]["stats"][1]), 3),
            "gemm_ms": round(float(out[f]["stats"][3]), 3),
            "walk_ms": round(float(out[f]["stats"][4]), 3),
            "fp64_redo_ms": round(float(out[f]["stats"][5]), 3),
            "fp64_redo_regions": int(out[f]["stats"][9]),
            "fp64_vertex_ms": round(float(out[f]["stats"][32]), 3),
            "fp64_vertex_regions": int(out[f]["stats"][33]),
            "fp64_chain_ms": round(float(out[f]["stats"][13]), 3),
            "fp64_ray_walk_ms": round(float(out[f]["stats"][28]), 3),
            "fp64_scan_walk_ms": round(float(out[f]["stats"][14]), 3),
            "fp64_ray_undecided": int(out[f]["stats"][30]),
            "fp64_scanned_regions": int(out[f]["stats"][29]),
            "insert_ms": round(float(out[f]["stats"][2]), 3),
            "newton_ms": round(float(out[f]["stats"][37]), 3),
            "newton_vertices": int(out[f]["stats"][38]),
            "combine_ms": round(float(getattr(out[f]["res"], "combine_ms", 0.0)), 3),
            "levels": int(out[f]["stats"][6]),
        }) for f in forms},
        "topology_vs_fp64": topo,
        "topology_vs_oracle_golden": topo_oracle,
        "roofline": roofline,
        "roofline_other": lines[1:] + walks,
        "traversal": traversal,
    }
    log("[bench] adjoint (backward) on the diff 3->128x4->1 net")
    line["adjoint"] = adjoint_extra()
    if not args.no_batch:
        log("[bench] 64-MLP batch")
        line["batch64"] = batch64_extra(64, 16, golden)
    if args.calib_copy_mb:
        line["copy_calibration"] = copy_calibration(args.calib_copy_mb)
    if not args.no_cpu:
        log(f"[bench] cpu baseline: oracle on the first {args.cpu_regions} regions")
        line["cpu_baseline"], cpu_fp = cpu_baseline(spec, pts, states, args.cpu_regions)
        if cpu_fp is not None:
            line["topology_vs_cpu_baseline_run"] = {
                f: ("identical" if same_fp(out[f]["fingerprint"], cpu_fp) else "DIFFERENT")
                for f in forms if out.get(f)}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
