"""Adjoint of mesh vertices w.r.t. weights: FP64 autograd reference and central differences."""
import sys
import time

import numpy as np
import pytest
import torch

from _adjoint_ref import canonical_vertices, unpack_bits, vertices_torch
from _common import workload

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name,seeds", [("sphere", 16), ("diff", 64)])
def test_adjoint_matches_autograd_and_fd(name, seeds):
    from analyticmesh_amd.engine import march
    spec, pts, st = workload(name, n_seeds=seeds)
    _say(f"[adjoint] {name} seeds={seeds}")
    res = march(spec, pts, st, precision="fp64")
    V, F = res.combine()
    rec = res.records()
    cv = canonical_vertices(rec)
    assert len(cv) == len(V)
    rng = np.random.Generator(np.random.PCG64(5))
    g = rng.normal(size=V.shape)
    shapes = [w.shape for w in spec.weights]
    t0 = time.perf_counter()
    dW, db = res.adjoint(g, shapes)
    t_adj = time.perf_counter() - t0

    regions = np.array([c[0] for c in cv])
    lo = torch.tensor([c[1] for c in cv])
    hi = torch.tensor([c[2] for c in cv])
    bits = torch.from_numpy(unpack_bits(rec.masks[regions], spec.L))
    params = []
    for w, b in zip(spec.weights, spec.biases):
        params += [torch.tensor(w, dtype=torch.float64, requires_grad=True),
                   torch.tensor(b, dtype=torch.float64, requires_grad=True)]
    vt = vertices_torch(params, spec.nodes, bits, lo, hi)
    assert np.abs(vt.detach().numpy() - V).max() < 1e-9
    loss = (vt * torch.from_numpy(g)).sum()
    grads = torch.autograd.grad(loss, params)
    worst = 0.0
    for i in range(len(dW)):
        for mine, ref in ((dW[i], grads[2 * i]), (db[i], grads[2 * i + 1])):
            ref = ref.numpy()
            rel = np.abs(mine - ref).max() / max(1e-30, np.abs(ref).max())
            worst = max(worst, rel)
    _say(f"  V={len(V)} adjoint {t_adj*1e3:.1f} ms, max rel err vs FP64 autograd {worst:.2e}")
    assert worst < 1e-9

    # central differences, FP64, h = 1e-6, regions held fixed
    h = 1e-6
    fd_worst = 0.0
    with torch.no_grad():
        for li in (0, len(spec.weights) - 2):
            W = params[2 * li]
            for (j, k) in [(0, 0), (W.shape[0] // 2, W.shape[1] - 1)]:
                old = W[j, k].item()
                W[j, k] = old + h
                lp = (vertices_torch(params, spec.nodes, bits, lo, hi) * torch.from_numpy(g)).sum().item()
                W[j, k] = old - h
                lm = (vertices_torch(params, spec.nodes, bits, lo, hi) * torch.from_numpy(g)).sum().item()
                W[j, k] = old
                fd = (lp - lm) / (2 * h)
                an = dW[li][j, k]
                fd_worst = max(fd_worst, abs(fd - an) / max(1e-12, abs(fd), abs(an)))
    _say(f"  max rel err vs central differences {fd_worst:.2e}")
    assert fd_worst < 1e-4


def test_autograd_function_backward():
    from analyticmesh_amd.diff import mesh_vertices
    spec, pts, st = workload("sphere", n_seeds=8)
    model = spec.to_mlp().double()
    V = mesh_vertices(model, pts, st)
    loss = (V ** 2).sum()
    loss.backward()
    assert all(torch.isfinite(l.weight.grad).all() for l in model.linears)
    assert model.linears[0].weight.grad.abs().sum() > 0


def test_adjoint_tiled_matches_per_region():
    """The tiled path (planes from the tiled chain, pull-back and weight
    gradients as GEMMs over all regions, on the FP64 matrix cores or as VALU
    fmas) against the one-kernel-per-region path (chain and pull-back in one
    kernel).  The sums run in different orders; the results agree to 1e-12 of
    each array's largest entry.  64 seeds: 4 chunks.  A second call gives the
    same gradients bit for bit (partial sums are added in a fixed order)."""
    from analyticmesh_amd.engine import march
    spec, pts, st = workload("diff", n_seeds=64)
    _say("[adjoint] tiled (MFMA, VALU) vs per-region, diff seeds=64")
    res = march(spec, pts, st, precision="fp64")
    V, _ = res.combine()
    g = np.random.Generator(np.random.PCG64(9)).normal(size=V.shape)
    shapes = [w.shape for w in spec.weights]
    runs = {}
    try:
        for name, tiled, mfma in (("mfma", 1, 1), ("valu", 1, 0), ("per-region", 0, 0)):
            res._eng.set_option("adjoint_tiled", tiled)
            res._eng.set_option("adjoint_mfma", mfma)
            runs[name] = res.adjoint(g, shapes)
    finally:
        res._eng.set_option("adjoint_tiled", 1)
        res._eng.set_option("adjoint_mfma", 1)
    again = res.adjoint(g, shapes)
    for a, b in zip(again[0] + again[1], runs["mfma"][0] + runs["mfma"][1]):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), "the adjoint is not deterministic"
    ref_W, ref_b = runs["per-region"]
    for name in ("mfma", "valu"):
        dW, db = runs[name]
        worst = 0.0
        for a, b in list(zip(dW, ref_W)) + list(zip(db, ref_b)):
            worst = max(worst, np.abs(a - b).max() / max(1e-30, np.abs(b).max()))
        _say(f"  {name}: V={len(V)} max rel diff vs per-region {worst:.2e}")
        assert worst < 1e-12
