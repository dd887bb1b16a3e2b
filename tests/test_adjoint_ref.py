"""The FP64 torch vertex reference that the adjoint GPU tests differentiate,
pinned on the CPU against the oracle.

Each canonical vertex is rebuilt with torch ops from its region's mask alone
(masked chain, then the 3x3 solve for its two tight neurons and f = iso). It
must equal the oracle's vertex, which the oracle finds by its own walk.
Dropping the mask, using a transposed weight or taking the wrong neuron pair
moves the vertex, so each of those mistakes fails here."""
import numpy as np
import torch

import oracle
from _adjoint_ref import canonical_vertices, unpack_bits, vertices_torch
from _common import workload


def _first_positions(rec, cv):
    """Tab position k of each canonical vertex (region r, pair lo, hi)."""
    pos = []
    for r, lo, hi in cv:
        nv = int(rec.info[r, 2])
        k = next(k for k in range(nv) if int(rec.vpairs[r, k, 0]) == lo and int(rec.vpairs[r, k, 1]) == hi)
        pos.append(k)
    return pos


def _params(spec):
    out = []
    for w, b in zip(spec.weights, spec.biases):
        out += [torch.tensor(w, dtype=torch.float64, requires_grad=True),
                torch.tensor(b, dtype=torch.float64, requires_grad=True)]
    return out


def test_vertices_torch_match_oracle():
    spec, pts, st = workload("sphere", n_seeds=4)
    rec = oracle.run_oracle(spec, pts, st, traversal="bfs")
    rec.L = spec.L
    cv = canonical_vertices(rec)
    assert len(cv) > 100
    regions = np.array([c[0] for c in cv])
    lo = torch.tensor([c[1] for c in cv])
    hi = torch.tensor([c[2] for c in cv])
    bits = torch.from_numpy(unpack_bits(rec.masks[regions], spec.L))
    vt = vertices_torch(_params(spec), spec.nodes, bits, lo, hi).detach().numpy()
    ref = np.array([rec.verts[r, k] for (r, _, _), k in zip(cv, _first_positions(rec, cv))])
    scale = max(1.0, np.abs(ref).max())
    assert np.abs(vt - ref).max() < 1e-10 * scale


def test_vertices_torch_gradient_matches_central_differences():
    """Autograd through the reference against central differences of the
    same function (FP64, h = 1e-6), for a weight of the first and of the last
    hidden layer."""
    spec, pts, st = workload("sphere", n_seeds=2)
    rec = oracle.run_oracle(spec, pts, st, traversal="bfs")
    rec.L = spec.L
    cv = canonical_vertices(rec)
    regions = np.array([c[0] for c in cv])
    lo = torch.tensor([c[1] for c in cv])
    hi = torch.tensor([c[2] for c in cv])
    bits = torch.from_numpy(unpack_bits(rec.masks[regions], spec.L))
    params = _params(spec)
    g = torch.from_numpy(np.random.Generator(np.random.PCG64(3)).normal(size=(len(cv), 3)))
    loss = (vertices_torch(params, spec.nodes, bits, lo, hi) * g).sum()
    grads = torch.autograd.grad(loss, params)
    h = 1e-6
    with torch.no_grad():
        for li in (0, len(spec.weights) - 2):
            W = params[2 * li]
            j, k = W.shape[0] // 2, W.shape[1] - 1
            old = W[j, k].item()
            W[j, k] = old + h
            lp = (vertices_torch(params, spec.nodes, bits, lo, hi) * g).sum().item()
            W[j, k] = old - h
            lm = (vertices_torch(params, spec.nodes, bits, lo, hi) * g).sum().item()
            W[j, k] = old
            fd = (lp - lm) / (2 * h)
            an = grads[2 * li][j, k].item()
            assert abs(fd - an) <= 1e-5 * max(1e-8, abs(fd), abs(an))
