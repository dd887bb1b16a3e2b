"""FP64 torch references for the vertex adjoint (test helper).

Each canonical mesh vertex (first occurrence in visiting order) is rebuilt
from its region's masked chain with torch ops, v = -A^{-1} b for its planes
(lo, hi, f = iso); autograd then gives dL/dtheta with the regions fixed."""
import numpy as np
import torch

from analyticmesh_amd.canonical import vertex_key


def canonical_vertices(rec):
    """[(region, lo, hi)] in the engine's mesh vertex order (first occurrence)."""
    seen = set()
    out = []
    for r in range(len(rec.masks)):
        nv = int(rec.info[r, 2])
        if nv < 3:
            continue
        for k in range(nv):
            a, b = int(rec.vpairs[r, k, 0]), int(rec.vpairs[r, k, 1])
            key = vertex_key(rec.masks[r], a, b, rec.L)
            if key in seen:
                continue
            seen.add(key)
            out.append((r, a, b))
    return out


def vertices_torch(params, nodes, masks_bits, lo, hi, iso=0.0):
    """params: [W0, b0, W1, b1, ...] float64 tensors; masks_bits: (V, L) float64 0/1."""
    W = params[0::2]
    B = params[1::2]
    V = masks_bits.shape[0]
    P = torch.cat([W[0], B[0][:, None]], dim=1)[None].expand(V, -1, -1)   # (V, n1, 4)
    planes = [P]
    off = 0
    eq = None
    for li in range(1, len(W)):
        nin = nodes[li]
        m = masks_bits[:, off:off + nin]
        off += nin
        X = P * m[:, :, None]
        P = torch.einsum("jk,vkq->vjq", W[li], X)
        P = P + torch.cat([torch.zeros_like(B[li])[:, None].expand(-1, 3), B[li][:, None]], dim=1)[None]
        if li < len(W) - 1:
            planes.append(P)
        else:
            eq = P[:, 0, :]
    allp = torch.cat(planes, dim=1)                                        # (V, L, 4)
    idx = torch.arange(V)
    p1 = allp[idx, lo]
    p2 = allp[idx, hi]
    e = eq.clone()
    e = torch.cat([e[:, :3], (e[:, 3] - iso)[:, None]], dim=1)
    A = torch.stack([p1[:, :3], p2[:, :3], e[:, :3]], dim=1)
    b = torch.stack([p1[:, 3], p2[:, 3], e[:, 3]], dim=1)
    return -torch.linalg.solve(A, b[:, :, None])[:, :, 0]


def unpack_bits(masks, L):
    m = np.asarray(masks, dtype=np.uint32)
    bits = ((m[:, :, None] >> np.arange(32, dtype=np.uint32)[None, None, :]) & 1).reshape(len(m), -1)
    return bits[:, :L].astype(np.float64)
