"""Independent FP64 checks for the oracle (test helper).

Planes are obtained without the masked-chain recursion: with the activation
pattern frozen, every pre-activation is an affine function of x, so its plane
is read off by evaluating the frozen network at 0, e_x, e_y, e_z.  The face of
a region is then the polygon {f = iso} clipped against every constraint
half-space (Sutherland-Hodgman on a large square in the f = iso plane).
"""
import itertools

import numpy as np


def frozen_preacts(spec, mask_bits, x):
    """Pre-activations of the network with ReLUs replaced by the fixed mask.

    Additive skips (``spec.skips_into``) add ``T @ node_k`` or ``node_k`` to a
    layer's pre-activation, node 0 being the input and node k > 0 hidden layer
    k after its (frozen) ReLU; skips into the output layer are ignored, as the
    reference's engine ignores them (backend/inc/process.h:42)."""
    x = np.atleast_2d(np.asarray(x, dtype=np.float64))
    h = x
    out, pre = None, []
    node = [x]
    off = 0
    n = len(spec.weights)
    skips = getattr(spec, "skips_into", lambda i: [])
    for i in range(n):
        z = h @ spec.weights[i].astype(np.float64).T + spec.biases[i].astype(np.float64)
        for frm, t in skips(i):
            src = node[frm]
            if t is None:
                k = min(src.shape[1], z.shape[1])
                z[:, :k] += src[:, :k]
            else:
                z = z + src @ t.astype(np.float64).T
        if i != n - 1:
            w = z.shape[1]
            m = np.asarray(mask_bits[off:off + w], dtype=np.float64)
            off += w
            pre.append(z)
            h = z * m
            node.append(h)
        else:
            out = z[:, 0]
    return np.concatenate(pre, axis=1), out


def region_planes(spec, mask_bits, iso=0.0):
    """(C, 4) signed constraint planes (a.x + c <= 0 inside) and the eq plane."""
    pts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float64)
    pre, out = frozen_preacts(spec, mask_bits, pts)
    a = (pre[1:] - pre[0]).T
    c = pre[0]
    sign = np.where(np.asarray(mask_bits, dtype=bool), -1.0, 1.0)
    A = np.concatenate([a * sign[:, None], (c * sign)[:, None]], axis=1)
    eq = np.array([out[1] - out[0], out[2] - out[0], out[3] - out[0], out[0] - iso])
    return A, eq


def clip_polygon(A, eq, half=1e3):
    """Clip a (2*half)-wide square of the plane eq against a.x + c <= 0."""
    n = eq[:3]
    nn = n @ n
    p0 = -eq[3] * n / nn
    u = np.cross(n, [1.0, 0, 0] if abs(n[0]) < 0.9 * np.sqrt(nn) else [0, 1.0, 0])
    u /= np.linalg.norm(u)
    v = np.cross(n, u) / np.sqrt(nn)
    poly = [p0 + half * (s * u + t * v) for s, t in ((-1, -1), (1, -1), (1, 1), (-1, 1))]
    for g in range(len(A)):
        a, c = A[g, :3], A[g, 3]
        if not poly:
            break
        vals = [a @ q + c for q in poly]
        new = []
        for k in range(len(poly)):
            q1, q2 = poly[k], poly[(k + 1) % len(poly)]
            s1, s2 = vals[k], vals[(k + 1) % len(poly)]
            if s1 <= 0:
                new.append(q1)
            if (s1 < 0 < s2) or (s2 < 0 < s1):
                t = s1 / (s1 - s2)
                new.append(q1 + t * (q2 - q1))
        poly = new
    return np.array(poly).reshape(-1, 3)


def polygon_area(P):
    if len(P) < 3:
        return 0.0
    s = np.zeros(3)
    for k in range(len(P)):
        s += np.cross(P[k], P[(k + 1) % len(P)])
    return 0.5 * np.linalg.norm(s)


def dedup_points(P, tol):
    out = []
    for p in P:
        if all(np.abs(p - q).sum() > tol for q in out):
            out.append(p)
    return np.array(out).reshape(-1, 3)


def unpack_mask(words, L):
    words = np.asarray(words, dtype=np.uint32)
    return np.array([(int(words[g >> 5]) >> (g & 31)) & 1 for g in range(L)], dtype=np.uint8)


def pack_mask(bits):
    L = len(bits)
    words = np.zeros((L + 31) // 32, dtype=np.uint32)
    for g, b in enumerate(bits):
        if b:
            words[g >> 5] |= np.uint32(1 << (g & 31))
    return words


def enumerate_regions(spec, iso=0.0, min_area=0.0):
    """All masks (tiny L only) whose clipped face has area > min_area."""
    L = spec.L
    found = {}
    for bits in itertools.product((0, 1), repeat=L):
        A, eq = region_planes(spec, bits, iso)
        P = clip_polygon(A, eq)
        if len(P) >= 3 and polygon_area(P) > min_area:
            found[pack_mask(bits).tobytes()] = P
    return found
