"""Deterministic workloads for analytic marching.

Each generator draws from ``numpy.random.Generator(PCG64(seed))`` in a fixed
order and returns float32 parameters, so a workload is reproducible on any
machine.  The architectures follow the reference's geometric initialisation
(backend/model.py:63-70) and its polytope example (examples/2_polytope.py:14-36).

Draw order for the geometric nets: hidden weights layer by layer
(N(0, 2/fan_out)), then the output weights (N(sqrt(pi)/sqrt(fan_in), 1e-8^2)),
then the perturbation of every hidden weight matrix, layer by layer.
"""
import math

import numpy as np

__all__ = ["NetSpec", "polytope", "geometric", "sphere", "deep", "diff", "batch64", "skip_mlp2",
           "skip_custom2", "wide", "forward_fp64", "preact_fp64", "bisect_seed", "seeds_fibonacci", "mask_of"]


class NetSpec:
    """ReLU MLP: ``weights[i]`` is (out, in) float32, ``biases[i]`` (out,).

    Optional additive skips follow the reference's MLP (backend/model.py:23-31,
    101-123): ``arc_table[i] = [count, from_0, tm_0, from_1, tm_1, ...]`` adds
    ``tms[tm_j] @ node[from_j]`` (node 0 = the input, node k > 0 = hidden layer
    k after its ReLU) to the output of linear layer i + 1 before its ReLU; a
    ``tms`` entry of None is the identity.  Skips into the output layer are
    ignored, as the reference's engine ignores them (backend/inc/process.h:42)."""

    def __init__(self, nodes, weights, biases, name="", arc_table=None, tms=None):
        self.nodes = list(nodes)
        self.weights = [np.ascontiguousarray(w, dtype=np.float32) for w in weights]
        self.biases = [np.ascontiguousarray(b, dtype=np.float32) for b in biases]
        self.name = name
        self.arc_table = None if arc_table is None else [list(map(int, r)) for r in arc_table]
        self.tms = [None if t is None else np.ascontiguousarray(t, dtype=np.float32) for t in (tms or [])]

    @property
    def has_skips(self):
        return self.arc_table is not None and any(r and r[0] > 0 for r in self.arc_table)

    def skips_into(self, li):
        """(from_node, transform or None) pairs added to linear layer li's output (li >= 1).

        Skips into the output layer are ignored (backend/inc/process.h:42)."""
        if self.arc_table is None or li < 1 or li >= len(self.nodes) - 2:
            return []
        row = self.arc_table[li - 1]
        return [(row[1 + 2 * j], self.tms[row[2 + 2 * j]]) for j in range(row[0] if row else 0)]

    def tm_arrays(self):
        """Transform matrices as the engine takes them: (0, 0) arrays for identities."""
        return [np.zeros((0, 0), dtype=np.float32) if t is None else t for t in self.tms]

    @property
    def L(self):
        return int(sum(self.nodes[1:-1]))

    @property
    def W(self):
        return (self.L + 31) // 32

    def to_mlp(self):
        import torch
        from .model import MLP
        arc = self.arc_table if self.arc_table is not None else [[0]] * (len(self.nodes) - 2)
        shapes = [[0, 0] if t is None else list(t.shape) for t in self.tms]
        m = MLP(self.nodes, arc, shapes, initialization=None, enable_print=False)
        with torch.no_grad():
            for lin, w, b in zip(m.linears, self.weights, self.biases):
                lin.weight.copy_(torch.from_numpy(w))
                lin.bias.copy_(torch.from_numpy(b))
            for tm, t in zip(m.tms, self.tms):
                if t is not None:
                    tm.weight.copy_(torch.from_numpy(t))
        return m

    def gemm_flops_per_region(self):
        """8 * sum_{i>=1} n_{i+1} n_i: 2 flops per MAC x 4 columns (x, y, z, bias)."""
        return 8 * sum(self.nodes[i + 1] * self.nodes[i] for i in range(1, len(self.nodes) - 1))


_POLY_W0 = np.array([[1, 1, 1], [-1, -1, -1], [0, 1, 1], [0, -1, -1], [1, 0, 1], [-1, 0, -1],
                     [1, 1, 0], [-1, -1, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0],
                     [0, 0, 1], [0, 0, -1]], dtype=np.float64)


def polytope(seed=0, noise=1e-7):
    """3->14->1 polytope; every parameter gets N(0, noise^2) (examples/2_polytope.py:14-36)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    w0 = _POLY_W0.copy()
    b0 = np.zeros(14)
    w1 = np.ones((1, 14))
    b1 = np.array([-2.0])
    ps = [w0, b0, w1, b1]
    if noise:
        ps = [p + rng.normal(0.0, noise, p.shape) for p in ps]
    return NetSpec([3, 14, 1], [ps[0], ps[2]], [ps[1], ps[3]], name=f"polytope(seed={seed})")


def geometric(nodes, radius, sigma, seed, name=""):
    rng = np.random.Generator(np.random.PCG64(seed))
    n = len(nodes) - 1
    weights, biases = [], []
    for i in range(n - 1):
        weights.append(rng.normal(0.0, math.sqrt(2.0 / nodes[i + 1]), (nodes[i + 1], nodes[i])))
        biases.append(np.zeros(nodes[i + 1]))
    weights.append(rng.normal(math.sqrt(math.pi) / math.sqrt(nodes[-2]), 1e-8, (1, nodes[-2])))
    biases.append(np.array([-radius]))
    for i in range(n - 1):
        weights[i] = weights[i] + rng.normal(0.0, sigma, weights[i].shape)
    return NetSpec(nodes, weights, biases, name=name)


def sphere(seed=0):
    """3->64->64->1, r=0.5, hidden perturbation 0.02."""
    return geometric([3, 64, 64, 1], 0.5, 0.02, seed, name=f"sphere(seed={seed})")


def deep(seed=0):
    """3->256x4->1, r=0.6, hidden perturbation 0.05."""
    return geometric([3, 256, 256, 256, 256, 1], 0.6, 0.05, seed, name=f"deep(seed={seed})")


def diff(seed=0):
    """3->128x4->1, r=0.5, hidden perturbation 0.03."""
    return geometric([3, 128, 128, 128, 128, 1], 0.5, 0.03, seed, name=f"diff(seed={seed})")


def _skipnet(nodes, arc_table, tm_shapes, radius, sigma, seed, tm_scale, name):
    """Geometric-init net plus transform matrices U(-s, s)/sqrt(in) (torch's Linear default, scaled)."""
    base = geometric(nodes, radius, sigma, seed)
    rng = np.random.Generator(np.random.PCG64(seed + 7919))
    tms = []
    for h, w in tm_shapes:
        if h == 0 and w == 0:
            tms.append(None)
        else:
            bound = tm_scale / math.sqrt(w)
            tms.append(rng.uniform(-bound, bound, (h, w)))
    return NetSpec(nodes, base.weights, base.biases, name=name, arc_table=arc_table, tms=tms)


def skip_mlp2(seed=0):
    """The reference's test_mlp_2 architecture (backend/test/test_onnx_io.py:145-151): five hidden
    layers of 80, a transform skip from the input, an identity skip from hidden 1, an 80x80
    transform skip from hidden 3, and an output skip from the input (ignored, Q1)."""
    return _skipnet([3, 80, 80, 80, 80, 80, 1], [[1, 0, 0], [1, 1, 1], [0], [1, 3, 2], [1, 0, 3]],
                    [[80, 3], [0, 0], [80, 80], [1, 3]], 0.5, 0.02, seed, 0.5, name=f"skip_mlp2(seed={seed})")


def skip_custom2(seed=0):
    """The reference's test_custom_2 architecture (backend/test/test_onnx_io.py:99-141): widths
    20, 50, 20, 80, 20; an identity skip from hidden 1 into hidden 3, an 80x20 transform skip from
    hidden 1 into hidden 4, and an output skip from hidden 2 (ignored, Q1)."""
    return _skipnet([3, 20, 50, 20, 80, 20, 1], [[0], [1, 1, 0], [1, 1, 1], [0], [1, 2, 2]],
                    [[0, 0], [80, 20], [1, 50]], 0.5, 0.02, seed, 0.5, name=f"skip_custom2(seed={seed})")


def wide(seed=0):
    """3->320->64->1: a first layer wider than the tiled chain's 256 rows."""
    return geometric([3, 320, 64, 1], 0.5, 0.02, seed, name=f"wide(seed={seed})")


def batch64(seed=0, count=64):
    """64 nets: depth U{2..6}, width U{32,64,128,256}, r~U(0.3,0.8), perturbation 0.03."""
    rng = np.random.Generator(np.random.PCG64(seed))
    specs = []
    for k in range(count):
        depth = int(rng.integers(2, 7))
        width = int(rng.choice([32, 64, 128, 256]))
        radius = float(rng.uniform(0.3, 0.8))
        sub_seed = int(rng.integers(0, 2**31 - 1))
        specs.append(geometric([3] + [width] * depth + [1], radius, 0.03, sub_seed,
                               name=f"batch64[{k}](d={depth},w={width},r={radius:.3f})"))
    return specs


# ----------------------------------------------------------------------------
# FP64 forward pass and seeds
# ----------------------------------------------------------------------------
def preact_fp64(spec, x):
    """Hidden pre-activations (concatenated) and the output, in float64 (skips included,
    output-layer skips ignored as the engine ignores them)."""
    x = np.atleast_2d(np.asarray(x, dtype=np.float64))
    pre = []
    node = [x]          # node k's output: the input, then each hidden layer after its ReLU
    h = x
    n = len(spec.weights)
    for i in range(n):
        z = h @ spec.weights[i].astype(np.float64).T + spec.biases[i].astype(np.float64)
        for frm, t in (spec.skips_into(i) if hasattr(spec, "skips_into") else []):
            src = node[frm]
            if t is None:
                z = z + (src if src.shape[1] == z.shape[1] else np.pad(src, ((0, 0), (0, z.shape[1] - src.shape[1]))))
            else:
                z = z + src @ t.astype(np.float64).T
        if i != n - 1:
            pre.append(z)
            h = np.maximum(z, 0.0)
            node.append(h)
        else:
            out = z[:, 0]
    return np.concatenate(pre, axis=1), out


def forward_fp64(spec, x):
    return preact_fp64(spec, x)[1]


def mask_of(spec, x):
    """Activation state (post-ReLU > 0) at points x as bool[N, L] (backend/main.py:408-411)."""
    pre, _ = preact_fp64(spec, x)
    return pre > 0


def bisect_seed(spec, direction, iso=0.0, iters=200):
    """Bisect f along a ray from the origin; f(0) < iso is required."""
    d = np.asarray(direction, dtype=np.float64)
    d = d / np.linalg.norm(d)
    f = lambda t: forward_fp64(spec, t * d)[0] - iso
    lo, hi = 0.0, 1.0
    if f(lo) >= 0:
        raise ValueError("f(origin) must be below the iso level")
    while f(hi) <= 0:
        hi *= 2.0
        if hi > 1e6:
            raise ValueError("no sign change along the ray")
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        if mid == lo or mid == hi:
            break
        if f(mid) < 0:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi) * d


def seeds_fibonacci(spec, count, iso=0.0):
    """Seeds by bisection along ``count`` Fibonacci-sphere directions."""
    pts = []
    g = math.pi * (3.0 - math.sqrt(5.0))
    for i in range(count):
        z = 1.0 - 2.0 * (i + 0.5) / count
        r = math.sqrt(max(0.0, 1.0 - z * z))
        pts.append(bisect_seed(spec, (r * math.cos(g * i), r * math.sin(g * i), z), iso))
    return np.array(pts)
