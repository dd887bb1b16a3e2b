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

__all__ = ["NetSpec", "polytope", "geometric", "sphere", "deep", "diff", "batch64",
           "forward_fp64", "preact_fp64", "bisect_seed", "seeds_fibonacci", "mask_of"]


class NetSpec:
    """Plain ReLU MLP: ``weights[i]`` is (out, in) float32, ``biases[i]`` (out,)."""

    def __init__(self, nodes, weights, biases, name=""):
        self.nodes = list(nodes)
        self.weights = [np.ascontiguousarray(w, dtype=np.float32) for w in weights]
        self.biases = [np.ascontiguousarray(b, dtype=np.float32) for b in biases]
        self.name = name

    @property
    def L(self):
        return int(sum(self.nodes[1:-1]))

    @property
    def W(self):
        return (self.L + 31) // 32

    def to_mlp(self):
        import torch
        from .model import MLP
        m = MLP(self.nodes, [[0]] * (len(self.nodes) - 2), [], initialization=None, enable_print=False)
        with torch.no_grad():
            for lin, w, b in zip(m.linears, self.weights, self.biases):
                lin.weight.copy_(torch.from_numpy(w))
                lin.bias.copy_(torch.from_numpy(b))
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
    """Hidden pre-activations (concatenated) and the output, in float64."""
    x = np.atleast_2d(np.asarray(x, dtype=np.float64))
    pre = []
    h = x
    n = len(spec.weights)
    for i in range(n):
        z = h @ spec.weights[i].astype(np.float64).T + spec.biases[i].astype(np.float64)
        if i != n - 1:
            pre.append(z)
            h = np.maximum(z, 0.0)
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
