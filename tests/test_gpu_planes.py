"""Face-kernel planes of the reduced-precision forms against the oracle's FP64 planes.

The checked walk certifies a predicate only when its margin exceeds
eps * max|plane coefficient| * (1 + |v|_1); these tests pin, on the sphere and
on the 4x256 net, that each form's plane error stays below the eps it is
walked with (FP32 MFMA 8e-5, split-BF16 2e-3).  A wrong MFMA operand layout, a dropped bias column or a
mask applied to the wrong layer gives errors of order 1 and fails here."""
import numpy as np
import pytest

from _common import workload

pytestmark = pytest.mark.gpu

EPS = {"fp32": 8e-5, "bf16x3": 2e-3}


def _oracle_planes(spec, masks):
    import oracle
    run = oracle.OracleRun(spec)
    P, E = [], []
    for m in masks:
        a, eq = run.planes(m)
        P.append(a)
        E.append(eq)
    return np.array(P), np.array(E)


@pytest.mark.parametrize("name,n_seeds,n_masks", [("sphere", 16, 1000), ("deep", 64, 2000)])
@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_plane_error_within_walk_eps(prec, name, n_seeds, n_masks):
    from analyticmesh_amd.engine import Engine, march
    spec, pts, st = workload(name, n_seeds=n_seeds)
    print(f"[planes {prec}] {name}", flush=True)
    ref = march(spec, pts, st, precision="fp64", max_regions=n_masks + 500)
    masks = ref.masks[:n_masks]
    P64, E64 = _oracle_planes(spec, masks)              # (n, C, 4), (n, 4)
    eng = Engine(spec.nodes, precision=prec)
    W = [np.ascontiguousarray(w, np.float32) for w in spec.weights]
    B = [np.ascontiguousarray(b, np.float32) for b in spec.biases]
    eng.run(W, B, pts.astype(np.float32), st)
    P, E = eng.planes(masks)
    P = np.transpose(P, (0, 2, 1)).astype(np.float64)
    rel = np.abs(P - P64).max(axis=2) / np.abs(P64).max(axis=2)
    rel_eq = np.abs(E.astype(np.float64) - E64).max(axis=1) / np.abs(E64).max(axis=1)
    print(f"  plane rel err median {np.median(rel):.2e} max {rel.max():.2e}; eq max {rel_eq.max():.2e}", flush=True)
    assert rel.max() < EPS[prec]
    assert rel_eq.max() < EPS[prec]
