"""Face-kernel planes of the reduced-precision forms against the oracle's FP64 planes.

FP32 form: the checked walk bounds each plane's coefficient error by
lam * sigma (lam = 3, plus a few ulps of its largest coefficient), sigma being
the face kernel's own error estimate for that plane.  The test pins, on the
sphere and on the 4x256 net, that the measured error of every plane
coefficient stays below sigma (a third of the bound), and that hidden layer
1's planes, whose estimate is 0, are exact.  Split-BF16 form: one bound for
every plane, 2e-3 of its largest coefficient (the estimate is not computed).
A wrong MFMA operand layout, a dropped bias column, a mask applied to the wrong
layer or an estimate that misses a term gives errors far above these."""
import numpy as np
import pytest

from _common import workload

pytestmark = pytest.mark.gpu

EPS_BF16 = 2e-3


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
def test_plane_error_within_estimate(prec, name, n_seeds, n_masks):
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
    P, E, S = eng.planes(masks, with_sigma=True)
    P = np.transpose(P, (0, 2, 1)).astype(np.float64)
    err = np.abs(P - P64).max(axis=2)                   # (n, C): largest coefficient error per plane
    err_eq = np.abs(E.astype(np.float64) - E64).max(axis=1)
    n1 = spec.nodes[1]
    rel = err[:, n1:] / np.abs(P64[:, n1:]).max(axis=2)
    rel_eq = err_eq / np.abs(E64).max(axis=1)
    assert np.all(err[:, :n1] == 0), "hidden layer 1's planes must be exact"
    if prec == "bf16x3":
        print(f"  relative error median {np.median(rel):.2e} max {rel.max():.2e}; face plane max {rel_eq.max():.2e}",
              flush=True)
        assert np.all(S == 0)
        assert rel.max() < EPS_BF16 and rel_eq.max() < EPS_BF16
        return
    sig = S[:, :-1].astype(np.float64)
    sig_eq = S[:, -1].astype(np.float64)
    assert np.all(sig[:, :n1] == 0)
    assert np.all(sig[:, n1:] > 0) and np.all(sig_eq > 0)
    ratio = err[:, n1:] / sig[:, n1:]
    ratio_eq = err_eq / sig_eq
    print(f"  error / sigma: median {np.median(ratio):.3f} max {ratio.max():.3f}; face plane max "
          f"{ratio_eq.max():.3f}; relative error max {rel.max():.2e}, sigma relative median "
          f"{np.median(sig[:, n1:] / np.abs(P64[:, n1:]).max(axis=2)):.2e}", flush=True)
    assert ratio.max() < 1.0
    assert ratio_eq.max() < 1.0
