"""The reference's native-module surface (``build.cuam``, backend/src/cuam.cpp:186-217)
over the MI355X engine: ``Init``, ``AnalyticMarching``, ``CombineMesh``,
``ExportMesh``, ``Destroy``.

Same argument names, meanings and checks as the reference: argument errors
raise ``RuntimeError`` (the reference's TORCH_CHECK), and calling out of order
prints the reference's message and returns (backend/src/cuam.cpp:63-67,107-111).
One process-wide context, as in the reference; create
:class:`analyticmesh_amd.engine.Engine` objects directly for several at once.

``float_type`` selects the arithmetic: ``"float64"`` runs the FP64 form (the
reference's default dtype); ``"float32"`` runs the FP32-MFMA form with FP64
re-decision; ``"bf16x3"`` runs the split-BF16 MFMA form (float32 tensors);
``"fp64ray"`` builds the FP64 planes and walks them with the checked FP64 ray
walk, leaving only undecided regions to the candidate scan (float64 tensors).
"""
import numpy as np
import torch

from .engine import Engine, MarchResult

_S = {"engine": None, "float_type": "", "nodes": None, "arc": None, "E": 0, "result": None}

_PRECISION = {"float64": "fp64", "float32": "fp32", "bf16x3": "bf16x3", "fp64ray": "fp64ray"}


def _check(cond, msg):
    if not cond:
        raise RuntimeError(msg)


def _check_input(t, name):
    _check(isinstance(t, torch.Tensor), f"{name} must be a tensor")
    _check(t.is_cuda, f"{name} must be a CUDA tensor")
    _check(t.is_contiguous(), f"{name} must be contiguous")


def Init(float_type, nodesnum, arc_table, num_extra_constraints):
    """backend/src/cuam.cpp:58-95."""
    if float_type not in _PRECISION:
        print("Error: `float_type` is either `float32` or `float64`!")
        return
    nodesnum = [int(n) for n in nodesnum]
    _check(nodesnum[0] == 3, "nodesnum.front() == 3")
    _check(nodesnum[-1] == 1, "nodesnum.back() == 1")
    _check(len(nodesnum) >= 3, "nodesnum.size() >= 3")
    _check(isinstance(arc_table, torch.Tensor) and not arc_table.is_cuda, "arc_table must be a CPU tensor")
    _check(arc_table.dim() == 2, "arc_table.dim() == 2")
    _check(arc_table.size(0) == len(nodesnum) - 2, "arc_table.size(0) == nodesnum.size() - 2")
    _check(arc_table.size(1) >= 1 and (arc_table.size(1) + 1) % 2 == 0, "arc_table must have 2k+1 columns")
    _check(arc_table.dtype == torch.int32, "arc_table.dtype() == torch::kInt32")
    for row in arc_table.tolist():
        _check(1 + 2 * row[0] <= arc_table.size(1), "arc_table row declares more skips than it holds")
    _check(int(num_extra_constraints) >= 0, "num_extra_constraints >= 0")
    if _S["engine"] is not None:
        Destroy()
    _S["engine"] = Engine(nodesnum, arc_table.tolist(), int(num_extra_constraints), precision=_PRECISION[float_type])
    _S["float_type"] = float_type
    _S["nodes"] = nodesnum
    _S["arc"] = arc_table.clone()
    _S["E"] = int(num_extra_constraints)


def AnalyticMarching(weights, biases, states, points, arc_tm, w_extra_constraints, b_extra_constraints, iso,
                     flip_insideout):
    """backend/src/cuam.cpp:97-184."""
    if _S["engine"] is None:
        print("Environment must be initialized first!")
        return
    ft = torch.float64 if _S["float_type"] in ("float64", "fp64ray") else torch.float32
    nodes = _S["nodes"]
    _check(len(weights) == len(biases), "weights.size() == biases.size()")
    _check(len(weights) == len(nodes) - 1, "fc_layers_num == nodesnum.size() - 1")
    L = 0
    for i, (w, b) in enumerate(zip(weights, biases)):
        _check_input(w, f"weights[{i}]")
        _check_input(b, f"biases[{i}]")
        _check(w.dim() == 2 and b.dim() == 1, "weights 2-D, biases 1-D")
        _check(w.size(0) == b.size(0), "weights[i].size(0) == biases[i].size(0)")
        if i:
            _check(w.size(1) == weights[i - 1].size(0), "weights[i].size(1) == weights[i-1].size(0)")
        _check(w.dtype == ft and b.dtype == ft, "weights/biases dtype must match float_type")
        _check(nodes[i + 1] == w.size(0), "nodesnum[i+1] == weights[i].size(0)")
        L += w.size(0)
    _check(weights[0].size(1) == 3 and weights[-1].size(0) == 1, "input width 3, output width 1")
    L -= 1
    _check_input(states, "states")
    _check(states.dim() == 2 and states.size(0) >= 1 and states.size(1) == L, "states must be [N >= 1, L]")
    _check(states.dtype == torch.bool, "states.dtype() == torch::kBool")
    _check_input(points, "points")
    _check(points.dim() == 2 and points.size(0) == states.size(0) and points.size(1) == 3, "points must be [N, 3]")
    _check(points.dtype == ft, "points dtype must match float_type")
    for t in arc_tm:
        _check_input(t, "arc_tm")
        _check(t.dim() == 2 and t.dtype == ft, "arc_tm entries are 2-D float_type tensors")
    _check_input(w_extra_constraints, "w_extra_constraints")
    _check_input(b_extra_constraints, "b_extra_constraints")
    _check(w_extra_constraints.dim() == 2 and w_extra_constraints.size(1) == 3, "w_extra_constraints [E, 3]")
    _check(b_extra_constraints.dim() == 1 and b_extra_constraints.size(0) == w_extra_constraints.size(0),
           "b_extra_constraints [E]")
    _check(w_extra_constraints.size(0) == _S["E"], "number of extra constraints differs from Init")
    torch.cuda.synchronize()
    _S["result"] = _S["engine"].run(list(weights), list(biases), points, states, tms=list(arc_tm),
                                    extra_w=w_extra_constraints, extra_b=b_extra_constraints, iso=float(iso),
                                    flip=bool(flip_insideout))


def CombineMesh(scale, center):
    """backend/src/cuam_kernel.cu:185-200."""
    if _S["result"] is None:
        print("Error: AnalyticMarching must be called first!")
        return
    _check(len(center) == 3, "center must have 3 entries")
    return _S["result"].combine(float(scale), [float(c) for c in center])


def ExportMesh(file_path, is_polymesh, is_float32):
    """backend/src/cuam_kernel.cu:205-246."""
    if _S["result"] is None:
        print("Error: CombineMesh must be called first!")
        return
    _S["result"].export_ply(file_path, bool(is_polymesh), bool(is_float32))


def Destroy():
    """backend/src/cuam_kernel.cu:153-179."""
    _S.update(engine=None, float_type="", nodes=None, arc=None, E=0, result=None)


def last_result():
    """The engine's records of the last AnalyticMarching call (None before it)."""
    return _S["result"]
