"""``AnalyticMarching(model, ...)``: seeds -> activation states -> engine -> PLY.

Same signature, defaults and returned timing dict as the reference's
``AnalyticMarching`` (backend/main.py:335-559), including the three seed
searches (dichotomy by default; sphere tracing; gradient descent) and the
voxel-partition mode.  The marching itself runs in the MI355X engine.
"""
import copy
import os
import random
import tempfile
import time

import numpy as np
import torch

from . import cuam
from .ply import read_ply, write_ply

SPHERE_TRACING_DICT = {
    "method": "sphere_tracing",
    "args": {"init_num": 1024, "step_size_max": 1.0, "step_size_mul": 0.5, "step_size_min": 1e-3,
             "init_ball_radius": 1.0, "avg_eps": 1e-3, "hist_len": 100, "corr_min": 0.1, "time_out": 60},
}
GRADIENT_DESCENT_DICT = {
    "method": "gradient_descent",
    "args": {"init_num": 1024, "lr_max": 1e-2, "lr_mul": 0.3, "lr_min": 1e-5, "corr_min": 0.1, "hist_len": 100,
             "loss_type": "l1", "optimizer_type": "adam", "accept_bad": True, "init_ball_radius": 1.0,
             "batch_mul": 2, "avg_eps": 1e-3, "time_out": 60},
}
DICHOTOMY_DICT = {
    "method": "dichotomy",
    "args": {"init_num": 1024, "try_pts_num": 4096, "init_ball_radius": 1.0, "iter_max": 100, "avg_eps": 1e-3,
             "time_out": 60, "provided_surfpts": None, "provided_surfstd": None},
}
VOXEL_DICT = {"voxel_size": 0.1}

CONSTRAINTS_JITTER = 1e-8   # added to extra-plane normals (backend/main.py:352,427)


def get_boundary(boundary_type="cube", **kwargs):
    """Six half-spaces of an axis-aligned box, w.x + b < 0 inside (backend/utils.py:5-29)."""
    if boundary_type != "cube":
        raise Exception(f"Error: No such {boundary_type}")
    lo, hi = kwargs["min_vert"], kwargs["max_vert"]
    w = torch.tensor([[1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1]], dtype=torch.float32)
    b = torch.tensor([-hi[0], -hi[1], -hi[2], lo[0], lo[1], lo[2]], dtype=torch.float32)
    return w, b


def _inside(x, w, b):
    if w is None or b is None or w.shape[0] == 0 or x.shape[0] == 0:
        return torch.ones(x.shape[0], dtype=torch.bool, device=x.device)
    t = x @ w.to(x).t() + b.to(x).reshape(1, -1)
    return (t < 0).all(dim=1)


def _ball(n, radius, w, b, device):
    pts = torch.zeros([0, 3], device=device)
    while pts.shape[0] < n:
        p = (torch.rand([n - pts.shape[0], 3], device=device) * 2 - 1) * radius
        p = p[(p ** 2).sum(1) < radius ** 2]
        p = p[_inside(p, w, b)]
        pts = torch.cat([pts, p])
    return pts


class _Trend:
    """Correlation of the last `hist_len` errors with time (the reference's stall test, backend/main.py:94-106)."""

    def __init__(self, n):
        self.n = n
        self.h = []

    def push(self, v):
        if not self.h:
            self.h = [v] * self.n
            return 1.0
        self.h = [v] + self.h[:-1]
        return float(np.corrcoef(np.array(self.h), np.arange(self.n))[0, 1])


def dichotomy(model, iso, init_num, w_extra_constraints, b_extra_constraints, try_pts_num, init_ball_radius,
              iter_max, avg_eps, time_out, provided_surfpts=None, provided_surfstd=None):
    """Random opposite-sign pairs in a ball, then bisection (backend/main.py:252-326)."""
    dev = next(model.parameters()).device
    t0 = time.time()
    neg = torch.zeros([0, 3], device=dev)
    pos = torch.zeros([0, 3], device=dev)
    while neg.shape[0] < init_num:
        if time.time() - t0 > time_out:
            raise Exception(f"Error: dichotomy cannot find solution within {time_out} seconds!")
        if provided_surfpts is not None and provided_surfstd is not None:
            pts = (provided_surfpts + provided_surfstd * torch.randn_like(provided_surfpts)).to(dev)
        else:
            pts = _ball(try_pts_num, init_ball_radius, w_extra_constraints, b_extra_constraints, dev)
        with torch.no_grad():
            v = model(pts).reshape(-1) - iso
        ip, ineg = torch.where(v > 0)[0], torch.where(v < 0)[0]
        if len(ip) and len(ineg):
            need = init_num - neg.shape[0]
            k = min(need, len(ip) * len(ineg))
            pick = torch.tensor(random.sample(range(len(ip) * len(ineg)), k), device=dev)
            pos = torch.cat([pos, pts[ip[pick % len(ip)]]])
            neg = torch.cat([neg, pts[ineg[pick // len(ip)]]])
    mid = (neg + pos) / 2
    for _ in range(iter_max):
        with torch.no_grad():
            v = model(mid).reshape(-1) - iso
        if v.abs().mean() < avg_eps:
            break
        below = v < 0
        neg[below] = mid[below]
        pos[~below] = mid[~below]
        mid = (neg + pos) / 2
    return mid


def sphere_tracing(model, iso, init_num, step_size_max, step_size_mul, step_size_min, w_extra_constraints,
                   b_extra_constraints, init_ball_radius, avg_eps, hist_len, corr_min, time_out):
    """Gradient steps x <- x - step * f(x) * grad f(x) (backend/main.py:112-177)."""
    dev = next(model.parameters()).device
    step = step_size_max
    t0 = time.time()
    trend = _Trend(hist_len)
    while True:
        pts = _ball(init_num, init_ball_radius, w_extra_constraints, b_extra_constraints, dev)
        prev = 1e10
        done = False
        while True:
            pts = pts.detach().requires_grad_(True)
            pred = model(pts) - iso
            err = pred.abs().mean()
            c = trend.push(err.item())
            if time.time() - t0 > time_out:
                raise Exception(f"Error: sphere_tracing cannot find solution within time={time_out}!")
            if err > prev * 2 or not torch.isfinite(err) or err > 100 or c < corr_min or not np.isfinite(c):
                step *= step_size_mul
                if step < step_size_min:
                    raise Exception("Error: sphere_tracing cannot find solution!")
                break
            prev = err
            if err < avg_eps:
                done = True
                break
            g = torch.autograd.grad(pred, pts, torch.ones_like(pred))[0]
            with torch.no_grad():
                pts = pts - step * pred * g
        if done:
            break
    pts = pts.detach()
    pts = pts[(pts ** 2).sum(1) < init_ball_radius ** 2]
    return pts[_inside(pts, w_extra_constraints, b_extra_constraints)]


def gradient_descent(model, iso, init_num, lr_max, lr_mul, lr_min, corr_min, hist_len, loss_type, optimizer_type,
                     accept_bad, w_extra_constraints, b_extra_constraints, init_ball_radius, batch_mul, avg_eps,
                     time_out):
    """Minimise |f(x) - iso| over points with Adam/SGD (backend/main.py:180-249)."""
    dev = next(model.parameters()).device
    losses = {"l1": lambda x: x.abs().mean(), "l2": lambda x: (x ** 2).mean(),
              "smooth_l1": lambda x: (torch.where((x / (avg_eps * 10)).abs() < 1, 0.5 * (x / (avg_eps * 10)) ** 2,
                                                  (x / (avg_eps * 10)).abs() - 0.5) * avg_eps * 10).mean()}
    if loss_type not in losses:
        raise Exception(f"Error: No such {loss_type}")
    t0 = time.time()
    out = torch.zeros([0, 3], device=dev)
    while out.shape[0] < init_num:
        p = _ball(int((init_num - out.shape[0]) * batch_mul), init_ball_radius, w_extra_constraints,
                  b_extra_constraints, dev).requires_grad_(True)
        if optimizer_type == "adam":
            opt = torch.optim.Adam([p], lr=lr_max)
        elif optimizer_type == "sgd":
            opt = torch.optim.SGD([p], lr=lr_max, momentum=0.9)
        else:
            raise Exception(f"Error: No such {optimizer_type}")
        trend = _Trend(hist_len)
        while True:
            pred = model(p) - iso
            err = pred.abs().mean().item()
            if err <= avg_eps:
                break
            if time.time() - t0 > time_out:
                raise Exception(f"Error: gradient_descent cannot find solution within {time_out} seconds!")
            if trend.push(err) < corr_min:
                trend = _Trend(hist_len)
                for g in opt.param_groups:
                    g["lr"] *= lr_mul
                if opt.param_groups[0]["lr"] < lr_min:
                    break
            opt.zero_grad()
            losses[loss_type](pred).backward()
            opt.step()
        if opt.param_groups[0]["lr"] >= lr_min or accept_bad:
            q = p.detach()
            q = q[_inside(q, w_extra_constraints, b_extra_constraints)][: init_num - out.shape[0]]
            out = torch.cat([out, q])
    return out


_SEARCH = {"dichotomy": dichotomy, "sphere_tracing": sphere_tracing, "gradient_descent": gradient_descent}
_ENV = {"key": ""}


def _states(model, pts):
    """Activation state of each point: post-ReLU output > 0 (backend/main.py:408-411)."""
    with torch.no_grad():
        model(pts, requires_outputs_list=True)
        st = torch.cat([(s > 0) for s in model.outputs_list], dim=1)
        model.outputs_list = []
    return st


def AnalyticMarching(model, save_ply_path="mesh.ply", iso=0.0, scale=1.0, center=(0.0, 0.0, 0.0),
                     w_extra_constraints_=torch.zeros([0, 3]), b_extra_constraints_=torch.zeros([0]),
                     init_configs=DICHOTOMY_DICT, save_polymesh=True, save_float32_verts=True,
                     flip_insideout=False, dtype=torch.float64, voxel_configs=None, precision=None):
    """Mesh the iso level set of ``model`` into ``save_ply_path`` (backend/main.py:335-559).

    ``precision`` (optional) overrides the arithmetic chosen by ``dtype``:
    "fp64", "fp32" (FP32 MFMA + FP64 re-decision), "bf16x3" (split-BF16 MFMA) or
    "fp64ray" (FP64 planes, checked FP64 ray walk, candidate scan where undecided).
    Returns {"init_point_time", "init_cuda_time", "am_time", "export_time"} in seconds.
    """
    assert w_extra_constraints_.shape[0] == b_extra_constraints_.shape[0]
    E = w_extra_constraints_.shape[0] + (0 if voxel_configs is None else 6)
    model = model.cuda().eval()
    for p in model.parameters():
        p.requires_grad_(False)
    cfg = copy.deepcopy(init_configs)   # the reference mutates its default dict in place; this copy does not
    cfg["args"]["w_extra_constraints"] = w_extra_constraints_
    cfg["args"]["b_extra_constraints"] = b_extra_constraints_
    out = {}
    t = time.time()
    if cfg["method"] not in _SEARCH:
        raise Exception(f"Error: No such {cfg['method']}")
    points = _SEARCH[cfg["method"]](model, iso, **cfg["args"])
    out["init_point_time"] = time.time() - t

    info = model.get_info()
    float_type = {"fp64": "float64", "fp32": "float32", "bf16x3": "bf16x3", "fp64ray": "fp64ray"}.get(
        precision, "float64" if dtype == torch.float64 else "float32")
    tdt = torch.float64 if float_type in ("float64", "fp64ray") else torch.float32
    # states at the seeds, in the marching precision's own arithmetic
    mdt = copy.deepcopy(model).to(tdt)
    states = _states(mdt, points.to(tdt))
    weights = [w.detach().to(tdt).cuda().contiguous() for w in info["weights"]]
    biases = [b.detach().to(tdt).cuda().contiguous() for b in info["biases"]]
    arc_tm = [t_.detach().to(tdt).cuda().contiguous() for t_ in info["arc_tm"]]
    arc_table = info["arc_table"].to(torch.int32).cpu()
    w_extra = w_extra_constraints_.to(tdt).cuda() + CONSTRAINTS_JITTER
    b_extra = b_extra_constraints_.to(tdt).cuda()
    key = f"{float_type}|{model.nodes}|{arc_table.tolist()}|{E}"
    t = time.time()
    if key != _ENV["key"]:
        cuam.Init(float_type=float_type, nodesnum=model.nodes, arc_table=arc_table, num_extra_constraints=E)
        _ENV["key"] = key
        out["init_cuda_time"] = time.time() - t
    else:
        out["init_cuda_time"] = 0.0

    if voxel_configs is None:
        pts = points.to(tdt).cuda().contiguous()
        st = states.cuda().contiguous()
        torch.cuda.synchronize()
        t = time.time()
        cuam.AnalyticMarching(weights=weights, biases=biases, states=st, points=pts, arc_tm=arc_tm,
                              w_extra_constraints=w_extra.contiguous(), b_extra_constraints=b_extra.contiguous(),
                              iso=iso, flip_insideout=flip_insideout)
        out["am_time"] = time.time() - t
        t = time.time()
        cuam.CombineMesh(scale=scale, center=list(center))
        cuam.ExportMesh(file_path=save_ply_path, is_polymesh=save_polymesh, is_float32=save_float32_verts)
        out["export_time"] = time.time() - t
        return out

    # voxel partition: one run per occupied voxel with its six box planes, meshes
    # concatenated without vertex merging (backend/main.py:475-556)
    vs = voxel_configs["voxel_size"]
    idx3 = torch.floor(points.double() / vs).long()
    lo3 = idx3.min(0).values
    idx3 = idx3 - lo3
    dim3 = idx3.max(0).values + 1
    idx1 = idx3[:, 0] * dim3[1] * dim3[2] + idx3[:, 1] * dim3[2] + idx3[:, 2]
    am_t, ex_t = 0.0, 0.0
    allV, allF = [], []
    nbase = 0
    tmp = tempfile.mkdtemp()
    for gid, v in enumerate(torch.unique(idx1).tolist()):
        sel = torch.where(idx1 == v)[0]
        i0 = v // int(dim3[1] * dim3[2])
        i1 = (v % int(dim3[1] * dim3[2])) // int(dim3[2])
        i2 = v % int(dim3[2])
        c = [int(i0 + lo3[0]), int(i1 + lo3[1]), int(i2 + lo3[2])]
        wb, bb = get_boundary("cube", min_vert=[x * vs for x in c], max_vert=[(x + 1) * vs for x in c])
        wv = torch.cat([wb.to(tdt).cuda() + CONSTRAINTS_JITTER, w_extra]).contiguous()
        bv = torch.cat([bb.to(tdt).cuda(), b_extra]).contiguous()
        t = time.time()
        cuam.AnalyticMarching(weights=weights, biases=biases, states=states[sel].cuda().contiguous(),
                              points=points[sel].to(tdt).cuda().contiguous(), arc_tm=arc_tm,
                              w_extra_constraints=wv, b_extra_constraints=bv, iso=iso, flip_insideout=flip_insideout)
        am_t += time.time() - t
        t = time.time()
        V, F = cuam.CombineMesh(scale=scale, center=list(center))
        ex_t += time.time() - t
        allV.append(np.asarray(V))
        allF.extend([np.asarray(f) + nbase for f in F])
        nbase += len(V)
    t = time.time()
    write_ply(save_ply_path, np.concatenate(allV) if allV else np.zeros((0, 3)), allF,
              is_float32=save_float32_verts, is_polymesh=save_polymesh)
    os.rmdir(tmp)
    out["am_time"] = am_t
    out["export_time"] = ex_t + time.time() - t
    return out
