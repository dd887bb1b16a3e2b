"""Differentiable meshing: vertex positions as a function of the MLP parameters.

``mesh_vertices(model, seed_points, seed_states)`` returns the (V, 3) vertex
tensor of the marched mesh; its backward runs the engine's adjoint kernel,
which differentiates each vertex through the implicit function of its three
tight planes (two neurons and f = iso) with the activation regions held
fixed.  Faces are returned alongside as index arrays (not differentiable).
"""
import numpy as np
import torch

from .engine import Engine

__all__ = ["MeshVertices", "mesh_vertices"]


class MeshVertices(torch.autograd.Function):
    @staticmethod
    def forward(ctx, spec_nodes, seed_points, seed_states, iso, *params):
        weights = [p.detach().double().cpu().numpy() for p in params[0::2]]
        biases = [p.detach().double().cpu().numpy() for p in params[1::2]]
        eng = Engine(spec_nodes, None, 0, precision="fp64")
        res = eng.run(weights, biases, np.asarray(seed_points, dtype=np.float64),
                      np.asarray(seed_states, dtype=np.uint8), iso=float(iso))
        V, F = res.combine()
        ctx.res = res
        ctx.shapes = [tuple(w.shape) for w in weights]
        ctx.dtypes = [p.dtype for p in params]
        ctx.devices = [p.device for p in params]
        ctx.faces = F
        ctx.engine = eng
        out = torch.from_numpy(np.array(V))
        return out.to(params[0].device)

    @staticmethod
    def backward(ctx, grad_v):
        g = grad_v.detach().double().cpu().numpy()
        dW, db = ctx.res.adjoint(g, ctx.shapes)
        grads = []
        for i in range(len(dW)):
            grads.append(torch.from_numpy(dW[i]).to(ctx.devices[2 * i], ctx.dtypes[2 * i]))
            grads.append(torch.from_numpy(db[i]).to(ctx.devices[2 * i + 1], ctx.dtypes[2 * i + 1]))
        return (None, None, None, None, *grads)


def mesh_vertices(model, seed_points, seed_states, iso=0.0):
    """(V, 3) float64 vertices (differentiable w.r.t. model.linears' parameters) and the faces."""
    params = []
    for lin in model.linears:
        params += [lin.weight, lin.bias]
    V = MeshVertices.apply(list(model.nodes), seed_points, seed_states, iso, *params)
    return V
