"""ReLU MLP descriptor with optional additive skip connections.

Same constructor, forward and ``get_info`` contract as the reference's
``MLP`` (backend/model.py:5-148): ``nodes`` lists every layer width
(input 3, output 1), ``arc_table[r]`` lists the skip connections added after
linear layer ``r+1`` as ``[count, src0, tm0, src1, tm1, ...]`` and
``arc_tm_shape`` gives each transform matrix's shape (``[0, 0]`` = identity).
"""
import numpy as np
import torch

__all__ = ["MLP", "geometric_init_"]


def geometric_init_(linears, nodes, radius, generator=None):
    """Geometric (SAL) initialisation (backend/model.py:63-70).

    Hidden layers: W ~ N(0, 2/fan_out), b = 0.  Output layer:
    W = sqrt(pi)/sqrt(fan_in), b = -radius.
    """
    n = len(linears)
    with torch.no_grad():
        for i, lin in enumerate(linears):
            if i != n - 1:
                lin.weight.normal_(0.0, float(np.sqrt(2) / np.sqrt(nodes[i + 1])), generator=generator)
                lin.bias.zero_()
            else:
                lin.weight.fill_(float(np.sqrt(np.pi) / np.sqrt(nodes[i])))
                lin.bias.fill_(-radius)


class MLP(torch.nn.Module):
    """MLP of arbitrary width/depth for analytic marching (backend/model.py:5-99)."""

    def __init__(self, nodes, arc_table, arc_tm_shape, initialization="geometric",
                 geometric_radius=1.0, enable_print=True):
        super().__init__()
        if nodes[0] != 3 or nodes[-1] != 1 or len(nodes) < 3:
            raise ValueError("nodes must start with 3, end with 1 and have at least 3 entries")
        if len(arc_table) != len(nodes) - 2:
            raise ValueError("arc_table needs one row per hidden layer")
        self.nodes = list(nodes)
        self.arc_table = [list(r) for r in arc_table]
        self.arc_tm_shape = [list(s) for s in arc_tm_shape]
        self.geometric_radius = geometric_radius

        self.num_of_linears = len(self.nodes) - 1
        self.linears = torch.nn.ModuleList(
            [torch.nn.Linear(self.nodes[i], self.nodes[i + 1], bias=True) for i in range(self.num_of_linears)])
        if initialization == "kaiming":
            with torch.no_grad():
                for lin in self.linears:
                    torch.nn.init.kaiming_normal_(lin.weight)
                    torch.nn.init.zeros_(lin.bias)
        elif initialization == "geometric":
            geometric_init_(self.linears, self.nodes, geometric_radius)
        elif initialization is not None:
            raise Exception(f"Error: No such initialization: {initialization}")
        elif enable_print:
            print("Warning: Initialization strategy for MLP is not specified.")

        self.num_of_acts = self.num_of_linears
        self.tms = torch.nn.ModuleList()
        for shp in self.arc_tm_shape:
            if shp[0] == 0 and shp[1] == 0:
                self.tms.append(torch.nn.Identity())
            else:
                self.tms.append(torch.nn.Linear(in_features=shp[1], out_features=shp[0], bias=False))
        self.srcs = {}
        for arc in self.arc_table:
            for i in range(arc[0]):
                self.srcs[arc[1 + 2 * i]] = None
        self.outputs_list = []

    def forward(self, x, requires_outputs_list=False):
        """Forward pass; keeps post-ReLU hidden outputs when asked (backend/model.py:101-123)."""
        if requires_outputs_list:
            self.outputs_list.clear()
        srcs = {}
        for i in range(self.num_of_linears):
            if i in self.srcs:
                srcs[i] = x
            x = self.linears[i](x)
            if i >= 1:
                row = self.arc_table[i - 1]
                for j in range(row[0]):
                    x = x + self.tms[row[2 + 2 * j]](srcs[row[1 + 2 * j]])
            if i != self.num_of_linears - 1:
                x = torch.relu(x)
                if requires_outputs_list:
                    self.outputs_list.append(x)
        return x

    def get_info(self):
        """Weights (out,in), biases, transform matrices and the int32 arc table (backend/model.py:125-148)."""
        weights = [lin.weight for lin in self.linears]
        biases = [lin.bias for lin in self.linears]
        arc_tm = []
        for tm in self.tms:
            arc_tm.append(torch.zeros([0, 0]) if isinstance(tm, torch.nn.Identity) else tm.weight)
        width = max(len(a) for a in self.arc_table)
        arc_table = torch.zeros([len(self.arc_table), width], dtype=torch.int32)
        for r, row in enumerate(self.arc_table):
            for c, v in enumerate(row):
                arc_table[r, c] = v
        return {"weights": weights, "biases": biases, "arc_tm": arc_tm, "arc_table": arc_table}

    def has_skips(self):
        return any(r[0] != 0 for r in self.arc_table)
