"""Python front end of the native engine (``libam.so`` through ``_am``).

``march()`` runs one network from explicit seeds and returns the engine's
records; it is what the reference-style entry points and the benchmark use.
The extension must be built in-tree (``python -m analyticmesh_amd.build``);
there is no fallback path.
"""
import os

import numpy as np

from .canonical import RegionRecords

PRECISIONS = {"fp32": 0, "bf16x3": 1, "fp64": 2, "fp64ray": 3}

_mod = None


def native():
    """Import the compiled module, failing loudly when it is missing."""
    global _mod
    if _mod is None:
        try:
            from . import _am  # noqa: F401
        except ImportError as e:
            here = os.path.dirname(os.path.abspath(__file__))
            raise ImportError(
                "analyticmesh_amd native engine is not built; run `python -m analyticmesh_amd.build` "
                f"(looked in {here}): {e}") from e
        _mod = _am
    return _mod


def _as_ptr(a):
    """Address of a torch tensor or numpy array (host or device)."""
    if hasattr(a, "data_ptr"):
        return int(a.data_ptr())
    return int(a.ctypes.data)


class Faces:
    """Polygon faces in CSR form; indexing yields one face's vertex indices."""

    def __init__(self, offsets, indices):
        self.offsets = offsets
        self.indices = indices

    def __len__(self):
        return len(self.offsets) - 1

    def __getitem__(self, i):
        if i < 0:
            i += len(self)
        return self.indices[self.offsets[i]:self.offsets[i + 1]]

    def __iter__(self):
        for i in range(len(self)):
            yield self[i]


class MarchResult:
    """Result of one ``Engine.run``.

    The per-region records (``masks``, ``info``, ``tabs``, ``verts``,
    ``vpairs``, ``points``) stay in device memory until first read; they
    belong to the context's latest run, so reading them after the context has
    run again raises RuntimeError."""

    _REGION_FIELDS = ("masks", "info", "tabs", "verts", "vpairs", "points")

    def __init__(self, eng, spec_L, Wd, vm, owner=None, generation=0):
        self._eng = eng
        self._owner = owner
        self._generation = generation
        self._regions = None
        self.L = spec_L
        self.Wd = Wd
        self.vm = vm
        self.stats = eng.stats()

    def _check_current(self):
        if self._owner is not None and self._owner._generation != self._generation:
            raise RuntimeError("the engine has run again since this result; its region records are gone")

    def __getattr__(self, name):
        if name in MarchResult._REGION_FIELDS:
            if self._regions is None:
                self._check_current()
                self._regions = self._eng.regions(self.Wd)
            return self._regions[name]
        raise AttributeError(name)

    def fingerprint(self):
        """Topology fingerprint computed on the device (see canonical.fingerprint)."""
        self._check_current()
        return dict(self._eng.fingerprint())

    def records(self):
        return RegionRecords(self.masks, self.info, self.tabs, self.verts, self.vpairs, L=self.L)

    @property
    def run_ms(self):
        return float(self.stats[0])

    def combine(self, scale=1.0, center=(0.0, 0.0, 0.0)):
        """Deduplicate vertices and index faces on the GPU.

        Returns (vertices (V, 3) float64, Faces) where Faces is a list-like view
        over the CSR arrays ``offsets`` (F + 1) and ``indices``."""
        self._check_current()
        self._eng.combine(float(scale), [float(c) for c in center])
        V, off, idx = self._eng.mesh()
        self.combine_ms = float(self._eng.stats()[10])
        return V, Faces(off, idx)

    def adjoint(self, grad_vertices, layer_shapes):
        """dL/d(weights, biases) for L = sum <grad_vertices, v> (regions held fixed).

        grad_vertices: (V, 3) array in the order of ``combine()``'s vertices.
        layer_shapes: [(out, in), ...] per linear layer.  Returns (dW list, db list), float64."""
        self._check_current()
        g = np.ascontiguousarray(grad_vertices, dtype=np.float64)
        nv = int(self._eng.counts()[2])
        if g.shape != (nv, 3):
            raise ValueError(f"grad_vertices must be ({nv}, 3), got {g.shape}")
        layer_shapes = [tuple(int(x) for x in s) for s in layer_shapes]
        if self._owner is not None:
            n = self._owner.nodes
            want = [(n[i + 1], n[i]) for i in range(len(n) - 1)]
            if layer_shapes != want:
                raise ValueError(f"layer_shapes must be {want}, got {layer_shapes}")
        dW = [np.zeros(s, dtype=np.float64) for s in layer_shapes]
        db = [np.zeros(s[0], dtype=np.float64) for s in layer_shapes]
        self._eng.adjoint(int(g.ctypes.data), 1, [int(a.ctypes.data) for a in dW], [int(a.ctypes.data) for a in db])
        return dW, db

    def export_ply(self, path, is_polymesh=True, is_float32=True):
        self._check_current()
        self._eng.export_ply(str(path), bool(is_polymesh), bool(is_float32))


class Engine:
    """One engine context for a fixed architecture (mirrors ``Init``)."""

    def __init__(self, nodes, arc_table=None, num_extra_constraints=0, precision="fp64", vm=20,
                 table_log2=None, walk="ray", ray_waves=None, scan_blocks=None, graph_levels=False,
                 reduced_vertices="newton_where_needed"):
        """ray_waves / scan_blocks: launch shapes (see ``autotune``); None takes the tuned
        shape stored for the nearest network shape.  graph_levels: launch each level as one
        captured hipGraph (every form; the stage times of the stats are then not taken).  reduced_vertices (fp32 form):
        "newton_where_needed" (default) gives one FP64 Newton step after the march to each vertex
        solved from the reduced planes that the face kernel's error estimate does not put within
        1e-5 of |v|; "newton" gives every such vertex the step; "fp64" re-solves every vertex from
        FP64 planes; "estimate" keeps the vertices of a region solved from the reduced planes
        where the estimate puts all of them within 1e-5 of |v|, else re-solves them from FP64."""
        m = native()
        nodes = [int(n) for n in nodes]
        D = len(nodes) - 2
        if arc_table is None:
            arc = np.zeros((D, 1), dtype=np.int32)
        else:
            arc_table = [list(r) for r in arc_table]
            width = max(len(r) for r in arc_table)
            arc = np.zeros((len(arc_table), width), dtype=np.int32)
            for i, r in enumerate(arc_table):
                arc[i, :len(r)] = r
        self.nodes = nodes
        self.L = sum(nodes[1:-1])
        self.Wd = (self.L + 31) // 32
        self.E = int(num_extra_constraints)
        self.vm = int(vm)
        self._eng = m.Engine(nodes, arc, self.E, PRECISIONS[precision])
        self._eng.set_option("vm", self.vm)
        if table_log2 is not None:
            self._eng.set_option("table_log2", int(table_log2))
        self._eng.set_option("walk", {"ray": 0, "scan": 1}[walk])
        self._eng.set_option("reduced_vertices", {"fp64": 0, "estimate": 1, "newton": 2, "newton_where_needed": 3}[reduced_vertices])
        from .autotune import lookup
        tuned = lookup(nodes)
        self.ray_waves = int(ray_waves if ray_waves is not None else tuned["ray_waves"])
        self.scan_blocks = int(scan_blocks if scan_blocks is not None else tuned["scan_blocks"])
        self._eng.set_option("ray_waves", self.ray_waves)
        self._eng.set_option("scan_blocks", self.scan_blocks)
        self._eng.set_option("graph_levels", 1 if graph_levels else 0)
        self._generation = 0

    def set_option(self, key, value):
        self._eng.set_option(key, int(value))

    def planes(self, masks, with_sigma=False):
        """Reduced-precision planes of the given packed masks: ((n, 4, C) f32 signed, (n, 4) eq), and
        with ``with_sigma`` the face kernel's error estimate of every plane, (n, C + 1) (the face
        plane's last)."""
        P, E, S = self._eng.planes(np.ascontiguousarray(masks, dtype=np.uint32), self.L + self.E)
        return (P, E, S) if with_sigma else (P, E)

    def run(self, weights, biases, points, states, tms=(), extra_w=None, extra_b=None, iso=0.0, flip=False,
            stream=0):
        """Buffers may be numpy arrays or torch tensors (CPU or GPU), all float32 or all float64."""
        first = weights[0]
        is64 = str(getattr(first, "dtype", "")) in ("float64", "torch.float64")
        dtype = 1 if is64 else 0
        keep = []

        def cont(a):
            if hasattr(a, "contiguous"):
                a = a.contiguous()
            else:
                a = np.ascontiguousarray(a)
            keep.append(a)
            return a

        wp = [_as_ptr(cont(w)) for w in weights]
        bp = [_as_ptr(cont(b)) for b in biases]
        def shape2(t):
            n = t.numel() if hasattr(t, "numel") else np.size(t)
            return (int(t.shape[0]), int(t.shape[1])) if n else (0, 0)

        tm_shapes = [shape2(t) for t in tms]
        tp = [_as_ptr(cont(t)) if s != (0, 0) else 0 for t, s in zip(tms, tm_shapes)]
        pts = cont(points)
        st = states
        if hasattr(st, "to"):
            import torch
            st = st.to(torch.uint8).contiguous()
        else:
            st = np.ascontiguousarray(st, dtype=np.uint8)
        keep.append(st)
        n = int(pts.shape[0])
        ew = cont(extra_w) if extra_w is not None and len(extra_w) else None
        eb = cont(extra_b) if extra_b is not None and len(extra_b) else None
        self._eng.run(wp, bp, tm_shapes, tp, dtype, _as_ptr(pts), _as_ptr(st), n,
                      _as_ptr(ew) if ew is not None else 0, _as_ptr(eb) if eb is not None else 0,
                      float(iso), bool(flip), int(stream))
        self._generation += 1
        return MarchResult(self._eng, self.L, self.Wd, self.vm, owner=self, generation=self._generation)


def march(spec, seed_points, seed_states, precision="fp64", vm=20, iso=0.0, flip=False, max_regions=None,
          table_log2=None, walk="ray", extra_w=None, extra_b=None, **engine_options):
    """Mesh one NetSpec (with its skips, if any) from explicit seeds (points (N,3) float64,
    states (N,L) bool).

    walk: reduced-precision forms only; "ray" (ray shooting, default) or
    "scan" (the reference's candidate scan with checked predicates).
    extra_w (E,3), extra_b (E,): extra half-spaces a.x + c <= 0, used as given."""
    E = 0 if extra_w is None else len(extra_w)
    eng = Engine(spec.nodes, getattr(spec, "arc_table", None), E, precision=precision, vm=vm,
                 table_log2=table_log2, walk=walk, **engine_options)
    if max_regions is not None:
        eng.set_option("max_regions", max_regions)
    pts = np.ascontiguousarray(seed_points, dtype=np.float64).reshape(-1, 3)
    st = np.ascontiguousarray(seed_states, dtype=np.uint8).reshape(len(pts), spec.L)
    W = [np.ascontiguousarray(w, dtype=np.float64) for w in spec.weights]
    B = [np.ascontiguousarray(b, dtype=np.float64) for b in spec.biases]
    tms = [np.ascontiguousarray(t, dtype=np.float64) for t in spec.tm_arrays()] if hasattr(spec, "tm_arrays") else []
    ew = None if not E else np.ascontiguousarray(extra_w, dtype=np.float64).reshape(E, 3)
    eb = None if not E else np.ascontiguousarray(extra_b, dtype=np.float64).reshape(E)
    res = eng.run(W, B, pts, st, tms=tms, extra_w=ew, extra_b=eb, iso=iso, flip=flip)
    res.engine = eng
    return res


def march_batch(specs, seeds, precision="fp32", workers=16, vm=20, walk="ray", combine=True, max_regions=None):
    """Mesh many independent networks in one job.

    Every network gets its own engine context and HIP stream; a pool of host
    threads drives them so their levels overlap on the GPU.  ``seeds`` is a
    list of (points, states) per network.  Returns the per-network results
    (with ``.mesh`` = (V, faces) when ``combine``) and the wall time in seconds.
    """
    import concurrent.futures
    import time

    def one(i):
        spec = specs[i]
        pts, st = seeds[i]
        r = march(spec, pts, st, precision=precision, vm=vm, walk=walk, max_regions=max_regions)
        if combine:
            r.mesh = r.combine()
        return r

    t0 = time.perf_counter()
    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as ex:
        out = list(ex.map(one, range(len(specs))))
    return out, time.perf_counter() - t0


class BatchResult:
    """Per-network views of one batch run (records are gathered from device memory on demand)."""

    def __init__(self, batch, specs, vm):
        self._b = batch
        self.specs = specs
        self.vm = vm
        self.stats = batch.stats()

    def __len__(self):
        return len(self.specs)

    def fingerprint(self, i):
        return dict(self._b.fingerprint(int(i)))

    def regions(self, i):
        return dict(self._b.regions(int(i)))

    def records(self, i):
        d = self.regions(i)
        return RegionRecords(d["masks"], d["info"], d["tabs"], d["verts"], d["vpairs"], L=self.specs[i].L)

    def combine(self, scale=1.0, center=(0.0, 0.0, 0.0)):
        """Vertex dedup and face indexing of every network at once; returns (V, F, face indices) counts."""
        self._b.combine(float(scale), [float(c) for c in center])
        r, f, v, n = self._b.counts()
        self.combine_ms = float(self._b.stats()[10])
        return int(v), int(f), int(n)

    @property
    def run_ms(self):
        return float(self.stats[0])


class BatchEngine:
    """Many networks marched in one job by one set of kernel launches per breadth-first level
    (``am_batch_*``).  precision: "fp64" (bit-exact scan) or "fp64ray"."""

    def __init__(self, specs, precision="fp64ray", vm=20, ray_waves=None, scan_blocks=None, graph_levels=False):
        m = native()
        self.specs = list(specs)
        for s in self.specs:
            if getattr(s, "has_skips", False):
                raise ValueError(f"{s.name}: a batch takes networks without skips")
        self.vm = int(vm)
        self._b = m.Batch([list(map(int, s.nodes)) for s in self.specs], PRECISIONS[precision])
        self._b.set_option("vm", self.vm)
        from .autotune import lookup
        widest = max(self.specs, key=lambda s: (max(s.nodes[1:-1]), len(s.nodes)))
        tuned = lookup(widest.nodes)
        self._b.set_option("ray_waves", int(ray_waves if ray_waves is not None else tuned["ray_waves"]))
        self._b.set_option("scan_blocks", int(scan_blocks if scan_blocks is not None else tuned["scan_blocks"]))
        self._b.set_option("graph_levels", 1 if graph_levels else 0)
        self._keep = []

    def run(self, seeds, iso=0.0, flip=False, dtype=np.float64):
        """seeds: list of (points (N,3), states (N,L)) per network."""
        keep = []

        def cont(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return int(a.ctypes.data)

        W, B, P, S, N = [], [], [], [], []
        for spec, (pts, st) in zip(self.specs, seeds):
            W += [cont(w, dtype) for w in spec.weights]
            B += [cont(b, dtype) for b in spec.biases]
            pts = np.asarray(pts, dtype=np.float64).reshape(-1, 3)
            P.append(cont(pts, dtype))
            S.append(cont(np.asarray(st).reshape(len(pts), spec.L), np.uint8))
            N.append(len(pts))
        self._b.run(W, B, 1 if dtype == np.float64 else 0, P, S, N, float(iso), bool(flip), 0)
        self._keep = keep
        return BatchResult(self._b, self.specs, self.vm)
