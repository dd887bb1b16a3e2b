"""Canonical, order-independent views of a marching result.

A run produces one record per visited region (activation mask, the tab of
constraint indices the walk recorded, the polygon's vertices and, for each
vertex, the two constraint planes whose intersection with the f = iso plane
gave it).  Traversal order, the choice of start plane and the direction of
the walk only change how a face is written down, so comparisons use:

* the visited set: the set of packed masks;
* faces: per region with >= 3 vertices, the cyclic sequence of vertex plane
  pairs in the stored (oriented) order, rotated to start at its smallest pair;
* vertices: global keys ``(lo, hi, mask without bits lo and hi)``.  The same
  mesh vertex appears in every region around it; clearing the two tight
  neurons' bits gives the same key from each of them.
"""
import numpy as np

__all__ = ["RegionRecords", "vertex_key"]


def vertex_key(mask_words, a, b, L):
    lo, hi = (a, b) if a < b else (b, a)
    m = np.array(mask_words, dtype=np.uint32, copy=True)
    for g in (lo, hi):
        if 0 <= g < L:
            m[g >> 5] &= np.uint32(~(1 << (g & 31)) & 0xFFFFFFFF)
    return (int(lo), int(hi), m.tobytes())


class RegionRecords:
    """Per-region arrays: masks (R, W) u32, info (R, 8) i32, tabs (R, vm+1),
    verts (R, vm, 3) f64, vpairs (R, vm, 2)."""

    def __init__(self, masks, info, tabs, verts, vpairs, L):
        self.masks = np.asarray(masks)
        self.info = np.asarray(info)
        self.tabs = np.asarray(tabs)
        self.verts = np.asarray(verts)
        self.vpairs = np.asarray(vpairs)
        self.L = int(L)

    def __len__(self):
        return len(self.masks)

    @property
    def nverts(self):
        return self.info[:, 2]

    def visited(self):
        return {m.tobytes() for m in self.masks}

    def face_count(self):
        return int((self.nverts >= 3).sum())

    def faces(self):
        out = {}
        for r in range(len(self.masks)):
            nv = int(self.info[r, 2])
            if nv < 3:
                continue
            pairs = [tuple(sorted((int(self.vpairs[r, k, 0]), int(self.vpairs[r, k, 1])))) for k in range(nv)]
            i0 = min(range(nv), key=lambda k: pairs[k])
            out[self.masks[r].tobytes()] = tuple(pairs[i0:] + pairs[:i0])
        return out

    def vertex_keys(self):
        keys = set()
        for r in range(len(self.masks)):
            nv = int(self.info[r, 2])
            if nv < 3:
                continue
            for k in range(nv):
                keys.add(vertex_key(self.masks[r], int(self.vpairs[r, k, 0]), int(self.vpairs[r, k, 1]), self.L))
        return keys

    def keyed_mesh(self):
        """Vertices deduplicated by key (first occurrence wins) and faces as index lists."""
        index = {}
        coords = []
        faces = []
        for r in range(len(self.masks)):
            nv = int(self.info[r, 2])
            if nv < 3:
                continue
            f = []
            for k in range(nv):
                key = vertex_key(self.masks[r], int(self.vpairs[r, k, 0]), int(self.vpairs[r, k, 1]), self.L)
                if key not in index:
                    index[key] = len(coords)
                    coords.append(self.verts[r, k])
                f.append(index[key])
            faces.append(f)
        return np.array(coords, dtype=np.float64).reshape(-1, 3), faces

    def vertex_coords(self):
        """key -> list of coordinates of that vertex from every region that has it."""
        out = {}
        for r in range(len(self.masks)):
            nv = int(self.info[r, 2])
            if nv < 3:
                continue
            for k in range(nv):
                key = vertex_key(self.masks[r], int(self.vpairs[r, k, 0]), int(self.vpairs[r, k, 1]), self.L)
                out.setdefault(key, []).append(self.verts[r, k])
        return out


def edges_of_faces(faces):
    """Undirected edge -> number of faces using it."""
    cnt = {}
    for f in faces:
        n = len(f)
        for i in range(n):
            a, b = f[i], f[(i + 1) % n]
            e = (a, b) if a < b else (b, a)
            cnt[e] = cnt.get(e, 0) + 1
    return cnt
