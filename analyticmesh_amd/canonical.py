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

__all__ = ["RegionRecords", "vertex_key", "same_topology", "vertex_error", "fingerprint"]


def _sorted_rows(a):
    """Rows of a 2-D integer array in lexicographic order."""
    a = np.ascontiguousarray(a)
    if len(a) == 0:
        return a
    return a[np.lexsort(a.T[::-1])]


def _unique_rows(a):
    a = _sorted_rows(a)
    if len(a) < 2:
        return a
    keep = np.ones(len(a), dtype=bool)
    keep[1:] = np.any(a[1:] != a[:-1], axis=1)
    return a[keep]


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

    def topology_rows(self):
        """Unsorted canonical rows: (mask rows, face rows, vertex-key rows, one per vertex occurrence)."""
        masks = self.masks.astype(np.int64)
        nv = self.info[:, 2].astype(np.int64)
        sel = nv >= 3
        m, n, vp = masks[sel], nv[sel], self.vpairs[sel].astype(np.int64)
        vm = vp.shape[1]
        lo, hi = np.minimum(vp[:, :, 0], vp[:, :, 1]), np.maximum(vp[:, :, 0], vp[:, :, 1])
        slot = np.arange(vm)[None, :]
        valid = slot < n[:, None]
        key = np.where(valid, (lo << 32) | hi, np.iinfo(np.int64).max)
        i0 = np.argmin(key, axis=1)
        src = (i0[:, None] + slot) % np.maximum(n, 1)[:, None]
        rot = np.where(valid, np.take_along_axis(key, src, axis=1), -1)
        faces = np.concatenate([m, n[:, None], rot], axis=1)
        r_idx, k_idx = np.nonzero(valid)
        vlo, vhi = lo[r_idx, k_idx], hi[r_idx, k_idx]
        vmk = m[r_idx].copy()
        rows = np.arange(len(vmk))
        for g in (vlo, vhi):
            ok = (g >= 0) & (g < self.L)
            w = np.where(ok, g >> 5, 0)
            bit = np.where(ok, np.int64(1) << (g & 31), 0)
            vmk[rows, w] &= ~bit
        vkeys = np.concatenate([vlo[:, None], vhi[:, None], vmk], axis=1)
        return masks, faces, vkeys

    def digest(self):
        """Vectorised, order-independent form of visited(), faces() and vertex_keys():
        (sorted unique mask rows, sorted face rows, sorted unique vertex-key rows), each an
        int64 array.  Two runs have the same topology iff the three arrays are equal."""
        masks, faces, vkeys = self.topology_rows()
        return _unique_rows(masks), _sorted_rows(faces), _unique_rows(vkeys)

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


def _row_hash_sum(rows):
    """Order-independent 64-bit fingerprint of a multiset of int64 rows: the sum mod 2^64 of a
    mixed per-row hash (multiply-xorshift over the columns)."""
    if len(rows) == 0:
        return 0
    h = np.full(len(rows), 0x243F6A8885A308D3, dtype=np.uint64)
    mul = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        for j in range(rows.shape[1]):
            h = (h ^ rows[:, j].astype(np.uint64)) * mul
            h ^= h >> np.uint64(31)
        h = (h ^ (h >> np.uint64(33))) * np.uint64(0xFF51AFD7ED558CCD)
        h ^= h >> np.uint64(33)
        return int(h.sum(dtype=np.uint64))


def fingerprint(rec):
    """Topology fingerprint of a run: region and face counts and order-independent hashes of the
    visited masks, the canonical faces and the vertex-key occurrences (as hex strings)."""
    masks, faces, vkeys = rec.topology_rows()
    return {"regions": int(len(masks)), "faces": int(len(faces)), "vertex_uses": int(len(vkeys)),
            "visited": f"{_row_hash_sum(masks):016x}", "faces_hash": f"{_row_hash_sum(faces):016x}",
            "vertex_keys": f"{_row_hash_sum(vkeys):016x}"}


def same_topology(a, b):
    """(equal?, reason) for two RegionRecords, compared through their digests."""
    da, db = a.digest(), b.digest()
    for name, x, y in zip(("visited set", "faces", "vertex keys"), da, db):
        if x.shape != y.shape or not np.array_equal(x, y):
            return False, f"{name} differ ({len(x)} vs {len(y)} rows)"
    return True, "identical"


def vertex_error(a, b):
    """Largest vertex difference of run a against run b, relative to max(|v_b|, 1e-12) per vertex.

    Regions are matched by mask and, within a region, vertices by their pair of tight
    constraints; returns None when a region of a is missing from b or its pairs differ."""
    idx = {m.tobytes(): i for i, m in enumerate(b.masks)}
    try:
        perm = np.fromiter((idx[m.tobytes()] for m in a.masks), dtype=np.int64, count=len(a.masks))
    except KeyError:
        return None
    big = np.int64(1) << 32

    def keyed(verts, pairs, nv):
        p = np.sort(pairs.astype(np.int64), axis=2)
        key = p[:, :, 0] * big + p[:, :, 1]
        slot = np.arange(key.shape[1])[None, :]
        key = np.where(slot < nv[:, None], key, np.iinfo(np.int64).max)
        o = np.argsort(key, axis=1)
        return np.take_along_axis(verts, o[:, :, None], axis=1), np.take_along_axis(key, o, axis=1)

    nv = a.info[:, 2].astype(np.int64)
    va, ka = keyed(a.verts, a.vpairs, nv)
    vb, kb = keyed(b.verts[perm], b.vpairs[perm], b.info[perm, 2].astype(np.int64))
    valid = (np.arange(va.shape[1])[None, :] < nv[:, None]) & (nv[:, None] >= 3)
    if not np.array_equal(np.where(valid, ka, 0), np.where(valid, kb, 0)):
        return None
    d = np.abs(va - vb).max(axis=2) / np.maximum(np.abs(vb).max(axis=2), 1e-12)
    return float(np.where(valid, d, 0.0).max()) if valid.any() else 0.0


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
