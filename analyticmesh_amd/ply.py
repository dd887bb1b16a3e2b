"""Binary little-endian PLY in the layout the engine writes
(backend/inc/polymesh.h:368-422): vertex x/y/z as float or double, faces as
``property list uchar int vertex_index``.
"""
import numpy as np

__all__ = ["read_ply", "write_ply"]


def write_ply(path, vertices, faces, is_float32=True, is_polymesh=True):
    V = np.asarray(vertices, dtype=np.float64).reshape(-1, 3)
    if not is_polymesh:
        tri = []
        for f in faces:
            for j in range(len(f) - 2):
                tri.append((f[0], f[j + 1], f[j + 2]))
        faces = tri
    ft = "float" if is_float32 else "double"
    head = ("ply\nformat binary_little_endian 1.0\n"
            f"element vertex {len(V)}\nproperty {ft} x\nproperty {ft} y\nproperty {ft} z\n"
            f"element face {len(faces)}\nproperty list uchar int vertex_index\nend_header\n")
    with open(path, "wb") as fh:
        fh.write(head.encode())
        fh.write(V.astype("<f4" if is_float32 else "<f8").tobytes())
        for f in faces:
            fh.write(np.uint8(len(f)).tobytes())
            fh.write(np.asarray(f, dtype="<i4").tobytes())


def read_ply(path):
    with open(path, "rb") as fh:
        data = fh.read()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    header = data[:end].decode().splitlines()
    if header[0] != "ply" or header[1] != "format binary_little_endian 1.0":
        raise ValueError("not a binary little-endian PLY")
    nv = nf = 0
    vtype = None
    for line in header:
        p = line.split()
        if p[:2] == ["element", "vertex"]:
            nv = int(p[2])
        elif p[:2] == ["element", "face"]:
            nf = int(p[2])
        elif p[:1] == ["property"] and p[-1] == "x":
            vtype = {"float": "<f4", "double": "<f8"}[p[1]]
    vb = np.dtype(vtype).itemsize * 3 * nv
    V = np.frombuffer(data[end:end + vb], dtype=vtype).reshape(nv, 3)
    pos = end + vb
    faces = []
    for _ in range(nf):
        n = data[pos]
        pos += 1
        faces.append(np.frombuffer(data[pos:pos + 4 * n], dtype="<i4").copy())
        pos += 4 * n
    if pos != len(data):
        raise ValueError("trailing bytes after the face list")
    return V, faces
