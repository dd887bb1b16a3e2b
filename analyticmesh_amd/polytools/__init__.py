"""Polygon-mesh post-processing: the surface of the reference's libpolytools
(backend/libpolytools/__init__.py): ``PolyMesh``, ``poly2tri``,
``get_faces_num`` and ``load_ply_header``.  PolyMesh is the compiled
``analyticmesh_amd._polylib`` module (csrc/polylib.cpp)."""
import os

__all__ = ["PolyMesh", "poly2tri", "get_faces_num", "load_ply_header"]

try:
    from .._polylib import PolyMesh
except ImportError as e:  # pragma: no cover - the build is part of the package
    raise ImportError("analyticmesh_amd._polylib is not built; run `python -m analyticmesh_amd.build`") from e


def poly2tri(src_file, dst_file):
    """Fan-triangulate a polygon mesh file into dst_file (backend/libpolytools/poly2tri.py)."""
    mesh = PolyMesh(src_file)
    mesh.poly2tri()
    mesh.save(dst_file)


def get_faces_num(mesh_file):
    """{'poly': polygon faces, 'tri': fan triangles} of a mesh file
    (backend/libpolytools/cntfaces.py)."""
    mesh = PolyMesh(mesh_file)
    return {"poly": mesh.num_polyfaces(), "tri": mesh.num_trifaces()}


def load_ply_header(ply_path):
    """Header of a PLY file as {'storing_type', 'storing_version', '<element>_num', ...}
    (backend/libpolytools/header.py)."""
    head = b""
    with open(ply_path, "rb") as f:
        while b"end_header" not in head.lower():
            chunk = f.read(4096)
            if not chunk:
                raise ValueError(f"{ply_path}: no end_header")
            head += chunk
    info = {}
    for raw in head.split(b"\n"):
        try:
            line = raw.decode("ascii").strip().lower()
        except UnicodeDecodeError:
            continue
        parts = line.split()
        if not parts:
            continue
        if parts[0] == "format" and len(parts) >= 3:
            info["storing_type"] = parts[1]
            info["storing_version"] = float(parts[2])
        elif parts[0] == "element" and len(parts) >= 3:
            info[f"{parts[1]}_num"] = int(parts[2])
        elif parts[0] == "end_header":
            break
    return info
