"""MI355X-native analytic marching: exact meshing of ReLU-MLP level sets.

Public surface of the reference package (backend/__init__.py:3-7):
``AnalyticMarching``, ``lib`` (the native module: here ``cuam``), ``load_model``,
``save_model``, ``MLP``, ``get_boundary``; the polygon-mesh tools of its
libpolytools live in ``analyticmesh_amd.polytools``.
"""
from .model import MLP  # noqa: F401
from .onnx_io import load_model, save_model  # noqa: F401


def __getattr__(name):
    if name in ("AnalyticMarching", "get_boundary", "DICHOTOMY_DICT", "SPHERE_TRACING_DICT",
                "GRADIENT_DESCENT_DICT", "VOXEL_DICT"):
        from . import main
        return getattr(main, name)
    if name == "lib":
        from . import cuam
        return cuam
    if name == "PolyMesh":
        from .polytools import PolyMesh
        return PolyMesh
    raise AttributeError(name)
