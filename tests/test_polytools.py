"""Polygon-mesh tools (the reference's libpolytools surface) on the CPU: PLY/OFF round trips
with the engine's own PLY layout, fan triangulation, face counts, header parsing, welding."""
import numpy as np
import pytest

from analyticmesh_amd.ply import read_ply, write_ply
from analyticmesh_amd.polytools import PolyMesh, get_faces_num, load_ply_header, poly2tri

V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [2, 0.5, 0], [0.5, 2, 0.25]], dtype=np.float64)
F = [np.array([0, 1, 2, 3]), np.array([1, 4, 2]), np.array([3, 2, 5, 5, 0])]


@pytest.mark.parametrize("f32", [True, False])
def test_engine_ply_loads(tmp_path, f32):
    """Files in the engine's PLY layout (float or double x/y/z) load with the same vertices
    (as float) and faces."""
    p = str(tmp_path / "m.ply")
    write_ply(p, V, F, is_float32=f32, is_polymesh=True)
    m = PolyMesh(p)
    np.testing.assert_array_equal(np.array(m.vertices()), V.astype(np.float32))
    assert m.faces() == [list(map(int, f)) for f in F]
    assert m.colors() == []
    assert load_ply_header(p) == {"storing_type": "binary_little_endian", "storing_version": 1.0,
                                  "vertex_num": 6, "face_num": 3}


def test_saved_ply_layout_and_roundtrip(tmp_path):
    """Saved PLY: float vertices, uchar-count int-index faces, optional uchar RGB per face."""
    m = PolyMesh(V.tolist(), [list(map(int, f)) for f in F], [[255, 0, 0], [0, 128, 0], [1, 2, 3]])
    p = str(tmp_path / "c.ply")
    m.save(p)
    raw = open(p, "rb").read()
    head = (b"ply\nformat binary_little_endian 1.0\nelement vertex 6\nproperty float x\nproperty float y\n"
            b"property float z\nelement face 3\nproperty list uchar int vertex_index\nproperty uchar red\n"
            b"property uchar green\nproperty uchar blue\nend_header\n")
    assert raw.startswith(head)
    body = raw[len(head):]
    assert len(body) == 6 * 12 + sum(1 + 4 * len(f) + 3 for f in F)
    np.testing.assert_array_equal(np.frombuffer(body[:72], dtype="<f4").reshape(6, 3), V.astype(np.float32))
    assert body[72] == 4 and np.frombuffer(body[73:89], dtype="<i4").tolist() == [0, 1, 2, 3]
    assert list(body[89:92]) == [255, 0, 0]
    m2 = PolyMesh(p)
    assert m2.faces() == m.faces() and m2.colors() == m.colors()
    # the engine's reader takes it too (colours aside, the same layout)
    p2 = str(tmp_path / "n.ply")
    PolyMesh(V.tolist(), [list(map(int, f)) for f in F], []).save(p2)
    V2, F2 = read_ply(p2)
    np.testing.assert_array_equal(V2, V.astype(np.float32))
    assert [list(f) for f in F2] == m.faces()


def test_off_roundtrip(tmp_path):
    m = PolyMesh(V.tolist(), [list(map(int, f)) for f in F], [[10, 20, 30], [0, 0, 0], [255, 255, 255]])
    p = str(tmp_path / "m.off")
    m.save(p)
    assert open(p).read().startswith("OFF\n6 3 0\n")
    m2 = PolyMesh(p)
    np.testing.assert_allclose(np.array(m2.vertices()), V, rtol=0, atol=1e-6)
    assert m2.faces() == m.faces() and m2.colors() == m.colors()
    # colours given as 0..1 floats (the reference's OFF reader's form)
    q = tmp_path / "f.off"
    q.write_text("OFF\n3 1 0\n0 0 0\n1 0 0\n0 1 0\n3 0 1 2 1.0 0.5 0.0\n")
    assert PolyMesh(str(q)).colors() == [[255, 128, 0]]


def test_poly2tri_fan_and_counts(tmp_path):
    """Fan (0, j+1, j+2) per polygon; triangles with a repeated index are dropped by poly2tri but
    counted by num_trifaces (backend/libpolytools/src/polylib.cpp:499-571)."""
    m = PolyMesh(V.tolist(), [list(map(int, f)) for f in F], [[1, 1, 1], [2, 2, 2], [3, 3, 3]])
    assert m.is_polymesh() and m.num_polyfaces() == 3 and m.num_trifaces() == 2 + 1 + 3
    m.poly2tri()
    assert m.faces() == [[0, 1, 2], [0, 2, 3], [1, 4, 2], [3, 2, 5], [3, 5, 0]]
    assert m.colors() == [[1, 1, 1], [1, 1, 1], [2, 2, 2], [3, 3, 3], [3, 3, 3]]
    assert not m.is_polymesh() and m.num_trifaces() == m.num_polyfaces() == 5
    src, dst = str(tmp_path / "p.ply"), str(tmp_path / "t.ply")
    write_ply(src, V, F, is_float32=True, is_polymesh=True)
    poly2tri(src, dst)
    assert get_faces_num(dst) == {"poly": 5, "tri": 5}
    assert get_faces_num(src) == {"poly": 3, "tri": 6}


def test_weld_joins_pieces():
    """Two copies of a square, the second moved by 1e-7: welding at 1e-6 leaves 4 vertices."""
    sq = V[:4]
    m = PolyMesh(np.concatenate([sq, sq + 1e-7]).tolist(), [[0, 1, 2, 3], [4, 5, 6, 7]], [])
    assert m.weld(1e-6) == 4
    assert len(m.vertices()) == 4 and m.faces() == [[0, 1, 2, 3], [0, 1, 2, 3]]


def test_errors():
    with pytest.raises(RuntimeError):
        PolyMesh([[0, 0]], [], [])
    with pytest.raises(RuntimeError):
        PolyMesh(V.tolist(), [[0, 1]], [])
    with pytest.raises(RuntimeError):
        PolyMesh(V.tolist(), [[0, 1, 2]], [[0, 0, 256]])
    with pytest.raises(RuntimeError):
        PolyMesh().load("mesh.obj")


def test_ascii_ply_rejected(tmp_path):
    p = tmp_path / "a.ply"
    p.write_text("ply\nformat ascii 1.0\nelement vertex 0\nelement face 0\nend_header\n")
    with pytest.raises(RuntimeError):
        PolyMesh(str(p))
