"""OBJ files and trace-line formats (host code of the library; no GPU).

The reader is checked against the vendored libigl igl::readOBJ (main.cpp:1153)
on the fixtures in tests/golden/obj (written and parsed by
tests/golden/make_golden.py); the writer's number format against std::ostream
with std::scientific (TriMesh.cpp:1758-1798) and the trace format ("%g")
against std::ostream with default flags (Optimizer.cpp:706-739), both from
tests/golden/stream_format.json. Values are compared exactly.
"""
import json
import os

import numpy as np
import pytest

from optcuts_amd import capi

HERE = os.path.dirname(os.path.abspath(__file__))
OBJ_DIR = os.path.join(HERE, "golden", "obj")


@pytest.fixture(scope="module")
def fmt():
    d = json.load(open(os.path.join(HERE, "golden", "stream_format.json")))
    d["values"] = [float.fromhex(h) for h in d["values_hex"]]
    return d


@pytest.mark.parametrize("name", ["basic", "mixed_syntax", "no_uv", "disc7"])
def test_obj_reader_matches_igl(golden, name):
    V, UV, F, FUV = capi.obj_read(os.path.join(OBJ_DIR, name + ".obj"))
    Vg, TCg, Fg, FTCg = (golden["obj_%s_%s" % (name, k)] for k in ("V", "TC", "F", "FTC"))
    np.testing.assert_array_equal(V, Vg)
    np.testing.assert_array_equal(F, Fg)
    if TCg.size:
        np.testing.assert_array_equal(UV, TCg[:, :2])
    else:
        assert UV.shape == (0, 2)
    if FTCg.shape[0]:
        np.testing.assert_array_equal(FUV, FTCg)
    else:
        assert FUV is None


def test_obj_writer_number_format_is_std_scientific(tmp_path, fmt):
    vals = np.array(fmt["values"])
    n3 = len(vals) // 3 * 3
    V = vals[:n3].reshape(-1, 3)
    n2 = len(vals) // 2 * 2
    UV = vals[:n2].reshape(-1, 2)
    F = np.array([[0, 1, 2]], np.int32)
    path = str(tmp_path / "w.obj")
    capi.obj_write(path, V, F, UV, FUV=np.array([[2, 1, 0]], np.int32))
    lines = open(path).read().split("\n")
    assert lines[-1] == ""
    vl = [l for l in lines if l.startswith("v ")]
    tl = [l for l in lines if l.startswith("vt ")]
    fl = [l for l in lines if l.startswith("f ")]
    assert len(vl) == V.shape[0] and len(tl) == UV.shape[0] and fl == ["f 1/3 2/2 3/1"]
    sci = fmt["scientific"]
    assert [w for l in vl for w in l.split()[1:]] == sci[:n3]
    assert [w for l in tl for w in l.split()[1:]] == sci[:n2]
    # without texture indices the vertex index is written for both (TriMesh.cpp:1788-1795)
    capi.obj_write(path, V[:3], F)
    assert open(path).read().split("\n")[-2] == "f 1/1 2/2 3/3"


def test_trace_number_format_is_stream_default(fmt):
    """The trace lines are written with C "%g"; the default std::ostream output agrees."""
    for v, s in zip(fmt["values"], fmt["default"]):
        assert "%g" % v == s


def test_obj_round_trip_through_reader(tmp_path):
    rng = np.random.default_rng(3)
    V = rng.standard_normal((30, 3))
    UV = rng.standard_normal((30, 2))
    F = rng.integers(0, 30, (40, 3)).astype(np.int32)
    path = str(tmp_path / "r.obj")
    capi.obj_write(path, V, F, UV)
    V2, UV2, F2, FUV2 = capi.obj_read(path)
    np.testing.assert_array_equal(F2, F)
    np.testing.assert_array_equal(FUV2, F)
    np.testing.assert_allclose(V2, V, rtol=6e-7, atol=0)  # 7 significant digits
    np.testing.assert_allclose(UV2, UV, rtol=6e-7, atol=0)


def test_obj_write_mesh_maps_cut_mesh_to_input_vertices(tmp_path):
    """TriMesh::saveAsMesh(path, F0, scaleUV): a square of two triangles cut along its
    diagonal; positions come from the cut mesh through F0, texture coordinates are the
    cut mesh's UV (normalised by the larger of the u and v ranges when scaleUV)."""
    P = np.array([[0, 0, 0], [2, 0, 0], [0, 2, 1], [2, 2, 1.0]])
    F0 = np.array([[0, 1, 2], [1, 3, 2]], np.int32)
    cut = np.array([0, 1, 2, 1, 3, 2])  # cut vertex -> input vertex
    Vrest = P[cut]
    F = np.array([[0, 1, 2], [3, 4, 5]], np.int32)
    UV = np.array([[1, 1], [3, 1], [1, 5], [4, 1], [6, 1], [4, 5.0]])
    path = str(tmp_path / "m.obj")
    capi.obj_write_mesh(path, Vrest, UV, F, F0, scale_uv=True)
    V2, UV2, F2, FUV2 = capi.obj_read(path)
    np.testing.assert_array_equal(V2, P)
    np.testing.assert_array_equal(F2, F0)
    np.testing.assert_array_equal(FUV2, F)
    scale = max(6 - 1, 5 - 1)
    np.testing.assert_allclose(UV2, (UV - [1, 1]) / scale, rtol=1e-6, atol=1e-7)
    capi.obj_write_mesh(path, Vrest, UV, F, F0, scale_uv=False)
    np.testing.assert_array_equal(capi.obj_read(path)[1], UV)


@pytest.mark.parametrize("text,match", [
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nf 1 2 3 4\n", "triangles"),
    ("v 0 0\nf 1 1 1\n", "3 coordinates"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 9\n", "out of range"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1/1 2/x 3/1\n", "invalid format"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nf 1/1 2/1 3\n", "invalid format"),
])
def test_obj_reader_rejects(tmp_path, text, match):
    path = str(tmp_path / "bad.obj")
    open(path, "w").write(text)
    with pytest.raises(capi.OCError, match=match):
        capi.obj_read(path)


def test_obj_reader_missing_file(tmp_path):
    with pytest.raises(capi.OCError, match="cannot open"):
        capi.obj_read(str(tmp_path / "none.obj"))
