"""Reference scene files: the OBJ loader (OBJLoader/OBJLoader.cpp:15-174)
on the reference's own models (tests/golden: teapot, box) and on files that
exercise its token rules.  Host only.

The vertices are checked against an independent parse of the same text with
libc strtof (the conversion fscanf's %f performs), the faces against a
line-by-line reading of the corners; together they pin the parse without
restating it."""
from __future__ import annotations

import ctypes
import ctypes.util
import math
from pathlib import Path

import numpy as np
import pytest

import hulktracer_amd as H

GOLDEN = Path(__file__).resolve().parent / "golden"
_libc = ctypes.CDLL(ctypes.util.find_library("c"))
_libc.strtof.restype = ctypes.c_float
_libc.strtof.argtypes = [ctypes.c_char_p, ctypes.c_void_p]


def strtof(s: str) -> np.float32:
    return np.float32(_libc.strtof(s.encode(), None))


def independent_parse(path: Path):
    """Vertices and triangles of a well-formed triangle OBJ, line by line."""
    V, T = [], []
    for line in path.read_text().splitlines():
        w = line.split()
        if not w:
            continue
        if w[0] == "v":
            V.append([strtof(x) for x in w[1:4]])
        elif w[0] == "f":
            T.append([int(c.split("/")[0]) - 1 for c in w[1:4]])
    return np.array(V, np.float32), np.array(T, np.int64)


@pytest.mark.parametrize("name,nv,nt", [("teapot.obj", 1461, 2882), ("Box.obj", 8, 12)])
def test_reference_models_load(name, nv, nt):
    s = H.Scene.from_obj(GOLDEN / name)
    assert s.vertices.shape == (nv, 4) and s.triangles.shape == (nt, 4)
    V, T = independent_parse(GOLDEN / name)
    # at scene time 0 the rotation is the identity: coordinates are the parsed floats
    # (a -0.0 would come out as +0.0 from the identity product; these files have none)
    assert np.array_equal(s.vertices[:, :3].view(np.uint32), V.view(np.uint32))
    assert np.all(s.vertices[:, 3] == 1.0)
    assert np.array_equal(s.triangles[:, :3].astype(np.int64), T) and np.all(s.triangles[:, 3] == 0)
    # box = per-lane min / max of the vertices
    assert np.array_equal(s.box[:3], s.vertices[:, :3].min(0)) and np.array_equal(s.box[4:7], s.vertices[:, :3].max(0))
    # the loader's one directional light: normalize(-1, 1, 1, 1) over four lanes
    assert s.lights.shape == (1, 8)
    assert np.array_equal(s.lights[0], np.array([0.9, 0.9, 0.9, 0.0, -0.5, 0.5, 0.5, 0.5], np.float32))
    assert not s.mirror.any()


def test_box_exponent_literal():
    s = H.Scene.from_obj(GOLDEN / "Box.obj")
    assert s.vertices[0, 1] == strtof("5.9264e-009")


def test_scene_time_rotates_about_y():
    base = H.Scene.from_obj(GOLDEN / "teapot.obj", 0.0)
    for t in (0.5, math.pi / 2, 3.0):
        s = H.Scene.from_obj(GOLDEN / "teapot.obj", t)
        v0, v1 = base.vertices[:, :3].astype(np.float64), s.vertices[:, :3].astype(np.float64)
        c, sn = math.cos(t), math.sin(t)
        # row vector times rotateMatrix(0, 1, 0, t): x' = x c - z s, z' = x s + z c, y kept
        want = np.stack([v0[:, 0] * c - v0[:, 2] * sn, v0[:, 1], v0[:, 0] * sn + v0[:, 2] * c], 1)
        assert np.abs(v1 - want).max() < 2e-7
        assert np.allclose(np.linalg.norm(v1[:, [0, 2]], axis=1), np.linalg.norm(v0[:, [0, 2]], axis=1), atol=2e-7)
    full = H.Scene.from_obj(GOLDEN / "teapot.obj", 2 * math.pi)
    assert np.abs(full.vertices - base.vertices).max() < 1e-6


def test_token_rules(tmp_path):
    """Quads keep their first three corners (the fourth is an ignored token),
    v/t/n corners keep the vertex index, keywords' arguments are skipped, a
    '#' swallows 99 characters only."""
    long_comment = "# " + "x" * 97 + " v 7 8 9"  # the tail after 99 characters goes on as tokens
    text = "\n".join([
        "mtllib scene.mtl", "o thing", "g group1", "usemtl red", "s 1",
        "v 0 0 0", "v 1 0 0", "v 1 1 0", "v 0 1 0",
        "vt 0.5 0.5", "vn 0 0 1",
        "f 1/1/1 2/1/1 3/1/1 4/1/1",
        "f 1//1 3//1 4//1",
        long_comment,
        "f 5 1 2",
    ]) + "\n"
    p = tmp_path / "quirks.obj"
    p.write_text(text)
    s = H.Scene.from_obj(p)
    assert s.vertices[:, :3].tolist() == [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [7, 8, 9]]
    assert s.triangles[:, :3].tolist() == [[0, 1, 2], [0, 2, 3], [4, 0, 1]]


@pytest.mark.parametrize("face", ["f 0 1 2", "f 1 2 9", "f -1 1 2"])
def test_face_outside_vertex_list_is_rejected(tmp_path, face):
    p = tmp_path / "bad.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\n" + face + "\n")
    with pytest.raises(H.HulkError) as e:
        H.Scene.from_obj(p)
    assert e.value.code == 4


def test_missing_file():
    with pytest.raises(H.HulkError) as e:
        H.Scene.from_obj(GOLDEN / "no_such_file.obj")
    assert e.value.code == 2


def test_host_tree_of_the_teapot_matches_the_restatement_oracle():
    """The host builder on the reference's own model against the pure-Python
    restatement of the SAH-NlogN build."""
    import oracle.kdtree_ref as K

    s = H.Scene.from_obj(GOLDEN / "teapot.obj")
    kd = H.KDTreeCPU()
    kd.buildKDTree(s)
    nodes, lp = K.build(s.vertices, s.triangles, s.box)
    assert np.array_equal(kd.getNodeArray(), nodes) and np.array_equal(kd.getPolygonArray(), lp)
