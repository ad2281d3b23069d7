"""The reference's on-disk KD dataset (paris.bin): writer
(GLVisitorGPUTrace::saveDataSetToFile, GLVisitorGPUTrace.cpp:638-674, with
buildPerLeafData, :163-234) and reader (GLVisitorGPUTraceSSD::buildTree,
GLVisitorGPUTraceSSD.cpp:43-107).  Host only; the trace check runs the CPU
oracle.

The file is parsed here independently with numpy from the layout the
reference's KdFileHeader / LeafExtData / TriangleIndices / SPVector define
(GLVisitorGPUTrace.hpp:44-63, SharedStructs.hpp:37-41), and each leaf's lists
are checked against the tree and the scene they came from."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import hulktracer_amd as H
import oracle as O
from tests.helpers import scene_and_tree

GOLDEN = Path(__file__).resolve().parent / "golden"


def parse(path: Path):
    raw = np.fromfile(path, np.uint8)
    h = raw[:72].view("<u8")
    keys = ("header", "tree_size", "tree_off", "ext_size", "ext_off", "index_size", "index_off", "vertex_size",
            "vertex_off")
    hd = dict(zip(keys, (int(x) for x in h)))
    nodes = raw[hd["tree_off"]:hd["tree_off"] + hd["tree_size"]].view("<u4").reshape(-1, 4)
    ext = raw[hd["ext_off"]:hd["ext_off"] + hd["ext_size"]].view("<u4").reshape(-1, 4)
    idx = raw[hd["index_off"]:hd["index_off"] + hd["index_size"]].view("<u4").reshape(-1, 4)
    vtx = raw[hd["vertex_off"]:hd["vertex_off"] + hd["vertex_size"]].view("<f4").reshape(-1, 4)
    return raw, hd, nodes, ext, idx, vtx


def _scenes():
    s1, n1, l1 = scene_and_tree(1, 1)
    t = H.Scene.from_obj(GOLDEN / "teapot.obj")
    kd = H.KDTreeCPU()
    kd.buildKDTree(t)
    return {"random1k": (s1, n1, l1), "teapot": (t, kd.getNodeArray(), kd.getPolygonArray())}


SCENES = _scenes()


def write(tmp_path, name):
    s, nodes, lp = SCENES[name]
    kd = H.KDTreeCPU()
    kd.m_treeNodes, kd.m_leafPolygons = nodes, lp
    path = tmp_path / f"{name}.bin"
    kd.saveDataSetToFile(s, path)
    return path, s, nodes, lp


@pytest.mark.parametrize("name", sorted(SCENES))
def test_file_layout(tmp_path, name):
    path, s, nodes, lp = write(tmp_path, name)
    raw, hd, fnodes, ext, idx, vtx = parse(path)
    leaves = np.nonzero(nodes[:, 0] == 4)[0]
    assert hd["header"] == 72 and hd["tree_off"] == 72 and hd["tree_size"] == 16 * len(nodes)
    assert hd["ext_off"] == 72 + 16 * len(nodes) and hd["ext_size"] == 16 * len(leaves)
    assert hd["index_off"] % 512 == 0 and 0 <= hd["index_off"] - (hd["ext_off"] + hd["ext_size"]) < 512
    assert not raw[hd["ext_off"] + hd["ext_size"]:hd["index_off"]].any()
    assert hd["vertex_off"] == hd["index_off"] + hd["index_size"]
    assert len(raw) == hd["vertex_off"] + hd["vertex_size"]
    assert np.array_equal(fnodes, nodes)
    # table: offsets are running sums of the padded sizes, sizes multiples of 32
    assert np.all(ext[:, 2] % 32 == 0) and np.all(ext[:, 3] % 32 == 0)
    assert np.array_equal(ext[:, 0], np.concatenate([[0], np.cumsum(ext[:, 3])[:-1]]))
    assert np.array_equal(ext[:, 1], np.concatenate([[0], np.cumsum(ext[:, 2])[:-1]]))
    assert ext[:, 3].sum() == len(idx) and ext[:, 2].sum() == len(vtx)


@pytest.mark.parametrize("name", sorted(SCENES))
def test_per_leaf_lists(tmp_path, name):
    """Each leaf's triangles, in the leaf's order, are its original triangles
    over a vertex list that holds each original vertex once, numbered by
    first use; the rest of both lists is zero padding."""
    path, s, nodes, lp = write(tmp_path, name)
    _, _, _, ext, idx, vtx = parse(path)
    leaves = np.nonzero(nodes[:, 0] == 4)[0]
    for k, node in enumerate(leaves):
        start, count = nodes[node, 1], nodes[node, 2]
        io, vo, vs, isz = ext[k]
        tris = s.triangles[lp[start:start + count], :3]
        local = idx[io:io + count, :3]
        assert np.all(idx[io:io + count, 3] == 0)
        # first-use numbering: local ids appear as 0, 1, 2, ... in reading order
        seen = []
        for v in local.ravel():
            if v not in seen:
                assert v == len(seen)
                seen.append(v)
        used = len(seen)
        # local -> original is a function and one to one
        pairs = {}
        for a, b in zip(local.ravel(), tris.ravel()):
            assert pairs.setdefault(int(a), int(b)) == int(b)
        assert len(set(pairs.values())) == used
        assert np.array_equal(vtx[vo:vo + used][local].view(np.uint32), s.vertices[tris].view(np.uint32))
        assert not idx[io + count:io + isz].any() and not vtx[vo + used:vo + vs].any()


@pytest.mark.parametrize("name", sorted(SCENES))
def test_round_trip_traces_the_same(tmp_path, name):
    """The loaded dataset traced by the oracle: same t, u, v, normal on every
    ray; triangle k of the dataset is the original lp[k]."""
    path, s, nodes, lp = write(tmp_path, name)
    s2, nodes2, lp2 = H.Scene.from_dataset(path)
    assert np.array_equal(nodes2, nodes) and np.array_equal(lp2, np.arange(len(lp)))
    assert np.array_equal(s2.box, s.box)
    cam = s.camera(0, 1.0) if s.kind else H.make_camera_lookat((0.9, 0.7, 1.1), (0.0, 0.15, 0.0), fov=50.0)
    a = O.OracleScene(s.vertices, s.triangles, s.box, nodes, lp, s.lights).trace_camera(O.V_FULLSTACK, cam, 128, 128)
    b = O.OracleScene(s2.vertices, s2.triangles, s2.box, nodes2, lp2, s.lights).trace_camera(O.V_FULLSTACK, cam, 128,
                                                                                              128)
    for f in ("t", "u", "v", "nx", "ny", "nz"):
        assert np.array_equal(a[f].view(np.uint32), b[f].view(np.uint32)), f
    hit = b["tri"] != H.NO_HIT
    assert np.array_equal(hit, a["tri"] != H.NO_HIT) and hit.sum() > 1000
    assert np.array_equal(lp[b["tri"][hit]], a["tri"][hit])


def test_malformed_files_are_rejected(tmp_path):
    path, s, nodes, lp = write(tmp_path, "random1k")
    raw = path.read_bytes()
    for name, data in (("truncated", raw[:len(raw) // 2]), ("short", raw[:40]),
                       ("bad_header", b"\x00" * 8 + raw[8:])):
        p = tmp_path / f"{name}.bin"
        p.write_bytes(data)
        with pytest.raises(H.HulkError) as e:
            H.Scene.from_dataset(p)
        assert e.value.code == 4, name
    with pytest.raises(H.HulkError) as e:
        H.Scene.from_dataset(tmp_path / "missing.bin")
    assert e.value.code == 2
