"""Device KD-tree build (hulk_build_kdtree_device) against the host builder.

The host builder (hulktracer_amd/csrc/kdtree_sah.cpp) is pinned to the
reference's SAH-NlogN build by tests/test_oracle.py (the pure-Python
restatement oracle/kdtree_ref.py and the reference's own cases).  The device
build must give the same arrays: node words and leaf polygon references
identical, except that a split plane at zero may carry the other sign
(std::sort leaves the order of -0.0 and +0.0 open; both classify and
traverse the same).  Then every trace over the device-built tree is bit for
bit the trace over the uploaded host tree.
"""
from __future__ import annotations

import numpy as np
import pytest

import hulktracer_amd as H
from tests.helpers import hit_mismatch, scene_and_tree

pytestmark = pytest.mark.gpu


def assert_same_tree(got_nodes, got_lp, ref_nodes, ref_lp, what=""):
    assert got_nodes.shape == ref_nodes.shape, f"{what}: {len(got_nodes)} nodes, host builder {len(ref_nodes)}"
    diff = got_nodes != ref_nodes
    junction = (ref_nodes[:, 0] & 8) != 0
    zero = junction & ((got_nodes[:, 3] & 0x7FFFFFFF) == 0) & ((ref_nodes[:, 3] & 0x7FFFFFFF) == 0)
    diff[zero, 3] = False
    rows = np.nonzero(diff.any(axis=1))[0]
    assert len(rows) == 0, f"{what}: {len(rows)} nodes differ, first {rows[:5]}: {got_nodes[rows[:3]]} vs {ref_nodes[rows[:3]]}"
    assert got_lp.shape == ref_lp.shape, f"{what}: {len(got_lp)} leaf refs, host builder {len(ref_lp)}"
    assert np.array_equal(got_lp, ref_lp), f"{what}: leaf polygon references differ"


def device_tree(scene, ctx=None, capacity=0):
    ctx = ctx or H.GPUTrace()
    ctx.upload_scene(scene)
    ctx.build_kdtree_device(capacity)
    nodes, lp = ctx.download_kdtree()
    return ctx, nodes, lp


def host_tree(scene):
    kd = H.KDTreeCPU()
    kd.buildKDTree(scene)
    return kd.getNodeArray(), kd.getPolygonArray()


@pytest.mark.parametrize("kind,seed,param,frame", [(1, 1, 0, 0), (1, 3, 300, 0), (1, 9, 5000, 0), (2, 1, 0, 0),
                                                   (3, 7, 0, 0), (3, 7, 0, 11)])
def test_device_build_matches_host_builder(kind, seed, param, frame):
    """Random triangles (1,000 / 300 / 5,000), terrain + spheres (163,650,
    depth 20, planes at zero) and two frames of the animated 1M-triangle
    scene."""
    s, ref_nodes, ref_lp = scene_and_tree(kind, seed, param, frame)
    print(f"device build kind={kind} seed={seed} param={param} frame={frame} tris={s.n_tris}", flush=True)
    ctx, nodes, lp = device_tree(s)
    assert_same_tree(nodes, lp, ref_nodes, ref_lp, f"kind {kind} frame {frame}")
    info = ctx.device_build_info()
    depth = _max_depth(ref_nodes)
    assert info["max_jdepth"] == (depth if len(ref_nodes) > 1 else 0)


def _max_depth(nodes):
    d = np.zeros(len(nodes), np.int64)
    for i in range(len(nodes)):
        if nodes[i, 0] & 8:
            for c in (nodes[i, 1], nodes[i, 2]):
                if c:
                    d[c] = d[i] + 1
    return int(d.max())


def _grid(n, z=0.5, jitter=False):
    """n x n quads (2 n^2 triangles) in the plane z (coplanar with every z split)."""
    xs = np.linspace(0.0, 1.0, n + 1, dtype=np.float32)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    V = np.stack([X.ravel(), Y.ravel(), np.full(X.size, z, np.float32)], 1)
    if jitter:
        V[:, 2] += (np.random.default_rng(5).random(len(V)).astype(np.float32) - 0.5) * 0.05
    idx = lambda i, j: i * (n + 1) + j
    T = []
    for i in range(n):
        for j in range(n):
            T.append((idx(i, j), idx(i + 1, j), idx(i, j + 1)))
            T.append((idx(i + 1, j), idx(i + 1, j + 1), idx(i, j + 1)))
    return V, np.array(T, np.uint32)


def edge_scenes():
    rng = np.random.default_rng(11)
    out = {}
    # fewer than 10 triangles: the root is a leaf
    V = rng.random((15, 3)).astype(np.float32)
    out["five"] = H.Scene.from_arrays(V, np.arange(15, dtype=np.uint32).reshape(5, 3))
    # exactly 10 triangles: the root is evaluated
    V = rng.random((30, 3)).astype(np.float32)
    out["ten"] = H.Scene.from_arrays(V, np.arange(30, dtype=np.uint32).reshape(10, 3))
    # a flat grid: zero box extent in z, every triangle coplanar with z planes
    out["flat"] = H.Scene.from_arrays(*_grid(40))
    out["terrain_like"] = H.Scene.from_arrays(*_grid(60, jitter=True))
    # duplicates: 300 copies of three triangles
    V = rng.random((9, 3)).astype(np.float32)
    T = np.tile(np.arange(9, dtype=np.uint32).reshape(3, 3), (300, 1))
    out["duplicates"] = H.Scene.from_arrays(V, T)
    # degenerate triangles (points, segments) among ordinary ones
    V = rng.random((3000, 3)).astype(np.float32)
    T = np.arange(3000, dtype=np.uint32).reshape(1000, 3)
    T[::7, 1] = T[::7, 0]
    T[::7, 2] = T[::7, 0]
    T[3::11, 2] = T[3::11, 1]
    out["degenerate"] = H.Scene.from_arrays(V, T)
    # coordinates at -0.0 and +0.0 around the box middle
    V = (rng.random((6000, 3)).astype(np.float32) - 0.5)
    V[::5, 0] = -0.0
    V[1::5, 0] = 0.0
    V[2::9, 1] = -0.0
    out["signed_zero"] = H.Scene.from_arrays(V, np.arange(6000, dtype=np.uint32).reshape(2000, 3))
    # shared vertices on an exact lattice: long runs of equal candidate values
    V = np.floor(rng.random((4000, 3)) * 8).astype(np.float32) / 8
    out["lattice"] = H.Scene.from_arrays(V, rng.integers(0, 4000, (3000, 3)).astype(np.uint32))
    return out


EDGE = edge_scenes()


@pytest.mark.parametrize("name", sorted(EDGE))
def test_device_build_edge_cases(name):
    s = EDGE[name]
    print(f"device build edge case {name} tris={s.n_tris}", flush=True)
    ref_nodes, ref_lp = host_tree(s)
    _, nodes, lp = device_tree(s)
    assert_same_tree(nodes, lp, ref_nodes, ref_lp, name)


def test_device_build_single_triangle_and_empty_leaf_root():
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    s = H.Scene.from_arrays(V, np.array([[0, 1, 2]], np.uint32))
    _, nodes, lp = device_tree(s)
    assert nodes.tolist() == [[4, 0, 1, 0]] and lp.tolist() == [0]
    s0 = H.Scene.from_arrays(V, np.zeros((0, 3), np.uint32))
    ref_nodes, ref_lp = host_tree(s0)
    _, nodes, lp = device_tree(s0)
    assert_same_tree(nodes, lp, ref_nodes, ref_lp, "no triangles")
    assert nodes.tolist() == [[4, 0, 0, 0]] and len(lp) == 0


def test_capacity_overflow_restarts_with_more_room():
    """A first attempt with room for 1,024 references (the 1,000-triangle
    tree holds 1,604 leaf references) overflows and doubles until the tree
    fits; the arrays do not change."""
    s, ref_nodes, ref_lp = scene_and_tree(1, 1)
    ctx, nodes, lp = device_tree(s, capacity=1024)
    assert_same_tree(nodes, lp, ref_nodes, ref_lp, "capacity 1024")
    info = ctx.device_build_info()
    assert info["retries"] >= 1 and info["capacity"] >= len(ref_lp) // 2


def test_rebuilds_in_one_context_follow_the_scene():
    """Frames of the animated scene, then a smaller scene, in one context:
    no state of an earlier build leaks into the next."""
    ctx = H.GPUTrace()
    for kind, seed, frame in ((3, 7, 3), (3, 7, 4), (1, 1, 0), (2, 1, 0)):
        s, ref_nodes, ref_lp = scene_and_tree(kind, seed, 0, frame)
        print(f"rebuild kind={kind} frame={frame}", flush=True)
        _, nodes, lp = device_tree(s, ctx)
        assert_same_tree(nodes, lp, ref_nodes, ref_lp, f"kind {kind} frame {frame}")


@pytest.mark.parametrize("variant", [H.FULLSTACK, H.LDSSTACK, H.RESTART, H.ROPES])
def test_trace_over_device_tree_matches_uploaded_tree(variant):
    """Terrain + spheres, pose 13: hits over the device-built tree equal the
    hits over the uploaded host tree bit for bit (the upload path is pinned
    to the oracle by test_gpu_parity.py)."""
    s, nodes, lp = scene_and_tree(2, 1)
    cam = s.camera(13, 256 / 384)
    up = H.GPUTrace(variant=variant, lds_top_levels=8)
    up.upload_scene(s)
    up.upload_kdtree(nodes, lp)
    ref = up.trace_camera(cam, 384, 256)
    dev = H.GPUTrace(variant=variant, lds_top_levels=8)
    dev.upload_scene(s)
    dev.build_kdtree_device()
    got = dev.trace_camera(cam, 384, 256)
    bad = hit_mismatch(got, ref)
    assert len(bad) == 0, f"{len(bad)} rays differ"
    assert (got["tri"] != H.NO_HIT).sum() > 10000


def test_render_frame_over_device_tree():
    """Animated frame 2: whole frame (shadow rays, two mirror generations)
    over the device-built tree equals the frame over the uploaded tree."""
    s, nodes, lp = scene_and_tree(3, 7, 0, 2)
    cam = s.camera(2, 1.0)
    up = H.GPUTrace(variant=H.FULLSTACK)
    up.upload_scene(s)
    up.upload_kdtree(nodes, lp)
    ref_px, ref_hits = up.render_frame(cam, 256, 256, max_bounce=2, mirror=s.mirror)
    dev = H.GPUTrace(variant=H.FULLSTACK)
    dev.upload_scene(s)
    dev.build_kdtree_device()
    px, hits = dev.render_frame(cam, 256, 256, max_bounce=2, mirror=s.mirror)
    assert np.array_equal(px, ref_px)
    assert len(hit_mismatch(hits, ref_hits)) == 0


def test_download_returns_uploaded_tree():
    s, nodes, lp = scene_and_tree(1, 1)
    ctx = H.GPUTrace()
    ctx.upload_scene(s)
    ctx.upload_kdtree(nodes, lp)
    n2, lp2 = ctx.download_kdtree()
    assert np.array_equal(n2, nodes) and np.array_equal(lp2, lp)


def test_nan_vertex_is_rejected():
    V = np.random.default_rng(2).random((60, 3)).astype(np.float32)
    V[17, 1] = np.nan
    s = H.Scene.from_arrays(V, np.arange(60, dtype=np.uint32).reshape(20, 3))
    ctx = H.GPUTrace()
    ctx.upload_scene(s)
    with pytest.raises(H.HulkError, match="NaN"):
        ctx.build_kdtree_device()


def test_visitor_builds_tree_on_the_device():
    """GLVisitorGPUTrace with --treeBuild=gpu renders the same framebuffer as
    with the host build."""
    s, _, _ = scene_and_tree(1, 1)
    frames = []
    for where in ("cpu", "gpu"):
        v = H.GLVisitorGPUTrace(s, H.Settings([f"--treeBuild={where}", "--kernelType=GPUTrace"]), 128, 128)
        v.initializeVisitor()
        v.traverseBegin()
        frames.append(np.array(v.framebuffer))
    assert np.array_equal(frames[0], frames[1])


def test_update_vertices_then_device_build():
    """Animated frames through hulk_update_vertices (vertices only) and a
    device build: each frame's tree and hits equal a fresh upload's."""
    s0, _, _ = scene_and_tree(3, 7, 0, 5)
    ctx = H.GPUTrace(variant=H.ROPES)
    ctx.upload_scene(s0)
    for frame in (6, 9):
        s, nodes, lp = scene_and_tree(3, 7, 0, frame)
        ctx.update_vertices(s.vertices, s.box)
        with pytest.raises(H.HulkError):
            ctx.trace_camera(s.camera(frame, 1.0), 32, 32)  # the old tree was dropped
        ctx.build_kdtree_device()
        n2, lp2 = ctx.download_kdtree()
        assert_same_tree(n2, lp2, nodes, lp, f"frame {frame}")
        ref = H.GPUTrace(variant=H.ROPES)
        ref.upload_scene(s)
        ref.upload_kdtree(nodes, lp)
        cam = s.camera(frame, 1.0)
        assert len(hit_mismatch(ctx.trace_camera(cam, 128, 128), ref.trace_camera(cam, 128, 128))) == 0
    bad = np.zeros((10, 4), np.float32)
    with pytest.raises(H.HulkError):
        ctx.update_vertices(bad, s0.box)  # vertex count differs
