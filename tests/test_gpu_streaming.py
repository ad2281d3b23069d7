"""Out-of-core trace (hulk_trace_camera_streamed): leaves paged into a small
device pool as rays request them, the reference's residency protocol and
relaunch loop (GPUTrace_stackless.cl:27-115, GLVisitorGPUTrace.cpp:412-502).
The hits must equal the in-core kd-restart variant's bit for bit (that
variant is pinned to the oracle by test_gpu_parity.py) for every pool size,
down to a pool that holds only the largest leaf."""
from __future__ import annotations

import numpy as np
import pytest

import hulktracer_amd as H
import oracle as O
from tests.helpers import describe, gpu_ctx, hit_mismatch, oracle_scene, scene_and_tree

pytestmark = pytest.mark.gpu


def leaf_sizes(nodes):
    leaves = nodes[:, 0] == 4
    return nodes[leaves, 2].astype(np.int64)


@pytest.mark.parametrize("kind,pose,width,height,fractions", [
    (1, 0, 128, 128, (1.0, 0.25, 0.0)),   # random 1k triangles; 0.0 = the largest leaf only
    (2, 13, 256, 192, (1.0, 0.2, 0.05)),  # terrain + spheres, 163,650 triangles
])
def test_streamed_hits_equal_in_core(kind, pose, width, height, fractions):
    s, nodes, lp = scene_and_tree(kind, 1)
    cam = s.camera(pose, height / width)
    ctx = gpu_ctx(s, nodes, lp, H.RESTART)
    ref = ctx.trace_camera(cam, width, height)
    sizes = leaf_sizes(nodes)
    for frac in fractions:
        pool = max(int(sizes.max()), int(frac * len(lp)))
        print(f"streamed kind={kind} pool={pool} of {len(lp)} records", flush=True)
        got, st = ctx.trace_camera_streamed(cam, width, height, pool)
        bad = hit_mismatch(got, ref)
        assert len(bad) == 0, f"pool {pool}: {len(bad)} rays differ\n{describe(got, ref, bad)}"
        assert st["pool_records"] == pool and st["launches"] >= 2
        assert st["bytes_loaded"] >= 64 * st["leaves_loaded"] and st["bytes_loaded"] % 64 == 0
        if pool >= len(lp):
            assert st["leaves_evicted"] == 0  # every leaf fits: each one is loaded once
            assert st["leaves_loaded"] <= int((nodes[:, 0] == 4).sum())
        else:
            assert st["leaves_evicted"] > 0
    assert (ref["tri"] != H.NO_HIT).sum() > 1000


def test_streamed_matches_oracle_restart():
    s, nodes, lp = scene_and_tree(1, 3, 300)
    cam = s.camera(0, 1.0)
    ctx = gpu_ctx(s, nodes, lp, H.FULLSTACK)
    got, st = ctx.trace_camera_streamed(cam, 96, 96, int(leaf_sizes(nodes).max()))
    ref = oracle_scene(1, 3, 300).trace_camera(O.V_RESTART, cam, 96, 96)
    bad = hit_mismatch(got, ref)
    assert len(bad) == 0, describe(got, ref, bad)


def test_streamed_after_device_build_and_repeated_calls():
    """Tables follow the tree: a device build after an upload, then calls
    with different pools on one context."""
    s, nodes, lp = scene_and_tree(3, 6, 5, 4)
    cam = s.camera(4, 1.0)
    ctx = H.GPUTrace(variant=H.RESTART)
    ctx.upload_scene(s)
    ctx.build_kdtree_device()
    ref = ctx.trace_camera(cam, 128, 128)
    for pool in (len(lp), len(lp) // 8, len(lp) // 3):
        pool = max(pool, int(leaf_sizes(nodes).max()))
        got, st = ctx.trace_camera_streamed(cam, 128, 128, pool)
        assert len(hit_mismatch(got, ref)) == 0, pool


def test_pool_smaller_than_a_leaf_is_rejected():
    s, nodes, lp = scene_and_tree(1, 1)
    ctx = gpu_ctx(s, nodes, lp, H.RESTART)
    with pytest.raises(H.HulkError) as e:
        ctx.trace_camera_streamed(s.camera(0, 1.0), 64, 64, int(leaf_sizes(nodes).max()) - 1)
    assert e.value.code == 4 and "largest leaf" in str(e.value)
