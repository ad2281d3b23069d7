"""Config 3 at its full size against the oracle.

The animated scene (50 tori x 20,000 triangles + a 2-triangle ground,
1,000,002 triangles, seed 7, 30 % of the tori mirrors), frames 0 and 11 (the
tori move rigidly per frame and the KD tree is rebuilt for each), 1024x1024,
primary rays + point-light shadow rays + 2 mirror generations, rendered by
the library with the frame pipelines the autotuner picks on this scene (lane
refill with rope traversal and cooperative leaves, two records in flight)
and the others.  Bar: every packed pixel and every primary hit record (tri,
t, u, v, normal) identical to the oracle of the same traversal variant
(orc_render, orc_trace_camera), zero ulp.  This is the reference's per-frame
flow of GLVisitorCPUTrace::traverseBegin (GLVisitorCPUTrace.cpp:52-59:
rebuild, camera, trace with shadow rays) with the mirror generations of
the animated workload added.
"""
from __future__ import annotations

import functools

import numpy as np
import pytest

import hulktracer_amd as H
import oracle as O
from tests.helpers import describe, gpu_ctx, hit_mismatch, scene_and_tree

pytestmark = pytest.mark.gpu

W = Hh = 1024
ORACLE_OF = {H.FULLSTACK: O.V_FULLSTACK, H.LDSSTACK: O.V_LDSSTACK, H.ROPES: O.V_ROPES}

OPTION_SETS = [
    # the autotuner's pick on this scene: lane refill, ropes, cooperative leaves at every leaf step
    dict(variant=H.ROPES, pipeline=3, block_size=256, leaf_batch=2, coop_lanes=64, persistent=1),
    dict(variant=H.ROPES, pipeline=3, block_size=64, leaf_batch=2, coop_lanes=16, persistent=1),
    dict(variant=H.ROPES, pipeline=3, block_size=256, leaf_batch=1, coop_lanes=4, persistent=1),
    dict(variant=H.LDSSTACK, pipeline=3, block_size=64, leaf_batch=2, coop_lanes=16, lds_top_levels=8),
    dict(variant=H.LDSSTACK, pipeline=2, block_size=64, leaf_batch=2, lds_top_levels=6, short_stack_depth=6),
    dict(variant=H.FULLSTACK, pipeline=1, block_size=128, persistent=1),
]


@functools.lru_cache(maxsize=None)
def _oracle_frame(frame: int, variant: int):
    s, nodes, lp = scene_and_tree(3, 7, 0, frame)
    osc = O.OracleScene(s.vertices, s.triangles, s.box, nodes, lp, s.lights, s.mirror)
    cam = s.camera(frame, Hh / W)
    img, counts = osc.render(variant, cam, W, Hh, max_bounce=2)
    hits = osc.trace_camera(variant, cam, W, Hh)
    return img, counts, hits


def test_config3_scene_shape():
    s, nodes, lp = scene_and_tree(3, 7, 0, 0)
    assert s.n_tris == 1_000_002
    assert 0.2 < s.mirror[:-2].mean() < 0.4  # per-mesh mirror flag, p = 0.3
    s11, _, _ = scene_and_tree(3, 7, 0, 11)
    assert not np.array_equal(s.vertices, s11.vertices)  # the tori moved
    assert np.array_equal(s.triangles, s11.triangles)


@pytest.mark.parametrize("frame", [0, 11])
@pytest.mark.parametrize("opts", OPTION_SETS, ids=lambda o: ",".join(f"{k}={v}" for k, v in o.items()))
def test_config3_frame_matches_oracle(frame, opts):
    s, nodes, lp = scene_and_tree(3, 7, 0, frame)
    cam = s.camera(frame, Hh / W)
    opts = dict(opts)
    variant = opts.pop("variant")
    print(f"config 3 frame {frame} variant {variant} {opts}", flush=True)
    ctx = gpu_ctx(s, nodes, lp, variant, **opts)
    try:
        img, hits = ctx.render_frame(cam, W, Hh, max_bounce=2, mirror=s.mirror, count=True)
        cnt = ctx.counters()
    finally:
        ctx.close()
    rimg, rcounts, rhits = _oracle_frame(frame, ORACLE_OF[variant])
    bad = hit_mismatch(hits, rhits)
    assert len(bad) == 0, f"{len(bad)} primary hit records differ\n{describe(hits, rhits, bad)}"
    diff = img != rimg
    assert not diff.any(), f"{int(diff.sum())} of {W * Hh} pixels differ, first at {np.nonzero(diff)[0][:5]}"
    # nearest-hit rays: primary plus both mirror generations, as the oracle counts them
    assert cnt["rays"] == int(rcounts[0]), (cnt["rays"], int(rcounts[0]))
    assert cnt["rays"] > W * Hh  # the mirror generations traced rays
    assert (hits["tri"] != H.NO_HIT).mean() > 0.5


def _expected_order(cost: np.ndarray) -> np.ndarray:
    """The tile order the library derives from tile costs: stable sort on a
    key of two buckets per octave (descending), rank r dealt to slice r % 8
    at position r / 8."""
    c = cost.astype(np.int64)
    msb = np.array([max(int(x).bit_length() - 1, 0) for x in c], np.int64)
    key = np.where(c < 2, c, 2 * msb + ((c >> np.maximum(msb - 1, 0)) & 1))
    ranked = np.argsort(-key, kind="stable")
    n = len(c)
    order = np.empty(n, np.int64)
    r = np.arange(n)
    order[(r % 8) * (n // 8) + r // 8] = ranked
    return order


def test_config3_tile_order_keeps_every_pixel():
    """Lane refill takes a frame's 8x8 tiles heaviest first by the previous
    frame's per-tile path costs (tile_order 0): frame 0 in pixel order, then
    frame 11 (rebuilt tree) in the order frame 0 measured, on one context.
    Frame 11 equals the oracle bit for bit, and the order is the sorted,
    slice-dealt permutation of frame 0's costs."""
    opts = dict(OPTION_SETS[0])
    variant = opts.pop("variant")
    s0, n0, lp0 = scene_and_tree(3, 7, 0, 0)
    s11, n11, lp11 = scene_and_tree(3, 7, 0, 11)
    print(f"config 3 tile order frames 0 -> 11 variant {variant} {opts}", flush=True)
    ctx = gpu_ctx(s0, n0, lp0, variant, **opts)
    try:
        ctx.render_frame(s0.camera(0, Hh / W), W, Hh, max_bounce=2, mirror=s0.mirror)
        cost0, order0 = ctx.tile_costs(W, Hh)
        ctx.upload_scene(s11)
        ctx.upload_kdtree(n11, lp11)
        img, hits = ctx.render_frame(s11.camera(11, Hh / W), W, Hh, max_bounce=2, mirror=s11.mirror, count=True)
        cnt = ctx.counters()
        cost11, _ = ctx.tile_costs(W, Hh)
    finally:
        ctx.close()
    n_tiles = (W // 8) * (Hh // 8)
    assert sorted(order0.tolist()) == list(range(n_tiles))
    assert np.array_equal(order0, _expected_order(cost0))
    per = n_tiles // 8
    heads = cost0[order0.reshape(8, per)[:, 0]]
    assert heads.min() >= np.percentile(cost0, 99)  # every slice starts on a heaviest tile
    assert cost0.sum() > 0 and cost11.sum() > 0
    rimg, rcounts, rhits = _oracle_frame(11, ORACLE_OF[variant])
    bad = hit_mismatch(hits, rhits)
    assert len(bad) == 0, f"{len(bad)} primary hit records differ\n{describe(hits, rhits, bad)}"
    diff = img != rimg
    assert not diff.any(), f"{int(diff.sum())} of {W * Hh} pixels differ, first at {np.nonzero(diff)[0][:5]}"
    assert cnt["rays"] == int(rcounts[0]), (cnt["rays"], int(rcounts[0]))


def test_tile_order_option_and_sizes():
    """tile_order 1 keeps pixel order (no cost map); a frame whose tile count
    is not a multiple of 8 measures none; the map belongs to the last size."""
    opts = dict(OPTION_SETS[0])
    variant = opts.pop("variant")
    s, nodes, lp = scene_and_tree(3, 7, 0, 0)
    cam = s.camera(0, 1.0)
    ctx = gpu_ctx(s, nodes, lp, variant, tile_order=1, **opts)
    try:
        a, _ = ctx.render_frame(cam, 128, 128, max_bounce=2, mirror=s.mirror)
        with pytest.raises(H.HulkError):
            ctx.tile_costs(128, 128)
    finally:
        ctx.close()
    ctx = gpu_ctx(s, nodes, lp, variant, **opts)
    try:
        ctx.render_frame(cam, 136, 120, max_bounce=2, mirror=s.mirror)  # 17 x 15 tiles: not a multiple of 8
        with pytest.raises(H.HulkError):
            ctx.tile_costs(136, 120)
        for _ in range(2):
            b, _ = ctx.render_frame(cam, 128, 128, max_bounce=2, mirror=s.mirror)
            assert np.array_equal(a, b)
        cost, order = ctx.tile_costs(128, 128)
        assert sorted(order.tolist()) == list(range(256)) and cost.sum() > 0
        with pytest.raises(H.HulkError):
            ctx.tile_costs(256, 256)
    finally:
        ctx.close()
