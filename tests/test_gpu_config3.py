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
        # the uncounted kernel, on the context's second trace of this tree
        # (ropes: with the pushed ropes built)
        img2, hits2 = ctx.render_frame(cam, W, Hh, max_bounce=2, mirror=s.mirror)
    finally:
        ctx.close()
    rimg, rcounts, rhits = _oracle_frame(frame, ORACLE_OF[variant])
    for im, ht in ((img, hits), (img2, hits2)):
        bad = hit_mismatch(ht, rhits)
        assert len(bad) == 0, f"{len(bad)} primary hit records differ\n{describe(ht, rhits, bad)}"
        diff = im != rimg
        assert not diff.any(), f"{int(diff.sum())} of {W * Hh} pixels differ, first at {np.nonzero(diff)[0][:5]}"
    # nearest-hit rays: primary plus both mirror generations, as the oracle counts them
    assert cnt["rays"] == int(rcounts[0]), (cnt["rays"], int(rcounts[0]))
    assert cnt["rays"] > W * Hh  # the mirror generations traced rays
    assert (hits["tri"] != H.NO_HIT).mean() > 0.5
