"""hulk_render_frame_static (the allocation-free frame) against
hulk_render_frame, through the C ABI: two calls in a row, the second with a
smaller frame, pixels, primary hit records and counters identical, and no
device memory allocated by the calls.  Also its refusals: frames above
HULK_STATIC_MAX_RAYS without a hit buffer, the queued pipeline, and rope
slots not built yet."""
from __future__ import annotations

import numpy as np
import pytest

import hulktracer_amd as H
from hulktracer_amd import _lib as L
from tests.helpers import gpu_ctx, scene_and_tree

pytestmark = pytest.mark.gpu

OPTION_SETS = [dict(variant=H.ROPES, pipeline=3, block_size=64, leaf_batch=2, coop_lanes=16, persistent=1),
               dict(variant=H.LDSSTACK, pipeline=2, block_size=128, lds_top_levels=6, persistent=1),
               dict(variant=H.FULLSTACK, pipeline=2, block_size=256)]


def _hits(t):
    return t.cpu().numpy().view(np.uint8).view(H.HIT_DTYPE)


@pytest.mark.parametrize("opts", OPTION_SETS, ids=lambda o: f"v{o['variant']}p{o['pipeline']}")
def test_static_frame_equals_render_frame(opts):
    import torch

    s, nodes, lp = scene_and_tree(3, 6, 5, 4)
    opts = dict(opts)
    ctx = gpu_ctx(s, nodes, lp, opts.pop("variant"), **opts)
    mirror = torch.from_numpy(s.mirror).cuda()
    big = 512 * 512
    rgba_ref = torch.zeros(big, dtype=torch.int32, device="cuda")
    hits_ref = torch.zeros(big * 32, dtype=torch.uint8, device="cuda")
    rgba = torch.zeros(big, dtype=torch.int32, device="cuda")
    hits = torch.zeros(big * 32, dtype=torch.uint8, device="cuda")
    # reference frames (this also builds the rope slots on a rope context)
    cases = [(s.camera(4, 1.0), 512, 512, None), (s.camera(7, 200 / 320), 320, 200, hits)]
    refs = []
    for cam, w, h, _ in cases:
        ctx.reset_counters()
        ctx.render_frame_dev(cam, w, h, 2, mirror.data_ptr(), rgba_ref.data_ptr(), hits_ref.data_ptr(),
                             flags=L.FLAG_COUNT)
        refs.append((rgba_ref[:w * h].clone(), _hits(hits_ref[:w * h * 32]), ctx.counters()))
    refs = [(a.cpu().numpy(), b, c) for a, b, c in refs]
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    outs = []  # host copies only between the two memory readings: no device temporaries
    for rep in range(2):
        for cam, w, h, hout in cases:
            rgba.fill_(-1)
            hits.fill_(0)
            torch.cuda.synchronize()
            ctx.reset_counters()
            ctx.render_frame_static_dev(cam, w, h, 2, mirror.data_ptr(), rgba.data_ptr(),
                                        None if hout is None else hout.data_ptr(), flags=L.FLAG_COUNT)
            outs.append((rgba.cpu().numpy(), _hits(hits[:w * h * 32]), ctx.counters()))
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0, "the static frame calls allocated device memory"
    for k, (img, got, cnt) in enumerate(outs):
        (cam, w, h, hout), (rimg, rhits, rcnt) = cases[k % 2], refs[k % 2]
        assert np.array_equal(img[:w * h], rimg), f"call {k}: {w}x{h} pixels differ"
        assert np.all(img[w * h:] == -1), "wrote past the frame"
        for f in ("rays", "node_visits", "tri_tests", "hits", "shadow_rays"):
            assert cnt[f] == rcnt[f], (f, cnt[f], rcnt[f])
        if hout is not None:
            for f in ("t", "u", "v", "tri", "nx", "ny", "nz", "flags"):
                assert np.array_equal(got[f].view(np.uint32), rhits[f].view(np.uint32)), f
    ctx.close()


def test_static_frame_refusals():
    import torch

    s, nodes, lp = scene_and_tree(3, 6, 5, 4)
    rgba = torch.zeros(513 * 512, dtype=torch.int32, device="cuda")
    hits = torch.zeros(513 * 512 * 32, dtype=torch.uint8, device="cuda")
    cam = s.camera(1, 1.0)
    ctx = gpu_ctx(s, nodes, lp, H.ROPES, pipeline=2)
    with pytest.raises(H.HulkError) as e:  # rope slots not built yet
        ctx.render_frame_static_dev(cam, 64, 64, 0, None, rgba.data_ptr())
    assert e.value.code == 103
    ctx.trace_camera(cam, 8, 8)  # builds them
    ctx.render_frame_static_dev(cam, 64, 64, 0, None, rgba.data_ptr())
    with pytest.raises(H.HulkError) as e:  # above the scratch's hit capacity
        ctx.render_frame_static_dev(cam, 513, 512, 0, None, rgba.data_ptr())
    assert e.value.code == 4 and "262144" in str(e.value)
    ctx.render_frame_static_dev(cam, 513, 512, 0, None, rgba.data_ptr(), hits.data_ptr())  # with a hit buffer
    ctx.set_options(pipeline=1)
    with pytest.raises(H.HulkError) as e:
        ctx.render_frame_static_dev(cam, 64, 64, 0, None, rgba.data_ptr())
    assert e.value.code == 103
    ctx.close()
