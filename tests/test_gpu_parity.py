"""GPU parity: every traversal variant through the C ABI against the CPU
oracle of the same variant, bit for bit (triangle id, t, u, v, normal).

Bar: integer / index results and every float of the hit record identical to
the oracle (the kernels restate the reference arithmetic with the same
operation order and no FMA contraction, so the tolerance is zero ulp).
"""
from __future__ import annotations

import numpy as np
import pytest

import hulktracer_amd as H
import oracle as O
from tests.helpers import describe, gpu_ctx, hit_mismatch, oracle_scene, random_rays, scene_and_tree

pytestmark = pytest.mark.gpu

VARIANTS = [H.FULLSTACK, H.SHORTSTACK, H.LDSSTACK, H.RESTART, H.ROPES]
ORACLE_OF = {H.BRUTEFORCE: O.V_BRUTE, H.FULLSTACK: O.V_FULLSTACK, H.SHORTSTACK: O.V_SHORTSTACK,
             H.LDSSTACK: O.V_LDSSTACK, H.RESTART: O.V_RESTART, H.ROPES: O.V_ROPES}


def _check(got, ref, what):
    bad = hit_mismatch(got, ref)
    assert len(bad) == 0, f"{what}: {len(bad)} of {len(ref)} rays differ\n{describe(got, ref, bad)}"


@pytest.mark.parametrize("variant", [H.BRUTEFORCE] + VARIANTS)
def test_primary_parity_random_tris(variant):
    """Config 1 scene (1,000 random triangles), 256x256 primary rays."""
    s, nodes, lp = scene_and_tree(1, 1)
    cam = s.camera(0, 1.0)
    ctx = gpu_ctx(s, nodes, lp, variant)
    got = ctx.trace_camera(cam, 256, 256)
    ref = oracle_scene(1, 1).trace_camera(ORACLE_OF[variant], cam, 256, 256, depth=6)
    _check(got, ref, f"variant {variant}")
    assert (got["tri"] != H.NO_HIT).sum() > 1000


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("pose", [0, 13])
def test_primary_parity_terrain(variant, pose):
    """Terrain + spheres scene (163,650 triangles), 384x256 primary rays."""
    s, nodes, lp = scene_and_tree(2, 1)
    cam = s.camera(pose, 256 / 384)
    ctx = gpu_ctx(s, nodes, lp, variant)
    got = ctx.trace_camera(cam, 384, 256)
    ref = oracle_scene(2, 1).trace_camera(ORACLE_OF[variant], cam, 384, 256, depth=6)
    _check(got, ref, f"variant {variant} pose {pose}")


@pytest.mark.parametrize("variant", [H.BRUTEFORCE] + VARIANTS)
def test_arbitrary_rays_parity(variant):
    """Rays from inside and outside the box, zero and axis-aligned direction
    components, origins at the box centre."""
    s, nodes, lp = scene_and_tree(1, 3, 300)
    rays = random_rays(20000, s.box, seed=variant)
    ctx = gpu_ctx(s, nodes, lp, variant)
    got = ctx.trace_rays(rays)
    ref = oracle_scene(1, 3, 300).trace_rays(ORACLE_OF[variant], rays, depth=6)
    _check(got, ref, f"variant {variant}")


@pytest.mark.parametrize("depth", [1, 4, 16])
@pytest.mark.parametrize("variant", [H.SHORTSTACK, H.LDSSTACK])
def test_short_stack_depths(variant, depth):
    s, nodes, lp = scene_and_tree(2, 1)
    cam = s.camera(5, 1.0)
    ctx = gpu_ctx(s, nodes, lp, variant, short_stack_depth=depth)
    got = ctx.trace_camera(cam, 192, 192)
    ref = oracle_scene(2, 1).trace_camera(ORACLE_OF[variant], cam, 192, 192, depth=depth)
    _check(got, ref, f"variant {variant} depth {depth}")


@pytest.mark.parametrize("opts", [dict(lds_top_levels=6), dict(lds_top_levels=10, block_size=128),
                                  dict(persistent=True), dict(persistent=True, lds_top_levels=8, block_size=128),
                                  dict(block_size=64), dict(leaf_batch=2), dict(leaf_batch=2, persistent=True)])
@pytest.mark.parametrize("variant", [H.FULLSTACK, H.LDSSTACK, H.RESTART, H.ROPES])
def test_launch_options_parity(variant, opts):
    """LDS top tree, persistent waves, block sizes and records in flight
    change only scheduling."""
    s, nodes, lp = scene_and_tree(2, 1)
    cam = s.camera(9, 1.0)
    ctx = gpu_ctx(s, nodes, lp, variant, **opts)
    got = ctx.trace_camera(cam, 256, 256)
    ref = oracle_scene(2, 1).trace_camera(ORACLE_OF[variant], cam, 256, 256, depth=6)
    _check(got, ref, f"variant {variant} {opts}")


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("variant", [H.FULLSTACK, H.RESTART])
def test_shade_parity(variant, mode):
    """Fused shading with point-light shadow rays (mode 0) and the live
    ShadePicture rule (mode 1): packed pixels identical to the oracle."""
    for kind in (1, 2):
        s, nodes, lp = scene_and_tree(kind, 1)
        cam = s.camera(3, 1.0)
        ctx = gpu_ctx(s, nodes, lp, variant)
        hits = ctx.trace_camera(cam, 256, 256)
        img = ctx.shade(cam, 256, 256, hits, mode)
        osc = oracle_scene(kind, 1)
        rh = osc.trace_camera(ORACLE_OF[variant], cam, 256, 256)
        rimg = osc.shade(ORACLE_OF[variant], cam, 256, 256, rh, mode=mode)
        diff = np.abs(img.view(np.uint8).astype(int) - rimg.view(np.uint8).astype(int))
        assert diff.max() == 0, f"scene {kind}: {int((diff > 0).sum())} channels differ, max {diff.max()} LSB"


@pytest.mark.parametrize("opts", [dict(variant=H.FULLSTACK), dict(variant=H.FULLSTACK, persistent=True),
                                  dict(variant=H.ROPES, persistent=True, block_size=64),
                                  dict(variant=H.FULLSTACK, pipeline=2),
                                  dict(variant=H.ROPES, pipeline=2, persistent=True, block_size=128),
                                  dict(variant=H.LDSSTACK, pipeline=2, leaf_batch=2, lds_top_levels=8),
                                  dict(variant=H.ROPES, pipeline=3, block_size=128),
                                  dict(variant=H.ROPES, pipeline=3, block_size=64, leaf_batch=2),
                                  dict(variant=H.ROPES, pipeline=3, block_size=128, coop_lanes=64),
                                  dict(variant=H.ROPES, pipeline=3, block_size=256, leaf_batch=2, coop_lanes=8),
                                  dict(variant=H.LDSSTACK, pipeline=3, block_size=128, lds_top_levels=8),
                                  dict(variant=H.LDSSTACK, pipeline=3, block_size=64, leaf_batch=2, coop_lanes=16,
                                       short_stack_depth=2)])
def test_render_bounces_parity(opts):
    """Animated tori (5 meshes, 100,002 triangles): primary + shadow + 2
    mirror generations with per-generation compaction; persistent waves
    also drain the device-counted shadow and mirror queues."""
    s, nodes, lp = scene_and_tree(3, 6, 5, 4)
    cam = s.camera(4, 1.0)
    opts = dict(opts)
    ctx = gpu_ctx(s, nodes, lp, opts.pop("variant"), **opts)
    img, hits = ctx.render_frame(cam, 160, 160, max_bounce=2, mirror=s.mirror)
    rimg, counts = oracle_scene(3, 6, 5, 4).render(O.V_FULLSTACK, cam, 160, 160, max_bounce=2)
    diff = np.abs(img.view(np.uint8).astype(int) - rimg.view(np.uint8).astype(int))
    assert diff.max() == 0, f"{int((diff > 0).sum())} channels differ, max {diff.max()} LSB"
    assert int(s.mirror.sum()) > 0
    flat, _ = ctx.render_frame(cam, 160, 160, max_bounce=0, mirror=s.mirror)
    assert (flat != img).sum() > 100, "mirror generations changed too few pixels"


@pytest.mark.parametrize("variant,pipeline", [(H.FULLSTACK, 2), (H.LDSSTACK, 2), (H.RESTART, 2), (H.ROPES, 2),
                                              (H.ROPES, 3), (H.LDSSTACK, 3)])
def test_fused_pipeline_equals_queued(variant, pipeline):
    """One fused launch per frame (2), or one with lane refill (3), against
    the queued launches: pixels, primary hit records and every counter
    identical, with mirror generations and a learned-style bound on the
    primary rays."""
    s, nodes, lp = scene_and_tree(3, 6, 5, 4)
    cam = s.camera(4, 1.0)
    W = H_ = 128
    plain = gpu_ctx(s, nodes, lp, variant).trace_camera(cam, W, H_)
    tb = np.where(plain["tri"] != H.NO_HIT, plain["t"] * np.float32(1.3), np.float32(0.5)).astype(np.float32)
    out = {}
    for pipe in (1, pipeline):
        ctx = gpu_ctx(s, nodes, lp, variant, pipeline=pipe, lds_top_levels=6, block_size=128)
        img, hits = ctx.render_frame(cam, W, H_, max_bounce=2, mirror=s.mirror, tbound=tb, count=True)
        out[pipe] = (img, hits, ctx.counters())
    (i1, h1, c1), (i2, h2, c2) = out[1], out[pipeline]
    assert np.array_equal(i1, i2)
    assert np.array_equal(h1.view(np.uint32), h2.view(np.uint32))
    for k in ("rays", "node_visits", "leaf_visits", "tri_tests", "hits", "fallbacks", "shadow_rays"):
        assert c1[k] == c2[k], (k, c1[k], c2[k])
    assert c1["shadow_rays"] > 0 and c1["rays"] > W * H_


@pytest.mark.parametrize("coop", [1, 8, 64])
def test_refill_cooperative_leaves(coop):
    """Lane refill (ropes and LDS short stack) with the wave testing the
    leaves of its few busy lanes together (coop_lanes = 64: on every step):
    pixels and primary hit records equal the queued pipeline's, with mirror
    generations and a bound on the primary rays, at two block sizes and both
    record batches."""
    s, nodes, lp = scene_and_tree(3, 6, 5, 4)
    cam = s.camera(4, 1.0)
    W = H_ = 128
    plain = gpu_ctx(s, nodes, lp, H.ROPES).trace_camera(cam, W, H_)
    tb = np.where(plain["tri"] != H.NO_HIT, plain["t"] * np.float32(0.9), np.float32(0.5)).astype(np.float32)
    ref = gpu_ctx(s, nodes, lp, H.ROPES, pipeline=1).render_frame(cam, W, H_, max_bounce=2, mirror=s.mirror,
                                                                  tbound=tb)
    for v, block, lb in ((H.ROPES, 64, 1), (H.ROPES, 256, 2), (H.LDSSTACK, 128, 1), (H.LDSSTACK, 256, 2)):
        ctx = gpu_ctx(s, nodes, lp, v, pipeline=3, block_size=block, leaf_batch=lb, coop_lanes=coop,
                      lds_top_levels=6)
        img, hits = ctx.render_frame(cam, W, H_, max_bounce=2, mirror=s.mirror, tbound=tb)
        assert np.array_equal(img, ref[0]), f"variant {v} block {block}: pixels differ"
        assert np.array_equal(hits.view(np.uint32), ref[1].view(np.uint32)), f"variant {v} block {block}: hits differ"
    with pytest.raises(H.HulkError):
        gpu_ctx(s, nodes, lp, H.ROPES, pipeline=3, coop_lanes=65)


@pytest.mark.parametrize("view", ["down", "outside", "inside"])
def test_refill_pushed_ropes(view):
    """The lane-refill rope traversal takes a leaf face's pushed rope only
    where the descent from the plain rope provably passes through it (from
    the second trace of a tree on, when the pushed ropes are built).  Views
    whose centre row or column of rays has exact zero direction components
    (no push for those rays), from above the scene box, from outside its
    side and from inside it (traversals from t = 0), with shadow rays and 2
    mirror generations: pixels and primary hit records equal the oracle's
    rope traversal (orc_render, orc_trace_camera), every bit."""
    s, nodes, lp = scene_and_tree(3, 6, 5, 4)
    from_, aa = {"down": ((0.0, 6.0, 0.0), (1.0, 0.0, 0.0, 90.0)),
                 "outside": ((0.0, 0.5, 6.0), (0.0, 1.0, 0.0, 0.0)),
                 "inside": ((0.25, 0.6, 0.0), (0.0, 1.0, 0.0, 0.0))}[view]
    cam = H.make_camera(np.array(from_, np.float32), aa, 45.0, 1.0)
    W = H_ = 128
    rays = O.cast_rays(cam, W, H_)
    assert (rays[:, 4:7] == 0).any(axis=1).sum() >= W  # rays the push never applies to
    osc = oracle_scene(3, 6, 5, 4)
    rimg, _ = osc.render(O.V_ROPES, cam, W, H_, max_bounce=2)
    rhits = osc.trace_camera(O.V_ROPES, cam, W, H_)
    for block, lb in ((64, 1), (256, 2)):
        ctx = gpu_ctx(s, nodes, lp, H.ROPES, pipeline=3, block_size=block, leaf_batch=lb, coop_lanes=64)
        # the first trace of a tree walks the plain ropes, the second builds
        # and takes the pushed ones
        for k in range(2):
            img, hits = ctx.render_frame(cam, W, H_, max_bounce=2, mirror=s.mirror)
            assert np.array_equal(img, rimg), f"block {block} trace {k}: {(img != rimg).sum()} pixels differ"
            bad = hit_mismatch(hits, rhits)
            assert len(bad) == 0, describe(hits, rhits, bad)


@pytest.mark.parametrize("n", [1, 63, 65, 1000, 40000])
def test_persistent_ray_counts(n):
    """Persistent waves over ray buffers of ragged sizes (fewer rays than
    one wave, than one chunk per XCD slice, than one fetch per wave)."""
    s, nodes, lp = scene_and_tree(1, 3, 300)
    rays = random_rays(n, s.box, seed=n)
    ref = oracle_scene(1, 3, 300).trace_rays(O.V_FULLSTACK, rays, depth=6)
    for block in (64, 256):
        ctx = gpu_ctx(s, nodes, lp, H.FULLSTACK, persistent=True, block_size=block)
        _check(ctx.trace_rays(rays), ref, f"persistent n={n} block {block}")


@pytest.mark.parametrize("variant", [H.FULLSTACK, H.SHORTSTACK, H.RESTART])
def test_bounded_trace_exact(variant):
    """Learned-bound traversal: bounds below, at and above the true hit
    distance (and for misses) give exactly the unbounded result; rays whose
    bounded pass missed are flagged as fallbacks."""
    s, nodes, lp = scene_and_tree(2, 1)
    cam = s.camera(11, 1.0)
    ctx = gpu_ctx(s, nodes, lp, variant)
    plain = ctx.trace_camera(cam, 256, 256)
    rng = np.random.default_rng(0)
    t = np.where(plain["tri"] != H.NO_HIT, plain["t"], 2.0).astype(np.float32)
    bound = (t * rng.choice([0.5, 0.999, 1.0, 1.01, 1.5, 4.0], size=t.shape)).astype(np.float32)
    got = ctx.trace_camera(cam, 256, 256, tbound=bound)
    _check(got, plain, f"variant {variant} bounded vs plain")
    ref = oracle_scene(2, 1).trace_camera(ORACLE_OF[variant], cam, 256, 256, tbound=bound)
    _check(got, ref, f"variant {variant} bounded vs oracle")
    fb = (got["flags"] & 1) != 0
    assert np.array_equal(fb, (ref["flags"] & 1) != 0)
    assert fb.sum() > 0 and (~fb).sum() > 0


def test_counters_match_oracle():
    """Node visits / leaf visits / triangle tests counted on the GPU equal
    the oracle's counts for the full-stack variant."""
    s, nodes, lp = scene_and_tree(2, 1)
    cam = s.camera(2, 1.0)
    ctx = gpu_ctx(s, nodes, lp, H.FULLSTACK)
    ctx.reset_counters()
    ctx.trace_camera(cam, 128, 128, count=True)
    c = ctx.counters()
    _, cnt = oracle_scene(2, 1).trace_camera(O.V_FULLSTACK, cam, 128, 128, counts=True)
    assert c["rays"] == 128 * 128
    assert c["node_visits"] == int(cnt[:, 0].sum())
    assert c["leaf_visits"] == int(cnt[:, 1].sum())
    assert c["tri_tests"] == int(cnt[:, 2].sum())


def test_full_size_terrain_1024():
    """Full size of the headline workload: terrain scene, 1024x1024 primary
    rays, full-stack variant, every ray against the oracle; plus the
    size-independent property that variants agree except at ties."""
    s, nodes, lp = scene_and_tree(2, 1)
    cam = s.camera(0, 1.0)
    ctx = gpu_ctx(s, nodes, lp, H.FULLSTACK)
    got = ctx.trace_camera(cam, 1024, 1024)
    ref = oracle_scene(2, 1).trace_camera(O.V_FULLSTACK, cam, 1024, 1024)
    _check(got, ref, "1024x1024 full stack")
    ctx.set_options(variant=H.SHORTSTACK)
    ss = ctx.trace_camera(cam, 1024, 1024)
    assert (ss["tri"] != got["tri"]).sum() <= 16, "short stack and full stack disagree beyond ties"


def test_reciprocal_selftest():
    """The kernels' 1 / x (v_rcp_f32 + one Newton step, IEEE division
    outside 2^-126 <= |x| < 2^126) equals 1.0f / x for every float."""
    import ctypes as C

    s, nodes, lp = scene_and_tree(1, 1)
    ctx = gpu_ctx(s, nodes, lp, H.FULLSTACK)
    out = (C.c_ulonglong * 4)()
    assert H.lib().hulk_selftest_reciprocal(ctx.h, out) == 0
    in_range, diff_in, diff_out, diff_exact = list(out)
    assert in_range == 2 * 252 * (1 << 23)
    assert diff_in == 0 and diff_exact == 0, (diff_in, diff_exact)
    assert diff_out > 0  # the refinement alone is not exact outside the range: the fallback is needed


# --------------------------------------------- Moller-Trumbore test mode --

@pytest.mark.parametrize("variant", [H.BRUTEFORCE] + VARIANTS)
def test_mt_primary_parity(variant):
    """tri_test = Moller-Trumbore: every hit field equals the oracle's MT
    mode bit for bit (config-1 scene, and arbitrary rays with degenerate
    directions); triangle ids agree with the reference test's except at a
    few grazes of shared edges."""
    s, nodes, lp = scene_and_tree(1, 1)
    cam = s.camera(0, 1.0)
    ctx = gpu_ctx(s, nodes, lp, variant, tri_test=H.TRI_MOLLER_TRUMBORE)
    got = ctx.trace_camera(cam, 256, 256)
    osc = oracle_scene(1, 1)
    _check(got, osc.trace_camera(ORACLE_OF[variant] | O.MT, cam, 256, 256, depth=6), f"MT variant {variant}")
    ref = osc.trace_camera(ORACLE_OF[variant], cam, 256, 256, depth=6)
    assert np.mean(got["tri"] == ref["tri"]) > 0.999
    s3, n3, l3 = scene_and_tree(1, 3, 300)
    rays = random_rays(20000, s3.box, seed=variant + 50)
    ctx3 = gpu_ctx(s3, n3, l3, variant, tri_test=H.TRI_MOLLER_TRUMBORE)
    _check(ctx3.trace_rays(rays), oracle_scene(1, 3, 300).trace_rays(ORACLE_OF[variant] | O.MT, rays, depth=6),
           f"MT rays variant {variant}")


@pytest.mark.parametrize("variant", VARIANTS)
def test_mt_terrain_parity(variant):
    """Terrain scene, MT mode, LDS top tree and persistent waves."""
    s, nodes, lp = scene_and_tree(2, 1)
    cam = s.camera(13, 1.0)
    ctx = gpu_ctx(s, nodes, lp, variant, tri_test=H.TRI_MOLLER_TRUMBORE, lds_top_levels=8, persistent=True)
    got = ctx.trace_camera(cam, 256, 256)
    _check(got, oracle_scene(2, 1).trace_camera(ORACLE_OF[variant] | O.MT, cam, 256, 256, depth=6),
           f"MT terrain variant {variant}")


@pytest.mark.parametrize("opts", [dict(variant=H.FULLSTACK, pipeline=1), dict(variant=H.LDSSTACK, pipeline=2),
                                  dict(variant=H.ROPES, pipeline=3, block_size=128),
                                  dict(variant=H.ROPES, pipeline=3, block_size=256, coop_lanes=64)])
def test_mt_render_parity(opts):
    """Whole frames with shadow rays and two mirror generations in MT mode
    (queued, fused, lane refill, cooperative leaves): pixels equal the
    oracle's MT render; the options that MT cannot take are refused."""
    s, nodes, lp = scene_and_tree(3, 6, 5, 4)
    cam = s.camera(4, 1.0)
    opts = dict(opts)
    v = opts.pop("variant")
    ctx = gpu_ctx(s, nodes, lp, v, tri_test=H.TRI_MOLLER_TRUMBORE, **opts)
    img, _ = ctx.render_frame(cam, 160, 160, max_bounce=2, mirror=s.mirror)
    rimg, _ = oracle_scene(3, 6, 5, 4).render(ORACLE_OF[v] | O.MT, cam, 160, 160, max_bounce=2)
    assert np.array_equal(img, rimg)
    with pytest.raises(H.HulkError):
        gpu_ctx(s, nodes, lp, v, tri_test=H.TRI_MOLLER_TRUMBORE, leaf_batch=2)
    with pytest.raises(H.HulkError):
        gpu_ctx(s, nodes, lp, v, tri_test=2)
    with pytest.raises(H.HulkError):
        gpu_ctx(s, nodes, lp, H.RESTART, tri_test=H.TRI_MOLLER_TRUMBORE).trace_camera_streamed(cam, 64, 64)


@pytest.mark.parametrize("pipeline", [1, 2])
def test_shadow_rays_only_toward_lit_lights(pipeline):
    """A point light the surface faces away from (n.L <= 0) adds 0 whether
    it is blocked or not, so no shadow ray is traced for it: the counted
    shadow rays equal the primary hits with n.L > 0, recomputed here with
    the kernel's float32 operations (P = o + d t, L = (pos - P) * (1 / |pos - P|),
    n.L = (nx Lx + ny Ly) + nz Lz), and the pixels still equal the oracle's,
    which traces every shadow ray."""
    s, nodes, lp = scene_and_tree(2, 1)
    cam = s.camera(11, 1.0)
    ctx = gpu_ctx(s, nodes, lp, H.LDSSTACK, pipeline=pipeline, lds_top_levels=8)
    img, hits = ctx.render_frame(cam, 192, 192, count=True)
    cnt = ctx.counters()
    rays = O.cast_rays(cam, 192, 192)
    f32 = np.float32
    hit = hits["tri"] != H.NO_HIT
    t = hits["t"].astype(f32)
    P = [rays[:, a] + rays[:, 4 + a] * t for a in range(3)]
    lit_total = 0
    for light in s.lights:
        if light[3] != 1.0:
            continue
        Lv = [f32(light[4 + a]) - P[a] for a in range(3)]
        dd = (Lv[0] * Lv[0] + Lv[1] * Lv[1]) + Lv[2] * Lv[2]
        inv = f32(1.0) / np.sqrt(dd)
        Ln = [Lv[a] * inv for a in range(3)]
        I = (hits["nx"] * Ln[0] + hits["ny"] * Ln[1]) + hits["nz"] * Ln[2]
        lit_total += int(np.sum(hit & (I > 0)))
    assert 0 < lit_total < int(hit.sum()) * sum(1 for l in s.lights if l[3] == 1.0)
    assert cnt["shadow_rays"] == lit_total
    rimg, _ = oracle_scene(2, 1).render(O.V_LDSSTACK, cam, 192, 192, max_bounce=0)
    assert np.array_equal(img, rimg)


def test_node_flag_bits():
    """Bits of a node's flags word above the axis, leaf and junction bits
    change no hit (the device copy keeps only those), and a node flagged both
    leaf and junction, which the reference's kernels read differently
    (GPUTrace.cl:34 tests the leaf bit first, GPUTrace_shortStack.cl:51 the
    junction bit), is refused."""
    s, nodes, lp = scene_and_tree(1, 3, 300)
    rays = random_rays(5000, s.box, seed=7)
    high = nodes.copy()
    high[:, 0] |= np.uint32(0xFFF0)
    for v in VARIANTS:
        ref = gpu_ctx(s, nodes, lp, v).trace_rays(rays)
        got = gpu_ctx(s, high, lp, v).trace_rays(rays)
        _check(got, ref, f"variant {v}, high flag bits")
    node_leaf, node_junction = 4, 8  # HULK_NODE_LEAF, HULK_NODE_JUNCTION
    both = nodes.copy()
    leaf = int(np.nonzero(nodes[:, 0] & node_leaf)[0][0])
    both[leaf, 0] |= np.uint32(node_junction)
    with pytest.raises(H.HulkError, match="both leaf and junction"):
        gpu_ctx(s, both, lp, H.LDSSTACK)
