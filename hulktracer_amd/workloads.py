"""Drivers of the benchmark workloads (one GPU, device-resident buffers).

  plain_frames    terrain scene, frame by frame, one traversal variant
  animated        1,000,002-triangle animated tori, 20 frames, CPU KD
                  rebuild per frame overlapped with the previous frame's
                  GPU render, primary + shadow + 2 mirror generations
  batched         64 frames of the random-triangle and terrain scenes,
                  resolutions 256^2-1024^2, bounce depth 0-3, one
                  graph-captured pipeline per scene, autotuned per scene
  learned         MLP bound trained on poses 0-9 of the terrain scene,
                  applied to poses 10-19, against the plain traversal
All timing uses HIP events / device synchronization around GPU work only;
CPU KD builds are timed separately, as the reference's Statistics does
(GLVisitorCPUTrace.cpp:45-47).
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib as L
from .engine import GPUTrace, KDTreeCPU, Scene


def _torch():
    import torch

    return torch


def _events(stream):
    torch = _torch()
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def make_ctx(scene, nodes, lp, stream, **opts) -> GPUTrace:
    ctx = GPUTrace(stream=stream.cuda_stream, **opts)
    ctx.upload_scene(scene)
    ctx.upload_kdtree(nodes, lp)
    return ctx


def build(scene) -> tuple[np.ndarray, np.ndarray, float]:
    kd = KDTreeCPU()
    t = time.perf_counter()
    kd.buildKDTree(scene)
    return kd.getNodeArray(), kd.getPolygonArray(), (time.perf_counter() - t) * 1e3


# ----------------------------------------------------------------- plain --

def plain_frames(ctx: GPUTrace, cams, W: int, H: int, stream, reps: int = 1, tbounds=None):
    """Render each camera (primary + shadow + shade) one frame at a time.
    Returns GPU ms for all frames x reps."""
    torch = _torch()
    n = W * H
    rgba = torch.empty(n, dtype=torch.int32, device="cuda")
    hits = torch.empty(n * 32, dtype=torch.uint8, device="cuda")
    for cam in cams[:1]:
        ctx.render_frame_dev(cam, W, H, 0, None, rgba.data_ptr(), hits.data_ptr(), flags=L.FLAG_NO_SYNC)
    torch.cuda.synchronize()
    e0, e1 = _events(stream)
    e0.record(stream)
    for _ in range(reps):
        for i, cam in enumerate(cams):
            tb = None if tbounds is None else tbounds[i].data_ptr()
            ctx.render_frame_dev(cam, W, H, 0, None, rgba.data_ptr(), hits.data_ptr(), tbound_dev=tb,
                                 flags=L.FLAG_NO_SYNC)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1)


# -------------------------------------------------------------- animated --

def animated(frames: int = 20, W: int = 1024, H: int = 1024, max_bounce: int = 2, seed: int = 7, stream=None,
             variant: int = L.FULLSTACK, log=print):
    """Config: 50 tori x 20,000 triangles + ground plane, rigid motion per
    frame, KD tree rebuilt on the CPU each frame; the build of frame k+1
    runs while the GPU renders frame k."""
    torch = _torch()
    stream = stream or torch.cuda.Stream()
    rgba = torch.empty(W * H, dtype=torch.int32, device="cuda")
    hits = torch.empty(W * H * 32, dtype=torch.uint8, device="cuda")
    scene = Scene.procedural(L.SCENE_ANIMATED_TORI, seed=seed, frame=0)
    nodes, lp, bms = build(scene)
    ctx = make_ctx(scene, nodes, lp, stream, variant=variant, persistent=True)
    mirror = torch.from_numpy(scene.mirror).cuda()
    # ray counts of frame 0 (counted run, not timed)
    cam = scene.camera(0, H / W)
    ctx.reset_counters()
    ctx.render_frame_dev(cam, W, H, max_bounce, mirror.data_ptr(), rgba.data_ptr(), hits.data_ptr(),
                         flags=L.FLAG_COUNT)
    c0 = ctx.counters()
    gpu_ms, build_ms, upload_ms, near, shadow = [], [bms], [], 0, 0
    t_wall = time.perf_counter()
    for f in range(frames):
        cam = scene.camera(f, H / W)
        e0, e1 = _events(stream)
        e0.record(stream)
        ctx.render_frame_dev(cam, W, H, max_bounce, mirror.data_ptr(), rgba.data_ptr(), hits.data_ptr(),
                             flags=L.FLAG_NO_SYNC)
        e1.record(stream)
        nxt = None
        if f + 1 < frames:  # CPU rebuild of the next frame overlaps the render
            nxt = Scene.procedural(L.SCENE_ANIMATED_TORI, seed=seed, frame=f + 1)
            nn, nl, bms = build(nxt)
            build_ms.append(bms)
        torch.cuda.synchronize()
        gpu_ms.append(e0.elapsed_time(e1))
        if nxt is not None:
            tu = time.perf_counter()
            scene = nxt
            ctx.upload_scene(scene)
            ctx.upload_kdtree(nn, nl)
            upload_ms.append((time.perf_counter() - tu) * 1e3)
    wall = time.perf_counter() - t_wall
    ctx.close()
    rays_frame = c0["rays"]
    res = {
        "triangles": int(scene.n_tris), "frames": frames, "res": f"{W}x{H}", "max_bounce": max_bounce,
        "nearest_hit_rays_frame0": int(rays_frame), "shadow_rays_frame0": int(c0["shadow_rays"]),
        "gpu_ms_per_frame": round(float(np.mean(gpu_ms)), 3),
        "cpu_kd_build_ms_per_frame": round(float(np.mean(build_ms)), 1),
        "upload_ms_per_frame": round(float(np.mean(upload_ms)) if upload_ms else 0.0, 1),
        "wall_ms_per_frame": round(wall * 1e3 / frames, 1),
        "gpu_mrays_s_nearest": round(rays_frame / np.mean(gpu_ms) / 1e3, 1),
        "gpu_mrays_s_total": round((rays_frame + c0["shadow_rays"]) / np.mean(gpu_ms) / 1e3, 1),
    }
    log(f"animated: {res}")
    return res


# --------------------------------------------------------------- batched --

def batch_specs(seed: int = 4, n: int = 64):
    """64 frames: scene (1 = random triangles, 2 = terrain), pose, square
    resolution 256..1024 (multiples of 128) and bounce depth 0..3."""
    rng = np.random.default_rng(seed)
    res = [256, 384, 512, 640, 768, 896, 1024]
    return [(int(rng.integers(1, 3)), int(rng.integers(0, 20)), int(rng.choice(res)), int(rng.integers(0, 4)))
            for _ in range(n)]


def batched(stream=None, specs=None, reps: int = 3, autotune: bool = True, log=print):
    torch = _torch()
    stream = stream or torch.cuda.Stream()
    specs = specs or batch_specs()
    out = {"frames": len(specs), "scenes": {}}
    total_plain, total_batch, total_rays = 0.0, 0.0, 0
    for kind in (1, 2):
        mine = [s for s in specs if s[0] == kind]
        if not mine:
            continue
        scene = Scene.procedural(kind, seed=1)
        nodes, lp, _ = build(scene)
        ctx = make_ctx(scene, nodes, lp, stream, variant=L.FULLSTACK)
        tuned = None
        if autotune:
            best, table = ctx.autotune(scene.camera(mine[0][1], 1.0), 512, 512, reps=2)
            tuned = best
        mirror = scene.mirror
        frames = [(scene.camera(p, 1.0), r, r, mb) for _, p, r, mb in mine]
        rays = sum(r * r for _, _, r, _ in mine)
        # plain: frame by frame through hulk_render_frame
        rgba = torch.empty(1024 * 1024, dtype=torch.int32, device="cuda")
        hits = torch.empty(1024 * 1024 * 32, dtype=torch.uint8, device="cuda")
        md = torch.from_numpy(mirror).cuda()
        for cam, w, h, mb in frames[:2]:
            ctx.render_frame_dev(cam, w, h, mb, md.data_ptr(), rgba.data_ptr(), hits.data_ptr(), flags=L.FLAG_NO_SYNC)
        torch.cuda.synchronize()
        e0, e1 = _events(stream)
        e0.record(stream)
        for _ in range(reps):
            for cam, w, h, mb in frames:
                ctx.render_frame_dev(cam, w, h, mb, md.data_ptr(), rgba.data_ptr(), hits.data_ptr(),
                                     flags=L.FLAG_NO_SYNC)
        e1.record(stream)
        torch.cuda.synchronize()
        plain_ms = e0.elapsed_time(e1) / reps
        # batched: one graph-captured pipeline for all frames of this scene
        brgba = torch.empty(rays, dtype=torch.int32, device="cuda")
        b = ctx.batch(frames, mirror, brgba.data_ptr(), use_graph=True)
        b.launch(L.FLAG_NO_SYNC)
        torch.cuda.synchronize()
        e0, e1 = _events(stream)
        e0.record(stream)
        for _ in range(reps):
            b.launch(L.FLAG_NO_SYNC)
        e1.record(stream)
        torch.cuda.synchronize()
        batch_ms = e0.elapsed_time(e1) / reps
        b.close()
        ctx.close()
        out["scenes"][kind] = {"frames": len(mine), "primary_rays": rays, "plain_ms": round(plain_ms, 3),
                               "batched_ms": round(batch_ms, 3), "tuned": tuned}
        total_plain += plain_ms
        total_batch += batch_ms
        total_rays += rays
    out["primary_rays"] = total_rays
    out["plain_mrays_s"] = round(total_rays / total_plain / 1e3, 1)
    out["batched_mrays_s"] = round(total_rays / total_batch / 1e3, 1)
    out["speedup"] = round(total_plain / total_batch, 3)
    log(f"batched: {out}")
    return out


# --------------------------------------------------------------- learned --

def learned(stream=None, W: int = 1024, H: int = 1024, n_train: int = 4_000_000, steps: int = 1500,
            quantile: float = 0.95, variants=(L.FULLSTACK, L.RESTART), log=print):
    """Train on (ray, log t) pairs of poses 0-9 of the terrain scene, bound
    poses 10-19; compare against the plain traversal of the same variant:
    hit agreement, fallback rate, node visits and Mrays/s."""
    from .learned import HitDistanceMLP, training_pairs

    torch = _torch()
    stream = stream or torch.cuda.Stream()
    scene = Scene.procedural(L.SCENE_TERRAIN_SPHERES, seed=1)
    nodes, lp, _ = build(scene)
    ctx = make_ctx(scene, nodes, lp, stream, variant=L.FULLSTACK)
    train_cams = [scene.camera(k, H / W) for k in range(10)]
    test_cams = [scene.camera(k, H / W) for k in range(10, 20)]
    t0 = time.perf_counter()
    rays, y = training_pairs(ctx, train_cams, W, H, n_samples=n_train)
    mlp = HitDistanceMLP(ctx, seed=1, max_batch=16384)
    t1 = time.perf_counter()
    losses = mlp.fit(rays, y, steps=steps, batch=16384, lr=2e-3, log_every=max(steps // 5, 1))
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    margin = mlp.calibrate(rays[: 1 << 20], y[: 1 << 20], quantile)
    n = W * H
    tb = [torch.empty(n, dtype=torch.float32, device="cuda") for _ in test_cams]
    # inference cost (all 10 poses)
    e0, e1 = _events(stream)
    e0.record(stream)
    for cam, t in zip(test_cams, tb):
        mlp.bound_camera(cam, W, H, t, sync=False)
    e1.record(stream)
    torch.cuda.synchronize()
    infer_ms = e0.elapsed_time(e1) / len(test_cams)
    res = {"train_pairs": int(rays.shape[0]), "steps": steps, "pairs_s_M": round(steps * 16384 / (t2 - t1) / 1e6, 2),
           "data_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 2), "loss": [[int(s), round(float(l), 5)] for s, l in losses],
           "margin": round(margin, 4), "infer_ms_per_frame": round(infer_ms, 3),
           "mlp_gflop_per_frame": round(n * 2 * (64 * 128 + 2 * 128 * 128 + 128) / 1e9, 2), "variants": {}}
    for v in variants:
        ctx.set_options(variant=v, persistent=1)
        agree, fb, nodes_plain, nodes_bound = 0, 0, 0, 0
        for cam, t in zip(test_cams, tb):
            plain = ctx.trace_camera(cam, W, H, count=True)
            cp = ctx.counters()
            ctx.reset_counters()
            got = ctx.trace_camera(cam, W, H, tbound=t.cpu().numpy(), count=True)
            cb = ctx.counters()
            ctx.reset_counters()
            agree += int((got["tri"] == plain["tri"]).sum() - 0)
            same = np.array_equal(got["t"].view(np.uint32), plain["t"].view(np.uint32)) and np.array_equal(
                got["tri"], plain["tri"])
            assert same, "learned bound changed a nearest hit"
            fb += int((got["flags"] & 1).sum())
            nodes_plain += cp["node_visits"]
            nodes_bound += cb["node_visits"]
        plain_ms = plain_frames(ctx, test_cams, W, H, stream)
        bound_ms = plain_frames(ctx, test_cams, W, H, stream, tbounds=tb)
        res["variants"][L.VARIANT_NAMES[v]] = {
            "hit_agreement": agree / (n * len(test_cams)), "fallback_rate": round(fb / (n * len(test_cams)), 4),
            "node_visits_saved": round(1 - nodes_bound / max(nodes_plain, 1), 4),
            "plain_frame_ms": round(plain_ms / len(test_cams), 3),
            "bounded_frame_ms": round(bound_ms / len(test_cams) + infer_ms, 3),
            "plain_mrays_s": round(n * len(test_cams) / plain_ms / 1e3, 1),
            "bounded_mrays_s": round(n * len(test_cams) / (bound_ms + infer_ms * len(test_cams)) / 1e3, 1),
        }
    log(f"learned: {res}")
    mlp.close()
    ctx.close()
    return res
