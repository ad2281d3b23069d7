"""Benchmark of the traverse-and-shade hot path on one MI355X.

Headline workload (one step): poses 10-19 of the terrain + spheres scene
(163,650 triangles, SAH-NlogN KD tree built on the CPU), each frame 1024x1024
pinhole primary rays (nearest hit), one nearest-hit shadow ray per hit and
point light, Lambert shading into RGBA8.  Inputs are resident in HBM before
the timed region; outputs stay in HBM.

Options (traversal variant, block size, stack / LDS depths, persistent
waves, fused or queued frame pipeline) come from the autotuner, timing
whole frames of pose 10 at the workload resolution; for the batched methods
its six fastest shapes are timed again on the 10-frame batched launch
(workloads.tune_batch).  Methods on the same 10 frames:
  plain    - frame after frame (hulk_render_frame)
  batched  - the 10 frames in one graph-captured launch (hulk_batch)
  learned  - batched, with per-pixel bounds from the MLP trained on poses
             0-9 (inference inside the timed region; training is not)
Every method's primary hits are compared with the plain trace's on every
ray.  value = nearest-hit (primary) rays per second of the fastest method
whose hits match 100%.  Extras: every variant's primary-trace rate, the
roofline of the batched frame kernel, the reference CPU path (port) on the
host, and the animated 1M-triangle and 64-frame batched workloads.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


OPT_KEYS = ("variant", "block_size", "short_stack_depth", "lds_top_levels", "persistent", "pipeline", "leaf_batch",
            "coop_lanes")


def kernel_name(opts: dict) -> str:
    """The batched launch's dominant kernel: the fused frame kernel, the
    lane-refill frame kernel, or the queued pipeline's graph of trace,
    compaction and shading kernels."""
    lb = 2 if int(opts.get("leaf_batch", 1)) == 2 else 1
    if int(opts.get("pipeline", 0)) == 1:
        return "queued pipeline graph"
    if int(opts.get("pipeline", 0)) == 3:
        return f"k_render_refill<{int(opts['variant'])},false,{lb}>"
    return f"k_trace<{int(opts['variant'])},false,true,{lb}>"


def profiled_traffic(kernel: str, opts: dict, W: int, H: int):
    """HBM bytes per launch of the batched frame kernel, as measured by the
    FETCH_SIZE and WRITE_SIZE counter runs whose summary is kept in
    profiles/ (line: traffic <kernel> <options> <W>x<H> <bytes>).  A profile
    of the same kernel instance (variant, frame kernel, records in flight) at
    the same resolution counts: its bytes are the hit records and pixels
    written plus the scene reads that miss L2 / MALL, which the launch shape
    (block size, LDS depth) hardly changes; the profiled options are named
    in the source field.  Else None."""
    want = ",".join(f"{k}={int(opts.get(k, 0))}" for k in OPT_KEYS)
    found = None
    for f in sorted((ROOT / "profiles").glob("*traffic*.txt")):
        for line in f.read_text().splitlines():
            parts = line.split()
            if len(parts) == 5 and parts[0] == "traffic" and parts[1] == kernel and parts[3] == f"{W}x{H}":
                if found is None or parts[2] == want:
                    found = (int(parts[4]), f"profiles/{f.name} (profiled options {parts[2]})")
    if found:
        return found
    return None, f"no FETCH_SIZE/WRITE_SIZE profile of {kernel} at {W}x{H}"


def kd_build_only(args) -> int:
    """--kd-build-only: warmup + steps device KD builds of one scene in one
    context; prints the wall time per build (the call returns once the tree
    is installed) and the host builder's time on the same scene."""
    import hulktracer_amd as H
    from hulktracer_amd import _lib as L
    from hulktracer_amd import workloads as WL

    if args.kernel_scene == "animated":
        scene = H.Scene.procedural(L.SCENE_ANIMATED_TORI, seed=7, frame=0)
    else:
        scene = H.Scene.procedural(L.SCENE_TERRAIN_SPHERES, seed=1)
    ctx = H.GPUTrace()
    ctx.upload_scene(scene)
    for _ in range(args.warmup):
        ctx.build_kdtree_device()
    t = time.perf_counter()
    for _ in range(args.steps):
        ctx.build_kdtree_device()
    dev_ms = (time.perf_counter() - t) * 1e3 / max(1, args.steps)
    info = ctx.device_build_info()
    nodes, lp = ctx.download_kdtree()
    ctx.close()
    _, _, host_ms = WL.build(scene)
    print(json.dumps({"kd_build": "device", "scene": args.kernel_scene, "triangles": scene.n_tris,
                      "nodes": int(len(nodes)), "leaf_refs": int(len(lp)), "device_ms_per_build": round(dev_ms, 3),
                      "host_builder_ms": round(host_ms, 1), "builds": args.steps, **info}), flush=True)
    return 0


def animated_kernel(args, W: int, H: int) -> int:
    """--kernel-only on frame 0 of the animated scene (50 tori + ground, 1,000,002
    triangles, 2 mirror generations): warmup + steps renders, HIP events."""
    import torch

    import hulktracer_amd as HT
    from hulktracer_amd import _lib as L
    from hulktracer_amd import workloads as WL

    opts = {k: int(v) for k, v in (kv.split("=") for kv in args.kernel_only.split(","))}
    scene = HT.Scene.procedural(L.SCENE_ANIMATED_TORI, seed=7, frame=0)
    nodes, lp, _ = WL.build(scene)
    stream = torch.cuda.Stream()
    ctx = WL.make_ctx(scene, nodes, lp, stream, **opts)
    mirror = torch.from_numpy(scene.mirror).cuda()
    rgba = torch.empty(W * H, dtype=torch.int32, device="cuda")
    hits = torch.empty(W * H * 32, dtype=torch.uint8, device="cuda")
    cam = scene.camera(0, H / W)

    def frame():
        ctx.render_frame_dev(cam, W, H, 2, mirror.data_ptr(), rgba.data_ptr(), hits.data_ptr(), flags=L.FLAG_NO_SYNC)

    for _ in range(max(args.warmup, 1)):
        frame()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        frame()
    e1.record(stream)
    torch.cuda.synchronize()
    print(json.dumps({"kernel": kernel_name(opts), "scene": "animated_1M_frame0", "options": opts,
                      "kernel_ms": round(e0.elapsed_time(e1) / args.steps, 4), "timed_launches": args.steps,
                      "launches": args.steps + max(args.warmup, 1), "primary_rays_per_launch": W * H}), flush=True)
    ctx.close()
    return 0


def compare_methods(ctx, scene, cams, train_cams, W: int, H: int, stream, args, best: dict, bbest: dict, log):
    """plain / batched / learned-bound + batched on `cams` (bounds from an MLP
    trained on `train_cams`), for the tuned options and for kd-restart.
    plain runs with `best` (tuned on single frames), the batched methods with
    `bbest` (tuned on the batched launch).  Returns (methods, learned info,
    batched-launch ms of the tuned options)."""
    import numpy as np
    import torch

    import hulktracer_amd as HT
    from hulktracer_amd import _lib as L

    n = W * H
    nf = len(cams)
    dev = torch.device("cuda:0")
    hits = torch.empty(n * 32, dtype=torch.uint8, device=dev)
    rgba = torch.empty(n, dtype=torch.int32, device=dev)
    frames = [(cam, W, H, 0) for cam in cams]
    brgba = torch.empty(nf * n, dtype=torch.int32, device=dev)
    ref_hits = [ctx.trace_camera(cam, W, H) for cam in cams]
    methods = {}

    def timed(fn, reps):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps

    def time_batch(batch, reps):
        batch.launch(L.FLAG_NO_SYNC)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            batch.launch(L.FLAG_NO_SYNC)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    def hits_match(h, ref=None):
        ref = ref_hits if ref is None else ref
        return float(np.mean([np.array_equal(h[k * n:(k + 1) * n]["tri"], ref[k]["tri"]) and
                              np.array_equal(h[k * n:(k + 1) * n]["t"].view(np.uint32),
                                             ref[k]["t"].view(np.uint32)) for k in range(nf)]))

    def entry(sec, match=1.0):
        return {"ms_per_step": round(sec * 1e3, 3), "mrays_s": round(nf * n / sec / 1e6, 1), "hit_match": match}

    # plain: frame after frame
    def plain_step():
        for cam in cams:
            ctx.render_frame_dev(cam, W, H, 0, None, rgba.data_ptr(), hits.data_ptr(), flags=L.FLAG_NO_SYNC)

    methods["plain"] = entry(timed(plain_step, args.steps))
    # batched: one graph for the frames
    bopts = {k: int(v) for k, v in bbest.items() if k != "device"}
    ctx.set_options(**bopts)
    cross = None
    if bopts["variant"] != int(best["variant"]):
        # another traversal variant: its own plain trace is the reference
        # (variants may break exact ties differently); the agreement of the
        # two variants' triangle ids is reported beside it
        plain_hits = ref_hits
        ref_hits = [ctx.trace_camera(cam, W, H) for cam in cams]
        cross = float(np.mean([np.mean(a["tri"] == b["tri"]) for a, b in zip(plain_hits, ref_hits)]))
    batch = ctx.batch(frames, None, brgba.data_ptr(), use_graph=True)
    methods["batched"] = entry(timed(lambda: batch.launch(L.FLAG_NO_SYNC), args.steps),
                               hits_match(batch.primary_hits()))
    if cross is not None:
        methods["batched"]["tri_agreement_with_plain_variant"] = round(cross, 7)
    # the batched launch is the dominant kernel (the whole step when the
    # pipeline is fused): live HIP events on its stream
    k_ms = time_batch(batch, max(args.steps, 10))
    batch.close()

    # learned + batched, on the tuned variant and on the reference's live
    # kernel (kd-restart, GPUTrace_stackless: no exit before the box exit,
    # so a bound is what lets it stop early)
    learned_info = None
    if not args.no_learned:
        from hulktracer_amd.learned import HitDistanceMLP, select_margin, training_pairs

        tl = time.time()
        # train on every other pose; the rest calibrate the margin (frames the network has not seen)
        fit_cams, calib_cams = train_cams[0::2], train_cams[1::2]
        rays, y = training_pairs(ctx, fit_cams, W, H, n_samples=4_000_000)
        n_pairs = int(rays.shape[0])
        mlp = HitDistanceMLP(ctx, seed=1, max_batch=16384)
        tt = time.time()
        losses = mlp.fit(rays, y, steps=args.train_steps, batch=16384, lr=2e-3,
                         log_every=max(args.train_steps // 4, 1))
        torch.cuda.synchronize()
        train_s = time.time() - tt
        del rays, y
        mlp.stride = 2  # one prediction per 2x2 pixels (select_margin keeps it)
        tb = torch.empty(nf * n, dtype=torch.float32, device=dev)

        def bounds():
            for k, cam in enumerate(cams):
                mlp.bound_camera(cam, W, H, tb[k * n:(k + 1) * n], sync=False)

        bounds()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        bounds()
        e1.record(stream)
        torch.cuda.synchronize()
        infer_ms = e0.elapsed_time(e1) / nf
        learned_info = {"train_pairs": n_pairs, "train_steps": args.train_steps, "train_s": round(train_s, 2),
                        "data_s": round(tt - tl, 2), "bound_stride": 2,
                        "inference_ms_per_frame": round(infer_ms, 4),
                        "inference_tflops": round(n / 4 * 82176 / (infer_ms * 1e-3) / 1e12, 2),
                        "loss": [[int(a), round(float(b), 4)] for a, b in losses], "margin_selection": {}}
        opts_now = bopts
        # kd-restart runs in the fused frame kernel (lane refill is a rope-traversal pipeline)
        restart_opts = dict(opts_now, variant=HT.RESTART,
                            pipeline=2 if opts_now.get("pipeline") == 3 else opts_now.get("pipeline", 2))
        for tag, vopts in (("", opts_now), ("_stackless", restart_opts)):
            ctx.set_options(**vopts)
            # margin (and whether to bound at all) chosen on training frames for this traversal
            sel = select_margin(mlp, ctx, calib_cams, W, H, stride=2)
            learned_info["margin_selection"][tag or "tuned"] = sel
            ref_v = [ctx.trace_camera(cam, W, H) for cam in cams] if tag else ref_hits
            if tag:  # the reference's live kernel: plain frame by frame, and batched without a bound
                methods["plain" + tag] = entry(timed(plain_step, max(1, args.steps // 5)))
                pbatch = ctx.batch(frames, None, brgba.data_ptr(), use_graph=True)
                methods["batched" + tag] = entry(timed(lambda: pbatch.launch(L.FLAG_NO_SYNC),
                                                       max(1, args.steps // 2)),
                                                 hits_match(pbatch.primary_hits(), ref_v))
                pbatch.close()
            use_bound = mlp.enabled
            lbatch = ctx.batch(frames, None, brgba.data_ptr(), tbound_dev=tb.data_ptr() if use_bound else None,
                               use_graph=True)

            def learned_step():
                if use_bound:  # a bound that saves no work is not computed (select_margin)
                    bounds()
                lbatch.launch(L.FLAG_NO_SYNC)

            if use_bound:
                bounds()
            s_learn = timed(learned_step, args.steps)
            lh = lbatch.primary_hits()
            fb = float(np.mean((lh["flags"] & 1) != 0))
            bounds()  # node visits with the selected margin, reported also where the bound is not used
            torch.cuda.synchronize()
            ctx.reset_counters()
            ctx.trace_camera(cams[0], W, H, count=True)
            cp = ctx.counters()["node_visits"]
            ctx.reset_counters()
            ctx.trace_camera(cams[0], W, H, tbound=tb[:n].cpu().numpy(), count=True)
            cb = ctx.counters()["node_visits"]
            methods["learned+batched" + tag] = dict(entry(s_learn, hits_match(lh, ref_v)),
                                                    bound_used=bool(use_bound), fallback_rate=round(fb, 4),
                                                    node_visits_saved=round(1 - cb / cp, 4))
            lbatch.close()
        ctx.set_options(**opts_now)
        mlp.close()
    return methods, learned_info, k_ms


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--res", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--train-steps", type=int, default=1500)
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only run a device copy of this many MiB (HBM traffic counter calibration)")
    ap.add_argument("--kernel-only", default="",
                    help="only time the primary trace kernel of pose 10 with these options "
                         "(e.g. variant=5,block_size=64,short_stack_depth=6,lds_top_levels=0,persistent=0): "
                         "warmup + steps launches, for profiler runs")
    ap.add_argument("--kd-build-only", action="store_true",
                    help="only time the device KD-tree build of --kernel-scene (terrain, or frame 0 of the "
                         "animated scene): warmup + steps builds, wall clock per build, for profiler runs")
    ap.add_argument("--kernel-scene", default="terrain", choices=("terrain", "animated"),
                    help="--kernel-only workload: the 10-frame terrain batch, or frame 0 of the animated "
                         "1M-triangle scene (1024x1024, 2 mirror generations, one render per launch)")
    args = ap.parse_args()
    if args.gpus != 1:
        log("bench.py measures one GPU only: the library has no multi-GPU path (rays would be split "
            "per GPU by image tiles or frames, with the scene replicated). Run with --gpus 1.")
        return 2

    import numpy as np
    import torch

    if args.copy_calibration:
        nbytes = args.copy_calibration << 20
        a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": nbytes, "launches": 5}))
        return 0

    import hulktracer_amd as H
    from hulktracer_amd import _lib as L
    from hulktracer_amd import workloads as WL

    W = Hh = args.res
    n = W * Hh
    if args.kd_build_only:
        return kd_build_only(args)
    if args.kernel_only and args.kernel_scene == "animated":
        return animated_kernel(args, W, Hh)
    t0 = time.time()
    scene = H.Scene.procedural(L.SCENE_TERRAIN_SPHERES, seed=1)
    nodes, lp, build_ms = WL.build(scene)
    cams = [scene.camera(k, Hh / W) for k in range(10, 20)]
    nf = len(cams)
    log(f"scene {scene.n_tris} tris, {len(nodes)} nodes, {len(lp)} leaf refs, KD build {build_ms:.0f} ms "
        f"(setup {time.time() - t0:.1f}s)")

    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)  # default priority; library + events on it
    hits = torch.empty(n * 32, dtype=torch.uint8, device=dev)
    rgba = torch.empty(n, dtype=torch.int32, device=dev)

    def time_trace(ctx, cam, reps, tb=None):
        ctx.trace_camera_dev(cam, W, Hh, hits.data_ptr(), tbound_dev=tb, flags=L.FLAG_NO_SYNC)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            ctx.trace_camera_dev(cam, W, Hh, hits.data_ptr(), tbound_dev=tb, flags=L.FLAG_NO_SYNC)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    frames = [(cam, W, Hh, 0) for cam in cams]
    brgba = torch.empty(nf * n, dtype=torch.int32, device=dev)

    def time_batch(batch, reps):
        batch.launch(L.FLAG_NO_SYNC)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            batch.launch(L.FLAG_NO_SYNC)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    if args.kernel_only:
        # the batched launch of the 10 frames: one kernel when the pipeline is fused
        opts = {k: int(v) for k, v in (kv.split("=") for kv in args.kernel_only.split(","))}
        ctx = WL.make_ctx(scene, nodes, lp, stream, **opts)
        batch = ctx.batch(frames, None, brgba.data_ptr(), use_graph=True)
        for _ in range(max(args.warmup - 1, 0)):
            batch.launch(L.FLAG_NO_SYNC)
        ms = time_batch(batch, args.steps)
        batch.close()
        ctx.close()
        print(json.dumps({"kernel": kernel_name(opts), "options": opts, "kernel_ms": round(ms, 4),
                          "timed_launches": args.steps, "launches": args.steps + max(args.warmup, 1),
                          "primary_rays_per_launch": nf * n}), flush=True)
        return 0

    # -- every variant at its default shape (primary trace of pose 10)
    shapes = {"GPUTrace": dict(variant=H.FULLSTACK), "GPUTrace_shortStack": dict(variant=H.SHORTSTACK),
              "GPUTrace_LDSshortStack": dict(variant=H.LDSSTACK, lds_top_levels=8),
              "GPUTrace_stackless": dict(variant=H.RESTART), "GPUTrace_ropes": dict(variant=H.ROPES)}
    variants = {}
    for name, kw in shapes.items():
        ctx = WL.make_ctx(scene, nodes, lp, stream, **kw)
        ms = time_trace(ctx, cams[0], 5 if kw["variant"] != H.RESTART else 2)
        variants[name] = {"primary_trace_ms": round(ms, 4), "mrays_s": round(n / ms / 1e3, 1)}
        log(f"  {name:24s} primary trace {ms:8.3f} ms  {n / ms / 1e3:9.1f} Mrays/s")
        ctx.close()

    # -- autotuned options for this scene: whole frames at the workload resolution
    ctx = WL.make_ctx(scene, nodes, lp, stream, variant=H.FULLSTACK)
    tt = time.time()
    best, table = ctx.autotune(cams[0], W, Hh, reps=2, frame=True)
    log(f"autotune {time.time() - tt:.1f}s over {len(table)} shapes -> {best}")
    # batch stage: the fastest shapes again, on the 10-frame graph launch itself
    tt = time.time()
    bbest, btimes = WL.tune_batch(ctx, frames, None, brgba.data_ptr(), table, stream)
    log(f"batch tune {time.time() - tt:.1f}s over {len(btimes)} shapes {btimes} -> {bbest}")
    ctx.set_options(**{k: int(v) for k, v in best.items() if k != "device"})

    # -- algorithmic bytes of one batched launch (10 frames): full-stack
    # counts of every traced ray (primary and shadow), 32 B ray + 16 B
    # result each, 16 B per node visit, 64 B per triangle test
    cnt = WL.make_ctx(scene, nodes, lp, stream, variant=H.FULLSTACK)
    cbatch = cnt.batch(frames, None, brgba.data_ptr(), use_graph=False)
    cbatch.launch(L.FLAG_COUNT)
    c = cnt.counters()
    cbatch.close()
    cnt.close()
    traced = c["rays"] + c["shadow_rays"]
    bytes_launch = 48 * traced + 16 * c["node_visits"] + 64 * c["tri_tests"]
    n_node, n_tri = c["node_visits"] / traced, c["tri_tests"] / traced

    ctx.reset_counters()
    ctx.render_frame(cams[0], W, Hh, count=True)
    shadow_per_frame = ctx.counters()["shadow_rays"]
    train_cams = [scene.camera(k, Hh / W) for k in range(10)]
    methods, learned_info, k_ms = compare_methods(ctx, scene, cams, train_cams, W, Hh, stream, args, best, bbest,
                                                  log)
    for k, v in methods.items():
        log(f"method {k:16s} {v}")

    ok = {k: v for k, v in methods.items() if v["hit_match"] == 1.0}
    best_m = max(ok, key=lambda k: ok[k]["mrays_s"])
    s_best = methods[best_m]["ms_per_step"] / 1e3

    achieved = bytes_launch / (k_ms * 1e-3) / 1e9
    log(f"batched launch {k_ms:.3f} ms, {bytes_launch / 1e9:.2f} GB algorithmic ({traced} rays, "
        f"nodes {n_node:.2f} / tris {n_tri:.2f} per ray), {achieved:.0f} GB/s")
    bat = best_m != "plain"
    tuned = bbest if bat else best
    tuned_name = L.VARIANT_NAMES[tuned["variant"]]
    kname = kernel_name(bbest)
    traffic, traffic_src = profiled_traffic(kname, bbest, W, Hh)

    result = {
        "metric": "Mrays/s nearest-hit @1024^2, learned-bound+batched vs plain, 100% hit match",
        "value": round(nf * n / s_best / 1e6, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(s_best * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": f"terrain_spheres_163650tris_poses10-19_{W}x{Hh}_primary+shadow+shade",
                   "frames_per_step": nf, "method": best_m, "variant": tuned_name,
                   "options": {k: int(v) for k, v in tuned.items()},
                   "options_single_frame": {k: int(v) for k, v in best.items()},
                   "batch_tune_ms": btimes, "kd_build_ms": round(build_ms, 1)},
        "methods": methods,
        "speedup_vs_plain": round(methods[best_m]["mrays_s"] / methods["plain"]["mrays_s"], 3),
        "shadow_rays_per_frame": int(shadow_per_frame),
        "variants_primary_trace": variants,
        "learned": learned_info,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": kname, "kernel_ms": round(k_ms, 4), "launch": "10-frame batch (graph)",
                     "primary_rays_per_launch": nf * n, "traced_rays_per_launch": int(traced),
                     "algorithmic_bytes_per_launch": int(bytes_launch),
                     "node_visits_per_ray": round(n_node, 3), "tri_tests_per_ray": round(n_tri, 3)},
        "cpu_baseline": None,
    }
    ctx.close()

    if not args.no_cpu:
        import oracle as O

        threads = min(16, os.cpu_count() or 1)
        osc = O.OracleScene(scene.vertices, scene.triangles, scene.box, nodes, lp, scene.lights)
        osc.cpu_trace_scene(cams[0], W, Hh, 0, 8, threads=threads, out=False)
        frames_done, rays_total, tc = 0, 0, time.perf_counter()
        while True:
            rays_total += osc.cpu_trace_scene(cams[frames_done % nf], W, Hh, threads=threads, out=False)[0]
            frames_done += 1
            if time.perf_counter() - tc > args.cpu_seconds or frames_done >= 2000:
                break
        cpu_s = time.perf_counter() - tc
        result["cpu_baseline"] = {
            "value": round(n * frames_done / cpu_s / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{frames_done} whole {W}x{Hh} frames of poses 10-19 (primary + shadow + shade), "
                      f"{cpu_s:.1f}s, {rays_total / cpu_s / 1e6:.2f} M total rays/s incl. shadow rays, OpenMP "
                      f"over rows, host has {os.cpu_count()} cpus",
        }
        log(f"cpu baseline {result['cpu_baseline']}")

    if not args.no_extras:
        extras = {}
        def methods_1m():
            # the same comparison at 1M triangles: geometry of frame 0 of the
            # animated scene, its camera path poses 0-9 train, 10-19 measured
            sc = H.Scene.procedural(L.SCENE_ANIMATED_TORI, seed=7, frame=0)
            nd, lpp, bms = WL.build(sc)
            cx = WL.make_ctx(sc, nd, lpp, stream, variant=H.FULLSTACK)
            cams1 = [sc.camera(k, Hh / W) for k in range(10, 20)]
            b1, t1 = cx.autotune(cams1[0], W, Hh, reps=2, frame=True)
            fr1 = [(cam, W, Hh, 0) for cam in cams1]
            bb1, _ = WL.tune_batch(cx, fr1, None, brgba.data_ptr(), t1, stream)
            cx.set_options(**{k: int(v) for k, v in b1.items() if k != "device"})
            m1, li1, _ = compare_methods(cx, sc, cams1, [sc.camera(k, Hh / W) for k in range(10)], W, Hh, stream,
                                         args, b1, bb1, log)
            cx.close()
            for k, v in m1.items():
                log(f"1M method {k:16s} {v}")
            return {"workload": f"animated_tori_{sc.n_tris}tris_frame0_poses10-19_{W}x{Hh}_primary+shadow+shade",
                    "kd_build_ms": round(bms, 1), "tuned": b1, "tuned_batch": bb1, "methods": m1, "learned": li1}

        def mt_compare():
            # the optional Moller-Trumbore test against the reference's on the
            # headline frames: batched frame time, and triangle ids of the
            # primary hits; every ray whose ids differ is classified with the
            # FP64 twin of the reference test: a tie (the two triangles' t
            # within 1e-4 relative, the FP32 error of t at grazing angles), or
            # a hit point within 1e-4 of an edge of either triangle
            import numpy as np

            import oracle as O

            opts = {k: int(v) for k, v in bbest.items() if k != "device"}
            if opts.get("leaf_batch") == 2:
                opts["leaf_batch"] = 1
            out = {}
            hits_by = {}
            for tag, tt in (("reference", H.TRI_REFERENCE), ("moller_trumbore", H.TRI_MOLLER_TRUMBORE)):
                cx = WL.make_ctx(scene, nodes, lp, stream, **dict(opts, tri_test=tt))
                b = cx.batch(frames, None, brgba.data_ptr(), use_graph=True)
                ms = time_batch(b, max(args.steps, 5))
                hits_by[tag] = b.primary_hits()
                b.close()
                cx.close()
                out[tag] = {"batch_ms": round(ms, 4), "mrays_s": round(nf * n / ms / 1e3, 1)}
            a, m = hits_by["reference"], hits_by["moller_trumbore"]
            diff = np.nonzero(a["tri"] != m["tri"])[0]
            V = scene.vertices[:, :3].astype(np.float32)
            T = scene.triangles
            tie = edge = other = 0
            rays_all = {}
            for i in diff:
                f, px = divmod(int(i), n)
                if f not in rays_all:
                    rays_all[f] = O.cast_rays(cams[f], W, Hh).reshape(-1, 8)
                r = rays_all[f][px]
                dd = []
                for tri in (a["tri"][i], m["tri"][i]):
                    if tri == H.NO_HIT:
                        dd.append(None)
                        continue
                    p = V[T[tri, :3]]
                    dd.append(O.tri_test_f64(r, p[0], p[1], p[2]))
                if all(d is not None for d in dd) and abs(dd[0][0] - dd[1][0]) <= 1e-4 * abs(dd[0][0]):
                    tie += 1
                elif any(d is not None and min(d[1], d[2], 1 - d[1] - d[2]) < 1e-4 for d in dd):
                    edge += 1
                else:
                    other += 1
            out["primary_rays"] = int(len(a))
            out["tri_id_differences"] = {"total": int(len(diff)), "tie_in_t": tie, "edge_graze": edge,
                                         "unclassified": other}
            both = (a["tri"] == m["tri"]) & (a["tri"] != H.NO_HIT)
            out["t_rel_diff_max_same_tri"] = float(np.max(np.abs(a["t"][both] - m["t"][both]) / a["t"][both]))
            out["options"] = opts
            log(f"triangle test: {out}")
            return out

        def streamed():
            # out-of-core: terrain pose 10 at 1024^2, kd-restart, record pool
            # of 100 / 25 / 5 % of the leaf references against the in-core kernel
            sc = scene
            cx = WL.make_ctx(sc, nodes, lp, stream, variant=H.RESTART)
            cam = cams[0]
            cx.trace_camera(cam, W, Hh)
            t0 = time.perf_counter()
            ref = cx.trace_camera(cam, W, Hh)
            core_ms = (time.perf_counter() - t0) * 1e3
            out = {"workload": f"terrain_{sc.n_tris}tris_pose10_{W}x{Hh}_primary_kd-restart",
                   "leaf_refs": int(len(lp)), "in_core_ms": round(core_ms, 2)}
            for frac in (1.0, 0.25, 0.05):
                pool = max(int(frac * len(lp)), int(nodes[nodes[:, 0] == 4, 2].max()))
                got, st = cx.trace_camera_streamed(cam, W, Hh, pool)
                same = bool(np.array_equal(got.view(np.uint32).reshape(-1, 8)[:, :7],
                                           ref.view(np.uint32).reshape(-1, 8)[:, :7]))
                out[f"pool_{int(frac * 100)}pct"] = {
                    "pool_records": pool, "ms": round(st["ms"], 2), "launches": st["launches"],
                    "leaves_loaded": st["leaves_loaded"], "leaves_evicted": st["leaves_evicted"],
                    "mb_loaded": round(st["bytes_loaded"] / 1e6, 1), "mrays_s": round(n / st["ms"] / 1e3, 1),
                    "hits_equal_in_core": same}
            cx.close()
            log(f"streamed: {out}")
            return out

        for name, fn in (("triangle_test_mt", mt_compare), ("streamed_out_of_core", streamed),
                         ("methods_1M", methods_1m),
                         ("animated_1M", lambda: WL.animated(frames=20, stream=stream, log=log)),
                         ("animated_1M_device_build",
                          lambda: WL.animated_device_build(frames=20, stream=stream, log=log)),
                         ("batched_64_frames", lambda: WL.batched(stream=stream, log=log))):
            try:
                extras[name] = fn()
            except Exception as e:  # an extra must not hide the headline line
                extras[name] = {"error": repr(e)}
                log(f"{name} failed: {e!r}")
        result["extras"] = extras

    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
