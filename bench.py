"""Benchmark of the traverse-and-shade hot path on one MI355X.

Headline workload (one step): the 1,000,002-triangle animated-tori scene
(50 tori x 20,000 triangles + a 2-triangle ground, seed 7) at the geometry of
its frame 0, camera poses 10-19, each frame 1024x1024 pinhole primary rays
(nearest hit), one nearest-hit shadow ray per hit and lit point light,
Lambert shading into RGBA8: the reference's per-frame trace
(GLVisitorCPUTrace.cpp:149-200; GPU twin GPUTrace.cl:92-161) at 1M triangles
and 1024^2 rays.  The SAH-NlogN KD tree is built on the host (timed
separately).  Inputs are resident in HBM before the timed region; outputs
stay in HBM.

Options (traversal variant, block size, stack / LDS depths, persistent
waves, frame pipeline, records in flight) come from the autotuner timing
whole frames of pose 10; for the batched methods its six fastest shapes are
timed again on the 10-frame batched launch (workloads.tune_batch).  Methods
on the same 10 frames:
  plain    - frame after frame (hulk_render_frame)
  batched  - the 10 frames in one graph-captured launch (hulk_batch)
  learned  - batched, with per-pixel bounds from the MLP trained on poses
             0-9 (inference inside the timed region; training is not)
and the same three on kd-restart, the reference's live kernel
(GPUTrace_stackless.cl).  Every method's primary hit records and pixels are
compared with the CPU oracle of the same traversal variant on every ray of
every frame (oracle_hit_match, oracle_pixel_match).  value = primary
(nearest-hit) rays per second of the fastest method whose every hit record
and pixel equals the oracle's.

The roofline object prices the dominant kernel (the batched launch) against
the resource its counters show closest to saturation (rocprofv3 summaries in
profiles/, matched to the exact kernel, options and workload): VALU issue
(wave64 instructions per CU per cycle) and HBM bytes, each divided by the
live HIP-event time of the launch measured here.  The algorithmic bytes per
ray, after the reference's traversal counters (Statistics.hpp:58-60,112-120:
32 + 16 + 16 per node visit + 64 per triangle test, counted by a full-stack
pass), are reported as a cache-served byte rate, not as an HBM fraction: the
scene stays in L2 / MALL.

The CPU baseline is the oracle's port of the reference CPU path
(GLVisitorCPUTrace::traceScene + KDTreeCPU::checkCollision, OpenMP over rows)
on the same frames, on the host cores granted to this process.

Secondary lines (extras): the terrain scene (163,650 triangles) with the
same methods, the animated 20-frame workload with mirror generations, the
device KD build, the 64-frame batched workload, the Moller-Trumbore test and
the out-of-core trace.
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
# one wave64 VALU instruction per CU per cycle (four SIMD16 units, four
# cycles each); 256 CUs at the 2.4 GHz peak engine clock
VALU_PEAK_GINSTR_S = 256 * 2.4
HEADLINE_POSES = list(range(10, 20))
TRAIN_POSES = list(range(10))


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


OPT_KEYS = ("variant", "block_size", "short_stack_depth", "lds_top_levels", "persistent", "pipeline", "leaf_batch",
            "coop_lanes")


def opt_string(opts: dict) -> str:
    return ",".join(f"{k}={int(opts.get(k, 0))}" for k in OPT_KEYS)


def kernel_name(opts: dict) -> str:
    """The batched launch's dominant kernel: the fused frame kernel, the
    lane-refill frame kernel, or the queued pipeline's graph of trace,
    compaction and shading kernels."""
    lb = 2 if int(opts.get("leaf_batch", 1)) == 2 else 1
    if int(opts.get("pipeline", 0)) == 1:
        return "queued pipeline graph"
    if int(opts.get("pipeline", 0)) == 3:
        if int(opts.get("coop_lanes", 0)) >= 64:  # every leaf step cooperative: its own instance
            return f"k_render_refill<{int(opts['variant'])},false,{3 if int(opts.get('tri_test', 0)) else 1},true>"
        return f"k_render_refill<{int(opts['variant'])},false,{lb},false>"
    return f"k_trace<{int(opts['variant'])},false,true,{lb}>"


def profiled_counters(kernel: str, opts: dict, workload: str):
    """Per-launch counters of a kernel instance from the rocprofv3 summaries
    in profiles/.  A summary line reads
        counters <kernel> <options> <workload> key=value ...
    and only a line of the same kernel, the same options and the same
    workload counts.  Returns (dict, source) or (None, reason)."""
    want = opt_string(opts)
    for f in sorted((ROOT / "profiles").glob("*.txt")) + sorted((ROOT / "profiles").glob("*.md")):
        for line in f.read_text().splitlines():
            parts = line.split()
            if len(parts) >= 5 and parts[0] == "counters" and parts[1] == kernel and parts[2] == want \
                    and parts[3] == workload:
                vals = {}
                for kv in parts[4:]:
                    k, _, v = kv.partition("=")
                    vals[k] = float(v)
                return vals, f"profiles/{f.name}"
    return None, f"no counter profile of {kernel} with {want} on {workload}"


def profiled_shapes(workload: str):
    """Option tuples (OPT_KEYS order) of the batched launches profiled on
    `workload` in profiles/."""
    out = set()
    for f in sorted((ROOT / "profiles").glob("*.txt")) + sorted((ROOT / "profiles").glob("*.md")):
        for line in f.read_text().splitlines():
            parts = line.split()
            if len(parts) >= 5 and parts[0] == "counters" and parts[3] == workload:
                kv = dict(x.split("=") for x in parts[2].split(","))
                out.add(tuple(int(kv[k]) for k in OPT_KEYS))
    return out


def host_cpus():
    """(threads to use, CPUs in this process's affinity set, CPU model).
    The threads are the host cores granted to the process: OMP_NUM_THREADS
    when the environment sets it (a GPU box grants a share of its cores),
    else every CPU of the affinity set."""
    aff = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(aff, int(env)) if env.isdigit() and int(env) > 0 else aff
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return threads, aff, model


class Timer:
    """K launches bracketed by device synchronizations, timed by the host
    clock (ms_per_step) and by HIP events recorded on the library's stream
    in the same bracket (kernel_ms)."""

    def __init__(self, stream):
        self.stream = stream

    def __call__(self, fn, reps: int, warmup: int):
        import torch

        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t = time.perf_counter()
        e0.record(self.stream)
        for _ in range(reps):
            fn()
        e1.record(self.stream)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t) / reps
        return wall * 1e3, e0.elapsed_time(e1) / reps


# ------------------------------------------------------------- oracle checks --

class OracleFrames:
    """The CPU oracle's primary hit records and pixels of the workload's
    frames, per traversal variant (computed once, on the host cores)."""

    ORACLE_OF = {1: 1, 2: 2, 3: 3, 4: 4, 5: 5}  # hulk_variant -> oracle variant (same numbering)

    def __init__(self, osc, cams, W, H, threads):
        self.osc, self.cams, self.W, self.H, self.threads = osc, cams, W, H, threads
        self.cache = {}
        self.seconds = 0.0

    def get(self, variant: int, depth: int):
        key = (variant, depth)
        if key not in self.cache:
            t = time.perf_counter()
            hits, imgs = [], []
            for cam in self.cams:
                h = self.osc.trace_camera(self.ORACLE_OF[variant], cam, self.W, self.H, depth=depth,
                                          threads=self.threads)
                img, _ = self.osc.render(self.ORACLE_OF[variant], cam, self.W, self.H, max_bounce=0, depth=depth,
                                         threads=self.threads)
                hits.append(h)
                imgs.append(img)
            self.cache[key] = (hits, imgs)
            self.seconds += time.perf_counter() - t
        return self.cache[key]

    def check(self, variant: int, depth: int, hits, imgs) -> dict:
        """Share of primary rays whose hit record (t, u, v, tri, normal)
        equals the oracle's bit for bit, and of pixels equal to the oracle's."""
        import numpy as np

        rh, ri = self.get(variant, depth)
        n = self.W * self.H
        same_hits = same_px = 0
        for k in range(len(self.cams)):
            a, b = hits[k * n:(k + 1) * n], rh[k]
            eq = np.ones(n, bool)
            for f in ("t", "u", "v", "tri", "nx", "ny", "nz"):
                eq &= a[f].view(np.uint32) == b[f].view(np.uint32)
            same_hits += int(eq.sum())
            same_px += int(np.sum(imgs[k * n:(k + 1) * n] == ri[k]))
        total = n * len(self.cams)
        return {"oracle_hit_match": same_hits / total, "oracle_pixel_match": same_px / total,
                "rays_checked": total}


def classify_vs_reference(osc, cams, W, H, hits, ref_hits, scene, tol: float = 1e-6):
    """Primary triangle ids of `hits` against reference hits (the oracle's
    full-stack traversal, the reference's GPUTrace.cl trace_ray): every ray
    whose id differs is classified with the FP64 twin of the triangle test
    as a tie (the two triangles' t within `tol` relative) or a hit point
    within `tol` of an edge of either triangle (shared edges and vertices);
    the rest are unexplained.  The same arithmetic breaks ties alike, so
    1e-6 is ample there; another triangle test (Moller-Trumbore) rounds t
    up to 1e-4 apart at grazing angles."""
    import numpy as np

    import oracle as O

    n = W * H
    V = scene.vertices[:, :3].astype(np.float32)
    T = scene.triangles
    total = diff_n = tie = edge = other = 0
    listed = []
    for k, cam in enumerate(cams):
        a, b = hits[k * n:(k + 1) * n]["tri"], ref_hits[k]["tri"]
        total += n
        idx = np.nonzero(a != b)[0]
        diff_n += len(idx)
        if len(idx) == 0:
            continue
        rays = O.cast_rays(cam, W, H)
        for i in idx:
            r = rays[i]
            dd = []
            for tri in (a[i], b[i]):
                if tri == O.NO_HIT:
                    dd.append(None)
                    continue
                p = V[T[tri, :3]]
                dd.append(O.tri_test_f64(r, p[0], p[1], p[2]))
            if all(d is not None for d in dd) and abs(dd[0][0] - dd[1][0]) <= tol * abs(dd[1][0]):
                tie += 1
                kind = "tie"
            elif any(d is not None and min(d[1], d[2], 1 - d[1] - d[2]) < tol for d in dd):
                edge += 1
                kind = "edge"
            else:
                other += 1
                kind = "unexplained"
            if len(listed) < 20:
                listed.append({"frame": k, "pixel": int(i), "tri": int(a[i]), "reference_tri": int(b[i]),
                               "kind": kind})
    return {"rays": total, "same_triangle": (total - diff_n) / total, "differ": diff_n, "ties": tie,
            "edge_or_vertex": edge, "unexplained": other, "listed": listed}


# ---------------------------------------------------------------- methods --

def compare_methods(ctx, scene, osc_frames, cams, train_cams, W: int, H: int, stream, args, best: dict,
                    bbest: dict, learned: bool = True):
    """plain / batched / learned-bound + batched on `cams` (bounds from an MLP
    trained on `train_cams`), with the tuned options and with kd-restart.
    plain runs with `best` (tuned on single frames), the batched methods with
    `bbest` (tuned on the batched launch).  Every method's hits and pixels are
    checked against the oracle of its variant.  Returns (methods, learned
    info, {method: (ms_per_step, kernel_ms)}, {method: hits})."""
    import numpy as np
    import torch

    import hulktracer_amd as HT
    from hulktracer_amd import _lib as L

    n = W * H
    nf = len(cams)
    dev = torch.device("cuda:0")
    timer = Timer(stream)
    hits_all = torch.empty(nf * n * 32, dtype=torch.uint8, device=dev)
    rgba_all = torch.empty(nf * n, dtype=torch.int32, device=dev)
    frames = [(cam, W, H, 0) for cam in cams]
    methods, timing, method_hits = {}, {}, {}

    def host_hits():
        return hits_all.cpu().numpy().view(np.uint8).view(HT.HIT_DTYPE)

    def record(name, opts, ms, kms, hits, imgs, **extra):
        chk = osc_frames.check(int(opts["variant"]), int(opts["short_stack_depth"]), hits, imgs)
        methods[name] = {"ms_per_step": round(ms, 4), "mrays_s": round(nf * n / ms / 1e3, 1),
                         "variant": L.VARIANT_NAMES[int(opts["variant"])], **chk, **extra}
        timing[name] = (ms, kms)
        method_hits[name] = hits
        log(f"method {name:26s} {methods[name]}")

    def plain(opts, name, reps):
        ctx.set_options(**opts)

        def step():
            for k, cam in enumerate(cams):
                ctx.render_frame_dev(cam, W, H, 0, None, rgba_all.data_ptr() + 4 * k * n,
                                     hits_all.data_ptr() + 32 * k * n, flags=L.FLAG_NO_SYNC)

        ms, kms = timer(step, reps, args.warmup)
        record(name, opts, ms, kms, host_hits(), rgba_all.cpu().numpy().view(np.uint32))

    def batched(opts, name, reps, tb=None, pre=None, **extra):
        ctx.set_options(**opts)
        b = ctx.batch(frames, None, rgba_all.data_ptr(), tbound_dev=tb, use_graph=True)

        def step():
            if pre is not None:
                pre()
            b.launch(L.FLAG_NO_SYNC)

        if pre is not None:
            pre()
        ms, kms = timer(step, reps, args.warmup)
        h = b.primary_hits()
        b.close()
        if tb is not None:
            extra["fallback_rate"] = round(float(np.mean((h["flags"] & 1) != 0)), 4)
        record(name, opts, ms, kms, h, rgba_all.cpu().numpy().view(np.uint32), **extra)
        return h

    topts = {k: int(v) for k, v in best.items() if k in OPT_KEYS}
    bopts = {k: int(v) for k, v in bbest.items() if k in OPT_KEYS}
    # kd-restart runs in the fused frame kernel (lane refill is a rope / LDS-stack pipeline)
    ropts = dict(bopts, variant=HT.RESTART, pipeline=2 if bopts.get("pipeline") == 3 else bopts.get("pipeline", 2))
    if bopts.get("pipeline") == 3:
        # the tuned refill shape leaves records in flight open (its cooperative
        # instance ignores them): take kd-restart's faster of one and two
        def batch_ms(opts):
            ctx.set_options(**opts)
            b = ctx.batch(frames, None, rgba_all.data_ptr(), use_graph=True)
            ms, _ = timer(lambda: b.launch(L.FLAG_NO_SYNC), 3, 1)
            b.close()
            return ms

        ropts = min((dict(ropts, leaf_batch=lb) for lb in (1, 2)), key=batch_ms)
    plain(topts, "plain", args.steps)
    batched(bopts, "batched", args.steps)
    learned_info = None
    if learned:
        from hulktracer_amd.learned import HitDistanceMLP, select_margin, training_pairs

        tl = time.time()
        # train on every other pose; the rest calibrate the margin (frames the network has not seen)
        fit_cams, calib_cams = train_cams[0::2], train_cams[1::2]
        ctx.set_options(**bopts)
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

        def bounds():  # every frame's bounds in one launch
            mlp.bound_cameras(cams, W, H, tb, sync=False)

        bounds()
        _, infer_ms = timer(bounds, 3, 1)
        infer_ms /= nf
        learned_info = {"train_pairs": n_pairs, "train_steps": args.train_steps, "train_s": round(train_s, 2),
                        "data_s": round(tt - tl, 2), "bound_stride": 2,
                        "inference_ms_per_frame": round(infer_ms, 4),
                        "inference_tflops": round(n / 4 * 82176 / (infer_ms * 1e-3) / 1e12, 2),
                        "loss": [[int(a), round(float(b), 4)] for a, b in losses], "margin_selection": {}}
        # what a perfect proposal could save on kd-restart (no exit before
        # the box exit): bounds 1 % and 20 % above each ray's own nearest hit
        # (misses unbounded), node visits against the unbounded trace
        ctx.set_options(**ropts)
        cam_c = calib_cams[0]
        hh = ctx.trace_camera(cam_c, W, H)
        ctx.reset_counters()
        ctx.trace_camera(cam_c, W, H, count=True)
        base_nv = ctx.counters()["node_visits"]
        ideal = {}
        for slack in (1.01, 1.2):
            tbh = np.where(hh["tri"] != HT.NO_HIT, np.minimum(hh["t"].astype(np.float64) * slack, 3.4e38),
                           3.4028234663852886e38).astype(np.float32)
            ctx.reset_counters()
            ctx.trace_camera(cam_c, W, H, tbound=tbh, count=True)
            ideal[f"bound_{slack}x_hit"] = round(1 - ctx.counters()["node_visits"] / base_nv, 4)
        learned_info["ideal_bound_node_visits_saved_kd-restart"] = ideal
        learned_info["primary_hit_fraction"] = round(float(np.mean(hh["tri"] != HT.NO_HIT)), 4)
        for tag, vopts in (("", bopts), ("_kd-restart", ropts)):
            ctx.set_options(**vopts)
            # margin (and whether to bound at all) chosen on calibration frames for this traversal
            sel = select_margin(mlp, ctx, calib_cams, W, H, stride=2)
            learned_info["margin_selection"][tag.lstrip("_") or "tuned"] = sel
            if tag:  # the reference's live kernel: plain frame by frame, and batched without a bound
                plain(vopts, "plain" + tag, max(1, args.steps // 5))
                batched(vopts, "batched" + tag, max(1, args.steps // 2))
            use_bound = mlp.enabled
            bounds()  # node visits with the selected margin, reported also where the bound is not used
            torch.cuda.synchronize()
            ctx.set_options(**vopts)
            ctx.reset_counters()
            ctx.trace_camera(cams[0], W, H, count=True)
            cp = ctx.counters()["node_visits"]
            ctx.reset_counters()
            ctx.trace_camera(cams[0], W, H, tbound=tb[:n].cpu().numpy(), count=True)
            cb = ctx.counters()["node_visits"]
            batched(vopts, "learned+batched" + tag, args.steps, tb=tb.data_ptr() if use_bound else None,
                    pre=bounds if use_bound else None, bound_used=bool(use_bound),
                    node_visits_saved_if_bounded=round(1 - cb / cp, 4))
        ctx.set_options(**topts)
        mlp.close()
    else:
        plain(ropts, "plain_kd-restart", max(1, args.steps // 5))
        batched(ropts, "batched_kd-restart", max(1, args.steps // 2))
    return methods, learned_info, timing, method_hits


# --------------------------------------------------------------- workloads --

def run_workload(name: str, key: str, scene, W: int, H: int, stream, args, learned: bool = True,
                 cpu: bool = True, counted: bool = True):
    """Autotune, the methods with oracle checks, the roofline inputs and the
    CPU baseline of one workload (poses 10-19 of `scene`)."""
    import numpy as np
    import torch

    import hulktracer_amd as HT
    import oracle as O
    from hulktracer_amd import _lib as L
    from hulktracer_amd import workloads as WL

    threads, aff, model = host_cpus()
    nodes, lp, build_ms = WL.build(scene)
    cams = [scene.camera(k, H / W) for k in HEADLINE_POSES]
    train_cams = [scene.camera(k, H / W) for k in TRAIN_POSES]
    nf, n = len(cams), W * H
    log(f"[{name}] scene {scene.n_tris} tris, {len(nodes)} nodes, {len(lp)} leaf refs, host KD build "
        f"{build_ms:.0f} ms")
    out = {"workload": name, "workload_key": key, "triangles": int(scene.n_tris), "kd_nodes": int(len(nodes)),
           "leaf_refs": int(len(lp)), "kd_build_ms_host": round(build_ms, 1)}
    rgba_tmp = torch.empty(nf * n, dtype=torch.int32, device="cuda")

    # every variant at its default shape (primary trace of pose 10)
    hits = torch.empty(n * 32, dtype=torch.uint8, device="cuda")
    timer = Timer(stream)
    shapes = {"GPUTrace": dict(variant=HT.FULLSTACK), "GPUTrace_shortStack": dict(variant=HT.SHORTSTACK),
              "GPUTrace_LDSshortStack": dict(variant=HT.LDSSTACK, lds_top_levels=8),
              "GPUTrace_stackless": dict(variant=HT.RESTART), "GPUTrace_ropes": dict(variant=HT.ROPES)}
    variants = {}
    for vname, kw in shapes.items():
        cx = WL.make_ctx(scene, nodes, lp, stream, **kw)
        _, ms = timer(lambda: cx.trace_camera_dev(cams[0], W, H, hits.data_ptr(), flags=L.FLAG_NO_SYNC),
                      5 if kw["variant"] != HT.RESTART else 2, 1)
        variants[vname] = {"primary_trace_ms": round(ms, 4), "mrays_s": round(n / ms / 1e3, 1)}
        cx.close()
    out["variants_primary_trace"] = variants
    log(f"[{name}] variants {variants}")

    ctx = WL.make_ctx(scene, nodes, lp, stream, variant=HT.FULLSTACK)
    tt = time.time()
    best, table = ctx.autotune(cams[0], W, H, reps=2, frame=True)
    frames = [(cam, W, H, 0) for cam in cams]
    # among shapes within 2 % (the stage's run-to-run noise) prefer a profiled one, so that
    # the roofline's counters are those of the launch that runs
    bbest, btimes = WL.tune_batch(ctx, frames, None, rgba_tmp.data_ptr(), table, stream,
                                  prefer=profiled_shapes(key), tol=0.02)
    log(f"[{name}] autotune {time.time() - tt:.1f}s -> {best}; batch stage {btimes} -> {bbest}")
    ctx.set_options(**{k: int(v) for k, v in best.items() if k in OPT_KEYS})
    out["options_single_frame"] = {k: int(best[k]) for k in OPT_KEYS}
    out["options_batched"] = {k: int(bbest[k]) for k in OPT_KEYS}
    out["batch_tune_ms"] = btimes

    if counted:
        # algorithmic bytes of one batched launch: full-stack counts of every
        # traced ray (primary and shadow), 32 B ray + 16 B result each, 16 B
        # per node visit, 64 B per triangle test
        cnt = WL.make_ctx(scene, nodes, lp, stream, variant=HT.FULLSTACK)
        cb = cnt.batch(frames, None, rgba_tmp.data_ptr(), use_graph=False)
        cb.launch(L.FLAG_COUNT)
        c = cnt.counters()
        cb.close()
        cnt.close()
        traced = c["rays"] + c["shadow_rays"]
        out["algorithmic"] = {"traced_rays_per_launch": int(traced), "shadow_rays_per_launch": int(c["shadow_rays"]),
                              "bytes_per_launch": int(48 * traced + 16 * c["node_visits"] + 64 * c["tri_tests"]),
                              "node_visits_per_ray": round(c["node_visits"] / traced, 3),
                              "tri_tests_per_ray": round(c["tri_tests"] / traced, 3)}

    osc = O.OracleScene(scene.vertices, scene.triangles, scene.box, nodes, lp, scene.lights)
    ofr = OracleFrames(osc, cams, W, H, threads)
    methods, learned_info, timing, mhits = compare_methods(ctx, scene, ofr, cams, train_cams, W, H, stream, args,
                                                           best, bbest, learned=learned)
    ctx.close()
    out["methods"] = methods
    out["learned"] = learned_info
    # eligible: every hit record and pixel equal to the oracle's; a learned
    # method whose calibration turned the bound off runs the batched launch
    # itself, so it is not a method of its own
    ok = {k: v for k, v in methods.items() if v["oracle_hit_match"] == 1.0 and v["oracle_pixel_match"] == 1.0
          and v.get("bound_used", True)}
    best_m = max(ok, key=lambda k: ok[k]["mrays_s"]) if ok else None
    out["best_method"] = best_m
    out["timing"] = timing
    out["speedup_vs_plain"] = {k: round(v["mrays_s"] / methods["plain"]["mrays_s"], 3) for k, v in methods.items()
                               if not k.endswith("kd-restart")}
    if "plain_kd-restart" in methods:
        out["speedup_vs_plain_kd-restart"] = {k: round(v["mrays_s"] / methods["plain_kd-restart"]["mrays_s"], 3)
                                              for k, v in methods.items() if k.endswith("kd-restart")}
    if best_m:
        ref_hits, _ = ofr.get(HT.FULLSTACK, 6)
        out["vs_reference_fullstack"] = classify_vs_reference(osc, cams, W, H, mhits[best_m], ref_hits, scene)
    out["oracle_seconds"] = round(ofr.seconds, 1)
    out["oracle_threads"] = threads

    if cpu:
        # the reference CPU path (port): per frame the tree is rebuilt
        # (GLVisitorCPUTrace.cpp:42-59, timed separately as Statistics does)
        # and traceScene traces primary + shadow rays and shades
        kd = HT.KDTreeCPU()
        tk = time.perf_counter()
        kd.buildKDTree(scene, threads=threads)
        kd_ms = (time.perf_counter() - tk) * 1e3
        osc.cpu_trace_scene(cams[0], W, H, 0, 8, threads=threads, out=False)
        frames_done, rays_total, tc = 0, 0, time.perf_counter()
        while True:
            rays_total += osc.cpu_trace_scene(cams[frames_done % nf], W, H, threads=threads, out=False)[0]
            frames_done += 1
            if time.perf_counter() - tc > args.cpu_seconds or frames_done >= 2000:
                break
        cpu_s = time.perf_counter() - tc
        out["cpu_baseline"] = {
            "value": round(n * frames_done / cpu_s / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{frames_done} whole {W}x{H} frames of poses 10-19 (primary + shadow + shade, "
                      f"{rays_total / cpu_s / 1e6:.2f} M total rays/s incl. shadow rays), {cpu_s:.1f} s, OpenMP "
                      f"over rows on {threads} threads; KD rebuild per frame {kd_ms:.0f} ms on {threads} threads "
                      f"(timed separately)",
            "cpu_model": model, "affinity_cpus": aff, "kd_build_ms": round(kd_ms, 1),
            "total_mrays_s_incl_shadow": round(rays_total / cpu_s / 1e6, 3)}
        log(f"[{name}] cpu baseline {out['cpu_baseline']}")
    return out


def roofline(wl: dict, W: int, H: int) -> dict:
    """The batched launch (method "batched": the tuned options' graph over
    the 10 frames, the launch the profiles cover), priced against the
    resource its profiled counters show closest to saturation."""
    meth = "batched"
    opts = wl["options_batched"]
    ms, kms = wl["timing"][meth]
    kname = kernel_name(opts)
    wname = wl["workload_key"]
    ctr, src = profiled_counters(kname, opts, wname)
    alg = wl.get("algorithmic", {})
    alg_gbs = alg.get("bytes_per_launch", 0) / (kms * 1e-3) / 1e9 if alg else None
    r = {"kernel": kname, "options": opts, "method": meth, "kernel_ms": round(kms, 4), "ms_per_step": round(ms, 4),
         "launch": f"{len(HEADLINE_POSES)}-frame batch (graph), HIP events on the library's stream",
         "primary_rays_per_launch": len(HEADLINE_POSES) * W * H,
         "cache_served_algorithmic_GBs": round(alg_gbs, 1) if alg_gbs else None,
         "algorithmic_bytes_per_launch": alg.get("bytes_per_launch"), "counter_source": src}
    if ctr is None:
        r.update({"bound": None, "achieved": None, "peak": None, "unit": None, "frac": None, "traffic": None})
        return r
    valu = ctr["valu_wave_instr"] / (kms * 1e-3) / 1e9
    hbm_bytes = ctr["fetch_bytes"] + ctr["write_bytes"]
    hbm = hbm_bytes / (kms * 1e-3) / 1e9
    res = {"valu_issue": (valu, VALU_PEAK_GINSTR_S, "Gwave-instr/s"), "hbm": (hbm, HBM_PEAK_GBS, "GB/s")}
    bound = max(res, key=lambda k: res[k][0] / res[k][1])
    a, p, u = res[bound]
    r.update({"bound": bound, "achieved": round(a, 1), "peak": p, "unit": u, "frac": round(a / p, 4),
              "traffic": int(hbm_bytes),
              "valu": {"wave_instr_per_launch": int(ctr["valu_wave_instr"]), "achieved_Ginstr_s": round(valu, 1),
                       "peak_Ginstr_s": VALU_PEAK_GINSTR_S, "frac": round(valu / VALU_PEAK_GINSTR_S, 4),
                       "busy_counter": ctr.get("valu_busy")},
              "hbm": {"bytes_per_launch": int(hbm_bytes), "achieved_GBs": round(hbm, 1), "peak_GBs": HBM_PEAK_GBS,
                      "frac": round(hbm / HBM_PEAK_GBS, 4)},
              "occupancy_counter": ctr.get("occupancy"), "l2_hit_counter": ctr.get("l2_hit")})
    return r


# ------------------------------------------------------------ single modes --

def kd_build_only(args) -> int:
    """--kd-build-only: warmup + steps device KD builds of one scene in one
    context; prints the wall time per build (the call returns once the tree
    is installed) and the host builder's time on the same scene."""
    import hulktracer_amd as H
    from hulktracer_amd import _lib as L
    from hulktracer_amd import workloads as WL

    if args.kernel_scene in ("animated", "headline"):
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


def kernel_only(args, W: int, H: int) -> int:
    """--kernel-only <options>: warmup + steps launches of one workload's
    dominant kernel with fixed options, timed with HIP events, for profiler
    runs.  --kernel-scene headline: the 10-frame batched launch of the
    1M-triangle scene (poses 10-19); terrain: the same on the terrain scene;
    animated: one render of frame 0 of the animated scene with 2 mirror
    generations."""
    import torch

    import hulktracer_amd as HT
    from hulktracer_amd import _lib as L
    from hulktracer_amd import workloads as WL

    opts = {k: int(v) for k, v in (kv.split("=") for kv in args.kernel_only.split(","))}
    if args.kernel_scene == "terrain":
        scene = HT.Scene.procedural(L.SCENE_TERRAIN_SPHERES, seed=1)
    else:
        scene = HT.Scene.procedural(L.SCENE_ANIMATED_TORI, seed=7, frame=0)
    nodes, lp, _ = WL.build(scene)
    stream = torch.cuda.Stream()
    ctx = WL.make_ctx(scene, nodes, lp, stream, **opts)
    timer = Timer(stream)
    if args.kernel_scene == "animated":
        mirror = torch.from_numpy(scene.mirror).cuda()
        rgba = torch.empty(W * H, dtype=torch.int32, device="cuda")
        hits = torch.empty(W * H * 32, dtype=torch.uint8, device="cuda")
        cam = scene.camera(0, H / W)

        def step():
            ctx.render_frame_dev(cam, W, H, 2, mirror.data_ptr(), rgba.data_ptr(), hits.data_ptr(),
                                 flags=L.FLAG_NO_SYNC)

        rays, wname, b = W * H, "animated1M_frame0_mirror2", None
    else:
        cams = [scene.camera(k, H / W) for k in HEADLINE_POSES]
        rgba = torch.empty(len(cams) * W * H, dtype=torch.int32, device="cuda")
        b = ctx.batch([(c, W, H, 0) for c in cams], None, rgba.data_ptr(), use_graph=True)

        def step():
            b.launch(L.FLAG_NO_SYNC)

        rays = len(cams) * W * H
        wname = ("animated1M" if args.kernel_scene == "headline" else "terrain163k") + f"_poses10-19_{W}x{H}"
    ms, kms = timer(step, args.steps, max(args.warmup, 1))
    if b is not None:
        b.close()
    ctx.close()
    print(json.dumps({"kernel": kernel_name(opts), "options": opt_string(opts), "workload": wname,
                      "kernel_ms": round(kms, 4), "ms_per_step": round(ms, 4), "timed_launches": args.steps,
                      "launches": args.steps + max(args.warmup, 1), "primary_rays_per_launch": rays}), flush=True)
    return 0


def mlp_only(args, W: int, H: int) -> int:
    """--mlp-only: the learned bound's kernels alone, for profiler runs: the
    per-pixel bounds of the 10 headline frames at stride 2 (inference, as
    the learned method runs it), warmup + steps times, then `--train-steps`
    training steps of 16,384 pairs; HIP events on the library's stream."""
    import torch

    import hulktracer_amd as HT
    from hulktracer_amd import _lib as L
    from hulktracer_amd import workloads as WL
    from hulktracer_amd.learned import HitDistanceMLP, camera_rays

    scene = HT.Scene.procedural(L.SCENE_ANIMATED_TORI, seed=7, frame=0)
    nodes, lp, _ = WL.build(scene)
    stream = torch.cuda.Stream()
    ctx = WL.make_ctx(scene, nodes, lp, stream)
    cams = [scene.camera(k, H / W) for k in HEADLINE_POSES]
    mlp = HitDistanceMLP(ctx, seed=1, max_batch=16384)
    mlp.stride = 2
    tb = torch.empty(len(cams) * W * H, dtype=torch.float32, device="cuda")
    timer = Timer(stream)

    def bounds():
        mlp.bound_cameras(cams, W, H, tb, sync=False)

    def bounds_per_frame():
        for k, cam in enumerate(cams):
            mlp.bound_camera(cam, W, H, tb[k * W * H:(k + 1) * W * H], sync=False)

    _, frame_ms = timer(bounds_per_frame, args.steps, max(args.warmup, 1))
    _, infer_ms = timer(bounds, args.steps, max(args.warmup, 1))
    rays, _ = camera_rays(ctx, cams[0], W, H)
    rays = rays[:16384].contiguous()
    y = torch.ones(16384, dtype=torch.float32, device="cuda")
    _, train_ms = timer(lambda: mlp.train_step(rays, y, 16384, 1e-3), max(args.train_steps, 1), 1)
    preds = len(cams) * ((W + 1) // 2) * ((H + 1) // 2)
    print(json.dumps({"mlp": "bound_camera stride 2 + train_step", "frames": len(cams),
                      "inference_ms_per_frame": round(infer_ms / len(cams), 4),
                      "inference_ms_per_frame_one_launch_per_frame": round(frame_ms / len(cams), 4),
                      "inference_tflops": round(preds * 82176 / (infer_ms * 1e-3) / 1e12, 2),
                      "train_step_ms": round(train_ms, 4), "predictions_per_launch": preds // len(cams)}),
          flush=True)
    mlp.close()
    ctx.close()
    return 0


# ------------------------------------------------------------------ extras --

def extras_run(args, W, H, stream, headline):
    import numpy as np
    import torch

    import hulktracer_amd as HT
    from hulktracer_amd import _lib as L
    from hulktracer_amd import workloads as WL

    extras = {}

    def terrain():
        sc = HT.Scene.procedural(L.SCENE_TERRAIN_SPHERES, seed=1)
        wl = run_workload("terrain", f"terrain163k_poses10-19_{W}x{H}", sc, W, H, stream, args,
                          learned=not args.no_learned, cpu=not args.no_cpu, counted=True)
        wl["roofline"] = roofline(wl, W, H)
        wl.pop("timing", None)
        best = wl["best_method"]
        wl["value_mrays_s"] = wl["methods"][best]["mrays_s"] if best else None
        return wl

    def mt_compare():
        # the optional Moller-Trumbore test against the reference's on the
        # headline frames: batched launch time, and triangle ids of the
        # primary hits, every difference classified with the FP64 twin
        import oracle as O

        sc = HT.Scene.procedural(L.SCENE_ANIMATED_TORI, seed=7, frame=0)
        nodes, lp, _ = WL.build(sc)
        cams = [sc.camera(k, H / W) for k in HEADLINE_POSES]
        frames = [(c, W, H, 0) for c in cams]
        rgba = torch.empty(len(cams) * W * H, dtype=torch.int32, device="cuda")
        opts = dict(headline["options_batched"])
        if opts.get("leaf_batch") == 2:
            opts["leaf_batch"] = 1
        out, hits_by = {}, {}
        timer = Timer(stream)
        for tag, tt in (("reference", HT.TRI_REFERENCE), ("moller_trumbore", HT.TRI_MOLLER_TRUMBORE)):
            cx = WL.make_ctx(sc, nodes, lp, stream, **dict(opts, tri_test=tt))
            b = cx.batch(frames, None, rgba.data_ptr(), use_graph=True)
            _, ms = timer(lambda: b.launch(L.FLAG_NO_SYNC), max(args.steps, 5), 1)
            hits_by[tag] = b.primary_hits()
            b.close()
            cx.close()
            out[tag] = {"batch_ms": round(ms, 4), "mrays_s": round(len(cams) * W * H / ms / 1e3, 1)}
        osc = O.OracleScene(sc.vertices, sc.triangles, sc.box, nodes, lp, sc.lights)
        out["tri_id_differences"] = classify_vs_reference(osc, cams, W, H, hits_by["moller_trumbore"],
                                                          [hits_by["reference"][k * W * H:(k + 1) * W * H]
                                                           for k in range(len(cams))], sc, tol=1e-4)
        out["options"] = opts
        log(f"triangle test: {out}")
        return out

    def streamed():
        # out-of-core: terrain pose 10 at 1024^2, kd-restart, record pool of
        # 100 / 25 / 5 % of the leaf references against the in-core kernel
        sc = HT.Scene.procedural(L.SCENE_TERRAIN_SPHERES, seed=1)
        nodes, lp, _ = WL.build(sc)
        cx = WL.make_ctx(sc, nodes, lp, stream, variant=HT.RESTART)
        cam = sc.camera(10, H / W)
        cx.trace_camera(cam, W, H)
        t0 = time.perf_counter()
        ref = cx.trace_camera(cam, W, H)
        core_ms = (time.perf_counter() - t0) * 1e3
        out = {"workload": f"terrain_{sc.n_tris}tris_pose10_{W}x{H}_primary_kd-restart",
               "leaf_refs": int(len(lp)), "in_core_ms": round(core_ms, 2)}
        for frac in (1.0, 0.25, 0.05):
            pool = max(int(frac * len(lp)), int(nodes[nodes[:, 0] == 4, 2].max()))
            got, st = cx.trace_camera_streamed(cam, W, H, pool)
            same = bool(np.array_equal(got.view(np.uint32).reshape(-1, 8)[:, :7],
                                       ref.view(np.uint32).reshape(-1, 8)[:, :7]))
            out[f"pool_{int(frac * 100)}pct"] = {
                "pool_records": pool, "ms": round(st["ms"], 2), "launches": st["launches"],
                "leaves_loaded": st["leaves_loaded"], "leaves_evicted": st["leaves_evicted"],
                "mb_loaded": round(st["bytes_loaded"] / 1e6, 1), "mrays_s": round(W * H / st["ms"] / 1e3, 1),
                "hits_equal_in_core": same}
        cx.close()
        log(f"streamed: {out}")
        return out

    jobs = [("terrain_163k", terrain), ("triangle_test_mt", mt_compare), ("streamed_out_of_core", streamed),
            ("animated_1M", lambda: WL.animated(frames=20, stream=stream, log=log)),
            ("animated_1M_device_build", lambda: WL.animated_device_build(frames=20, stream=stream, log=log)),
            ("batched_64_frames", lambda: WL.batched(stream=stream, log=log))]
    for name, fn in jobs:
        try:
            extras[name] = fn()
        except Exception as e:  # an extra must not hide the headline line
            extras[name] = {"error": repr(e)}
            log(f"{name} failed: {e!r}")
    return extras


# -------------------------------------------------------------------- main --

def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--res", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--train-steps", type=int, default=1500)
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only run a device copy of this many MiB (HBM traffic counter calibration)")
    ap.add_argument("--kernel-only", default="",
                    help="only time the dominant kernel of --kernel-scene with these options "
                         "(e.g. variant=3,block_size=64,short_stack_depth=6,lds_top_levels=6,persistent=0,"
                         "pipeline=2,leaf_batch=2,coop_lanes=0): warmup + steps launches, for profiler runs")
    ap.add_argument("--mlp-only", action="store_true",
                    help="only time the learned bound's inference (10 headline frames, stride 2) and "
                         "--train-steps training steps, for profiler runs")
    ap.add_argument("--kd-build-only", action="store_true",
                    help="only time the device KD-tree build of --kernel-scene: warmup + steps builds")
    ap.add_argument("--kernel-scene", default="headline", choices=("headline", "terrain", "animated"),
                    help="--kernel-only / --kd-build-only workload: headline = the 10-frame batch of the 1M-triangle "
                         "scene, terrain = the 10-frame terrain batch, animated = frame 0 of the animated scene "
                         "with 2 mirror generations")
    args = ap.parse_args()
    if args.gpus != 1:
        log("bench.py measures one GPU only: the library has no multi-GPU path (rays would be split "
            "per GPU by image tiles or frames, with the scene replicated). Run with --gpus 1.")
        return 2

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

    import hulktracer_amd as HT
    from hulktracer_amd import _lib as L

    W = Hh = args.res
    if args.kd_build_only:
        return kd_build_only(args)
    if args.kernel_only:
        return kernel_only(args, W, Hh)
    if args.mlp_only:
        return mlp_only(args, W, Hh)

    stream = torch.cuda.Stream(device=torch.device("cuda:0"))  # default priority; library + events on it
    t0 = time.time()
    scene = HT.Scene.procedural(L.SCENE_ANIMATED_TORI, seed=7, frame=0)
    hl = run_workload("animated1M", f"animated1M_poses10-19_{W}x{Hh}", scene, W, Hh, stream, args,
                      learned=not args.no_learned, cpu=not args.no_cpu)
    rl = roofline(hl, W, Hh)
    best_m = hl["best_method"]
    if best_m is None:
        log("no method matched the oracle on every ray and pixel")
        return 1
    ms, _ = hl["timing"][best_m]
    nf = len(HEADLINE_POSES)
    cpu = hl.pop("cpu_baseline", None)
    hl.pop("timing", None)
    result = {
        "metric": "Mrays/s nearest-hit @1024^2, learned-bound+batched vs plain, 100% hit match",
        "value": round(nf * W * Hh / ms / 1e3, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": f"animated_tori_{scene.n_tris}tris_frame0_poses10-19_{W}x{Hh}_primary+shadow+shade",
                   "frames_per_step": nf, "method": best_m, "variant": hl["methods"][best_m]["variant"],
                   "options": hl["options_batched"] if "batched" in best_m else hl["options_single_frame"],
                   "kd_build_ms_host": hl["kd_build_ms_host"]},
        "hit_match": {"oracle_hit_match": hl["methods"][best_m]["oracle_hit_match"],
                      "oracle_pixel_match": hl["methods"][best_m]["oracle_pixel_match"],
                      "rays_checked": hl["methods"][best_m]["rays_checked"],
                      "vs_reference_fullstack": hl.get("vs_reference_fullstack")},
        "roofline": rl,
        "cpu_baseline": cpu,
        "headline": hl,
    }
    log(f"headline done in {time.time() - t0:.0f}s: {result['value']} Mrays/s ({best_m})")
    if not args.no_extras:
        result["extras"] = extras_run(args, W, Hh, stream, hl)
    print(json.dumps(result, default=float), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
