"""Learned hit-distance MLP: FP64 oracle pinned by finite differences (CPU);
BF16 MFMA forward / backward kernels against the oracle (GPU); training
reduces the loss; the learned bound keeps every nearest hit exact."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp_ref


def _params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(mlp_ref.N_PARAMS)
    k = 0
    for name, shape in mlp_ref.SIZES:
        n = int(np.prod(shape))
        fan = shape[1] if len(shape) == 2 else 128
        p[k:k + n] = rng.normal(0, scale * np.sqrt(2.0 / fan), n) if name[0] in "Ww" else rng.normal(0, 0.1, n)
        k += n
    return p


def test_oracle_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = _params(1)
    e = rng.normal(0, 1, (24, 64))
    y = rng.normal(0, 1, 24)
    L, g = mlp_ref.loss_and_grad(p, e, y, use_bf16=False)
    idx = rng.choice(len(p), 60, replace=False)
    for i in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mlp_ref.loss_and_grad(pp, e, y, use_bf16=False)[0] - mlp_ref.loss_and_grad(pm, e, y, use_bf16=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.01171875, -2.0 ** -130, 3.0e38])
    r = mlp_ref.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.015625


# ----------------------------------------------------------------- GPU --

def _gpu_setup(seed=1, n=4096):
    import torch

    import hulktracer_amd as H
    from hulktracer_amd.learned import HitDistanceMLP, camera_rays
    from tests.helpers import gpu_ctx, scene_and_tree

    s, nodes, lp = scene_and_tree(2, 1)
    ctx = gpu_ctx(s, nodes, lp, H.FULLSTACK)
    mlp = HitDistanceMLP(ctx, seed=seed, max_batch=16384)
    cam = s.camera(3, 1.0)
    rays, texit = camera_rays(ctx, cam, 128, 128)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    idx = torch.randperm(rays.shape[0], device="cuda", generator=g)[:n]
    return s, ctx, mlp, rays[idx].contiguous(), texit[idx].contiguous()


@pytest.mark.gpu
def test_mlp_forward_matches_oracle():
    import torch

    s, ctx, mlp, rays, _ = _gpu_setup()
    p, F = mlp.get_params()
    p[-1] = 0.3  # b4
    mlp.set_params(p)
    n = rays.shape[0]
    pred = torch.empty(n, dtype=torch.float32, device="cuda")
    mlp.forward(rays, n, pred)
    e = mlp_ref.features(rays.cpu().numpy(), s.box, F)
    ref, _ = mlp_ref.forward(p, e, use_bf16=True)
    got = pred.cpu().numpy()
    err = np.abs(got - ref)
    scale = np.abs(ref).mean() + 1e-3
    assert np.quantile(err, 0.99) < 1e-2 * scale and err.max() < 5e-2 * scale, (np.quantile(err, 0.99), err.max(), scale)


@pytest.mark.gpu
def test_mlp_gradients_match_oracle():
    import torch

    s, ctx, mlp, rays, texit = _gpu_setup(n=4096)
    p, F = mlp.get_params()
    y = torch.log(torch.clamp(texit, min=1e-3)) * 0.5
    loss, g = mlp.loss_grad(rays, y.contiguous(), rays.shape[0])
    e = mlp_ref.features(rays.cpu().numpy(), s.box, F)
    L, gref = mlp_ref.loss_and_grad(p, e, y.cpu().numpy(), use_bf16=True)
    assert abs(loss - L) <= 2e-3 * abs(L) + 1e-6
    k = 0
    for name, shape in mlp_ref.SIZES:
        m = int(np.prod(shape))
        a, b = g[k:k + m], gref[k:k + m]
        rel = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-12)
        assert rel < 2e-2, (name, rel)
        k += m


@pytest.mark.gpu
def test_mlp_training_reduces_loss():
    import torch

    from hulktracer_amd.learned import training_pairs

    s, ctx, mlp, _, _ = _gpu_setup()
    cams = [s.camera(k, 1.0) for k in (0, 2, 4)]
    rays, y = training_pairs(ctx, cams, 128, 128)
    l0, _ = mlp.loss_grad(rays[:8192].contiguous(), y[:8192].contiguous(), 8192)
    mlp.fit(rays, y, steps=300, batch=8192, lr=2e-3)
    l1, _ = mlp.loss_grad(rays[:8192].contiguous(), y[:8192].contiguous(), 8192)
    assert l1 < 0.2 * l0, (l0, l1)


@pytest.mark.gpu
@pytest.mark.parametrize("variant", [1, 4])
def test_learned_bound_keeps_hits_exact(variant):
    """Train on some poses, bound other poses: every hit record equals the
    plain traversal's (the fallback re-runs rays the bound cut short)."""
    import torch

    from hulktracer_amd.learned import calibrate_frames, training_pairs

    s, ctx, mlp, _, _ = _gpu_setup()
    ctx.set_options(variant=variant)
    cams = [s.camera(k, 1.0) for k in (0, 4, 8)]
    rays, y = training_pairs(ctx, cams, 256, 256)
    mlp.fit(rays, y, steps=400, batch=16384, lr=2e-3)
    calibrate_frames(mlp, ctx, cams, 256, 256, quantile=0.9, stride=2)
    visits_plain = visits_bound = 0
    for pose in (2, 6):
        cam = s.camera(pose, 1.0)
        ctx.reset_counters()
        plain = ctx.trace_camera(cam, 256, 256, count=True)
        visits_plain += ctx.counters()["node_visits"]
        tb = torch.empty(256 * 256, dtype=torch.float32, device="cuda")
        mlp.bound_camera(cam, 256, 256, tb)
        ctx.reset_counters()
        got = ctx.trace_camera(cam, 256, 256, tbound=tb.cpu().numpy(), count=True)
        visits_bound += ctx.counters()["node_visits"]
        assert np.array_equal(got["tri"], plain["tri"])
        assert np.array_equal(got["t"].view(np.uint32), plain["t"].view(np.uint32))
        fb = (got["flags"] & 1).mean()
        assert fb < 0.5, fb
    saved = 1.0 - visits_bound / visits_plain
    print(f"variant {variant}: node visits saved by the learned bound {saved:.3f}")
    if variant == 4:
        # kd-restart (the reference's live kernel) walks to the box exit:
        # the trained bound is what lets it stop early
        assert saved > 0.05, saved


@pytest.mark.gpu
def test_margin_selection_by_time():
    """select_margin picks, per traversal, the fastest quantile on
    calibration frames, and enables the bound only when it saves time: it
    does for kd-restart (no early exit)."""
    from hulktracer_amd.learned import select_margin, training_pairs

    s, ctx, mlp, _, _ = _gpu_setup()
    ctx.set_options(variant=4)
    cams = [s.camera(k, 1.0) for k in (0, 4, 8)]
    rays, y = training_pairs(ctx, cams, 256, 256)
    mlp.fit(rays, y, steps=400, batch=16384, lr=2e-3)
    sel = select_margin(mlp, ctx, [s.camera(k, 1.0) for k in (2, 6)], 256, 256)
    times = [v[1] for v in sel["candidates"].values()]
    assert sel["bounded_ms"] == min(times)
    assert sel["enabled"] and mlp.enabled and sel["time_saved"] > 0.05, sel


@pytest.mark.gpu
def test_fit_generalizes_on_held_out_pairs():
    """Random batches from a large pair set (gathered on the library's
    stream): the held-out loss falls well below the labels' variance."""
    import torch

    from hulktracer_amd.learned import training_pairs

    s, ctx, mlp, _, _ = _gpu_setup()
    cams = [s.camera(k, 1.0) for k in range(0, 10, 2)]
    rays, y = training_pairs(ctx, cams, 512, 512)
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    perm = torch.randperm(rays.shape[0], device="cuda", generator=g)
    held, train = perm[:16384], perm[16384:]
    hr, hy = rays[held].contiguous(), y[held].contiguous()
    tr, ty = rays[train].contiguous(), y[train].contiguous()
    mlp.fit(tr, ty, steps=400, batch=16384, lr=2e-3)
    loss, _ = mlp.loss_grad(hr, hy, 16384)
    var = float(hy.var())
    assert loss < 0.1 * var, (loss, var)


@pytest.mark.gpu
@pytest.mark.parametrize("stride,W,H", [(2, 256, 192), (1, 96, 64), (3, 100, 70)])
def test_bound_cameras_equals_per_frame_calls(stride, W, H):
    """hulk_mlp_bound_cameras (every frame in one launch) writes each frame's
    bounds exactly as hulk_mlp_bound_camera does for that frame alone,
    ragged frame sizes and strides included."""
    import torch

    s, ctx, mlp, _, _ = _gpu_setup()
    cams = [s.camera(k, H / W) for k in (1, 5, 9, 13)]
    n = W * H
    one = torch.full((len(cams) * n,), -1.0, dtype=torch.float32, device="cuda")
    sep = torch.full((len(cams) * n,), -2.0, dtype=torch.float32, device="cuda")
    mlp.bound_cameras(cams, W, H, one, margin=0.1, stride=stride)
    for k, cam in enumerate(cams):
        mlp.bound_camera(cam, W, H, sep[k * n:(k + 1) * n], margin=0.1, stride=stride)
    a, b = one.cpu().numpy(), sep.cpu().numpy()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert (a > 0).all()  # every pixel of every frame written
