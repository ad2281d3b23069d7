"""Learned hit-distance bound: the MLP of hulk_mlp.hip driven from Python.

Training data are (ray, log t) pairs the tracer itself produces: primary
rays of earlier frames, t = nearest-hit distance, or the scene-box exit for
rays that miss (so a miss predicts "no useful bound" instead of a bound
that would force the fallback).  Device arrays are torch tensors (plumbing
only); every forward, backward and optimizer step runs in libhulk.so.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib as L
from ._lib import HulkError, lib, ptr
from .engine import GPUTrace

FLT_MAX = 3.4028234663852886e38


class HitDistanceMLP:
    def __init__(self, ctx: GPUTrace, seed: int = 1, max_batch: int = 16384):
        self.ctx = ctx
        self.max_batch = max_batch
        self.h = C.c_void_p()
        rc = lib().hulk_mlp_create(ctx.h, seed, max_batch, C.byref(self.h))
        if rc:
            raise HulkError(rc, "hulk_mlp_create failed")
        self.margin = 0.0
        self.stride = 1
        self.enabled = True  # select_margin turns the bound off where it saves no traversal work
        self._lib_stream = None

    def _c(self, rc):
        if rc:
            raise HulkError(rc, (lib().hulk_mlp_last_error(self.h) or b"").decode())

    def _order(self):
        """Device inputs come from torch's current stream: the library's
        stream waits for it before reading them."""
        import torch

        if self._lib_stream is None:
            self._lib_stream = torch.cuda.ExternalStream(self.ctx.stream)
        cur = torch.cuda.current_stream()
        if cur.cuda_stream != self._lib_stream.cuda_stream:
            self._lib_stream.wait_stream(cur)

    def close(self):
        if self.h:
            lib().hulk_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def param_count() -> int:
        return lib().hulk_mlp_param_count()

    def get_params(self):
        p = np.zeros(self.param_count(), np.float32)
        f = np.zeros(32 * 6, np.float32)
        self._c(lib().hulk_mlp_get_params(self.h, ptr(p), ptr(f)))
        return p, f

    def set_params(self, params=None, fourier=None):
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        f = None if fourier is None else np.ascontiguousarray(fourier, np.float32)
        self._c(lib().hulk_mlp_set_params(self.h, ptr(p), ptr(f)))

    def forward(self, rays_dev, n: int, pred_dev):
        self._order()
        self._c(lib().hulk_mlp_forward(self.h, ptr(rays_dev), n, ptr(pred_dev)))

    def loss_grad(self, rays_dev, y_dev, n: int):
        self._order()
        loss = C.c_float()
        g = np.zeros(self.param_count(), np.float32)
        self._c(lib().hulk_mlp_loss_grad(self.h, ptr(rays_dev), ptr(y_dev), n, C.byref(loss), ptr(g)))
        return loss.value, g

    def train_step(self, rays_dev, y_dev, n: int, lr: float, want_loss: bool = False):
        self._order()
        loss = C.c_float()
        self._c(lib().hulk_mlp_train_step(self.h, ptr(rays_dev), ptr(y_dev), n, lr,
                                          C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def bound_camera(self, cam, W: int, H: int, tbound_dev, margin: float | None = None, max_log: float = 60.0,
                     stride: int | None = None, sync: bool = True):
        self._order()
        m = self.margin if margin is None else margin
        st = self.stride if stride is None else stride
        self._c(lib().hulk_mlp_bound_camera(self.h, ptr(np.ascontiguousarray(cam, np.float32)), W, H, m, max_log,
                                            st, ptr(tbound_dev), 0 if sync else L.FLAG_NO_SYNC))

    def bound_cameras(self, cams, W: int, H: int, tbound_dev, margin: float | None = None, max_log: float = 60.0,
                      stride: int | None = None, sync: bool = True):
        """Bounds of several frames (same size) in one launch: frame f's at
        tbound_dev[f * W * H:]."""
        self._order()
        m = self.margin if margin is None else margin
        st = self.stride if stride is None else stride
        arr = np.ascontiguousarray(np.stack([np.asarray(c, np.float32) for c in cams]), np.float32)
        self._c(lib().hulk_mlp_bound_cameras(self.h, ptr(arr), len(cams), W, H, m, max_log, st, ptr(tbound_dev),
                                             0 if sync else L.FLAG_NO_SYNC))

    # ------------------------------------------------------------ training --

    def fit(self, rays, y, steps: int, batch: int = 16384, lr: float = 2e-3, seed: int = 0, log_every: int = 0):
        """rays: (N, 8) float32 cuda tensor, y: (N,) float32 cuda tensor.

        The batches are gathered on the library's stream: the kernels of a
        step read them there, and the caching allocator may hand a batch's
        memory to the next gather only once that stream is past the step."""
        import torch

        n = rays.shape[0]
        batch = min(batch, self.max_batch, n - n % 64)
        g = torch.Generator(device=rays.device)
        g.manual_seed(seed)
        losses = []
        lib_stream = torch.cuda.ExternalStream(self.ctx.stream, device=rays.device)
        lib_stream.wait_stream(torch.cuda.current_stream(rays.device))
        with torch.cuda.stream(lib_stream):
            for s in range(steps):
                idx = torch.randint(0, n, (batch,), device=rays.device, generator=g)
                rb = rays.index_select(0, idx).contiguous()
                yb = y.index_select(0, idx).contiguous()
                cur_lr = lr * 0.5 * (1 + math.cos(math.pi * s / steps))  # cosine decay
                want = bool(log_every) and (s % log_every == 0 or s == steps - 1)
                loss = self.train_step(rb, yb, batch, cur_lr, want_loss=want)
                if want:
                    losses.append((s, loss))
        torch.cuda.current_stream(rays.device).wait_stream(lib_stream)
        return losses

    def calibrate(self, rays, y, quantile: float = 0.95, chunk: int = 1 << 20) -> float:
        """margin = quantile of (y - y^) on held-out pairs, so that about
        `quantile` of the rays get a bound at or above their hit."""
        import torch

        res = []
        for a in range(0, rays.shape[0], chunk):
            r = rays[a:a + chunk].contiguous()
            pred = torch.empty(r.shape[0], dtype=torch.float32, device=r.device)
            self.forward(r, r.shape[0], pred)
            res.append(y[a:a + chunk] - pred)
        res = torch.cat(res)
        self.margin = float(torch.quantile(res.float()[:1 << 24], quantile).item())
        return self.margin


def calibrate_frames(mlp: HitDistanceMLP, ctx: GPUTrace, cams, W: int, H: int, quantile: float = 0.95,
                     stride: int = 1) -> float:
    """margin from the deployed bound itself: on training frames, the
    quantile of log(t) - log(bound at margin 0) over rays inside the box
    (t = nearest hit, or the box exit for a miss)."""
    import torch

    res = []
    tb = torch.empty(W * H, dtype=torch.float32, device="cuda")
    hits = torch.empty(W * H * 32, dtype=torch.uint8, device="cuda")
    for cam in cams:
        _, texit = camera_rays(ctx, cam, W, H)
        mlp.bound_camera(cam, W, H, tb, margin=0.0, max_log=80.0, stride=stride)
        ctx.trace_camera_dev(cam, W, H, hits.data_ptr())
        h = hits.view(torch.float32).view(-1, 8)
        hit = h[:, 3].contiguous().view(torch.int32) != -1
        lab = torch.where(hit, h[:, 0], texit)
        keep = lab > 0
        res.append(torch.log(lab[keep]) - torch.log(tb[keep]))
    r = torch.cat(res)
    mlp.stride = stride
    mlp.margin = float(torch.quantile(r[: 1 << 24], quantile).item())
    return mlp.margin


def select_margin(mlp: HitDistanceMLP, ctx: GPUTrace, cams, W: int, H: int,
                  quantiles=(0.5, 0.65, 0.8, 0.9, 0.95, 0.98, 0.995), stride: int = 2,
                  min_saving: float = 0.02, reps: int = 3) -> dict:
    """Margin chosen on calibration frames (frames the network was not
    trained on) by the traversal itself: for each quantile of
    log(t) - log(bound at margin 0), the time of the bounded primary traces
    (fallbacks included) with the context's current options; the fastest
    quantile wins.  The bound is enabled (mlp.enabled) only if bounded
    traces plus inference beat the unbounded traces by `min_saving`: a
    front-to-back traversal that stops at its first hit gains nothing from an
    upper bound, a traversal without early exit (kd-restart) does."""
    import torch

    n = W * H
    tbs = [torch.empty(n, dtype=torch.float32, device="cuda") for _ in cams]
    hits = torch.empty(n * 32, dtype=torch.uint8, device="cuda")
    lib_stream = torch.cuda.ExternalStream(ctx.stream)
    res = []
    for cam, tb in zip(cams, tbs):
        _, texit = camera_rays(ctx, cam, W, H)
        mlp.bound_camera(cam, W, H, tb, margin=0.0, max_log=80.0, stride=stride)
        ctx.trace_camera_dev(cam, W, H, hits.data_ptr())
        h = hits.view(torch.float32).view(-1, 8)
        hit = h[:, 3].contiguous().view(torch.int32) != -1
        lab = torch.where(hit, h[:, 0], texit)
        keep = lab > 0
        res.append(torch.log(lab[keep]) - torch.log(tb[keep]))
    r = torch.cat(res)[: 1 << 24]

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(lib_stream)
        for _ in range(reps):
            fn()
        e1.record(lib_stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    def traces(bounded):
        for cam, tb in zip(cams, tbs):
            ctx.trace_camera_dev(cam, W, H, hits.data_ptr(), tbound_dev=tb.data_ptr() if bounded else None,
                                 flags=L.FLAG_NO_SYNC)

    def infer(margin):
        for cam, tb in zip(cams, tbs):
            mlp.bound_camera(cam, W, H, tb, margin=margin, stride=stride, sync=False)

    base_ms = timed(lambda: traces(False))
    cands = {}
    for q in quantiles:
        m = float(torch.quantile(r, q).item())
        infer(m)
        cands[q] = (m, timed(lambda: traces(True)))
    q_best = min(cands, key=lambda q: cands[q][1])
    m_best, t_best = cands[q_best]
    infer_ms = timed(lambda: infer(m_best))
    saving = 1.0 - (t_best + infer_ms) / base_ms
    mlp.stride = stride
    mlp.margin = m_best
    mlp.enabled = saving >= min_saving
    return {"quantile": q_best, "margin": round(m_best, 4), "unbounded_ms": round(base_ms, 4),
            "bounded_ms": round(t_best, 4), "inference_ms": round(infer_ms, 4), "time_saved": round(saving, 4),
            "enabled": bool(mlp.enabled),
            "candidates": {str(q): [round(m, 4), round(t, 4)] for q, (m, t) in cands.items()}}


def camera_rays(ctx: GPUTrace, cam, W: int, H: int):
    """Primary rays (W*H, 8) and scene-box exit times (W*H,) as cuda tensors."""
    import torch

    n = W * H
    rays = torch.empty((n, 8), dtype=torch.float32, device="cuda")
    texit = torch.empty(n, dtype=torch.float32, device="cuda")
    rc = lib().hulk_camera_rays(ctx.h, ptr(np.ascontiguousarray(cam, np.float32)), W, H, ptr(rays), ptr(texit), 0)
    if rc:
        raise HulkError(rc, (lib().hulk_last_error(ctx.h) or b"").decode())
    return rays, texit


def training_pairs(ctx: GPUTrace, cams, W: int, H: int, n_samples: int | None = None, seed: int = 0):
    """(ray, log t) pairs from the tracer's own nearest hits on `cams`.
    Misses are labelled with the scene-box exit; rays outside the box are
    dropped."""
    import torch

    rays_all, y_all = [], []
    hits = torch.empty(W * H * 32, dtype=torch.uint8, device="cuda")
    for cam in cams:
        rays, texit = camera_rays(ctx, cam, W, H)
        ctx.trace_camera_dev(cam, W, H, hits.data_ptr())
        h = hits.view(torch.float32).view(-1, 8)
        t = h[:, 0]
        tri = h[:, 3].contiguous().view(torch.int32)
        hit = tri != -1
        lab = torch.where(hit, t, texit)
        keep = lab > 0
        rays_all.append(rays[keep])
        y_all.append(torch.log(lab[keep]))
    rays = torch.cat(rays_all)
    y = torch.cat(y_all)
    if n_samples is not None and n_samples < rays.shape[0]:
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        idx = torch.randperm(rays.shape[0], device="cuda", generator=g)[:n_samples]
        rays, y = rays[idx].contiguous(), y[idx].contiguous()
    return rays.contiguous(), y.contiguous()
