"""FP64 reference of the learned hit-distance MLP - TEST INFRASTRUCTURE ONLY.

The network (no reference counterpart; fixed by the build):
  x   = ((o - c) / h, d / |d|)                  6 inputs; c, h = scene box centre / half size
  z   = F x                                     F: 32 x 6 fixed Fourier matrix
  e   = (sin z, cos z)                          64-dim encoding
  h1  = relu(W1 e  + b1)                        W1: 128 x 64
  h2  = relu(W2 h1 + b2)                        W2: 128 x 128
  h3  = relu(W3 h2 + b3)                        W3: 128 x 128
  y^  = w4 . h3 + b4                            predicted log t
  L   = mean((y^ - y)^2)
With bf16=True the forward rounds weights, e, h1, h2, h3 and the backward
rounds dz3, dz2, dz1 to bfloat16 at the same points as the GPU kernels,
so the comparison with the GPU only sees FP32 accumulation-order
differences.  Parameter vector order: W1, b1, W2, b2, W3, b3, w4, b4.
"""
from __future__ import annotations

import numpy as np

SIZES = [("W1", (128, 64)), ("b1", (128,)), ("W2", (128, 128)), ("b2", (128,)), ("W3", (128, 128)),
         ("b3", (128,)), ("w4", (128,)), ("b4", (1,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def bf16(a: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest, ties to even) through float32, as float64."""
    f = np.asarray(a, np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p: np.ndarray) -> dict:
    out, k = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[k:k + n], np.float64).reshape(shape)
        k += n
    return out


def features(rays: np.ndarray, box: np.ndarray, fourier: np.ndarray) -> np.ndarray:
    """Encoding e (N x 64) in float32 arithmetic order of the kernel:
    x_i = (o_i - c_i) * (1 / h_i); d normalized by 1 / sqrt(d.d);
    z_j = ((((F_j0 x0 + F_j1 x1) + F_j2 x2) + F_j3 x3) + F_j4 x4) + F_j5 x5."""
    f32 = np.float32
    r = np.asarray(rays, f32).reshape(-1, 8)
    lo, hi = np.asarray(box[:3], f32), np.asarray(box[4:7], f32)
    c = (lo + hi) * f32(0.5)
    ih = f32(1) / ((hi - lo) * f32(0.5))
    o = (r[:, 0:3] - c) * ih
    d = r[:, 4:7]
    dd = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    d = d * (f32(1) / np.sqrt(dd))[:, None]
    x = np.concatenate([o, d], 1).astype(f32)
    F = np.asarray(fourier, f32).reshape(32, 6)
    z = np.zeros((len(x), 32), f32)
    for i in range(6):
        z = z + F[None, :, i] * x[:, i:i + 1]
    return np.concatenate([np.sin(z.astype(np.float64)), np.cos(z.astype(np.float64))], 1)


def forward(p: np.ndarray, e: np.ndarray, use_bf16: bool = True):
    P = unpack(p)
    q = bf16 if use_bf16 else (lambda a: np.asarray(a, np.float64))
    e = q(e)
    W1, W2, W3, w4 = (q(P[k]) for k in ("W1", "W2", "W3", "w4"))
    a1 = e @ W1.T + P["b1"]
    h1 = q(np.maximum(a1, 0))
    a2 = h1 @ W2.T + P["b2"]
    h2 = q(np.maximum(a2, 0))
    a3 = h2 @ W3.T + P["b3"]
    h3 = q(np.maximum(a3, 0))
    y = h3 @ w4 + P["b4"][0]
    return y, (e, h1, h2, h3, W1, W2, W3, w4)


def loss_and_grad(p: np.ndarray, e: np.ndarray, target: np.ndarray, use_bf16: bool = True):
    q = bf16 if use_bf16 else (lambda a: np.asarray(a, np.float64))
    y, (e, h1, h2, h3, W1, W2, W3, w4) = forward(p, e, use_bf16)
    n = len(y)
    r = y - np.asarray(target, np.float64)
    L = float(np.mean(r * r))
    dy = 2.0 * r / n
    g = {}
    g["w4"] = h3.T @ dy
    g["b4"] = np.array([dy.sum()])
    dz3 = q(dy[:, None] * w4[None, :] * (h3 > 0))
    g["W3"] = dz3.T @ h2
    g["b3"] = dz3.sum(0)
    dz2 = q((dz3 @ W3) * (h2 > 0))
    g["W2"] = dz2.T @ h1
    g["b2"] = dz2.sum(0)
    dz1 = q((dz2 @ W2) * (h1 > 0))
    g["W1"] = dz1.T @ e
    g["b1"] = dz1.sum(0)
    flat = np.concatenate([g[name].reshape(-1) for name, _ in SIZES])
    return L, flat
