"""Shared scene / tree / oracle fixtures for the tests (cached per process)."""
from __future__ import annotations

import functools

import numpy as np

import hulktracer_amd as H
import oracle as O

HIT_FIELDS = ("t", "u", "v", "tri", "nx", "ny", "nz")


@functools.lru_cache(maxsize=None)
def scene_and_tree(kind: int, seed: int = 1, param: int = 0, frame: int = 0):
    s = H.Scene.procedural(kind, seed=seed, param=param, frame=frame)
    kd = H.KDTreeCPU()
    kd.buildKDTree(s)
    return s, kd.getNodeArray(), kd.getPolygonArray()


@functools.lru_cache(maxsize=None)
def oracle_scene(kind: int, seed: int = 1, param: int = 0, frame: int = 0):
    s, nodes, lp = scene_and_tree(kind, seed, param, frame)
    return O.OracleScene(s.vertices, s.triangles, s.box, nodes, lp, s.lights, s.mirror)


def hit_mismatch(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Indices of rays whose hit records differ in any bit (flags excluded)."""
    bad = np.zeros(len(a), bool)
    for f in HIT_FIELDS:
        bad |= a[f].view(np.uint32) != b[f].view(np.uint32)
    return np.nonzero(bad)[0]


def describe(a: np.ndarray, b: np.ndarray, idx: np.ndarray, n: int = 5) -> str:
    rows = []
    for i in idx[:n]:
        rows.append(f"ray {i}: got tri={a['tri'][i]} t={a['t'][i]!r} u={a['u'][i]!r} v={a['v'][i]!r} | "
                    f"oracle tri={b['tri'][i]} t={b['t'][i]!r} u={b['u'][i]!r} v={b['v'][i]!r}")
    return "\n".join(rows)


def gpu_ctx(scene, nodes, lp, variant, **kw) -> H.GPUTrace:
    ctx = H.GPUTrace(variant=variant, **kw)
    ctx.upload_scene(scene)
    ctx.upload_kdtree(nodes, lp)
    return ctx


def random_rays(n: int, box: np.ndarray, seed: int = 0, degenerate: bool = True) -> np.ndarray:
    """Rays from inside / outside the box, with some axis-aligned and
    zero-component directions."""
    rng = np.random.default_rng(seed)
    lo, hi = box[:3].astype(np.float64), box[4:7].astype(np.float64)
    c, ext = (lo + hi) / 2, (hi - lo)
    o = c + (rng.random((n, 3)) - 0.5) * ext * 2.0
    tgt = c + (rng.random((n, 3)) - 0.5) * ext
    d = tgt - o
    if degenerate:
        k = n // 8
        d[np.arange(k), rng.integers(0, 3, k)] = 0.0     # one zero component
        d[k:2 * k, :2] = 0.0                              # axis aligned
        o[2 * k:3 * k] = c                                # origin at the box centre
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 4:7] = d
    bad = ~np.any(rays[:, 4:7] != 0, axis=1)
    rays[bad, 6] = -1.0
    return rays
