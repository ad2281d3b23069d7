"""CPU oracle of the traverse-and-shade hot path - TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
this package; the product library never does.  It wraps oracle/hulk_oracle.c
(a C restatement of the reference arithmetic, built by oracle/Makefile into
oracle/_build/liboracle.so) and oracle/kdtree_ref.py (a pure-Python
restatement of the SAH-NlogN builder for small scenes).

Parity status of each oracle function (how it is pinned) is listed in
tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"

V_BRUTE, V_FULLSTACK, V_SHORTSTACK, V_LDSSTACK, V_RESTART, V_ROPES = range(6)
MT = 0x100  # or-ed into a variant: trace with the Moller-Trumbore test (ORC_MT)
NO_HIT = 0xFFFFFFFF
FLT_MAX = np.float32(3.4028234663852886e38)

HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("tri", "<u4"),
                      ("nx", "<f4"), ("ny", "<f4"), ("nz", "<f4"), ("flags", "<u4")])


class RopeView(C.Structure):
    _fields_ = [("ropes", C.c_void_p), ("cells", C.c_void_p), ("kids", C.c_void_p), ("n_slots", C.c_uint32)]


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        vp, u32, i32 = C.c_void_p, C.c_uint32, C.c_int
        scene_args = [vp, vp, u32, vp, vp, u32, vp, C.POINTER(RopeView)]
        L.orc_trace_camera.argtypes = [i32, u32, *scene_args, vp, u32, u32, vp, vp, vp, i32]
        L.orc_trace_rays.argtypes = [i32, u32, *scene_args, vp, u32, vp, vp, vp, i32]
        L.orc_shade.argtypes = [i32, u32, *scene_args, vp, u32, u32, vp, vp, u32, i32, vp, i32]
        L.orc_render.argtypes = [i32, u32, *scene_args, vp, u32, u32, vp, u32, u32, vp, vp, vp, i32]
        L.orc_cast_rays.argtypes = [vp, u32, u32, vp]
        L.orc_ray_box.argtypes = [vp, vp, vp, vp]
        L.orc_tri_test_ex.argtypes = [vp, vp, vp, vp, vp]
        L.orc_tri_test_mt_ex.argtypes = [vp, vp, vp, vp, vp]
        L.orc_tri_test_f64.argtypes = [vp, vp, vp, vp, vp]
        L.orc_build_ropes.argtypes = [vp, u32, vp, vp, vp, vp]
        L.orc_build_ropes.restype = u32
        L.orc_pack_rgb.argtypes = [C.c_float, C.c_float, C.c_float]
        L.orc_pack_rgb.restype = u32
        L.orc_pcache_size.restype = C.c_size_t
        L.orc_cpu_build_cache.argtypes = [vp, vp, u32, vp]
        L.orc_cpu_trace_scene.argtypes = [vp, vp, u32, vp, vp, u32, vp, vp, vp, u32, u32, u32, u32, vp, u32, vp, i32]
        L.orc_cpu_trace_scene.restype = C.c_uint64
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


class OracleScene:
    """Scene arrays + KD tree as the reference holds them (Scene::m_vertices,
    m_vertexIndices, m_sceneBox; KDTreeCPU::m_treeNodes, m_leafPolygons)."""

    def __init__(self, vertices, triangles, box, nodes=None, leaf_polys=None, lights=None, mirror=None):
        self.v = np.ascontiguousarray(vertices, np.float32)
        self.t = np.ascontiguousarray(triangles, np.uint32)
        self.box = np.ascontiguousarray(box, np.float32)
        self.nodes = np.ascontiguousarray(nodes if nodes is not None else np.zeros((1, 4), np.uint32), np.uint32)
        self.lp = np.ascontiguousarray(leaf_polys if leaf_polys is not None else np.zeros(1, np.uint32), np.uint32)
        if len(self.lp) == 0:
            self.lp = np.zeros(1, np.uint32)
        self.lights = np.ascontiguousarray(lights if lights is not None else np.zeros((0, 8)), np.float32)
        self.mirror = None if mirror is None else np.ascontiguousarray(mirror, np.uint8)
        self._rv = None

    def rope_view(self) -> RopeView:
        if self._rv is None:
            n = len(self.nodes)
            cap = 2 * n + 2
            self._ropes = np.zeros(6 * cap, np.uint32)
            self._cells = np.zeros((cap, 8), np.float32)
            self._kids = np.zeros(2 * cap, np.uint32)
            slots = lib().orc_build_ropes(_p(self.nodes), n, _p(self.box), _p(self._ropes), _p(self._cells),
                                          _p(self._kids))
            self._rv = RopeView(self._ropes.ctypes.data, self._cells.ctypes.data, self._kids.ctypes.data, slots)
        return self._rv

    def _args(self):
        return (_p(self.v), _p(self.t), len(self.t), _p(self.box), _p(self.nodes), len(self.nodes), _p(self.lp),
                C.byref(self.rope_view()))

    def trace_camera(self, variant, cam, W, H, depth=6, tbound=None, counts=False, threads=0):
        out = np.zeros(W * H, HIT_DTYPE)
        cnt = np.zeros(3 * W * H, np.uint32) if counts else None
        tb = None if tbound is None else np.ascontiguousarray(tbound, np.float32)
        lib().orc_trace_camera(variant, depth, *self._args(), _p(np.ascontiguousarray(cam, np.float32)), W, H,
                               _p(tb), _p(out), _p(cnt), threads)
        return (out, cnt.reshape(-1, 3)) if counts else out

    def trace_rays(self, variant, rays, depth=6, tbound=None, counts=False, threads=0):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.zeros(len(r), HIT_DTYPE)
        cnt = np.zeros(3 * len(r), np.uint32) if counts else None
        tb = None if tbound is None else np.ascontiguousarray(tbound, np.float32)
        lib().orc_trace_rays(variant, depth, *self._args(), _p(r), len(r), _p(tb), _p(out), _p(cnt), threads)
        return (out, cnt.reshape(-1, 3)) if counts else out

    def shade(self, variant, cam, W, H, hits, mode=0, depth=6, threads=0):
        out = np.zeros(W * H, np.uint32)
        lib().orc_shade(variant, depth, *self._args(), _p(np.ascontiguousarray(cam, np.float32)), W, H,
                        _p(np.ascontiguousarray(hits)), _p(self.lights), len(self.lights), mode, _p(out), threads)
        return out

    def render(self, variant, cam, W, H, max_bounce=0, depth=6, threads=0):
        out = np.zeros(W * H, np.uint32)
        counts = np.zeros(2, np.uint64)
        lib().orc_render(variant, depth, *self._args(), _p(np.ascontiguousarray(cam, np.float32)), W, H,
                         _p(self.lights), len(self.lights), max_bounce, _p(self.mirror), _p(out), _p(counts), threads)
        return out, counts

    # reference CPU path (GLVisitorCPUTrace::traceScene port), the CPU baseline
    def cpu_trace_scene(self, cam, W, H, row0=0, rows=None, threads=0, out=True):
        if not hasattr(self, "_pcache"):
            sz = lib().orc_pcache_size()
            self._pcache = np.zeros(len(self.t) * sz, np.uint8)
            lib().orc_cpu_build_cache(_p(self.v), _p(self.t), len(self.t), _p(self._pcache))
        rows = H if rows is None else rows
        img = np.zeros(W * H, np.uint32) if out else None
        n = lib().orc_cpu_trace_scene(_p(self.v), _p(self.t), len(self.t), _p(self.box), _p(self.nodes),
                                      len(self.nodes), _p(self.lp), _p(self._pcache),
                                      _p(np.ascontiguousarray(cam, np.float32)), W, H, row0, rows, _p(self.lights),
                                      len(self.lights), _p(img), threads)
        return n, img


def cast_rays(cam, W, H) -> np.ndarray:
    out = np.zeros((W * H, 8), np.float32)
    lib().orc_cast_rays(_p(np.ascontiguousarray(cam, np.float32)), W, H, _p(out))
    return out


def tri_test(ray, p1, p2, p3):
    """checkRayToPoly: returns None or (t, s1, s2, det, n)."""
    r = np.ascontiguousarray(ray, np.float32).reshape(8)
    ps = [np.ascontiguousarray(np.r_[np.asarray(p, np.float32)[:3], np.float32(1)], np.float32) for p in (p1, p2, p3)]
    out = np.zeros(7, np.float32)
    if lib().orc_tri_test_ex(_p(r), *[_p(p) for p in ps], _p(out)):
        return out[0], out[1], out[2], out[3], out[4:7].copy()
    return None


def tri_test_mt(ray, p1, p2, p3):
    """Moller-Trumbore (ORC_MT mode): returns None or (t, u, v, 1, n)."""
    r = np.ascontiguousarray(ray, np.float32).reshape(8)
    ps = [np.ascontiguousarray(np.r_[np.asarray(p, np.float32)[:3], np.float32(1)], np.float32) for p in (p1, p2, p3)]
    out = np.zeros(7, np.float32)
    if lib().orc_tri_test_mt_ex(_p(r), *[_p(p) for p in ps], _p(out)):
        return out[0], out[1], out[2], out[3], out[4:7].copy()
    return None


def tri_test_f64(ray, p1, p2, p3):
    r = np.ascontiguousarray(ray, np.float32).reshape(8)
    ps = [np.ascontiguousarray(np.r_[np.asarray(p, np.float32)[:3], np.float32(1)], np.float32) for p in (p1, p2, p3)]
    out = np.zeros(3, np.float64)
    if lib().orc_tri_test_f64(_p(r), *[_p(p) for p in ps], _p(out)):
        return out
    return None


def ray_box(box, ray):
    b = np.ascontiguousarray(box, np.float32)
    r = np.ascontiguousarray(ray, np.float32).reshape(8)
    ts = np.zeros(1, np.float32)
    te = np.zeros(1, np.float32)
    ok = lib().orc_ray_box(_p(b), _p(r), _p(ts), _p(te))
    return (float(ts[0]), float(te[0])) if ok else None


def pack_rgb(r, g, b) -> int:
    return int(lib().orc_pack_rgb(r, g, b))
