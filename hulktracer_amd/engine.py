"""Host-side mirror of the HulkEngine interfaces on the traverse-and-shade path.

Names follow the reference so a caller of the reference finds the same
seams:
  Scene                 - Scene arrays (HulkEngine/public/Scene.hpp:18-96)
  KDTreeCPU             - buildKDTree / getNodeArray / getPolygonArray
                          (HulkEngine/src/KDTreeCPU.hpp:53-72)
  GPUTrace              - one libhulk context (the OpenCL objects of
                          GLVisitorGPUTrace, GLVisitorGPUTrace.hpp:73-109)
  GLVisitorGPUTrace     - the renderer plugin: initializeVisitor / buildTree /
                          updateCamera / traceScene / launchTraceIterations
                          (GLVisitorGPUTrace.hpp:22-35,72)
  Settings              - --key=value flags with the reference defaults
                          (HulkEngine/src/Settings.cpp:11-53)
Every compute call goes through libhulk.so; nothing here traces rays.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from ._lib import HulkError, HIT_DTYPE, RAY_DTYPE, ptr, lib


def _check(ctx_handle, code: int) -> None:
    if code != 0:
        msg = lib().hulk_last_error(ctx_handle) if ctx_handle else b""
        raise HulkError(code, (msg or b"").decode())


# ------------------------------------------------------------- settings --

class Settings:
    """--key=value flags (Settings.cpp:11-53).  Unknown keys raise, as the
    reference aborts on them."""

    DEFAULTS = {
        "--renderType": "GPUTrace", "--fixedTic": "0", "--singleFrame": "0", "--saveImages": "0",
        "--dumpStats": "0", "--treeType": "SAHNlogN", "--kernelType": "GPUTrace_stackless",
        "--saveBin": "0", "--p2p": "0",
        # MI355X knobs of this library (autotuner space)
        "--blockSize": "256", "--shortStackDepth": "6", "--ldsTopLevels": "0", "--persistent": "0",
        "--shadeMode": "fused",
        # where buildTree builds the SAH-NlogN tree: "cpu" (KDTreeCPU, then
        # upload) or "gpu" (KDTreeGPU, built in HBM); same arrays either way
        "--treeBuild": "cpu",
    }

    def __init__(self, argv: list[str] | None = None):
        self.values = dict(self.DEFAULTS)
        for a in argv or []:
            if "=" not in a or a.split("=", 1)[0] not in self.DEFAULTS:
                raise ValueError(f"unknown setting {a!r}")
            k, v = a.split("=", 1)
            self.values[k] = v

    def __getitem__(self, k: str) -> str:
        return self.values[k]


# ---------------------------------------------------------------- scene --

@dataclass
class Scene:
    """Flat scene arrays: vertices (V,4) float32 with w = 1, triangles (T,4)
    uint32 (TriangleIndices), lights (L,8) float32 (color.xyz + type,
    data.xyzw), scene box (8,) float32 (min.xyzw, max.xyzw)."""

    vertices: np.ndarray
    triangles: np.ndarray
    lights: np.ndarray
    box: np.ndarray
    mirror: np.ndarray | None = None
    kind: int = 0
    seed: int = 0
    fov: float = 45.0
    name: str = "scene"
    _handle: int | None = None

    @classmethod
    def procedural(cls, kind: int, seed: int = 1, param: int = 0, frame: int = 0) -> "Scene":
        h = C.c_void_p()
        _check(None, lib().hulk_scene_create(kind, seed, param, frame, C.byref(h)))
        s = cls._from_handle(h, kind, seed)
        s._frame = frame
        return s

    @classmethod
    def from_obj(cls, path, time: float = 0.0) -> "Scene":
        """An OBJ file as the reference's OBJLoader reads it (OBJLoader.cpp:
        15-174), at scene time `time` (rotation about y, OBJTransformVisitor.cpp:
        22-41), with the loader's directional light."""
        h = C.c_void_p()
        _check(None, lib().hulk_scene_load_obj(str(path).encode(), float(time), C.byref(h)))
        s = cls._from_handle(h, 0, 0)
        s.name = str(path)
        return s

    @classmethod
    def from_aff(cls, path, time: float = 0.0) -> "Scene":
        """A BART scene file as the reference's AFFParser reads it
        (AFFParser.cpp:1469-1539) at scene time `time`: meshes under static and
        key-framed transforms, point lights; camera() gives the camera of its
        "camera" transform (aspect None: the file's resolution)."""
        h = C.c_void_p()
        _check(None, lib().hulk_scene_load_aff(str(path).encode(), float(time), C.byref(h)))
        cam = np.zeros(16, np.float32)
        try:
            _check(None, lib().hulk_scene_camera(h, 0, 0.0, ptr(cam)))
        except Exception:
            lib().hulk_scene_free(h)
            raise
        s = cls._from_handle(h, 0, 0)
        s.name = str(path)
        s._file_camera = (str(path), float(time))
        s._file_cam0 = cam
        return s

    @classmethod
    def from_dataset(cls, path) -> tuple["Scene", np.ndarray, np.ndarray]:
        """A paris.bin KD dataset (GLVisitorGPUTraceSSD::buildTree,
        GLVisitorGPUTraceSSD.cpp:43-107) as (scene, nodes, leaf polygons): the
        per-leaf vertices and triangles (triangle k = the k-th triangle of
        the leaves in node order; no lights) and the tree over them."""
        hs, ht = C.c_void_p(), C.c_void_p()
        _check(None, lib().hulk_dataset_load(str(path).encode(), C.byref(hs), C.byref(ht)))
        try:
            nodes, lp = KDTreeCPU._arrays(ht)
        finally:
            lib().hulk_kdtree_free(ht)
        s = cls._from_handle(hs, 0, 0)
        s.name = str(path)
        return s, nodes, lp

    @classmethod
    def _from_handle(cls, h, kind: int, seed: int) -> "Scene":
        def view(ctype, n, addr, shape, dtype):
            if n == 0:
                return np.zeros(shape, dtype)
            return np.ctypeslib.as_array((ctype * (int(np.prod(shape)))).from_address(addr)).reshape(shape).copy()

        try:
            nv = lib().hulk_scene_vertex_count(h)
            nt = lib().hulk_scene_triangle_count(h)
            nl = lib().hulk_scene_light_count(h)
            verts = view(C.c_float, nv, lib().hulk_scene_vertices(h), (nv, 4), np.float32)
            tris = view(C.c_uint32, nt, lib().hulk_scene_triangles(h), (nt, 4), np.uint32)
            lights = view(C.c_float, nl, lib().hulk_scene_lights(h), (nl, 8), np.float32)
            mirror = view(C.c_uint8, nt, lib().hulk_scene_mirror_mask(h), (nt,), np.uint8)
            box = np.zeros(8, np.float32)
            if nv:
                _check(None, lib().hulk_scene_get_box(h, ptr(box)))
        finally:
            lib().hulk_scene_free(h)
        return cls(verts, tris, lights, box, mirror, kind, seed)

    @classmethod
    def from_arrays(cls, vertices, triangles, lights=None) -> "Scene":
        v = np.ascontiguousarray(vertices, np.float32)
        if v.shape[1] == 3:
            v = np.concatenate([v, np.ones((len(v), 1), np.float32)], 1)
        t = np.ascontiguousarray(triangles, np.uint32)
        if t.shape[1] == 3:
            t = np.concatenate([t, np.zeros((len(t), 1), np.uint32)], 1)
        lt = np.zeros((0, 8), np.float32) if lights is None else np.ascontiguousarray(lights, np.float32)
        return cls(v, np.ascontiguousarray(t), lt, scene_box(v))

    def camera(self, frame: int = 0, aspect: float | None = 1.0) -> np.ndarray:
        """CameraPos (64 B) of a procedural scene's pose / frame, or of a
        BART file's camera transform (aspect None: the file's)."""
        if getattr(self, "_file_camera", None):
            if aspect is None:
                return self._file_cam0.copy()
            h = C.c_void_p()
            path, t = self._file_camera
            _check(None, lib().hulk_scene_load_aff(path.encode(), t, C.byref(h)))
            try:
                cam = np.zeros(16, np.float32)
                _check(None, lib().hulk_scene_camera(h, 0, float(aspect), ptr(cam)))
            finally:
                lib().hulk_scene_free(h)
            return cam
        if not self.kind:
            raise ValueError("camera poses exist only for procedural scenes; use make_camera")
        cam = np.zeros(16, np.float32)
        _check(None, lib().hulk_scene_kind_camera(self.kind, frame, aspect, ptr(cam)))
        return cam

    @property
    def n_tris(self) -> int:
        return int(self.triangles.shape[0])


def scene_box(vertices: np.ndarray) -> np.ndarray:
    """BoundingBoxVisitor box (BoundingBoxVisitor.cpp:27-79)."""
    v = np.ascontiguousarray(vertices, np.float32)
    box = np.zeros(8, np.float32)
    _check(None, lib().hulk_scene_box(ptr(v), len(v), ptr(box)))
    return box


def make_camera(from_, axis_angle=(0.0, 1.0, 0.0, 0.0), fov: float = 45.0, aspect: float = 1.0) -> np.ndarray:
    """updateCamera (GLVisitorCPUTrace.cpp:105-148) -> CameraPos as (16,) float32."""
    f = np.asarray(from_, np.float32)
    aa = np.asarray(axis_angle, np.float32)
    cam = np.zeros(16, np.float32)
    _check(None, lib().hulk_make_camera(ptr(f), ptr(aa), fov, aspect, ptr(cam)))
    return cam


def make_camera_lookat(from_, target, up=(0.0, 1.0, 0.0), fov: float = 45.0, aspect: float = 1.0) -> np.ndarray:
    f, t, u = (np.asarray(x, np.float32) for x in (from_, target, up))
    cam = np.zeros(16, np.float32)
    _check(None, lib().hulk_make_camera_lookat(ptr(f), ptr(t), ptr(u), fov, aspect, ptr(cam)))
    return cam


# -------------------------------------------------------------- kd tree --

class KDTreeCPU:
    """SAH-NlogN KD tree (KDTreeCPU::buildKDTree, KDTreeCPU.cpp:61-94)."""

    def __init__(self):
        self.m_treeNodes = np.zeros((0, 4), np.uint32)
        self.m_leafPolygons = np.zeros(0, np.uint32)

    def buildKDTree(self, scene: Scene, threads: int = 0) -> int:
        """threads: 0 = automatic, 1 = single-threaded; the arrays are the
        same for every value."""
        h = C.c_void_p()
        _check(None, lib().hulk_kdtree_build_threads(ptr(scene.vertices), len(scene.vertices),
                                                     ptr(scene.triangles), scene.n_tris, ptr(scene.box),
                                                     threads, C.byref(h)))
        try:
            self.m_treeNodes, self.m_leafPolygons = self._arrays(h)
        finally:
            lib().hulk_kdtree_free(h)
        return 0

    @staticmethod
    def _arrays(h) -> tuple[np.ndarray, np.ndarray]:
        nn = lib().hulk_kdtree_node_count(h)
        nl = lib().hulk_kdtree_leaf_poly_count(h)
        nodes = np.ctypeslib.as_array(
            (C.c_uint32 * (4 * nn)).from_address(lib().hulk_kdtree_nodes(h))).reshape(nn, 4).copy()
        lp = (np.ctypeslib.as_array((C.c_uint32 * nl).from_address(lib().hulk_kdtree_leaf_polys(h))).copy()
              if nl else np.zeros(0, np.uint32))
        return nodes, lp

    def saveDataSetToFile(self, scene: Scene, path="paris.bin"):
        """The tree and its per-leaf data as the reference's KD dataset
        (GLVisitorGPUTrace::saveDataSetToFile, GLVisitorGPUTrace.cpp:638-674)."""
        nodes = np.ascontiguousarray(self.getNodeArray(), np.uint32)
        lp = np.ascontiguousarray(self.getPolygonArray(), np.uint32)
        _check(None, lib().hulk_dataset_write(str(path).encode(), ptr(scene.vertices), len(scene.vertices),
                                              ptr(scene.triangles), scene.n_tris, ptr(nodes), len(nodes),
                                              ptr(lp) if len(lp) else None, len(lp)))

    def getNodeArray(self) -> np.ndarray:
        return self.m_treeNodes

    def getPolygonArray(self) -> np.ndarray:
        return self.m_leafPolygons


class KDTreeGPU(KDTreeCPU):
    """KD tree built on the GPU (KDTreeGPU, KDTreeGPU.hpp:12-23, a subclass
    of the CPU tree that the reference leaves empty).  buildKDTree uploads
    the scene to a context, builds there and keeps the tree installed in that
    context; the arrays are downloaded only when asked for."""

    def __init__(self, ctx: "GPUTrace | None" = None):
        super().__init__()
        self.ctx = ctx
        self._stale = True

    def buildKDTree(self, scene: Scene, upload: bool = True, capacity: int = 0) -> int:
        if self.ctx is None:
            self.ctx = GPUTrace()
        if upload:
            self.ctx.upload_scene(scene)
        self.ctx.build_kdtree_device(capacity)
        self._stale = True
        return 0

    def _fetch(self):
        if self._stale:
            self.m_treeNodes, self.m_leafPolygons = self.ctx.download_kdtree()
            self._stale = False

    def getNodeArray(self) -> np.ndarray:
        self._fetch()
        return self.m_treeNodes

    def getPolygonArray(self) -> np.ndarray:
        self._fetch()
        return self.m_leafPolygons


# -------------------------------------------------------------- context --

class GPUTrace:
    """One libhulk context (one HIP stream, device buffers owned here)."""

    def __init__(self, variant: int | str = L.FULLSTACK, block_size: int = 256, short_stack_depth: int = 6,
                 lds_top_levels: int = 0, persistent: bool = False, device: int = 0, stream: int | None = None,
                 pipeline: int = 0, leaf_batch: int = 1, coop_lanes: int = 0, tri_test: int = 0):
        """pipeline: 0 = automatic, 1 = queued shadow / mirror launches,
        2 = one fused launch per frame, 3 = fused with lane refill
        (hulk_pipeline).  leaf_batch: triangle records in flight per lane
        (1 or 2).  coop_lanes: lane-refill pipeline only; when at most this
        many lanes of a wave are at a leaf, the whole wave tests their
        leaves (0 = never).  tri_test: 0 = the reference's triangle test,
        1 = Moller-Trumbore (hulk_tri_test)."""
        if isinstance(variant, str):
            variant = L.VARIANTS[variant]
        self.opts = L.Opts(device=device, variant=variant, block_size=block_size,
                           short_stack_depth=short_stack_depth, lds_top_levels=lds_top_levels,
                           persistent=int(persistent), pipeline=int(pipeline), leaf_batch=int(leaf_batch),
                           coop_lanes=int(coop_lanes), tri_test=int(tri_test), stream=stream)
        self.h = C.c_void_p()
        _check(None, lib().hulk_create(C.byref(self.h), C.byref(self.opts)))

    def close(self):
        if self.h:
            lib().hulk_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, code):
        _check(self.h, code)

    def set_options(self, **kw):
        for k, v in kw.items():
            if k == "variant" and isinstance(v, str):
                v = L.VARIANTS[v]
            setattr(self.opts, k, int(v))
        self._c(lib().hulk_set_options(self.h, C.byref(self.opts)))

    @property
    def stream(self) -> int:
        return lib().hulk_stream(self.h)

    def upload_scene(self, scene: Scene):
        self._c(lib().hulk_upload_scene(self.h, ptr(scene.vertices), len(scene.vertices), ptr(scene.triangles),
                                        scene.n_tris, ptr(scene.box)))
        if scene.lights is not None:
            self.set_lights(scene.lights)

    def update_vertices(self, vertices: np.ndarray, box: np.ndarray):
        """New vertex positions of the uploaded scene (same count and
        triangles); the tree must be rebuilt or uploaded afterwards."""
        v = np.ascontiguousarray(vertices, np.float32)
        self._c(lib().hulk_update_vertices(self.h, ptr(v), len(v), ptr(np.ascontiguousarray(box, np.float32)), 0))

    def upload_kdtree(self, nodes: np.ndarray, leaf_polys: np.ndarray):
        nodes = np.ascontiguousarray(nodes, np.uint32)
        lp = np.ascontiguousarray(leaf_polys, np.uint32)
        self._c(lib().hulk_upload_kdtree(self.h, ptr(nodes), len(nodes), ptr(lp) if len(lp) else None, len(lp)))

    def trace_camera_streamed(self, cam: np.ndarray, W: int, H: int, pool_records: int = 0) -> tuple[np.ndarray, dict]:
        """Primary rays with the leaves' records paged into a device pool of
        `pool_records` records (hulk_trace_camera_streamed): kd-restart
        traversal, the reference's residency protocol and relaunch loop.
        Returns (hits, stats)."""
        hits = np.zeros(W * H, HIT_DTYPE)
        st = L.StreamStats()
        self._c(lib().hulk_trace_camera_streamed(self.h, ptr(np.ascontiguousarray(cam, np.float32)), W, H,
                                                 int(pool_records), ptr(hits), 0, C.byref(st)))
        return hits, {k: getattr(st, k) for k, _ in L.StreamStats._fields_}

    def build_kdtree_device(self, capacity: int = 0):
        """SAH-NlogN KD tree built on the GPU from the uploaded scene and
        installed as this context's tree (hulk_build_kdtree_device); same
        arrays as KDTreeCPU.buildKDTree."""
        self._c(lib().hulk_build_kdtree_device(self.h, int(capacity)))

    def download_kdtree(self) -> tuple[np.ndarray, np.ndarray]:
        """The current tree in the reference layout: (nodes (N,4) uint32,
        leaf polygon references (L,) uint32)."""
        nn, nl = C.c_uint32(), C.c_uint32()
        self._c(lib().hulk_download_kdtree(self.h, None, 0, C.byref(nn), None, 0, C.byref(nl)))
        nodes = np.zeros((nn.value, 4), np.uint32)
        lp = np.zeros(nl.value, np.uint32)
        self._c(lib().hulk_download_kdtree(self.h, ptr(nodes), nn.value, C.byref(nn),
                                           ptr(lp) if nl.value else None, nl.value, C.byref(nl)))
        return nodes, lp

    def device_build_info(self) -> dict:
        out = np.zeros(4, np.uint32)
        self._c(lib().hulk_device_build_info(self.h, ptr(out)))
        return dict(capacity=int(out[0]), retries=int(out[1]), levels=int(out[2]), max_jdepth=int(out[3]))

    def set_lights(self, lights: np.ndarray):
        lt = np.ascontiguousarray(lights, np.float32)
        self._c(lib().hulk_set_lights(self.h, ptr(lt) if len(lt) else None, len(lt)))

    def trace_camera(self, cam: np.ndarray, W: int, H: int, tbound: np.ndarray | None = None,
                     count: bool = False) -> np.ndarray:
        hits = np.zeros(W * H, HIT_DTYPE)
        tb = None if tbound is None else np.ascontiguousarray(tbound, np.float32)
        self._c(lib().hulk_trace_camera(self.h, ptr(np.ascontiguousarray(cam, np.float32)), W, H, ptr(tb),
                                        ptr(hits), L.FLAG_COUNT if count else 0))
        return hits

    def trace_rays(self, rays: np.ndarray, tbound: np.ndarray | None = None, count: bool = False) -> np.ndarray:
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = np.zeros(len(r), HIT_DTYPE)
        tb = None if tbound is None else np.ascontiguousarray(tbound, np.float32)
        self._c(lib().hulk_trace_rays(self.h, ptr(r), len(r), ptr(tb), ptr(hits), L.FLAG_COUNT if count else 0))
        return hits

    def shade(self, cam: np.ndarray, W: int, H: int, hits: np.ndarray, mode: int = L.SHADE_FUSED) -> np.ndarray:
        out = np.zeros(W * H, np.uint32)
        self._c(lib().hulk_shade(self.h, ptr(np.ascontiguousarray(cam, np.float32)), W, H, ptr(hits), ptr(out),
                                 mode, 0))
        return out

    def render_frame(self, cam: np.ndarray, W: int, H: int, max_bounce: int = 0, mirror: np.ndarray | None = None,
                     tbound: np.ndarray | None = None, count: bool = False) -> tuple[np.ndarray, np.ndarray]:
        out = np.zeros(W * H, np.uint32)
        hits = np.zeros(W * H, HIT_DTYPE)
        m = None if mirror is None else np.ascontiguousarray(mirror, np.uint8)
        tb = None if tbound is None else np.ascontiguousarray(tbound, np.float32)
        self._c(lib().hulk_render_frame(self.h, ptr(np.ascontiguousarray(cam, np.float32)), W, H, max_bounce,
                                        ptr(m), ptr(tb), ptr(out), ptr(hits), L.FLAG_COUNT if count else 0))
        return out, hits

    # device-pointer forms (inputs resident in HBM, e.g. torch tensors)
    def trace_camera_dev(self, cam: np.ndarray, W: int, H: int, hits_dev: int, tbound_dev: int | None = None,
                         flags: int = 0):
        self._c(lib().hulk_trace_camera(self.h, ptr(np.ascontiguousarray(cam, np.float32)), W, H,
                                        ptr(tbound_dev), ptr(hits_dev), flags | L.FLAG_DEVICE_PTRS))

    def render_frame_dev(self, cam: np.ndarray, W: int, H: int, max_bounce: int, mirror_dev: int | None,
                         rgba_dev: int, hits_dev: int | None = None, tbound_dev: int | None = None, flags: int = 0):
        self._c(lib().hulk_render_frame(self.h, ptr(np.ascontiguousarray(cam, np.float32)), W, H, max_bounce,
                                        ptr(mirror_dev), ptr(tbound_dev), ptr(rgba_dev), ptr(hits_dev),
                                        flags | L.FLAG_DEVICE_PTRS))

    def render_frame_static_dev(self, cam: np.ndarray, W: int, H: int, max_bounce: int, mirror_dev: int | None,
                                rgba_dev: int, hits_dev: int | None = None, tbound_dev: int | None = None,
                                flags: int = 0):
        """hulk_render_frame_static: the frame of render_frame_dev, allocating
        nothing (device pointers only; without hits_dev at most
        STATIC_MAX_RAYS rays)."""
        self._c(lib().hulk_render_frame_static(self.h, ptr(np.ascontiguousarray(cam, np.float32)), W, H, max_bounce,
                                               ptr(mirror_dev), ptr(tbound_dev), ptr(rgba_dev), ptr(hits_dev),
                                               flags))

    def sync(self):
        self._c(lib().hulk_sync(self.h))

    def batch(self, frames, mirror: np.ndarray | None, rgba_dev: int, tbound_dev: int | None = None,
              use_graph: bool = True) -> "Batch":
        """frames: list of (camera (16,) float32, width, height, max_bounce)."""
        arr = (L.FrameDesc * len(frames))()
        for i, (cam, w, h, mb) in enumerate(frames):
            arr[i].cam[:] = [float(x) for x in np.asarray(cam, np.float32)]
            arr[i].width, arr[i].height, arr[i].max_bounce = int(w), int(h), int(mb)
        m = None if mirror is None else np.ascontiguousarray(mirror, np.uint8)
        hb = C.c_void_p()
        self._c(lib().hulk_batch_create(self.h, arr, len(frames), ptr(m), ptr(tbound_dev), ptr(rgba_dev),
                                        int(use_graph), C.byref(hb)))
        return Batch(self, hb)

    def autotune(self, cam: np.ndarray, W: int, H: int, reps: int = 3, variants=None, frame: bool = False,
                 max_bounce: int = 0, mirror: np.ndarray | None = None):
        """Per-scene search (hulk_autotune); returns (best options dict, table).
        frame=True times whole frames (primary + shadow + shading, with
        max_bounce mirror generations when `mirror` is given; fused and queued
        pipelines) instead of the primary trace."""
        mask = 0
        for v in variants or []:
            mask |= 1 << (L.VARIANTS[v] if isinstance(v, str) else int(v))
        cap = 512
        table = (L.TuneEntry * cap)()
        n = C.c_uint32()
        best = L.Opts()
        m = None if mirror is None else np.ascontiguousarray(mirror, np.uint8)
        req = L.TuneRequest()
        req.cam[:] = [float(x) for x in np.ascontiguousarray(cam, np.float32).ravel()[:16]]
        req.width, req.height, req.reps, req.variant_mask = W, H, reps, mask
        req.objective = L.TUNE_FRAME if frame else L.TUNE_TRACE
        req.max_bounce = max_bounce
        req.mirror_mask = None if m is None else m.ctypes.data
        self._c(lib().hulk_autotune(self.h, C.byref(req), table, cap, C.byref(n), C.byref(best)))
        self.opts = best
        rows = [{f: getattr(table[i], f) for f, _ in L.TuneEntry._fields_} for i in range(n.value)]
        return {f: getattr(best, f) for f, _ in L.Opts._fields_ if f != "stream"}, rows

    def counters(self) -> dict:
        c = L.Counters()
        self._c(lib().hulk_get_counters(self.h, C.byref(c)))
        return {f: getattr(c, f) for f, _ in L.Counters._fields_}

    def reset_counters(self):
        self._c(lib().hulk_reset_counters(self.h))


def batch_scenes(ctxs, frames, mirrors, rgba_dev: int, tbound_dev: int | None = None,
                 use_graph: bool = True) -> "Batch":
    """hulk_batch_create_scenes: frames of several scenes in one launch.
    ctxs: one GPUTrace per scene (ctxs[0]'s options and stream run the
    batch); frames: list of (scene index, camera, width, height,
    max_bounce); mirrors: per scene a mirror flag array or None."""
    arr = (L.FrameDesc * len(frames))()
    for i, (k, cam, w, h, mb) in enumerate(frames):
        arr[i].cam[:] = [float(x) for x in np.asarray(cam, np.float32)]
        arr[i].width, arr[i].height, arr[i].max_bounce, arr[i].reserved = int(w), int(h), int(mb), int(k)
    handles = (C.c_void_p * len(ctxs))(*[cx.h.value for cx in ctxs])
    keep = [None if m is None else np.ascontiguousarray(m, np.uint8) for m in mirrors]
    mptr = (C.c_void_p * len(ctxs))(*[None if m is None else m.ctypes.data for m in keep])
    hb = C.c_void_p()
    ctxs[0]._c(lib().hulk_batch_create_scenes(handles, len(ctxs), arr, len(frames), mptr, ptr(tbound_dev),
                                              ptr(rgba_dev), int(use_graph), C.byref(hb)))
    b = Batch(ctxs[0], hb)
    b.scene_ctxs = list(ctxs)  # the batch reads every context's buffers
    return b


class Batch:
    """A hulk_batch: many frames in one (graph-captured) pipeline."""

    def __init__(self, ctx: GPUTrace, handle: C.c_void_p):
        self.ctx, self.h = ctx, handle

    @property
    def rays(self) -> int:
        return lib().hulk_batch_ray_count(self.h)

    def launch(self, flags: int = 0):
        _check(self.ctx.h, lib().hulk_batch_launch(self.h, flags))

    def primary_hits(self) -> np.ndarray:
        out = np.zeros(self.rays, HIT_DTYPE)
        _check(self.ctx.h, lib().hulk_batch_primary_hits(self.h, ptr(out), 0))
        return out

    def close(self):
        if self.h:
            lib().hulk_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------- renderer plugin --

class GLVisitorGPUTrace:
    """Headless renderer plugin with the reference seam (GLVisitorGPUTrace.hpp:
    22-35,72): initializeVisitor once, then per frame traverseBegin =
    buildTree -> updateCamera -> traceScene (GLVisitorGPUTrace.cpp:598-631).
    Unlike the reference (static firstTime, :258-266) the tree is rebuilt on
    every frame, as the CPU path does (GLVisitorCPUTrace.cpp:52-55)."""

    def __init__(self, scene: Scene, settings: Settings | None = None, width: int = 256, height: int = 256):
        self.m_pScene = scene
        self.settings = settings or Settings()
        self.m_resx, self.m_resy = width, height
        self.m_pKdtree: KDTreeCPU | None = None
        self.ctx: GPUTrace | None = None
        self.m_pCamPos: np.ndarray | None = None
        self.frame = 0
        self.framebuffer: np.ndarray | None = None
        self.hits: np.ndarray | None = None

    def initializeVisitor(self):
        s = self.settings
        self.ctx = GPUTrace(variant=s["--kernelType"], block_size=int(s["--blockSize"]),
                            short_stack_depth=int(s["--shortStackDepth"]), lds_top_levels=int(s["--ldsTopLevels"]),
                            persistent=s["--persistent"] == "1")
        if s["--treeBuild"] not in ("cpu", "gpu"):
            raise ValueError("--treeBuild must be cpu or gpu")
        self.m_pKdtree = KDTreeGPU(self.ctx) if s["--treeBuild"] == "gpu" else KDTreeCPU()

    def buildTree(self):
        if isinstance(self.m_pKdtree, KDTreeGPU):
            self.m_pKdtree.buildKDTree(self.m_pScene)  # uploads the scene, builds in place
            return
        self.m_pKdtree.buildKDTree(self.m_pScene)
        self.ctx.upload_scene(self.m_pScene)
        self.ctx.upload_kdtree(self.m_pKdtree.getNodeArray(), self.m_pKdtree.getPolygonArray())

    def updateCamera(self):
        self.m_pCamPos = self.m_pScene.camera(self.frame, self.m_resy / self.m_resx)

    def launchTraceIterations(self):
        # every leaf is resident, so one launch finishes every ray
        self.hits = self.ctx.trace_camera(self.m_pCamPos, self.m_resx, self.m_resy)

    def traceScene(self) -> np.ndarray:
        self.launchTraceIterations()
        mode = L.SHADE_PICTURE if self.settings["--shadeMode"] == "picture" else L.SHADE_FUSED
        self.framebuffer = self.ctx.shade(self.m_pCamPos, self.m_resx, self.m_resy, self.hits, mode)
        return self.framebuffer

    def traverseBegin(self, node=None) -> bool:
        self.buildTree()
        self.updateCamera()
        self.traceScene()
        self.frame += 1
        return False
