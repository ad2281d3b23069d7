"""ctypes binding of libhulk.so (the C ABI in include/hulk_abi.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback for any traversal or shading entry point.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libhulk.so"

# Variants (hulk_variant) under the reference --kernelType names
# (HulkEngine/src/Settings.cpp:50).
VARIANTS = {
    "GPUTrace_bruteforce": 0,
    "GPUTrace": 1,
    "GPUTrace_shortStack": 2,
    "GPUTrace_LDSshortStack": 3,
    "GPUTrace_stackless": 4,
    "GPUTrace_ropes": 5,
}
VARIANT_NAMES = {v: k for k, v in VARIANTS.items()}
BRUTEFORCE, FULLSTACK, SHORTSTACK, LDSSTACK, RESTART, ROPES = range(6)
TRI_REFERENCE, TRI_MOLLER_TRUMBORE = 0, 1  # hulk_tri_test

FLAG_DEVICE_PTRS = 1
FLAG_ANYHIT = 2
FLAG_COUNT = 4
FLAG_NO_SYNC = 8
PIPELINE_AUTO, PIPELINE_QUEUED, PIPELINE_FUSED = range(3)
TUNE_TRACE, TUNE_FRAME = 0, 1
SHADE_FUSED = 0
SHADE_PICTURE = 1
NO_HIT = 0xFFFFFFFF
HITF_BOUNDED_FALLBACK = 1
STATIC_MAX_RAYS = 262144  # HULK_STATIC_MAX_RAYS

SCENE_RANDOM_TRIS = 1
SCENE_TERRAIN_SPHERES = 2
SCENE_ANIMATED_TORI = 3

STATUS = {
    0: "ok", 1: "e_allocFail", 2: "e_fileOpenFail", 3: "e_fileCloseFail", 4: "e_illegalInput",
    5: "e_nodeNotFound", 100: "hip error", 101: "no scene", 102: "no tree", 103: "unsupported",
}

# numpy views of the ABI structs
HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("tri", "<u4"),
                      ("nx", "<f4"), ("ny", "<f4"), ("nz", "<f4"), ("flags", "<u4")])
RAY_DTYPE = np.dtype([("o", "<f4", (4,)), ("d", "<f4", (4,))])


class HulkError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class Opts(C.Structure):
    _fields_ = [("device", C.c_int32), ("variant", C.c_int32), ("block_size", C.c_uint32),
                ("short_stack_depth", C.c_uint32), ("lds_top_levels", C.c_uint32),
                ("persistent", C.c_uint32), ("pipeline", C.c_uint32), ("leaf_batch", C.c_uint32),
                ("coop_lanes", C.c_uint32), ("tri_test", C.c_uint32), ("stream", C.c_void_p)]


class Counters(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("node_visits", C.c_uint64), ("leaf_visits", C.c_uint64),
                ("tri_tests", C.c_uint64), ("hits", C.c_uint64), ("fallbacks", C.c_uint64),
                ("shadow_rays", C.c_uint64), ("last_kernel_ms", C.c_float), ("last_shade_ms", C.c_float)]


class FrameDesc(C.Structure):
    _fields_ = [("cam", C.c_float * 16), ("width", C.c_uint32), ("height", C.c_uint32),
                ("max_bounce", C.c_uint32), ("reserved", C.c_uint32)]


class TuneEntry(C.Structure):
    _fields_ = [("variant", C.c_int32), ("block_size", C.c_uint32), ("short_stack_depth", C.c_uint32),
                ("lds_top_levels", C.c_uint32), ("persistent", C.c_uint32), ("pipeline", C.c_uint32),
                ("leaf_batch", C.c_uint32), ("coop_lanes", C.c_uint32), ("ms", C.c_float)]


class TuneRequest(C.Structure):
    _fields_ = [("cam", C.c_float * 16), ("width", C.c_uint32), ("height", C.c_uint32), ("reps", C.c_uint32),
                ("variant_mask", C.c_uint32), ("objective", C.c_uint32), ("max_bounce", C.c_uint32),
                ("mirror_mask", C.c_void_p)]


class StreamStats(C.Structure):
    _fields_ = [("launches", C.c_uint32), ("leaves_loaded", C.c_uint32), ("leaves_evicted", C.c_uint32),
                ("max_queue", C.c_uint32), ("bytes_loaded", C.c_uint64), ("ms", C.c_float),
                ("pool_records", C.c_uint32)]


EXPORTS = {
    # name: (restype, argtypes)
    "hulk_create": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(Opts)]),
    "hulk_destroy": (None, [C.c_void_p]),
    "hulk_last_error": (C.c_char_p, [C.c_void_p]),
    "hulk_set_options": (C.c_int, [C.c_void_p, C.POINTER(Opts)]),
    "hulk_get_options": (C.c_int, [C.c_void_p, C.POINTER(Opts)]),
    "hulk_stream": (C.c_void_p, [C.c_void_p]),
    "hulk_sync": (C.c_int, [C.c_void_p]),
    "hulk_upload_scene": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p]),
    "hulk_upload_kdtree": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32]),
    "hulk_update_vertices": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32]),
    "hulk_set_lights": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "hulk_trace_camera": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                                    C.c_uint32]),
    "hulk_trace_rays": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32]),
    "hulk_shade": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32,
                             C.c_uint32]),
    "hulk_render_frame": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]),
    "hulk_render_frame_static": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                           C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]),
    "hulk_get_counters": (C.c_int, [C.c_void_p, C.POINTER(Counters)]),
    "hulk_reset_counters": (C.c_int, [C.c_void_p]),
    "hulk_batch_create": (C.c_int, [C.c_void_p, C.POINTER(FrameDesc), C.c_uint32, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p)]),
    "hulk_batch_create_scenes": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint32, C.POINTER(FrameDesc), C.c_uint32,
                                           C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p, C.c_uint32,
                                           C.POINTER(C.c_void_p)]),
    "hulk_batch_launch": (C.c_int, [C.c_void_p, C.c_uint32]),
    "hulk_batch_ray_count": (C.c_uint32, [C.c_void_p]),
    "hulk_batch_destroy": (None, [C.c_void_p]),
    "hulk_batch_primary_hits": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "hulk_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(TuneEntry), C.c_uint32, C.POINTER(C.c_uint32),
                                C.POINTER(Opts)]),
    "hulk_mlp_create": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.POINTER(C.c_void_p)]),
    "hulk_mlp_destroy": (None, [C.c_void_p]),
    "hulk_mlp_last_error": (C.c_char_p, [C.c_void_p]),
    "hulk_mlp_param_count": (C.c_uint32, []),
    "hulk_mlp_get_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "hulk_mlp_set_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "hulk_mlp_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "hulk_mlp_loss_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_float),
                                     C.c_void_p]),
    "hulk_mlp_train_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_float,
                                      C.POINTER(C.c_float)]),
    "hulk_mlp_bound_cameras": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float,
                                         C.c_float, C.c_uint32, C.c_void_p, C.c_uint32]),
    "hulk_mlp_bound_camera": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_float, C.c_float,
                                        C.c_uint32, C.c_void_p, C.c_uint32]),
    "hulk_ctx_scene_box": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hulk_selftest_reciprocal": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hulk_camera_rays": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                                   C.c_uint32]),
    "hulk_trace_camera_streamed": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                             C.c_void_p, C.c_uint32, C.POINTER(StreamStats)]),
    "hulk_build_kdtree_device": (C.c_int, [C.c_void_p, C.c_uint32]),
    "hulk_download_kdtree": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32), C.c_void_p,
                                       C.c_uint32, C.POINTER(C.c_uint32)]),
    "hulk_device_build_info": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hulk_kdtree_build": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p,
                                    C.POINTER(C.c_void_p)]),
    "hulk_kdtree_build_threads": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p,
                                            C.c_uint32, C.POINTER(C.c_void_p)]),
    "hulk_kdtree_node_count": (C.c_uint32, [C.c_void_p]),
    "hulk_kdtree_leaf_poly_count": (C.c_uint32, [C.c_void_p]),
    "hulk_kdtree_nodes": (C.c_void_p, [C.c_void_p]),
    "hulk_kdtree_leaf_polys": (C.c_void_p, [C.c_void_p]),
    "hulk_kdtree_free": (None, [C.c_void_p]),
    "hulk_scene_box": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "hulk_make_camera": (C.c_int, [C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_void_p]),
    "hulk_make_camera_lookat": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_void_p]),
    "hulk_scene_create": (C.c_int, [C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)]),
    "hulk_scene_load_obj": (C.c_int, [C.c_char_p, C.c_float, C.POINTER(C.c_void_p)]),
    "hulk_scene_load_aff": (C.c_int, [C.c_char_p, C.c_float, C.POINTER(C.c_void_p)]),
    "hulk_dataset_write": (C.c_int, [C.c_char_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p,
                                     C.c_uint32, C.c_void_p, C.c_uint32]),
    "hulk_dataset_load": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "hulk_scene_free": (None, [C.c_void_p]),
    "hulk_scene_vertex_count": (C.c_uint32, [C.c_void_p]),
    "hulk_scene_triangle_count": (C.c_uint32, [C.c_void_p]),
    "hulk_scene_light_count": (C.c_uint32, [C.c_void_p]),
    "hulk_scene_vertices": (C.c_void_p, [C.c_void_p]),
    "hulk_scene_triangles": (C.c_void_p, [C.c_void_p]),
    "hulk_scene_lights": (C.c_void_p, [C.c_void_p]),
    "hulk_scene_mirror_mask": (C.c_void_p, [C.c_void_p]),
    "hulk_scene_get_box": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hulk_scene_camera": (C.c_int, [C.c_void_p, C.c_uint32, C.c_float, C.c_void_p]),
    "hulk_scene_kind_camera": (C.c_int, [C.c_int, C.c_uint32, C.c_float, C.c_void_p]),
}
# MLP entry points are bound when present (hulk_mlp.hip).
OPTIONAL_EXPORTS: dict = {}

_lib = None


def lib() -> C.CDLL:
    """Load libhulk.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m hulktracer_amd.build` (hipcc, gfx950) first")
    # PyTorch ships its own HIP / HSA runtime (same sonames as /opt/rocm's).
    # Loaded after libhulk, it would map a second runtime next to the one
    # libhulk pulled in, and that second runtime finds no GPU.  Loading torch
    # first lets libhulk bind to torch's runtime: one runtime per process.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(str(LIB_PATH))
    for name, (res, args) in {**EXPORTS, **OPTIONAL_EXPORTS}.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def ptr(a) -> C.c_void_p | None:
    """Pointer of a numpy array (host) or an integer device address."""
    if a is None:
        return None
    if isinstance(a, int):
        return C.c_void_p(a)
    if hasattr(a, "data_ptr"):
        return C.c_void_p(a.data_ptr())
    assert a.flags["C_CONTIGUOUS"], "arrays passed to libhulk must be C-contiguous"
    return C.c_void_p(a.ctypes.data)
