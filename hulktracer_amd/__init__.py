"""hulktracer_amd - MI355X (gfx950) traverse-and-shade hot path of hulkTracer.

The compute path is libhulk.so (hand-written HIP kernels + C ABI,
include/hulk_abi.h); this package is its host-side binding and a mirror of
the reference's renderer interface.
"""
from ._lib import (BRUTEFORCE, FULLSTACK, HIT_DTYPE, LDSSTACK, NO_HIT, RAY_DTYPE, RESTART, ROPES, SHORTSTACK,
                   TRI_MOLLER_TRUMBORE, TRI_REFERENCE,
                   VARIANTS, HulkError, lib)
from .engine import (GLVisitorGPUTrace, GPUTrace, KDTreeCPU, KDTreeGPU, Scene, Settings, batch_scenes, make_camera,
                     make_camera_lookat, scene_box)

__all__ = ["BRUTEFORCE", "FULLSTACK", "SHORTSTACK", "LDSSTACK", "RESTART", "ROPES", "VARIANTS", "HIT_DTYPE",
           "TRI_REFERENCE", "TRI_MOLLER_TRUMBORE",
           "RAY_DTYPE", "NO_HIT", "HulkError", "lib", "GLVisitorGPUTrace", "GPUTrace", "KDTreeCPU", "KDTreeGPU", "Scene",
           "Settings", "batch_scenes", "make_camera", "make_camera_lookat", "scene_box"]
