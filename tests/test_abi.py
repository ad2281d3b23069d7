"""The C-ABI library loads and exports every symbol include/*.h declares;
host-only entry points (scene synthesis, KD build, cameras) behave; no
GPU calls here."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import hulktracer_amd as H
from hulktracer_amd import _lib

INCLUDE = Path(__file__).resolve().parent.parent / "include"


def declared_functions() -> list[str]:
    names = []
    for h in sorted(INCLUDE.glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names += re.findall(r"\b(hulk_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_headers_declare_entry_points():
    names = declared_functions()
    for must in ("hulk_create", "hulk_upload_scene", "hulk_upload_kdtree", "hulk_trace_camera", "hulk_shade",
                 "hulk_render_frame", "hulk_kdtree_build"):
        assert must in names


def test_library_exports_every_declared_symbol():
    try:  # one HIP runtime per process: torch's, as _lib.lib() arranges
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_python_binding_covers_declarations():
    bound = set(_lib.EXPORTS) | set(_lib.OPTIONAL_EXPORTS)
    missing = [n for n in declared_functions() if n not in bound]
    assert not missing, f"declared but not bound in _lib.py: {missing}"


def test_struct_sizes_match_reference_layout():
    # SharedStructs.hpp: KDTreeNode 16 B, TriangleIndices 16 B, Box 32 B,
    # CameraPos 64 B, Light 32 B, Ray 32 B
    assert H.HIT_DTYPE.itemsize == 32
    assert H.RAY_DTYPE.itemsize == 32
    assert C.sizeof(_lib.Opts) == 48  # 10 x 4 B + stream pointer
    assert C.sizeof(_lib.TuneRequest) == 96  # camera 64 B + 6 x 4 B + mirror pointer
    assert C.sizeof(_lib.TuneEntry) == 36


def test_procedural_scenes_and_counts():
    s1 = H.Scene.procedural(1, seed=1)
    assert s1.triangles.shape == (1000, 4) and s1.vertices.shape == (3000, 4)
    assert np.all(s1.vertices[:, 3] == 1.0)
    edges = [np.linalg.norm(s1.vertices[s1.triangles[:, a], :3] - s1.vertices[s1.triangles[:, b], :3], axis=1)
             for a, b in ((0, 1), (1, 2), (0, 2))]
    assert max(e.max() for e in edges) <= 0.1 + 1e-6
    s2 = H.Scene.procedural(2, seed=1)
    assert s2.n_tris == 130050 + 200 * 168
    s3 = H.Scene.procedural(3, seed=7, param=2, frame=3)
    assert s3.n_tris == 2 * 20000 + 2
    assert s3.mirror.shape == (s3.n_tris,)


def test_scene_generation_is_deterministic():
    a = H.Scene.procedural(2, seed=3)
    b = H.Scene.procedural(2, seed=3)
    assert np.array_equal(a.vertices, b.vertices) and np.array_equal(a.triangles, b.triangles)


def test_scene_box_is_vertex_min_max():
    s = H.Scene.procedural(1, seed=4, param=100)
    assert np.array_equal(s.box[:3], s.vertices[:, :3].min(0))
    assert np.array_equal(s.box[4:7], s.vertices[:, :3].max(0))


def test_camera_identity_rotation():
    cam = H.make_camera((0.5, 0.5, 2.2), fov=45.0, aspect=1.0).reshape(4, 4)
    vf = np.float32(np.tan(np.float32(45.0 * np.float32(0.017453292519943)) / 2))
    assert np.allclose(cam[1, :3], [2 * vf, 0, 0], atol=1e-7)     # verticalFactor = left*hf*-2
    assert np.allclose(cam[2, :3], [0, 2 * vf, 0], atol=1e-7)     # horizontalFactor = down*vf*-2
    assert np.allclose(cam[3, :3], [0.5 - vf, 0.5 - vf, 1.2], atol=1e-6)


def test_kdtree_rejects_bad_input():
    v = np.zeros((3, 4), np.float32)
    t = np.array([[0, 1, 5, 0]], np.uint32)  # vertex index out of range
    h = C.c_void_p()
    rc = H.lib().hulk_kdtree_build(_lib.ptr(v), 3, _lib.ptr(t), 1, _lib.ptr(np.zeros(8, np.float32)), C.byref(h))
    assert rc == 4


def test_settings_mirror_reference_defaults():
    s = H.Settings(["--kernelType=GPUTrace"])
    assert s["--kernelType"] == "GPUTrace" and s["--treeType"] == "SAHNlogN"
    with pytest.raises(ValueError):
        H.Settings(["--bogus=1"])


def test_package_modules_import():
    """Every module of the package and the bench import without a GPU."""
    import importlib

    for name in ("hulktracer_amd.engine", "hulktracer_amd.learned", "hulktracer_amd.workloads",
                 "hulktracer_amd.build", "oracle", "oracle.mlp_ref", "oracle.kdtree_ref"):
        importlib.import_module(name)
