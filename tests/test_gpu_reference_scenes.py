"""GPU parity on the reference's own models (OBJLoader/Models: teapot,
box; copied to tests/golden): every traversal variant, shading and the
device KD build, through the C ABI, against the CPU oracle."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import hulktracer_amd as H
import oracle as O
from hulktracer_amd import _lib as L
from tests.helpers import describe, hit_mismatch

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
VARIANTS = [H.BRUTEFORCE, H.FULLSTACK, H.SHORTSTACK, H.LDSSTACK, H.RESTART, H.ROPES]
ORACLE_OF = {H.BRUTEFORCE: O.V_BRUTE, H.FULLSTACK: O.V_FULLSTACK, H.SHORTSTACK: O.V_SHORTSTACK,
             H.LDSSTACK: O.V_LDSSTACK, H.RESTART: O.V_RESTART, H.ROPES: O.V_ROPES}
VIEWS = {
    "teapot.obj": ((0.9, 0.7, 1.1), (0.0, 0.15, 0.0)),
    "Box.obj": ((0.8, 0.9, 1.0), (0.0, 0.2, 0.0)),
}


def _setup(name, time=0.0):
    s = H.Scene.from_obj(GOLDEN / name, time)
    kd = H.KDTreeCPU()
    kd.buildKDTree(s)
    nodes, lp = kd.getNodeArray(), kd.getPolygonArray()
    frm, tgt = VIEWS[name]
    cam = H.make_camera_lookat(frm, tgt, fov=50.0, aspect=1.0)
    osc = O.OracleScene(s.vertices, s.triangles, s.box, nodes, lp, s.lights, s.mirror)
    return s, nodes, lp, cam, osc


@pytest.mark.parametrize("name", sorted(VIEWS))
@pytest.mark.parametrize("variant", VARIANTS)
def test_reference_model_trace_and_shade(name, variant):
    s, nodes, lp, cam, osc = _setup(name, 0.7)
    print(f"reference model {name} variant {variant}", flush=True)
    ctx = H.GPUTrace(variant=variant, lds_top_levels=6)
    ctx.upload_scene(s)
    ctx.upload_kdtree(nodes, lp)
    got = ctx.trace_camera(cam, 256, 256)
    ref = osc.trace_camera(ORACLE_OF[variant], cam, 256, 256, depth=6)
    bad = hit_mismatch(got, ref)
    assert len(bad) == 0, f"{len(bad)} rays differ\n{describe(got, ref, bad)}"
    assert (got["tri"] != H.NO_HIT).sum() > 3000
    for mode in (L.SHADE_FUSED, L.SHADE_PICTURE):
        img = ctx.shade(cam, 256, 256, got, mode)
        rimg = osc.shade(ORACLE_OF[variant], cam, 256, 256, ref, mode=mode)
        assert np.array_equal(img, rimg), f"shade mode {mode}: {(img != rimg).sum()} pixels differ"


@pytest.mark.parametrize("name", sorted(VIEWS))
@pytest.mark.parametrize("inside", [False, True])
def test_reference_model_refill_ropes(name, inside):
    """The lane-refill rope kernel on the reference's models, from the view
    above and from inside the model's box (the box model's tree splits on
    planes of the scene box, which the pushed ropes never pass): the first
    trace of the tree on the plain ropes, the second on the pushed ones;
    pixels (with shadow rays) and primary hits equal the oracle's."""
    s, nodes, lp, cam, osc = _setup(name, 0.7)
    if inside:
        b = s.box
        c = [(b[k] + b[4 + k]) * 0.5 for k in range(3)]
        cam = H.make_camera_lookat(c, (b[4], b[1] + 0.25 * (b[5] - b[1]), b[6]), fov=70.0, aspect=1.0)
    print(f"reference model {name} refill ropes inside={inside}", flush=True)
    ctx = H.GPUTrace(variant=H.ROPES, pipeline=3, coop_lanes=64, block_size=64)
    ctx.upload_scene(s)
    ctx.upload_kdtree(nodes, lp)
    rimg, _ = osc.render(O.V_ROPES, cam, 192, 192, max_bounce=0)
    rhits = osc.trace_camera(O.V_ROPES, cam, 192, 192)
    for k in range(2):
        img, hits = ctx.render_frame(cam, 192, 192)
        bad = hit_mismatch(hits, rhits)
        assert len(bad) == 0, f"trace {k}: {len(bad)} rays differ\n{describe(hits, rhits, bad)}"
        assert np.array_equal(img, rimg), f"trace {k}: {(img != rimg).sum()} pixels differ"
    assert (hits["tri"] != H.NO_HIT).sum() > 1000


@pytest.mark.parametrize("name", sorted(VIEWS))
def test_reference_model_device_build(name):
    s, nodes, lp, cam, osc = _setup(name, 0.7)
    ctx = H.GPUTrace(variant=H.FULLSTACK)
    ctx.upload_scene(s)
    ctx.build_kdtree_device()
    dn, dlp = ctx.download_kdtree()
    assert np.array_equal(dn, nodes) and np.array_equal(dlp, lp)
    img, hits = ctx.render_frame(cam, 200, 200)
    rimg, _ = osc.render(O.V_FULLSTACK, cam, 200, 200)
    assert np.array_equal(img, rimg)


def test_dataset_round_trip_on_the_gpu(tmp_path):
    """Terrain + spheres written as a paris.bin dataset, loaded back and
    traced on the GPU: the same hits as the original scene, triangle k of the
    dataset being the original leaf reference k."""
    from tests.helpers import scene_and_tree

    s, nodes, lp = scene_and_tree(2, 1)
    kd = H.KDTreeCPU()
    kd.m_treeNodes, kd.m_leafPolygons = nodes, lp
    kd.saveDataSetToFile(s, tmp_path / "terrain.bin")
    s2, nodes2, lp2 = H.Scene.from_dataset(tmp_path / "terrain.bin")
    cam = s.camera(13, 1.0)
    out = []
    for sc, nd, l in ((s, nodes, lp), (s2, nodes2, lp2)):
        ctx = H.GPUTrace(variant=H.LDSSTACK, lds_top_levels=8)
        ctx.upload_scene(sc)
        ctx.upload_kdtree(nd, l)
        out.append(ctx.trace_camera(cam, 256, 256))
    a, b = out
    for f in ("t", "u", "v", "nx", "ny", "nz"):
        assert np.array_equal(a[f].view(np.uint32), b[f].view(np.uint32)), f
    hit = b["tri"] != H.NO_HIT
    assert np.array_equal(lp[b["tri"][hit]], a["tri"][hit]) and hit.sum() > 10000


@pytest.mark.parametrize("time", [0.5, 3.0])
@pytest.mark.parametrize("variant", VARIANTS)
def test_bart_museum_trace_and_device_build(variant, time):
    """The BART museum scene (museum3, from the reference's BartLoader) through
    its own camera path: every variant against the oracle, shading, and the
    device-built tree against the host builder."""
    s = H.Scene.from_aff(GOLDEN / "museum" / "museum3.aff", time)
    cam = s.camera(0, 1.0)
    kd = H.KDTreeCPU()
    kd.buildKDTree(s)
    nodes, lp = kd.getNodeArray(), kd.getPolygonArray()
    osc = O.OracleScene(s.vertices, s.triangles, s.box, nodes, lp, s.lights, s.mirror)
    print(f"museum3 t={time} variant {variant}", flush=True)
    ctx = H.GPUTrace(variant=variant, lds_top_levels=6)
    ctx.upload_scene(s)
    ctx.build_kdtree_device()
    dn, dlp = ctx.download_kdtree()
    assert np.array_equal(dn, nodes) and np.array_equal(dlp, lp)
    got = ctx.trace_camera(cam, 192, 192)
    ref = osc.trace_camera(ORACLE_OF[variant], cam, 192, 192, depth=6)
    bad = hit_mismatch(got, ref)
    assert len(bad) == 0, f"{len(bad)} rays differ\n{describe(got, ref, bad)}"
    assert (got["tri"] != H.NO_HIT).sum() > 500
    img = ctx.shade(cam, 192, 192, got, L.SHADE_FUSED)
    assert np.array_equal(img, osc.shade(ORACLE_OF[variant], cam, 192, 192, ref, mode=L.SHADE_FUSED))
