"""Pins the CPU oracle (CPU only, no GPU).

How each oracle function is pinned (parity status):
  orc_tri_test (checkRayToPoly, SharedFunctions.hpp:47-102)
      known answers at the centroid / vertices / edges (closed forms that a
      wrong sign, index or dropped term breaks), the parallel and behind
      cases, t == 0 accepted, and the FP64 twin on random rays.
  orc_ray_box (checkRayToBoundingBoxGPU, SharedFunctions.hpp:245-264)
      slab closed form incl. d == 0 branches, origin inside / outside.
  traversal variants (GPUTrace*.cl, KDTreeCPU.cpp:285-447)
      nearest hit is a property of scene + ray: every variant agrees with
      brute force except at classified ties, on every config-1 ray and on
      random rays with degenerate directions.
  kdtree_ref.build vs the product builder (KDTreeSAH.cpp:247-521)
      identical node arrays on several scenes + structural invariants; the
      only golden vectors of the reference tests that touch the path
      (getBoxSurface / getBoxVolume, test/sharedFunctionTest.cpp:7-56).
  orc_tri_test_mt (Moller-Trumbore, JGT 1997; optional mode, no reference
      counterpart) known answers on dyadic inputs (t, u, v exact, both
      determinant signs, edges inclusive, the parallel cut and behind cases), hit /
      miss agreement with checkRayToPoly except within 1e-5 of an edge, the
      FP64 twin, and brute-force agreement of every traversal in MT mode.
  orc_cast_ray / orc_cast_rays (castRay, GPU_common.cl:10-30)
      closed form on an asymmetric camera and a non-square image with
      dyadic values (every term exact): vF scales x / W, hF scales y / H,
      index y * W + x; the x <-> y, W <-> H and vF <-> hF swaps are shown
      to differ.
  orc_shade mode 0 / orc_pack_rgb (GPUTrace.cl:122-160, FrameBuffer.hpp:196-217)
      closed-form Lambert + shadow halving on a one-triangle scene;
      round-half-even packing.
  orc_shade mode 1 (ShadePicture, shadePicture.cl:20-54)
      closed form on a quad with a blocker: base 0.1 everywhere, point
      light c * max(n.L, 0) / 2 with no shadow ray (blocked points are not
      halved), directional c * max(n.dir, 0), hit iff t != FLT_MAX and t > 0
      (t = 0, t < 0 and FLT_MAX shade 0.1 grey whatever the triangle id).
  orc_render (mirror generations; no reference counterpart)
      a mirror facing the camera reflects rays back: second-generation hit
      on the triangle behind the camera, weights 0.2 / 0.8.
  orc_cpu_trace_scene (reference CPU path port)
      agrees with the full-stack oracle on hit pixels up to ties (its rays
      are rsqrt-normalized, so t is in other units).
The reference holds no golden hit / image files for this path, so the hit
records at scale are pinned only by these properties: parity unpinned
beyond them.
"""
from __future__ import annotations

import numpy as np
import pytest

import hulktracer_amd as H
import oracle as O
from oracle import kdtree_ref
from tests.helpers import oracle_scene, random_rays, scene_and_tree

F32 = np.float32


def ray(o, d):
    r = np.zeros(8, F32)
    r[0:3] = o
    r[4:7] = d
    return r


# ------------------------------------------------------- triangle test --

TRI = [np.array(p, F32) for p in ((0, 0, 0), (1, 0, 0), (0, 1, 0))]


def test_tri_vertices_and_edges_inclusive():
    # vertex p1: s1 = s2 = 0
    h = O.tri_test(ray((0, 0, 1), (0, 0, -1)), *TRI)
    assert h is not None and h[0] == 1.0 and h[1] == 0.0 and h[2] == 0.0 and h[3] == 1.0
    # vertex p2: s1 = det, s2 = 0  (s1 + s2 == det is inside)
    h = O.tri_test(ray((1, 0, 1), (0, 0, -1)), *TRI)
    assert h is not None and h[1] == h[3] and h[2] == 0.0
    # vertex p3: s2 = det
    h = O.tri_test(ray((0, 1, 1), (0, 0, -1)), *TRI)
    assert h is not None and h[2] == h[3] and h[1] == 0.0
    # hypotenuse midpoint: s1 = s2 = det / 2
    h = O.tri_test(ray((0.5, 0.5, 1), (0, 0, -1)), *TRI)
    assert h is not None and h[1] == h[2] == h[3] / 2
    # just outside each edge
    for o in ((-1e-6, 0.3, 1), (0.3, -1e-6, 1), (0.5 + 1e-6, 0.5 + 1e-6, 1)):
        assert O.tri_test(ray(o, (0, 0, -1)), *TRI) is None


def test_tri_centroid_barycentrics_and_normal():
    p = [np.array(v, F32) for v in ((0.0, 0.0, 0.0), (3.0, 0.0, 0.0), (0.0, 3.0, 0.0))]
    h = O.tri_test(ray((1.0, 1.0, 2.0), (0, 0, -1)), *p)
    t, s1, s2, det, n = h
    assert t == 2.0
    assert s1 == s2 == det / 3
    # normal: (cross(e0,e1) + cross(e1,e2) + cross(e2,e0)) normalized = +z
    assert np.array_equal(n, np.array([0, 0, 1], F32))


def test_tri_parallel_behind_and_zero_t():
    assert O.tri_test(ray((0.2, 0.2, 1), (1, 0, 0)), *TRI) is None           # parallel
    assert O.tri_test(ray((0.2, 0.2, 1), (1, 0, 1e-8)), *TRI) is None        # |n.d| < FLT_EPSILON
    assert O.tri_test(ray((0.2, 0.2, -1), (0, 0, -1)), *TRI) is None         # behind the origin
    h = O.tri_test(ray((0.2, 0.2, 0), (0, 0, 1)), *TRI)                      # origin on the triangle
    assert h is not None and h[0] == 0.0
    # unnormalized direction: t is in units of |d| (GPU_common.cl:27)
    h = O.tri_test(ray((0.2, 0.2, 4), (0, 0, -2)), *TRI)
    assert h[0] == 2.0


def test_tri_against_fp64():
    """FP32 error of the reference arithmetic against the FP64 twin: t
    relative, u / v absolute (barycentrics live in [0, 1])."""
    rng = np.random.default_rng(7)
    et, eu = [], []
    for _ in range(3000):
        p = [rng.random(3).astype(F32) for _ in range(3)]
        c = (p[0] + p[1] + p[2]) / 3
        o = (rng.random(3) * 2 - 0.5).astype(F32)
        tgt = c + (rng.random(3) - 0.5).astype(F32) * 0.2
        r = ray(o, (tgt - o).astype(F32))
        h = O.tri_test(r, *p)
        d = O.tri_test_f64(r, *p)
        if h is None or d is None:
            continue
        et.append(abs(h[0] - d[0]) / abs(d[0]))
        eu.append(max(abs(h[1] / h[3] - d[1]), abs(h[2] / h[3] - d[2])))
    et, eu = np.array(et), np.array(eu)
    assert len(et) > 500
    assert np.quantile(et, 0.999) < 1e-5 and et.max() < 1e-4, (np.quantile(et, 0.999), et.max())
    assert np.quantile(eu, 0.99) < 1e-5 and eu.max() < 1e-3, (np.quantile(eu, 0.99), eu.max())


def test_mt_known_answers():
    """Moller-Trumbore (optional mode): on the unit right triangle with the
    ray along -z, u = x and v = y of the crossing point and t the height,
    exactly for dyadic inputs; edges and vertices are inside (u, v >= 0,
    u + v <= 1); just outside, parallel (|n.d| < FLT_EPSILON, the reference's
    cut), behind: miss;
    t == 0 is a hit; the normal is the reference's."""
    for (x, y, z) in ((0.25, 0.5, 1.0), (0.0, 0.0, 2.0), (1.0, 0.0, 0.5), (0.0, 1.0, 4.0), (0.5, 0.5, 1.0),
                      (0.125, 0.75, 3.0)):
        h = O.tri_test_mt(ray((x, y, z), (0, 0, -1)), *TRI)
        assert h is not None, (x, y)
        t, u, v, det, n = h
        assert (t, u, v, det) == (z, x, y, 1.0)
        assert np.array_equal(n, np.array([0, 0, 1], F32))
    for o in ((-1e-6, 0.3, 1), (0.3, -1e-6, 1), (0.5 + 1e-6, 0.5 + 1e-6, 1)):
        assert O.tri_test_mt(ray(o, (0, 0, -1)), *TRI) is None
    assert O.tri_test_mt(ray((0.2, 0.2, 1), (1, 0, 0)), *TRI) is None          # det = 0
    assert O.tri_test_mt(ray((0.2, 0.2, 1), (1, 0, -1e-7)), *TRI) is None      # |n.d| < FLT_EPSILON
    # the cut is the reference's (scale-free): a tiny triangle seen at a
    # shallow angle (|n.d| = 1e-3) is a hit in both tests
    tiny = [np.array(p, F32) * F32(1.0 / 1024) for p in ((0, 0, 0), (1, 0, 0), (0, 1, 0))]
    r = ray((0.25 / 1024 - 1.0, 0.25 / 1024, 1e-3), (1, 0, -1e-3))
    a, b = O.tri_test(r, *tiny), O.tri_test_mt(r, *tiny)
    assert a is not None and b is not None and abs(a[0] - b[0]) <= 1e-6 * a[0]
    assert O.tri_test_mt(ray((0.2, 0.2, -1), (0, 0, -1)), *TRI) is None       # behind
    h = O.tri_test_mt(ray((0.2, 0.2, 0), (0, 0, 1)), *TRI)
    assert h is not None and h[0] == 0.0
    # unnormalized direction: t in units of |d|; a wrong sign of qvec or of
    # the determinant would flip t or u here
    h = O.tri_test_mt(ray((0.25, 0.5, 4), (0, 0, -2)), *TRI)
    assert (h[0], h[1], h[2]) == (2.0, 0.25, 0.5)
    h = O.tri_test_mt(ray((0.25, 0.5, -4), (0, 0, 2)), *TRI)                   # from below: det < 0
    assert (h[0], h[1], h[2]) == (2.0, 0.25, 0.5)


def test_mt_agrees_with_reference_and_fp64():
    """On random rays through random triangles, Moller-Trumbore and the
    reference test agree on hit / miss except within 1e-5 of an edge (FP64
    barycentrics), and both are within the FP32 error of the FP64 twin."""
    rng = np.random.default_rng(11)
    disagree, et, eu = 0, [], []
    for _ in range(4000):
        p = [rng.random(3).astype(F32) for _ in range(3)]
        c = (p[0] + p[1] + p[2]) / 3
        o = (rng.random(3) * 2 - 0.5).astype(F32)
        tgt = c + (rng.random(3) - 0.5).astype(F32) * 0.6
        r = ray(o, (tgt - o).astype(F32))
        a, b, d = O.tri_test(r, *p), O.tri_test_mt(r, *p), O.tri_test_f64(r, *p)
        if (a is None) != (b is None):
            disagree += 1
            assert d is not None
            u, v = d[1], d[2]
            assert min(u, v, 1 - u - v) > -1e-5 and min(u, v, 1 - u - v) < 1e-5, (u, v)
            continue
        if b is None:
            continue
        et.append(abs(b[0] - d[0]) / abs(d[0]))
        eu.append(max(abs(b[1] - d[1]), abs(b[2] - d[2])))
    assert len(et) > 500 and disagree < 20
    assert np.percentile(et, 99.9) < 1e-5 and np.percentile(eu, 99) < 1e-5


@pytest.mark.parametrize("variant", [O.V_FULLSTACK, O.V_SHORTSTACK, O.V_RESTART, O.V_ROPES])
def test_mt_variants_match_brute_force(variant):
    """Traversals with the Moller-Trumbore test equal brute force with it on
    every config-1 ray (nearest hit is independent of the tree), and agree
    with the reference test's triangle ids except at grazes of shared edges."""
    s, nodes, lp = scene_and_tree(1, 1)
    osc = oracle_scene(1, 1)
    cam = s.camera(0, 1.0)
    bf = osc.trace_camera(O.V_BRUTE | O.MT, cam, 128, 128)
    got = osc.trace_camera(variant | O.MT, cam, 128, 128, depth=6)
    assert np.array_equal(got["tri"], bf["tri"])
    assert np.array_equal(got["t"].view(np.uint32), bf["t"].view(np.uint32))
    ref = osc.trace_camera(variant, cam, 128, 128, depth=6)
    hit = ref["tri"] != O.NO_HIT
    assert (bf["tri"] != O.NO_HIT).sum() > 1000
    assert np.mean(got["tri"] == ref["tri"]) > 0.999
    both = hit & (got["tri"] == ref["tri"])
    rel = np.abs(got["t"][both] - ref["t"][both]) / ref["t"][both]
    # the two tests round differently; t is ill-conditioned at grazing angles
    assert np.quantile(rel, 0.999) < 1e-5 and rel.max() < 1e-3, (np.quantile(rel, 0.999), rel.max())


# ------------------------------------------------------------ box test --

BOX = np.array([0, 0, 0, 1, 1, 2, 3, 1], F32)


def test_ray_box_closed_form():
    assert O.ray_box(BOX, ray((-1, 1, 1), (1, 0, 0))) == (1.0, 2.0)
    assert O.ray_box(BOX, ray((0.5, 1, 1), (1, 0, 0))) == (0.0, 0.5)           # inside: t0 = 0
    assert O.ray_box(BOX, ray((0.5, 3, 1), (1, 0, 0))) is None                 # d.y == 0, above
    assert O.ray_box(BOX, ray((0.5, 1, 1), (0, 0, 1)))[1] == 2.0               # d.x == d.y == 0
    assert O.ray_box(BOX, ray((2, 1, 1), (1, 0, 0))) is None                   # box behind
    t0, t1 = O.ray_box(BOX, ray((-1, -1, -1), (1, 1, 1)))
    assert t0 == 1.0 and t1 == 2.0


# ------------------------------------------------------ traversal / brute --

def _classify(osc, rays, i, ta, tb):
    """Tie when the two triangles give t within 2 ulp or share a vertex."""
    a, b = osc.t[ta], osc.t[tb]
    if len(set(a[:3]) & set(b[:3])):
        return "shared vertex"
    ha = O.tri_test(rays[i], *[osc.v[k] for k in a[:3]])
    hb = O.tri_test(rays[i], *[osc.v[k] for k in b[:3]])
    if ha is not None and hb is not None and abs(int(np.float32(ha[0]).view(np.int32)) -
                                                 int(np.float32(hb[0]).view(np.int32))) <= 2:
        return "equal t"
    return "other"


@pytest.mark.parametrize("variant", [O.V_FULLSTACK, O.V_SHORTSTACK, O.V_RESTART, O.V_ROPES])
def test_variants_match_brute_force_config1(variant):
    s, _, _ = scene_and_tree(1, 1)
    osc = oracle_scene(1, 1)
    cam = s.camera(0, 1.0)
    brute = osc.trace_camera(O.V_BRUTE, cam, 256, 256)
    got = osc.trace_camera(variant, cam, 256, 256)
    assert np.array_equal(got["tri"], brute["tri"])
    assert np.array_equal(got["t"], brute["t"])


@pytest.mark.parametrize("variant", [O.V_FULLSTACK, O.V_SHORTSTACK, O.V_RESTART, O.V_ROPES])
def test_variants_match_brute_force_random_rays(variant):
    s, _, _ = scene_and_tree(1, 5, 400)
    osc = oracle_scene(1, 5, 400)
    rays = random_rays(6000, s.box, seed=11)
    brute = osc.trace_rays(O.V_BRUTE, rays)
    got = osc.trace_rays(variant, rays, depth=3)
    bad = np.nonzero(got["tri"] != brute["tri"])[0]
    kinds = [_classify(osc, rays, i, got["tri"][i], brute["tri"][i]) for i in bad
             if got["tri"][i] != O.NO_HIT and brute["tri"][i] != O.NO_HIT]
    assert len(bad) == len(kinds) and all(k != "other" for k in kinds), (len(bad), kinds[:5])


def test_counts_are_monotone_in_variant_cost():
    s, _, _ = scene_and_tree(2, 1)
    osc = oracle_scene(2, 1)
    cam = s.camera(0, 1.0)
    _, full = osc.trace_camera(O.V_FULLSTACK, cam, 64, 64, counts=True)
    _, rst = osc.trace_camera(O.V_RESTART, cam, 64, 64, counts=True)
    assert rst[:, 0].sum() > full[:, 0].sum()  # kd-restart re-descends and never exits early
    assert rst[:, 2].sum() >= full[:, 2].sum()


def test_bounded_oracle_equals_unbounded():
    s, _, _ = scene_and_tree(2, 1)
    osc = oracle_scene(2, 1)
    cam = s.camera(4, 1.0)
    for v in (O.V_FULLSTACK, O.V_SHORTSTACK, O.V_RESTART, O.V_ROPES):
        plain = osc.trace_camera(v, cam, 96, 96)
        rng = np.random.default_rng(v)
        t = np.where(plain["tri"] != O.NO_HIT, plain["t"], 1.0)
        b = (t * rng.choice([0.3, 1.0, 1.001, 2.0], size=t.shape)).astype(F32)
        got = osc.trace_camera(v, cam, 96, 96, tbound=b)
        assert np.array_equal(got["tri"], plain["tri"]) and np.array_equal(got["t"], plain["t"])


# --------------------------------------------------------- KD builder --

def test_box_surface_golden():
    """test/sharedFunctionTest.cpp: box (0,0,0)-(1,2,3) has surface 22, volume 6."""
    assert kdtree_ref.box_surface(np.array([0, 0, 0, 1, 1, 2, 3, 1], F32)) == 22.0
    assert kdtree_ref.box_volume(np.array([0, 0, 0, 1, 1, 2, 3, 1], F32)) == 6.0


def _grid_scene(n=12, z=0.0):
    """Axis-aligned grid: many coplanar triangles and equal split candidates."""
    vs, ts = [], []
    for i in range(n + 1):
        for j in range(n + 1):
            vs.append((i / n, j / n, z + 0.1 * ((i * j) % 3), 1.0))
    for i in range(n):
        for j in range(n):
            a = i * (n + 1) + j
            ts.append((a, a + 1, a + n + 2, 0))
            ts.append((a, a + n + 2, a + n + 1, 0))
    return np.array(vs, F32), np.array(ts, np.uint32)


@pytest.mark.parametrize("case", ["random1000", "random60", "grid", "flat", "dup"])
def test_builder_matches_python_restatement(case):
    if case == "random1000":
        s = H.Scene.procedural(1, seed=1)
        v, t = s.vertices, s.triangles
    elif case == "random60":
        s = H.Scene.procedural(1, seed=9, param=60)
        v, t = s.vertices, s.triangles
    elif case == "grid":
        v, t = _grid_scene()
    elif case == "flat":
        v, t = _grid_scene(10)
        v[:, 2] = 0.0
    else:
        s = H.Scene.procedural(1, seed=2, param=200)
        v, t = s.vertices, np.concatenate([s.triangles, s.triangles[:50]])
    box = H.scene_box(v)
    nodes, lp = kdtree_ref.build(v, t, box)
    # single-threaded, and with subtrees built on up to 16 threads
    for threads in (1, 16):
        kd = H.KDTreeCPU()
        kd.buildKDTree(H.Scene(v, t, np.zeros((0, 8), F32), box), threads=threads)
        assert np.array_equal(kd.getNodeArray(), nodes), threads
        assert np.array_equal(kd.getPolygonArray(), lp), threads


@pytest.mark.parametrize("kind,seed", [(2, 1), (3, 7)])
def test_parallel_builder_equals_single_threaded(kind, seed):
    """Large scenes: concurrent subtree builds give the same arrays as the
    single-threaded build (node indices, leaf starts and leaf ids shifted
    on concatenation)."""
    s = H.Scene.procedural(kind, seed=seed)
    ref = H.KDTreeCPU()
    ref.buildKDTree(s, threads=1)
    for threads in (0, 3, 8):
        kd = H.KDTreeCPU()
        kd.buildKDTree(s, threads=threads)
        assert np.array_equal(kd.getNodeArray(), ref.getNodeArray()), threads
        assert np.array_equal(kd.getPolygonArray(), ref.getPolygonArray()), threads


def _walk(nodes):
    """(node, depth, cell lo, cell hi) for every node reachable from the root."""
    out = []
    stack = [(0, 0)]
    while stack:
        i, d = stack.pop()
        out.append((i, d))
        f, a, b, _ = nodes[i]
        if f & 8:
            for c in (a, b):
                if c:
                    stack.append((c, d + 1))
    return out


def test_builder_invariants_terrain():
    s, nodes, lp = scene_and_tree(2, 1)
    seen = _walk(nodes)
    idx = [i for i, _ in seen]
    assert sorted(idx) == list(range(len(nodes))), "every node reached exactly once"
    assert max(d for _, d in seen) <= 20
    flags = nodes[:, 0]
    assert np.all((flags == 4) | ((flags & 8) != 0) & ((flags & 3) < 3))
    leaves = nodes[flags == 4]
    assert np.all(leaves[:, 1].astype(np.int64) + leaves[:, 2] <= len(lp))
    assert np.array_equal(np.sort(leaves[:, 3]), np.arange(len(leaves)))   # leaf ids sequential
    # leaf ranges tile m_leafPolygons in DFS order
    order = [i for i, _ in sorted(seen) if flags[i] == 4]
    starts = sorted(int(nodes[i][1]) for i in order)
    assert starts[0] == 0
    assert set(lp.tolist()) == set(range(s.n_tris))


def test_builder_conservative_classification():
    """A triangle whose vertices all lie strictly inside a leaf cell is
    stored in that leaf (KDTreeCPU.cpp:450-477 sends straddlers both ways)."""
    s, nodes, lp = scene_and_tree(1, 1)
    cells = {0: (s.box[:3].copy(), s.box[4:7].copy())}
    stack = [0]
    while stack:
        i = stack.pop()
        f, a, b, c = nodes[i]
        if f & 8:
            ax, v = f & 3, np.array([c], np.uint32).view(F32)[0]
            lo, hi = cells[i]
            if a:
                l2 = lo.copy(); l2[ax] = v; cells[a] = (l2, hi.copy()); stack.append(a)
            if b:
                h2 = hi.copy(); h2[ax] = v; cells[b] = (lo.copy(), h2); stack.append(b)
    V = s.vertices[:, :3]
    for i, (lo, hi) in cells.items():
        if nodes[i][0] != 4:
            continue
        start, cnt = nodes[i][1], nodes[i][2]
        members = set(lp[start:start + cnt].tolist())
        tv = V[s.triangles[:, :3]]
        inside = np.all((tv > lo) & (tv < hi), axis=(1, 2))
        assert set(np.nonzero(inside)[0].tolist()) <= members


# -------------------------------------------------------------- shading --

def test_pack_round_half_even():
    assert O.pack_rgb(1.0, 0.0, 2.0) == 255 | (0 << 8) | (255 << 16)
    assert O.pack_rgb(0.5, 0.0, 0.0) == 128           # 127.5 -> 128 (even)
    assert O.pack_rgb(np.float32(1.5 / 255), 0, 0) == 2  # 1.5 -> 2 (even)
    assert O.pack_rgb(np.float32(2.5 / 255), 0, 0) == 2  # 2.5 -> 2 (even)
    assert O.pack_rgb(-1.0, 0.0, 0.0) == 0


def _one_quad_scene(with_blocker: bool):
    v = [(-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0)]
    t = [(0, 1, 2, 0), (0, 2, 3, 0)]
    if with_blocker:
        v += [(-0.5, -0.5, 1.0), (0.5, -0.5, 1.0), (0.0, 0.5, 1.0)]
        t += [(4, 5, 6, 0)]
    v = np.array([(*p, 1.0) for p in v], F32)
    lights = np.array([[0.9, 0.9, 0.9, 1.0, 0.0, 0.0, 3.0, 1.0]], F32)
    box = H.scene_box(v)
    kd = H.KDTreeCPU()
    sc = H.Scene(v, np.array(t, np.uint32), lights, box)
    kd.buildKDTree(sc)
    return O.OracleScene(v, sc.triangles, box, kd.getNodeArray(), kd.getPolygonArray(), lights)


def test_shade_closed_form():
    cam = H.make_camera((0.0, 0.0, 5.0), fov=30.0)
    for blocker in (False, True):
        osc = _one_quad_scene(blocker)
        hits = osc.trace_camera(O.V_FULLSTACK, cam, 64, 64)
        img = osc.shade(O.V_FULLSTACK, cam, 64, 64, hits)
        i = 32 * 64 + 32  # centre pixel looks straight down -z at the light's foot
        h = hits[i]
        assert h["tri"] in (0, 1) or blocker
        rays = O.cast_rays(cam, 64, 64)
        P = rays[i, :3] + rays[i, 4:7] * h["t"]
        L = np.array([0, 0, 3], F32) - P
        cosang = float(np.dot([h["nx"], h["ny"], h["nz"]], L / np.linalg.norm(L)))
        if blocker and h["tri"] == 2:
            continue
        want = 0.1 + 0.9 * max(cosang, 0) * (0.5 if blocker else 1.0)
        got = img[i] & 0xFF
        assert abs(got - round(min(want, 1) * 255)) <= 1


def test_shade_picture_closed_form():
    """ShadePicture (shadePicture.cl:20-54): every pixel starts at 0.1 grey;
    a pixel counts as hit iff t != FLT_MAX and t > 0; a point light adds
    c * max(n.L^, 0) / 2 with no shadow ray; a directional light adds
    c * max(n.dir, 0).  A dropped /2, a shadow ray, a missing base or a
    miss shaded black breaks it."""
    osc = _one_quad_scene(True)  # the blocker sits between the quad and the point light
    pc = np.array([0.9, 0.6, 0.3], F32)
    dc = np.array([0.25, 0.5, 0.125], F32)
    d_dir = np.array([0.0, 0.6, 0.8], F32)  # n.dir = 0.8 on the quad (n = +z)
    osc.lights = np.array([[*pc, 1.0, 0.0, 0.0, 3.0, 1.0], [*dc, 0.0, *d_dir, 0.0]], F32)
    cam = H.make_camera((0.0, 0.0, 5.0), fov=30.0)
    W, Hh = 48, 40
    hits = osc.trace_camera(O.V_FULLSTACK, cam, W, Hh)
    img = osc.shade(O.V_FULLSTACK, cam, W, Hh, hits, mode=1)
    rays = O.cast_rays(cam, W, Hh).astype(np.float64)
    rgb = img.view(np.uint8).reshape(-1, 4)[:, :3].astype(int)
    grey = O.pack_rgb(0.1, 0.1, 0.1)
    checked_quad = shadowed = 0
    for i in range(W * Hh):
        h = hits[i]
        if h["tri"] == O.NO_HIT:
            assert img[i] == grey, f"miss pixel {i} is not 0.1 grey"
            continue
        n = np.array([h["nx"], h["ny"], h["nz"]], np.float64)
        P = rays[i, :3] + rays[i, 4:7] * float(h["t"])
        L = np.array([0.0, 0.0, 3.0]) - P
        want = 0.1 + pc * max(float(n @ (L / np.linalg.norm(L))), 0.0) / 2 + dc * max(float(n @ d_dir), 0.0)
        assert np.all(np.abs(rgb[i] - np.round(np.minimum(want, 1.0) * 255)) <= 1), (i, rgb[i], want * 255)
        if h["tri"] in (0, 1):
            checked_quad += 1
            # a quad point whose segment to the light crosses the blocker (z = 1, where the
            # segment is at 2/3 P + 1/3 light): a shadow ray would halve the point term
            qx, qy = 2.0 / 3.0 * P[0], 2.0 / 3.0 * P[1]
            if qy > -0.5 and qy < 0.5 - 2.0 * abs(qx):
                shadowed += 1
    assert checked_quad > 100 and shadowed > 0
    # the hit rule looks at t only: t == 0, t < 0 and t == FLT_MAX give 0.1 grey, a hit id notwithstanding
    i = next(k for k in range(W * Hh) if hits[k]["tri"] in (0, 1))
    for t in (0.0, -1.0, float(O.FLT_MAX)):
        hh = hits.copy()
        hh["t"][i] = t
        assert osc.shade(O.V_FULLSTACK, cam, W, Hh, hh, mode=1)[i] == grey, t
    # and a tiny positive t shades the pixel (no lower cut above zero)
    hh = hits.copy()
    hh["t"][i] = np.float32(1e-30)
    assert osc.shade(O.V_FULLSTACK, cam, W, Hh, hh, mode=1)[i] != grey


def test_cast_rays_orientation_closed_form():
    """castRay (GPU_common.cl:10-30, called as castRay(row = y, column = x)
    from GPUTrace.cl:105-116): d = lowerLeft + vF * (x / W) + hF * (y / H) -
    from, ray index y * W + x.  An asymmetric camera on a non-square image
    with dyadic values makes every term exact, so an x <-> y, W <-> H or
    vF <-> hF swap fails."""
    frm = np.array([0.5, 0.25, 1.0, 0.0])
    vf = np.array([4.0, 0.5, -0.25, 0.0])
    hf = np.array([0.25, 2.0, 0.125, 0.0])
    ll = np.array([-2.0, -1.0, -4.0, 0.0])
    cam = np.concatenate([frm, vf, hf, ll]).astype(F32)
    W, Hh = 8, 4
    rays = O.cast_rays(cam, W, Hh).astype(np.float64)
    assert np.array_equal(rays[0, 4:7], (ll - frm)[:3]), "pixel (0, 0) must be lowerLeft - from"
    x, y = np.meshgrid(np.arange(W), np.arange(Hh))  # index y * W + x
    x, y = x.ravel()[:, None], y.ravel()[:, None]
    want = ll[:3] + vf[:3] * (x / W) + hf[:3] * (y / Hh) - frm[:3]
    assert np.array_equal(rays[:, 4:7], want)
    assert np.array_equal(rays[:, 0:3], np.broadcast_to(frm[:3], (W * Hh, 3)))
    # the swapped readings really differ on this camera
    for swapped in (ll[:3] + vf[:3] * (y / Hh) + hf[:3] * (x / W) - frm[:3],
                    ll[:3] + hf[:3] * (x / W) + vf[:3] * (y / Hh) - frm[:3],
                    ll[:3] + vf[:3] * (x / Hh) + hf[:3] * (y / W) - frm[:3]):
        assert not np.array_equal(rays[:, 4:7], swapped)
    # pixel (x = 3, y = 1): d = ll + vf * 3/8 + hf * 1/4 - from
    assert np.array_equal(rays[1 * W + 3, 4:7], [-2.0 + 1.5 + 0.0625 - 0.5, -1.0 + 0.1875 + 0.5 - 0.25,
                                               -4.0 - 0.09375 + 0.03125 - 1.0])


def test_mirror_generation_reflects_back():
    """Mirror quad facing the camera: the reflected ray returns along -d and
    the pixel keeps 0.2 of the mirror's shading plus 0.8 of what it sees."""
    osc = _one_quad_scene(False)
    osc.mirror = np.array([1, 1], np.uint8)
    cam = H.make_camera((0.0, 0.0, 5.0), fov=30.0)
    img1, c1 = osc.render(O.V_FULLSTACK, cam, 32, 32, max_bounce=1)
    img0, c0 = osc.render(O.V_FULLSTACK, cam, 32, 32, max_bounce=0)
    # the reflection escapes (nothing behind the camera): 0.2 of the local shading
    i = 16 * 32 + 16
    assert abs(int(img1[i] & 0xFF) - round(0.2 * (img0[i] & 0xFF))) <= 1
    assert c1[0] > c0[0]


def test_cpu_baseline_port_agrees():
    s, nodes, lp = scene_and_tree(1, 1)
    osc = oracle_scene(1, 1)
    cam = s.camera(0, 1.0)
    n, img = osc.cpu_trace_scene(cam, 128, 128)
    hits = osc.trace_camera(O.V_FULLSTACK, cam, 128, 128)
    ref = osc.shade(O.V_FULLSTACK, cam, 128, 128, hits)
    hit_px = (hits["tri"] != O.NO_HIT)
    assert n >= 128 * 128
    d = np.abs(img.view(np.uint8).astype(int) - ref.view(np.uint8).astype(int)).reshape(-1, 4)[:, :3]
    close = (d.max(1) <= 2)
    assert close[hit_px].mean() > 0.995
