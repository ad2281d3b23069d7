"""Pure-Python restatement of the reference SAH-NlogN KD-tree builder.

TEST INFRASTRUCTURE ONLY (the checker of the product's C++ builder,
hulktracer_amd/csrc/kdtree_sah.cpp).  Written independently from the same
reference lines, with numpy float32 scalars so every float operation rounds
as the reference's single-precision code does:
  KDTreeCPU::buildKDTree            KDTreeCPU.cpp:61-94
  buildKDTreeSAH_NlogN              KDTreeSAH.cpp:247-311
  buildNodeSAHNlogN                 KDTreeSAH.cpp:313-521
  checkPolygonToAAPlane             KDTreeCPU.cpp:450-477
  makeLeafNode                      KDTreeCPU.cpp:46-57
  getBoxSurface                     SharedFunctions.hpp:32-36
Use it on small scenes only (pure-Python loops).
"""
from __future__ import annotations

import numpy as np

F = np.float32
FLT_MAX = F(3.4028234663852886e38)
M32 = 0xFFFFFFFF
BEGIN, COPLANAR, END = 0, 1, 2
ABOVE, BELOW, STRADDLE, COP = 0, 1, 2, 3


def _minps(a, b):
    return a if a < b else b


def _maxps(a, b):
    return a if a > b else b


def build(vertices: np.ndarray, triangles: np.ndarray, box: np.ndarray):
    """Returns (nodes (N,4) uint32, leaf_polys (L,) uint32)."""
    V = np.asarray(vertices, np.float32)
    T = np.asarray(triangles, np.uint32)
    n_tris = len(T)
    nodes: list[list[int]] = [[0, 0, 0, 0]]
    leaf_polys: list[int] = []
    leaf_count = [0]
    rel = [0] * n_tris

    def coord(poly, axis, k):
        return V[T[poly][k]][axis]

    def make_leaf(node, polys):
        nodes[node][0] = 4
        nodes[node][3] = leaf_count[0]
        leaf_count[0] += 1
        nodes[node][2] = len(polys)
        nodes[node][1] = len(leaf_polys)
        leaf_polys.extend(polys)

    def classify(axis, v, poly):
        c = [coord(poly, axis, k) for k in range(3)]
        if c[0] == v and c[1] == v and c[2] == v:
            return COP
        if c[0] >= v and c[1] >= v and c[2] >= v:
            return ABOVE
        if c[0] <= v and c[1] <= v and c[2] <= v:
            return BELOW
        return STRADDLE

    def node_step(node, depth, bmin, bmax, surface, ev, polys, cand):
        if depth >= 20 or len(polys) < 10:
            make_leaf(node, polys)
            return
        best, best_up, best_lo, best_val, best_dim = FLT_MAX, F(0), F(0), F(0), 0
        dim = [F(bmax[a] - bmin[a]) for a in range(3)]
        inv = F(F(1) / surface)
        for j in range(3):
            below = cop = above = 0
            plus_prev = cop_prev = 0
            x, y = dim[(j + 1) % 3], dim[(j + 2) % 3]
            xpy, area = F(x + y), F(x * y)
            first = True
            c = cand[j]
            i = 0
            while i < len(c):
                val = c[i][0]
                cnt = [0, 0, 0]
                while i < len(c) and c[i][0] == val:
                    cnt[c[i][1]] += 1
                    i += 1
                below = (below + cop_prev + plus_prev) & M32
                cop = cop_prev
                above = (above - cnt[COPLANAR] - cnt[END]) & M32
                if first:
                    below = 0
                    cop = cnt[COPLANAR]
                    above = ((ev & M32) - cnt[COPLANAR]) & M32
                    first = False
                plus_prev, cop_prev = cnt[BEGIN], cnt[COPLANAR]
                up = F(area + F(F(bmax[j] - val) * xpy))
                lo = F(area + F(F(val - bmin[j]) * xpy))
                cost = F(F(F(area * F(cop)) + F(lo * F(below))) + F(up * F(above)))
                sah = F(F(1) + F(F(5) * F(cost * inv)))
                if sah < best:
                    best, best_val, best_dim, best_up, best_lo = sah, val, j, up, lo
        if best >= F(F(5) * F(ev)):
            make_leaf(node, polys)
            return
        nodes[node][0] = best_dim | 8
        nodes[node][3] = int(np.array([best_val], np.float32).view(np.uint32)[0])
        a_polys, b_polys = [], []
        for p in polys:
            r = classify(best_dim, best_val, p)
            rel[p] = r
            if r == ABOVE:
                a_polys.append(p)
            elif r == BELOW:
                b_polys.append(p)
            else:
                a_polys.append(p)
                b_polys.append(p)
        a_cand, b_cand = [[], [], []], [[], [], []]
        a_ev = b_ev = 0
        for j in range(3):
            for (val, mk, p) in cand[j]:
                r = rel[p]
                if r == ABOVE:
                    a_cand[j].append((val, mk, p))
                elif r == BELOW:
                    b_cand[j].append((val, mk, p))
                elif r == STRADDLE:
                    if j == best_dim:
                        a_cand[j].append((best_val if best_val > val else val, mk, p))
                        b_cand[j].append((best_val if best_val < val else val, mk, p))
                    else:
                        a_cand[j].append((val, mk, p))
                        b_cand[j].append((val, mk, p))
            a_ev += len(a_cand[j])
            b_ev += len(b_cand[j])
        a_ev //= 6
        b_ev //= 6
        if a_polys:
            amin = list(bmin)
            amin[best_dim] = best_val
            idx = len(nodes)
            nodes.append([0, 0, 0, 0])
            nodes[node][1] = idx
            node_step(idx, depth + 1, amin, bmax, best_up, a_ev, a_polys, a_cand)
        if b_polys:
            bmx = list(bmax)
            bmx[best_dim] = best_val
            idx = len(nodes)
            nodes.append([0, 0, 0, 0])
            nodes[node][2] = idx
            node_step(idx, depth + 1, bmin, bmx, best_lo, b_ev, b_polys, b_cand)

    bmin = [F(box[0]), F(box[1]), F(box[2])]
    bmax = [F(box[4]), F(box[5]), F(box[6])]
    surface = box_surface(box)
    cand = [[], [], []]
    ev = 0
    for i in range(0, n_tris, 5):
        ev += 1
        for j in range(3):
            c = [coord(i, j, k) for k in range(3)]
            mn = _minps(_minps(c[0], c[1]), c[2])
            mx = _maxps(_maxps(c[0], c[1]), c[2])
            cand[j].append((mn, BEGIN, i))
            cand[j].append((mx, END, i))
    for j in range(3):
        cand[j].sort(key=lambda e: e[0])
    node_step(0, 0, bmin, bmax, surface, ev, list(range(n_tris)), cand)
    return np.array(nodes, np.uint32).reshape(-1, 4), np.array(leaf_polys, np.uint32)


def box_surface(box) -> np.float32:
    """getBoxSurface (SharedFunctions.hpp:32-36) on a (8,) min.xyzw/max.xyzw box."""
    sx, sy, sz = (F(F(box[4 + a]) - F(box[a])) for a in range(3))
    return F(F(2) * F(F(F(sx * sy) + F(sz * sy)) + F(sx * sz)))


def box_volume(box) -> np.float32:
    """getBoxVolume (SharedFunctions.hpp:24-28)."""
    sx, sy, sz = (F(F(box[4 + a]) - F(box[a])) for a in range(3))
    return F(F(sx * sy) * sz)
