// kdtree_sah.cpp - SAH-NlogN KD-tree builder (host, C++).
//
// The hot path consumes the reference's CPU-built tree unchanged.  This file
// is the C++ restatement of that builder so the library can be driven
// headless: KDTreeCPU::buildKDTree (HulkEngine/src/KDTreeCPU.cpp:61-94) with
// the default --treeType=SAHNlogN (HulkEngine/src/Settings.cpp:49):
// buildKDTreeSAH_NlogN (HulkEngine/src/KDTreeSAH.cpp:247-311),
// buildNodeSAHNlogN (KDTreeSAH.cpp:313-521), checkPolygonToAAPlane
// (KDTreeCPU.cpp:450-477) and makeLeafNode (KDTreeCPU.cpp:46-57).
//
// The output arrays (nodes in DFS pre-order, above child first; leaf
// polygon references) must equal the reference's bit for bit, so every
// float operation keeps the reference's order and width, unsigned counters
// keep their 32-bit wrap, and min / max keep the SSE operand rule
// (minps(a, b) = a < b ? a : b).  Built with -ffp-contract=off.
#include <algorithm>
#include <cfloat>
#include <cstring>
#include <new>
#include <system_error>
#include <thread>
#include <vector>

#include "hulk_abi.h"
#include "kdtree_data.hpp"

namespace {

constexpr uint32_t kMaxDepth = 20;          // c_maxKDTreeDepth, SharedStructs.hpp:16
constexpr uint32_t kReductionFactor = 5;    // KDTreeCPU.hpp:23
constexpr float kTraverseCost = 1.0f;       // KDTreeCPU.cpp:17
constexpr float kIntersectionCost = 5.0f;   // KDTreeCPU.cpp:18
constexpr uint32_t kMinLeafPolys = 10;      // KDTreeSAH.cpp:317

enum Relation : uint8_t { kAbove, kBelow, kStraddle, kCoplanar };  // EPolyToPlane
enum Marker : uint32_t { kBegin = 0, kCoplanarMark = 1, kEnd = 2 }; // EMarker

struct Candidate {
    float value;
    uint32_t marker;
    uint32_t poly;
};

inline float sse_min(float a, float b) { return a < b ? a : b; }
inline float sse_max(float a, float b) { return a > b ? a : b; }

struct Split {
    hulk_box box;
    float surface = 0.0f;        // full surface at the root, half surface below
    uint32_t depth = 0;
    uint32_t node = 0;
    size_t eval = 0;             // polygonsEval
    std::vector<uint32_t> polys;
    std::vector<Candidate> cand[3];
};

inline float lane(const hulk_float4& v, uint32_t a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
inline void set_lane(hulk_float4& v, uint32_t a, float x) {
    if (a == 0) v.x = x; else if (a == 1) v.y = x; else v.z = x;
}

// Output of one subtree build: nodes in depth-first pre-order (the subtree
// root at local index 0, empty child = 0), its leaf-polygon run and its leaf
// count.  Subtrees built in parallel are concatenated in the sequential
// order (above subtree, then below subtree) with their indices, leaf starts
// and leaf ids shifted, which reproduces the sequential arrays exactly.
struct Out {
    std::vector<hulk_kd_node> nodes;
    std::vector<uint32_t> leaf_polys;
    uint32_t leaf_count = 0;
};

class Builder {
public:
    Builder(const hulk_float4* v, const hulk_tri* t, uint32_t nt, uint32_t spawn_depth)
        : verts_(v), tris_(t), n_tris_(nt), spawn_depth_(spawn_depth) {}

    void build(const hulk_box& scene_box, Out& out) {
        Split s;
        s.polys.resize(n_tris_);
        for (uint32_t i = 0; i < n_tris_; i++) s.polys[i] = i;
        s.node = 0;
        s.depth = 0;
        s.box = scene_box;
        s.surface = box_surface(scene_box);
        s.eval = 0;
        for (uint32_t i = 0; i < n_tris_; i += kReductionFactor) {
            s.eval++;
            const uint32_t p = s.polys[i];
            const hulk_float4& p1 = verts_[tris_[p].idx[0]];
            const hulk_float4& p2 = verts_[tris_[p].idx[1]];
            const hulk_float4& p3 = verts_[tris_[p].idx[2]];
            for (uint32_t j = 0; j < 3; j++) {
                float mn = sse_min(sse_min(lane(p1, j), lane(p2, j)), lane(p3, j));
                float mx = sse_max(sse_max(lane(p1, j), lane(p2, j)), lane(p3, j));
                s.cand[j].push_back({mn, kBegin, p});
                s.cand[j].push_back({mx, kEnd, p});
            }
        }
        // the three candidate sorts are independent
        std::vector<std::thread> sorts;
        for (uint32_t j = 0; j < 3; j++)
            sorts.emplace_back([&s, j] {
                std::sort(s.cand[j].begin(), s.cand[j].end(),
                          [](const Candidate& a, const Candidate& b) { return a.value < b.value; });
            });
        for (auto& t : sorts) t.join();
        hulk_kd_node root;
        std::memset(&root, 0, sizeof(root));
        out.nodes.reserve(5000);
        out.nodes.push_back(root);  // KDTreeCPU.cpp:74-75
        std::vector<Relation> rel(n_tris_);
        build_node(s, out, rel);
    }

private:
    // getBoxSurface, SharedFunctions.hpp:32-36.
    static float box_surface(const hulk_box& b) {
        float sx = b.max.x - b.min.x, sy = b.max.y - b.min.y, sz = b.max.z - b.min.z;
        return 2.0f * (sx * sy + sz * sy + sx * sz);
    }

    Relation classify(uint32_t axis, float v, uint32_t poly) const {
        const hulk_tri& t = tris_[poly];
        float c1 = lane(verts_[t.idx[0]], axis);
        float c2 = lane(verts_[t.idx[1]], axis);
        float c3 = lane(verts_[t.idx[2]], axis);
        if ((c1 == v) & (c2 == v) & (c3 == v)) return kCoplanar;
        if ((c1 >= v) & (c2 >= v) & (c3 >= v)) return kAbove;
        if ((c1 <= v) & (c2 <= v) & (c3 <= v)) return kBelow;
        return kStraddle;
    }

    static void make_leaf(const Split& s, Out& out) {
        hulk_kd_node& n = out.nodes[s.node];
        n.flags = HULK_NODE_LEAF;
        n.u.leaf.leaf_id = out.leaf_count++;
        n.u.leaf.count = static_cast<uint32_t>(s.polys.size());
        n.u.leaf.start = static_cast<uint32_t>(out.leaf_polys.size());
        out.leaf_polys.insert(out.leaf_polys.end(), s.polys.begin(), s.polys.end());
    }

    // Append a subtree built on its own (root at local 0) after the nodes
    // already in `out`; returns the subtree root's index in `out`.
    static uint32_t append(Out& out, const Out& sub) {
        const uint32_t base = static_cast<uint32_t>(out.nodes.size());
        const uint32_t lbase = static_cast<uint32_t>(out.leaf_polys.size());
        for (hulk_kd_node n : sub.nodes) {
            if (n.flags & HULK_NODE_LEAF) {
                n.u.leaf.start += lbase;
                n.u.leaf.leaf_id += out.leaf_count;
            } else if (n.flags & HULK_NODE_JUNCTION) {
                if (n.u.junction.above) n.u.junction.above += base;
                if (n.u.junction.below) n.u.junction.below += base;
            }
            out.nodes.push_back(n);
        }
        out.leaf_polys.insert(out.leaf_polys.end(), sub.leaf_polys.begin(), sub.leaf_polys.end());
        out.leaf_count += sub.leaf_count;
        return base;
    }

    void build_node(Split& s, Out& out, std::vector<Relation>& rel) {
        if (s.depth >= kMaxDepth || s.polys.size() < kMinLeafPolys) {
            make_leaf(s, out);
            return;
        }
        float best = FLT_MAX, best_upper = 0.0f, best_lower = 0.0f, best_value = 0.0f;
        uint32_t best_dim = 0;
        const float dim[3] = {s.box.max.x - s.box.min.x, s.box.max.y - s.box.min.y,
                              s.box.max.z - s.box.min.z};
        const float inv_surface = 1.0f / s.surface;
        for (uint32_t j = 0; j < 3; j++) {
            uint32_t below = 0, cop = 0, above = 0;
            uint32_t counters[3] = {0, 0, 0};
            uint32_t plus_prev = 0, cop_prev = 0;
            const float x = dim[(j + 1) % 3];
            const float y = dim[(j + 2) % 3];
            const float xpy = x + y;
            const float area = x * y;
            bool first = true;
            const std::vector<Candidate>& c = s.cand[j];
            size_t i = 0;
            while (i < c.size()) {
                const float value = c[i].value;
                while (i < c.size() && c[i].value == value) {
                    counters[c[i].marker]++;
                    i++;
                }
                below = below + cop_prev + plus_prev;
                cop = cop_prev;
                above = above - counters[kCoplanarMark] - counters[kEnd];
                if (first) {
                    below = 0;
                    cop = counters[kCoplanarMark];
                    above = static_cast<uint32_t>(s.eval) - counters[kCoplanarMark];
                    first = false;
                }
                plus_prev = counters[kBegin];
                cop_prev = counters[kCoplanarMark];
                counters[0] = counters[1] = counters[2] = 0;

                const float upper = area + (lane(s.box.max, j) - value) * xpy;
                const float lower = area + (value - lane(s.box.min, j)) * xpy;
                const float cost_below = lower * static_cast<float>(below);
                const float cost_above = upper * static_cast<float>(above);
                const float cost_cop = area * static_cast<float>(cop);
                const float sah =
                    kTraverseCost + kIntersectionCost * (((cost_cop + cost_below) + cost_above) * inv_surface);
                if (sah < best) {
                    best = sah;
                    best_value = value;
                    best_dim = j;
                    best_upper = upper;
                    best_lower = lower;
                }
            }
        }
        const float leaf_cost = kIntersectionCost * static_cast<float>(s.eval);
        if (best >= leaf_cost) {
            make_leaf(s, out);
            return;
        }
        {
            hulk_kd_node& n = out.nodes[s.node];
            n.u.junction.plane = best_value;
            n.flags = best_dim | HULK_NODE_JUNCTION;
        }
        Split a, b;
        for (uint32_t p : s.polys) {
            Relation r = classify(best_dim, best_value, p);
            rel[p] = r;
            if (r == kAbove) a.polys.push_back(p);
            else if (r == kBelow) b.polys.push_back(p);
            else { a.polys.push_back(p); b.polys.push_back(p); }
        }
        for (uint32_t j = 0; j < 3; j++) {
            a.cand[j].reserve(s.cand[j].size());
            b.cand[j].reserve(s.cand[j].size());
            for (const Candidate& c : s.cand[j]) {
                switch (rel[c.poly]) {
                    case kAbove: a.cand[j].push_back(c); break;
                    case kBelow: b.cand[j].push_back(c); break;
                    case kStraddle:
                        if (j == best_dim) {
                            Candidate t = c;
                            t.value = best_value > c.value ? best_value : c.value;  // MAX, util.hpp:66
                            a.cand[j].push_back(t);
                            t.value = best_value < c.value ? best_value : c.value;  // MIN, util.hpp:67
                            b.cand[j].push_back(t);
                        } else {
                            a.cand[j].push_back(c);
                            b.cand[j].push_back(c);
                        }
                        break;
                    case kCoplanar: break;  // candidates of coplanar polygons are dropped
                }
            }
            a.eval += a.cand[j].size();
            b.eval += b.cand[j].size();
        }
        a.eval /= 6;
        b.eval /= 6;
        // The parent's lists are no longer needed: free them before
        // recursing so peak memory stays O(depth * level size).
        for (uint32_t j = 0; j < 3; j++) std::vector<Candidate>().swap(s.cand[j]);
        std::vector<uint32_t>().swap(s.polys);

        const uint32_t parent = s.node;
        const bool has_a = !a.polys.empty(), has_b = !b.polys.empty();
        if (has_a) {
            a.depth = s.depth + 1;
            a.box = s.box;
            a.surface = best_upper;
            set_lane(a.box.min, best_dim, best_value);
        }
        if (has_b) {
            b.depth = s.depth + 1;
            b.box = s.box;
            b.surface = best_lower;
            set_lane(b.box.max, best_dim, best_value);
        }
        if (has_a && has_b && s.depth < spawn_depth_) {
            // both subtrees in parallel, each into its own arrays
            Out oa, ob;
            hulk_kd_node empty;
            std::memset(&empty, 0, sizeof(empty));
            a.node = 0;
            b.node = 0;
            oa.nodes.push_back(empty);
            ob.nodes.push_back(empty);
            std::thread ta([this, &a, &oa] {
                std::vector<Relation> rel_a(n_tris_);
                build_node(a, oa, rel_a);
            });
            build_node(b, ob, rel);
            ta.join();
            const uint32_t ia = append(out, oa);
            out.nodes[parent].u.junction.above = ia;
            const uint32_t ib = append(out, ob);
            out.nodes[parent].u.junction.below = ib;
            return;
        }
        if (has_a) {
            const uint32_t idx = static_cast<uint32_t>(out.nodes.size());
            a.node = idx;
            hulk_kd_node empty;
            std::memset(&empty, 0, sizeof(empty));
            out.nodes.push_back(empty);
            out.nodes[parent].u.junction.above = idx;
            build_node(a, out, rel);
        }
        if (has_b) {
            const uint32_t idx = static_cast<uint32_t>(out.nodes.size());
            b.node = idx;
            hulk_kd_node empty;
            std::memset(&empty, 0, sizeof(empty));
            out.nodes.push_back(empty);
            out.nodes[parent].u.junction.below = idx;
            build_node(b, out, rel);
        }
    }

    const hulk_float4* verts_;
    const hulk_tri* tris_;
    uint32_t n_tris_;
    uint32_t spawn_depth_;
};

}  // namespace


extern "C" int hulk_kdtree_build_threads(const hulk_float4* verts, uint32_t n_verts, const hulk_tri* tris,
                                         uint32_t n_tris, const hulk_box* scene_box, uint32_t max_threads,
                                         hulk_kdtree** out) {
    if (!out || !scene_box || (n_tris && (!verts || !tris))) return HULK_E_ILLEGAL_INPUT;
    for (uint32_t i = 0; i < n_tris; i++)
        for (int k = 0; k < 3; k++)
            if (tris[i].idx[k] >= n_verts) return HULK_E_ILLEGAL_INPUT;
    hulk_kdtree* t = new (std::nothrow) hulk_kdtree;
    if (!t) return HULK_E_ALLOC;
    try {
        // Subtrees below the top `spawn` levels run on their own threads
        // (up to 2^spawn concurrent builds).  0 = automatic: up to 16
        // threads, one thread for scenes under 20k triangles.
        uint32_t threads = max_threads;
        if (threads == 0) {
            threads = std::min(std::max(1u, std::thread::hardware_concurrency()), 16u);
            if (n_tris < 20000) threads = 1;
        }
        uint32_t spawn = 0;
        while (spawn < 10 && (1u << (spawn + 1)) <= threads) spawn++;
        Builder b(verts, tris, n_tris, spawn);
        Out res;
        b.build(*scene_box, res);
        t->nodes.swap(res.nodes);
        t->leaf_polys.swap(res.leaf_polys);
    } catch (const std::bad_alloc&) {
        delete t;
        return HULK_E_ALLOC;
    } catch (const std::system_error&) {
        delete t;
        return HULK_E_ALLOC;
    }
    *out = t;
    return HULK_OK;
}

extern "C" int hulk_kdtree_build(const hulk_float4* verts, uint32_t n_verts, const hulk_tri* tris,
                                 uint32_t n_tris, const hulk_box* scene_box, hulk_kdtree** out) {
    return hulk_kdtree_build_threads(verts, n_verts, tris, n_tris, scene_box, 0, out);
}

extern "C" uint32_t hulk_kdtree_node_count(const hulk_kdtree* t) {
    return t ? static_cast<uint32_t>(t->nodes.size()) : 0;
}
extern "C" uint32_t hulk_kdtree_leaf_poly_count(const hulk_kdtree* t) {
    return t ? static_cast<uint32_t>(t->leaf_polys.size()) : 0;
}
extern "C" const hulk_kd_node* hulk_kdtree_nodes(const hulk_kdtree* t) { return t ? t->nodes.data() : nullptr; }
extern "C" const uint32_t* hulk_kdtree_leaf_polys(const hulk_kdtree* t) {
    return t ? t->leaf_polys.data() : nullptr;
}
extern "C" void hulk_kdtree_free(hulk_kdtree* t) { delete t; }

// BoundingBoxVisitor::postAct / traverseEnd (BoundingBoxVisitor.cpp:27-79)
// over one flat vertex array: start from the first vertex, then
// minps(box, v) / maxps(box, v) per lane.
extern "C" int hulk_scene_box(const hulk_float4* verts, uint32_t n_verts, hulk_box* out) {
    if (!verts || !out || n_verts == 0) return HULK_E_ILLEGAL_INPUT;
    hulk_box b;
    b.min = verts[0];
    b.max = verts[0];
    for (uint32_t i = 1; i < n_verts; i++) {
        const hulk_float4& v = verts[i];
        b.min.x = sse_min(b.min.x, v.x); b.min.y = sse_min(b.min.y, v.y);
        b.min.z = sse_min(b.min.z, v.z); b.min.w = sse_min(b.min.w, v.w);
        b.max.x = sse_max(b.max.x, v.x); b.max.y = sse_max(b.max.y, v.y);
        b.max.z = sse_max(b.max.z, v.z); b.max.w = sse_max(b.max.w, v.w);
    }
    *out = b;
    return HULK_OK;
}
