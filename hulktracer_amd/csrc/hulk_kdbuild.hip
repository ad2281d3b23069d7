// hulk_kdbuild.hip - SAH-NlogN KD-tree build on the device, level by level.
//
// Builds the tree of buildKDTreeSAH_NlogN / buildNodeSAHNlogN
// (KDTreeSAH.cpp:247-521, called by KDTreeCPU::buildKDTree,
// KDTreeCPU.cpp:61-94) with the same node and leaf-polygon arrays, from the
// vertices and triangles already in HBM.  The reference recurses depth first
// over one node at a time; here every node of one depth is processed at once
// and a depth costs a fixed sequence of launches:
//
//   group scan   the split candidates of every node sit in one array, node
//                after node and axis after axis ([x | y | z] per node), each
//                axis run sorted by value.  One scan gives, per candidate,
//                the BEGIN markers before it, the END markers up to it and
//                the start of its run of equal values (max-scan of run heads).
//   SAH          the last candidate of every run of equal values evaluates
//                the reference's sweep counters in closed form:
//                  below = BEGIN markers of the earlier runs of the node axis,
//                  above = eval - END markers of runs 1..this one
//                (the sweep never subtracts the END markers of the first run
//                and never sees COPLANAR markers: the root creates only BEGIN
//                and END, KDTreeSAH.cpp:288-295, and splits copy markers);
//                the SAH is computed with the reference's float operations
//                and reduced per node with a 64-bit atomicMin of
//                (SAH, candidate index), which keeps the sweep's first-minimum
//                rule (strict <, axis 0..2, ascending value).
//   decide       leaf if depth >= 20, fewer than 10 polygons or
//                SAH >= 5 * eval (KDTreeSAH.cpp:316-321, 408-413); else
//                split and set up the child boxes and surfaces.
//   split scan   classify every polygon and candidate against the plane
//                (checkPolygonToAAPlane, KDTreeCPU.cpp:450-477); a scan of
//                (above, below, leaf) flags gives each element its place in
//                the next depth's arrays, which keeps the reference's list
//                order (stable partition, KDTreeSAH.cpp:456-489).  Straddling
//                candidates on the split axis are clamped to the plane;
//                candidates of coplanar polygons are dropped.
//   scatter      write the next depth's polygons and candidates and the
//                children's ranges; leaf polygons go to a staging array.
// Children live in heap slots (depth d + 1, slots 2i and 2i + 1), so no
// scan over nodes is needed.  After depth 20 a bottom-up pass counts
// subtree nodes, leaves and leaf polygons, a top-down pass turns them into
// the depth-first pre-order index (above child first), leaf id and leaf
// start of the reference arrays, and into the device layout the traversal
// kernels read (breadth-first above depth 12, depth-first below).
//
// Sorting the root candidates (one rocPRIM radix sort of (axis, value)
// keys) replaces the three std::sort calls of KDTreeSAH.cpp:303-306.  Only the order of
// values matters: candidates of equal value form one sweep step.  -0.0 and
// +0.0 sort as equal keys, as under std::sort; which of them comes first in
// such a run is left open by std::sort, so a split plane at zero may carry
// either sign (the classification and traversal compare equal either way).
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <new>

#include "hulk_kdbuild.hpp"

namespace hulk {
namespace {

constexpr uint32_t kMaxDepth = 20;         // c_maxKDTreeDepth, SharedStructs.hpp:16
constexpr uint32_t kMinLeafPolys = 10;     // KDTreeSAH.cpp:317
constexpr uint32_t kReduction = 5;         // c_reductionFactor, KDTreeCPU.hpp:23
constexpr uint32_t kThreads = 256;
constexpr uint32_t kItems = 4;
constexpr uint32_t kTile = kThreads * kItems;
constexpr uint32_t kSlots = (1u << kKdLevels) - 1u;  // heap slots of depths 0..20
constexpr uint32_t kPolyMask = 0x3FFFFFFFu;
constexpr uint32_t kEndBit = 1u << 30;      // candidate marker: END (else BEGIN)
constexpr uint32_t kNegZeroBit = 1u << 31;  // sort payload: the value was -0.0
constexpr uint32_t kMaxGrid = 2048;

enum : uint32_t { kSlotLeaf = 1u, kSlotSplit = 2u };
enum : uint8_t { kAbove = 0, kBelow = 1, kStraddle = 2, kCoplanar = 3, kToLeaf = 4 };

// One node of one depth (heap slot).  128 B.
struct alignas(16) LNode {
    float bmin[3], bmax[3];
    float surf;                // full box surface at the root, the split's face term below
    uint32_t ev;               // polygonsEval
    uint32_t pstart, pcount;   // polygon references of the depth's array
    uint32_t cstart[3], ccount[3];
    unsigned long long key;    // (ordered SAH << 32) | candidate index, min over the node
    uint32_t g1[3];            // first candidate of the second run of equal values, per axis
    uint32_t flags;            // kSlotLeaf / kSlotSplit | axis << 4 (valid for the nodes of the lists)
    float plane;
    uint32_t stage;            // leaf: first staging entry
    uint32_t size, nleaf, nlp; // subtree: nodes, leaves, leaf polygon references
    uint32_t pre, lid, lps;    // reference layout: pre-order index, leaf id, leaf start
    uint32_t dev;              // device layout index
    uint32_t pad;
};
static_assert(sizeof(LNode) == 128, "LNode layout");

struct LevelInfo {
    uint32_t n_polys, n_cands, stage_base;
    uint32_t n_nodes;  // nodes of this depth (entries of its node list)
};

struct Status {
    uint32_t overflow, bad_input, max_jdepth, n_nodes, n_lp, stage_total, levels;
    uint32_t internal;  // an index failed a consistency check (the kernel skipped it)
    uint32_t prefix[16];
};

struct Caps {
    uint32_t polys, cands, stage;
};

struct KdDev {
    const float4* verts;
    const uint4* tris;
    uint32_t n_tris;
    float2* tbox;  // per axis and triangle: (min, max) of the three vertex coordinates
    LNode* nodes;
    LevelInfo* lvl;  // kKdLevels + 1 entries
    Status* st;
    uint32_t* poly[2];
    uint32_t* pslot[2];
    uint2* cand[2];     // value bits, poly | END bit
    uint32_t* cseg[2];  // slot << 2 | axis
    uint32_t *bx, *ei, *hd;
    uint32_t *pa, *pb, *pl;
    uint32_t *ca, *cb;
    uint8_t *rp, *rc;
    uint4 *part_p, *part_c;
    uint32_t* nlist;  // per depth L at slot_base(L): the slots of the depth's nodes, in no order
    uint32_t *stage_poly, *stage_owner;
    uint4* ref_nodes;
    uint4* dev_nodes;
    uint32_t* lp;
    Caps cap;
};

__device__ __forceinline__ uint32_t slot_base(uint32_t L) { return (1u << L) - 1u; }

__device__ __forceinline__ float lane3(const float4& v, uint32_t a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

// minps / maxps operand rule (util.hpp:66-67 MIN / MAX use the same form).
__device__ __forceinline__ float sse_min(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float sse_max(float a, float b) { return a > b ? a : b; }

// Float -> unsigned key with the same order as float <; -0.0 maps to +0.0.
__device__ __forceinline__ uint32_t order_key(float v) {
    uint32_t u = __float_as_uint(v);
    if (u == 0x80000000u) u = 0u;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ uint32_t key_bits(uint32_t k) { return (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k; }

__device__ __forceinline__ float cval(const uint2& c) { return __uint_as_float(c.x); }

// checkPolygonToAAPlane (KDTreeCPU.cpp:450-477) from the triangle's extent
// on the axis: all three coordinates >= v exactly when their minimum is,
// <= v when their maximum is, == v when both are (coordinates are not NaN:
// the root pass rejects NaN).  One 8-byte load instead of three vertices.
__device__ __forceinline__ uint8_t classify(const KdDev& D, uint32_t poly, uint32_t axis, float v) {
    const float2 b = D.tbox[size_t(axis) * D.n_tris + poly];
    if ((b.x == v) & (b.y == v)) return kCoplanar;
    if (b.x >= v) return kAbove;
    if (b.y <= v) return kBelow;
    return kStraddle;
}

// The sweep's face terms for a plane at `value` on axis j
// (KDTreeSAH.cpp:338-405): x, y = the other two box extents.
__device__ __forceinline__ void faces(const LNode& n, uint32_t j, float value, float& area, float& upper,
                                      float& lower) {
    const float d0 = n.bmax[0] - n.bmin[0], d1 = n.bmax[1] - n.bmin[1], d2 = n.bmax[2] - n.bmin[2];
    const float x = j == 0 ? d1 : (j == 1 ? d2 : d0);
    const float y = j == 0 ? d2 : (j == 1 ? d0 : d1);
    const float xpy = x + y;
    area = x * y;
    const float hi = j == 0 ? n.bmax[0] : (j == 1 ? n.bmax[1] : n.bmax[2]);
    const float lo = j == 0 ? n.bmin[0] : (j == 1 ? n.bmin[1] : n.bmin[2]);
    upper = area + (hi - value) * xpy;
    lower = area + (value - lo) * xpy;
}

__device__ __forceinline__ bool level_empty(const KdDev& D, uint32_t L) { return L > 0 && D.lvl[L].n_polys == 0; }

// Consistency guard on an index read back from a device array: false (and
// the internal flag set) when it is out of range, so a build error surfaces
// as an error code instead of a stray access.
__device__ __forceinline__ bool in_range(const KdDev& D, uint32_t v, uint32_t n) {
    if (v < n) return true;
    D.st->internal = 1u;
    return false;
}

// ------------------------------------------------------------ block scans --

struct SumOp {
    __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return a + b; }
};
struct MaxOp {
    __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return a > b ? a : b; }
};

template <class Op>
__device__ __forceinline__ uint32_t wave_incl(uint32_t v, Op op) {
    const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v = op(v, t);
    }
    return v;
}

// Exclusive scan over the NT threads of a block (identity 0 for both
// operators: every value is >= 0); `total` = the block's reduction.
template <uint32_t NT, class Op>
__device__ __forceinline__ uint32_t block_excl(uint32_t v, Op op, uint32_t* lds, uint32_t& total) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t inc = wave_incl(v, op);
    if (lane == 63u) lds[wave] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (uint32_t w = 0; w < NT / 64; w++) {
        const uint32_t s = lds[w];
        if (w < wave) pre = op(pre, s);
        tot = op(tot, s);
    }
    __syncthreads();
    uint32_t ex = __shfl_up(inc, 1, 64);
    if (lane == 0) ex = 0;
    total = tot;
    return op(pre, ex);
}

// Three scans in one pass (two sums and a sum or max in the third lane):
// the same steps as block_excl with one barrier pair for all three.
template <uint32_t NT, bool MAX3>
__device__ __forceinline__ uint3 block_excl3(uint3 v, uint3* lds, uint3& total) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint3 inc = v;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t ta = __shfl_up(inc.x, d, 64), tb = __shfl_up(inc.y, d, 64), tc = __shfl_up(inc.z, d, 64);
        if (lane >= d) {
            inc.x += ta;
            inc.y += tb;
            inc.z = MAX3 ? (inc.z > tc ? inc.z : tc) : inc.z + tc;
        }
    }
    if (lane == 63u) lds[wave] = inc;
    __syncthreads();
    uint3 pre = make_uint3(0, 0, 0), tot = make_uint3(0, 0, 0);
#pragma unroll
    for (uint32_t w = 0; w < NT / 64; w++) {
        const uint3 q = lds[w];
        if (w < wave) {
            pre.x += q.x;
            pre.y += q.y;
            pre.z = MAX3 ? (pre.z > q.z ? pre.z : q.z) : pre.z + q.z;
        }
        tot.x += q.x;
        tot.y += q.y;
        tot.z = MAX3 ? (tot.z > q.z ? tot.z : q.z) : tot.z + q.z;
    }
    __syncthreads();
    uint3 ex = make_uint3(__shfl_up(inc.x, 1, 64), __shfl_up(inc.y, 1, 64), __shfl_up(inc.z, 1, 64));
    if (lane == 0) ex = make_uint3(0, 0, 0);
    total = tot;
    return make_uint3(pre.x + ex.x, pre.y + ex.y, MAX3 ? (pre.z > ex.z ? pre.z : ex.z) : pre.z + ex.z);
}

// ------------------------------------------------------------------ root --

__global__ void k_kdb_init(KdDev D, SceneBox box, uint32_t n_tris, uint32_t ns) {
    const uint32_t t = threadIdx.x;
    if (t < kKdLevels + 1) {
        D.lvl[t].n_polys = t == 0 ? n_tris : 0u;
        D.lvl[t].n_cands = t == 0 ? 6u * ns : 0u;
        D.lvl[t].stage_base = 0;
        D.lvl[t].n_nodes = t == 0 ? 1u : 0u;
    }
    if (t < sizeof(Status) / 4) reinterpret_cast<uint32_t*>(D.st)[t] = 0;
    if (t == 0) {
        D.nlist[0] = 0;
        LNode& r = D.nodes[0];
        for (int a = 0; a < 3; a++) {
            r.bmin[a] = box.mn[a];
            r.bmax[a] = box.mx[a];
        }
        // getBoxSurface (SharedFunctions.hpp:32-36)
        const float sx = box.mx[0] - box.mn[0], sy = box.mx[1] - box.mn[1], sz = box.mx[2] - box.mn[2];
        r.surf = 2.0f * (sx * sy + sz * sy + sx * sz);
        r.ev = ns;
        r.pstart = 0;
        r.pcount = n_tris;
        for (uint32_t a = 0; a < 3; a++) {
            r.cstart[a] = 2u * ns * a;
            r.ccount[a] = 2u * ns;
            r.g1[a] = 0xFFFFFFFFu;
        }
        r.key = ~0ull;
        r.flags = 0;
        r.plane = 0.0f;
        r.stage = 0;
    }
}

// Root polygon list, input check and the unsorted candidates: every 5th
// triangle, per axis a BEGIN at its min and an END at its max
// (KDTreeSAH.cpp:273-295).  Key = axis << 32 | ordered value, so one sort
// of all three axes leaves them as 3 runs of 2 * ns.
__global__ void k_kdb_root(KdDev D, uint32_t n_tris, uint32_t ns, unsigned long long* keys, uint32_t* vals) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n_tris) {
        D.poly[0][g] = g;
        D.pslot[0][g] = 0;
        const uint4 t = D.tris[g];
        const float4 a = D.verts[t.x], b = D.verts[t.y], c = D.verts[t.z];
        const bool nan = (a.x != a.x) | (a.y != a.y) | (a.z != a.z) | (b.x != b.x) | (b.y != b.y) | (b.z != b.z) |
                         (c.x != c.x) | (c.y != c.y) | (c.z != c.z);
        if (nan) D.st->bad_input = 1u;
#pragma unroll
        for (uint32_t j = 0; j < 3; j++)
            D.tbox[size_t(j) * n_tris + g] =
                make_float2(sse_min(sse_min(lane3(a, j), lane3(b, j)), lane3(c, j)),
                            sse_max(sse_max(lane3(a, j), lane3(b, j)), lane3(c, j)));
    }
    if (g < ns) {
        const uint32_t p = g * kReduction;
        const uint4 t = D.tris[p];
        const float4 p1 = D.verts[t.x], p2 = D.verts[t.y], p3 = D.verts[t.z];
#pragma unroll
        for (uint32_t j = 0; j < 3; j++) {
            const float mn = sse_min(sse_min(lane3(p1, j), lane3(p2, j)), lane3(p3, j));
            const float mx = sse_max(sse_max(lane3(p1, j), lane3(p2, j)), lane3(p3, j));
            const uint32_t o = j * 2u * ns + 2u * g;
            const unsigned long long axis = static_cast<unsigned long long>(j) << 32;
            keys[o] = axis | order_key(mn);
            vals[o] = p | (__float_as_uint(mn) == 0x80000000u ? kNegZeroBit : 0u);
            keys[o + 1] = axis | order_key(mx);
            vals[o + 1] = p | kEndBit | (__float_as_uint(mx) == 0x80000000u ? kNegZeroBit : 0u);
        }
    }
}

__global__ void k_kdb_root_cands(KdDev D, uint32_t ns, const unsigned long long* keys, const uint32_t* vals) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 6u * ns) return;
    const uint32_t v = vals[e];
    const uint32_t bits = (v & kNegZeroBit) ? 0x80000000u : key_bits(static_cast<uint32_t>(keys[e]));
    D.cand[0][e] = make_uint2(bits, v & ~kNegZeroBit);
    D.cseg[0][e] = static_cast<uint32_t>(keys[e] >> 32);  // axis run e / (2 ns)
}

// ------------------------------------------------------------ group scan --

// One candidate's scan terms: BEGIN, END, run head (its index when it starts
// a run of equal values of its node axis, else 0).  `first_run_end`: it is
// the last candidate of its node axis's first run.
__device__ __forceinline__ void group_item(const KdDev& D, uint32_t L, uint32_t i, uint32_t& b, uint32_t& e,
                                           uint32_t& h, uint32_t* seg_start = nullptr, uint32_t* seg_end = nullptr) {
    const uint2* cand = D.cand[L & 1];
    const uint2 c = cand[i];
    const uint32_t seg = D.cseg[L & 1][i];
    b = e = h = 0;
    if (!in_range(D, seg >> 2, 1u << L) || !in_range(D, seg & 3u, 3u)) return;
    const LNode& nd = D.nodes[slot_base(L) + (seg >> 2)];
    const uint32_t j = seg & 3u;
    const uint32_t s = nd.cstart[j];
    const bool head = i == s || i == 0 || cval(c) != cval(cand[i - 1]);
    b = (c.y & kEndBit) ? 0u : 1u;
    e = (c.y & kEndBit) ? 1u : 0u;
    h = head ? i : 0u;
    if (seg_start) {
        *seg_start = s;
        *seg_end = s + nd.ccount[j];
    }
}

__global__ void __launch_bounds__(kThreads) k_kdb_group_reduce(KdDev D, uint32_t L) {
    __shared__ uint3 s3[kThreads / 64];
    if (D.st->overflow) return;
    const uint32_t n = D.lvl[L].n_cands;
    const uint32_t tiles = (n + kTile - 1) / kTile;
    for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        uint32_t b = 0, e = 0, h = 0;
        const uint32_t i0 = t * kTile + threadIdx.x * kItems;
        for (uint32_t k = 0; k < kItems; k++) {
            const uint32_t i = i0 + k;
            if (i >= n) break;
            uint32_t bi, ei, hi;
            group_item(D, L, i, bi, ei, hi);
            b += bi;
            e += ei;
            h = max(h, hi);
        }
        uint3 tot;
        block_excl3<kThreads, true>(make_uint3(b, e, h), s3, tot);
        if (threadIdx.x == 0) D.part_c[t] = make_uint4(tot.x, tot.y, tot.z, 0);
    }
}

// Exclusive scan of tile partials in place, one block of 1024 threads per
// array.  mode 0: the candidates' (BEGIN sum, END sum, head max).
// mode 1: block 0 scans the polygons' (above, below, leaf) sums, block 1
// the candidates' (above, below); both write the totals after the last
// element and the next depth's sizes, and flag an overflow.
__global__ void __launch_bounds__(1024) k_kdb_scan_parts(KdDev D, uint32_t L, uint32_t mode) {
    __shared__ uint32_t s0[16], s1[16], s2[16];
    if (D.st->overflow) return;
    const bool polys = mode == 1 && blockIdx.x == 0;
    const uint32_t n = polys ? D.lvl[L].n_polys : D.lvl[L].n_cands;
    uint4* part = polys ? D.part_p : D.part_c;
    const uint32_t T = (n + kTile - 1) / kTile;
    const uint32_t t = threadIdx.x;
    const uint32_t per = (T + 1023u) / 1024u;
    const uint32_t a = min(T, t * per), b = min(T, a + per);
    uint32_t x = 0, y = 0, z = 0;
    for (uint32_t k = a; k < b; k++) {
        const uint4 v = part[k];
        x += v.x;
        y += v.y;
        z = mode == 0 ? max(z, v.z) : z + v.z;
    }
    uint32_t tx, ty, tz, px, py, pz;
    px = block_excl<1024>(x, SumOp(), s0, tx);
    py = block_excl<1024>(y, SumOp(), s1, ty);
    if (mode == 0)
        pz = block_excl<1024>(z, MaxOp(), s2, tz);
    else
        pz = block_excl<1024>(z, SumOp(), s2, tz);
    for (uint32_t k = a; k < b; k++) {
        const uint4 v = part[k];
        part[k] = make_uint4(px, py, pz, 0);
        px += v.x;
        py += v.y;
        pz = mode == 0 ? max(pz, v.z) : pz + v.z;
    }
    if (mode == 0 || t != 0) return;
    LevelInfo& next = D.lvl[L + 1];
    if (polys) {
        D.pa[n] = tx;
        D.pb[n] = ty;
        D.pl[n] = tz;
        const uint64_t np = (uint64_t)tx + ty;
        const uint64_t sb = (uint64_t)D.lvl[L].stage_base + tz;
        if (np > D.cap.polys || sb > D.cap.stage) D.st->overflow = 1u;
        next.n_polys = static_cast<uint32_t>(np);
        next.stage_base = static_cast<uint32_t>(sb);
    } else {
        D.ca[n] = tx;
        D.cb[n] = ty;
        const uint64_t nc = (uint64_t)tx + ty;
        if (nc > D.cap.cands) D.st->overflow = 1u;
        next.n_cands = static_cast<uint32_t>(nc);
    }
}

__global__ void __launch_bounds__(kThreads) k_kdb_group_apply(KdDev D, uint32_t L) {
    __shared__ uint3 s3[kThreads / 64];
    if (D.st->overflow) return;
    const uint32_t n = D.lvl[L].n_cands;
    const uint32_t tiles = (n + kTile - 1) / kTile;
    for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        uint32_t bk[kItems], ek[kItems], hk[kItems], sk[kItems], nk[kItems];
        uint32_t b = 0, e = 0, h = 0;
        const uint32_t i0 = t * kTile + threadIdx.x * kItems;
#pragma unroll
        for (uint32_t k = 0; k < kItems; k++) {
            bk[k] = ek[k] = hk[k] = 0;
            sk[k] = nk[k] = 0;
            if (i0 + k < n) group_item(D, L, i0 + k, bk[k], ek[k], hk[k], &sk[k], &nk[k]);
            b += bk[k];
            e += ek[k];
            h = max(h, hk[k]);
        }
        const uint4 P = D.part_c[t];
        uint3 tot;
        const uint3 ex = block_excl3<kThreads, true>(make_uint3(b, e, h), s3, tot);
        uint32_t B = P.x + ex.x;
        uint32_t E = P.y + ex.y;
        uint32_t H = max(P.z, ex.z);
#pragma unroll
        for (uint32_t k = 0; k < kItems; k++) {
            const uint32_t i = i0 + k;
            if (i >= n) break;
            D.bx[i] = B;
            B += bk[k];
            E += ek[k];
            D.ei[i] = E;
            H = max(H, hk[k]);
            D.hd[i] = H;
            // the last candidate of the node axis's first run records where
            // the second run starts (one writer per node axis)
            if (H == sk[k] && (i + 1 >= nk[k] || cval(D.cand[L & 1][i + 1]) != cval(D.cand[L & 1][i]))) {
                const uint32_t seg = D.cseg[L & 1][i];
                D.nodes[slot_base(L) + (seg >> 2)].g1[seg & 3u] = i + 1;
            }
        }
    }
}

// ------------------------------------------------------------------- SAH --

// SAH of one candidate if it ends a run of equal values of an open node
// (else false).
__device__ __forceinline__ bool run_sah(const KdDev& D, uint32_t L, uint32_t i, uint32_t n, uint32_t& slot,
                                        unsigned long long& key) {
    const uint2* cand = D.cand[L & 1];
    const uint32_t seg = D.cseg[L & 1][i];
    if (!in_range(D, seg >> 2, 1u << L) || !in_range(D, seg & 3u, 3u)) return false;
    slot = seg >> 2;
    const LNode& nd = D.nodes[slot_base(L) + slot];
    if (nd.pcount < kMinLeafPolys) return false;
    const uint32_t j = seg & 3u;
    const uint32_t s = nd.cstart[j], end = s + nd.ccount[j];
    if (!in_range(D, s, i + 1) || !in_range(D, end, n + 1)) return false;
    if (i + 1 < end && cval(cand[i + 1]) == cval(cand[i])) return false;  // not the run's last candidate
    const uint32_t f = D.hd[i];
    if (!in_range(D, f, i + 1) || !in_range(D, s, f + 1)) return false;
    uint32_t below, above;
    if (f == s) {
        below = 0;
        above = nd.ev;
    } else {
        const uint32_t g1 = nd.g1[j];
        if (!in_range(D, g1, f + 1) || !in_range(D, s, g1)) return false;
        below = D.bx[f] - D.bx[s];
        above = nd.ev - (D.ei[i] - D.ei[g1 - 1u]);
    }
    const float value = cval(cand[f]);
    float area, upper, lower;
    faces(nd, j, value, area, upper, lower);
    const float inv_surface = 1.0f / nd.surf;
    const float cost_below = lower * static_cast<float>(below);
    const float cost_above = upper * static_cast<float>(above);
    const float cost_cop = area * static_cast<float>(0u);
    const float sah = 1.0f + 5.0f * (((cost_cop + cost_below) + cost_above) * inv_surface);
    if (!(sah < FLT_MAX)) return false;
    key = (static_cast<unsigned long long>(order_key(sah)) << 32) | f;
    return true;
}

// Per node minimum of the run keys.  The lanes of a wave hold consecutive
// candidates, so they belong to few nodes: the wave reduces each node's keys
// first (one pass per distinct node, leader lane first) and issues one
// atomic per node, which keeps the large nodes of the top depths from
// serialising a million atomics on one address.
//
// Depths up to kSahLdsDepth have at most 64 nodes: there the block also
// folds the wave results into an LDS slot table and flushes it once, and
// the host launches at most 1,024 blocks, so a node sees at most 1,024
// global atomics.
constexpr uint32_t kSahLdsDepth = 6;

__global__ void __launch_bounds__(kThreads) k_kdb_sah(KdDev D, uint32_t L) {
    __shared__ unsigned long long table[1u << kSahLdsDepth];
    if (D.st->overflow || L >= kMaxDepth) return;
    const uint32_t n = D.lvl[L].n_cands;
    const uint32_t lane = threadIdx.x & 63u;
    const bool shallow = L <= kSahLdsDepth;
    LNode* nodes = D.nodes + slot_base(L);
    if (shallow) {
        if (threadIdx.x < (1u << L)) table[threadIdx.x] = ~0ull;
        __syncthreads();
    }
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        uint32_t slot = 0;
        unsigned long long key = ~0ull;
        const bool valid = i < n && run_sah(D, L, i, n, slot, key);
        unsigned long long pending = __ballot(valid);
        while (pending) {
            const int leader = __ffsll(static_cast<long long>(pending)) - 1;
            const uint32_t ls = __shfl(slot, leader, 64);
            const unsigned long long match = __ballot(valid && slot == ls) & pending;
            unsigned long long v = ((match >> lane) & 1ull) ? key : ~0ull;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const unsigned long long o = __shfl_xor(v, off, 64);
                v = o < v ? o : v;
            }
            if (static_cast<int>(lane) == leader) {
                if (shallow)
                    atomicMin(&table[ls], v);
                else
                    atomicMin(&nodes[ls].key, v);
            }
            pending &= ~match;
        }
    }
    if (shallow) {
        __syncthreads();
        if (threadIdx.x < (1u << L) && table[threadIdx.x] != ~0ull) atomicMin(&nodes[threadIdx.x].key, table[threadIdx.x]);
    }
}

// ---------------------------------------------------------------- decide --

// The nodes of a depth, one thread each, from the depth's node list, so that
// a pass costs the depth's nodes, not its 2^L heap slots.  The list holds
// both children of every split of the depth above (decide appends them);
// a child that got no polygons is no node (the root always is one).
__device__ __forceinline__ bool is_node(const KdDev& D, uint32_t L, uint32_t i) {
    return in_range(D, i, 1u << L) && (L == 0 || D.nodes[slot_base(L) + i].pcount > 0);
}
#define KDB_FOR_NODES(D, L, i)                                                                          \
    for (uint32_t k_ = blockIdx.x * blockDim.x + threadIdx.x, n_ = (D).lvl[L].n_nodes; k_ < n_;         \
         k_ += gridDim.x * blockDim.x)                                                                   \
        if (const uint32_t i = (D).nlist[slot_base(L) + k_]; is_node((D), (L), i))

__device__ bool decide_node(const KdDev& D, uint32_t L, uint32_t i) {
    LNode& nd = D.nodes[slot_base(L) + i];
    uint32_t flags = kSlotLeaf;
    if (L < kMaxDepth && nd.pcount >= kMinLeafPolys) {
        const unsigned long long key = nd.key;
        const float best = key == ~0ull ? FLT_MAX : __uint_as_float(key_bits(static_cast<uint32_t>(key >> 32)));
        const float leaf_cost = 5.0f * static_cast<float>(nd.ev);
        if (!(best >= leaf_cost)) {
            const uint32_t f = static_cast<uint32_t>(key);
            if (!in_range(D, f, D.lvl[L].n_cands)) return false;
            const uint32_t j = D.cseg[L & 1][f] & 3u;
            const float value = cval(D.cand[L & 1][f]);
            float area, upper, lower;
            faces(nd, j, value, area, upper, lower);
            flags = kSlotSplit | (j << 4);
            nd.plane = value;
            LNode& a = D.nodes[slot_base(L + 1) + 2u * i];
            LNode& b = D.nodes[slot_base(L + 1) + 2u * i + 1u];
            for (uint32_t k = 0; k < 3; k++) {
                a.bmin[k] = b.bmin[k] = nd.bmin[k];
                a.bmax[k] = b.bmax[k] = nd.bmax[k];
                a.g1[k] = b.g1[k] = 0xFFFFFFFFu;
            }
            a.bmin[j] = value;
            b.bmax[j] = value;
            a.surf = upper;
            b.surf = lower;
            a.key = b.key = ~0ull;
            a.stage = b.stage = 0;
            a.pcount = b.pcount = 0;
        }
    }
    nd.flags = flags;
    return (flags & kSlotSplit) != 0;
}

// Decide every node of the depth; both children of each split go to the
// next depth's node list, one atomic per wave.
__global__ void __launch_bounds__(kThreads) k_kdb_decide(KdDev D, uint32_t L) {
    if (D.st->overflow) return;
    const uint32_t n = D.lvl[L].n_nodes;
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const uint32_t k = base + threadIdx.x;
        uint32_t i = 0;
        bool split = false;
        if (k < n) {
            i = D.nlist[slot_base(L) + k];
            if (is_node(D, L, i)) split = decide_node(D, L, i);
        }
        const unsigned long long m = __ballot(split);
        if (m == 0ull) continue;
        const uint32_t leader = static_cast<uint32_t>(__builtin_ctzll(m));
        uint32_t first = 0;
        if (lane == leader) first = atomicAdd(&D.lvl[L + 1].n_nodes, 2u * static_cast<uint32_t>(__builtin_popcountll(m)));
        first = static_cast<uint32_t>(__shfl(static_cast<int>(first), static_cast<int>(leader), 64));
        if (split) {
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                                            __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
            uint32_t* next = D.nlist + slot_base(L + 1) + first + 2u * rank;
            next[0] = 2u * i;
            next[1] = 2u * i + 1u;
        }
    }
}

// ------------------------------------------------------------ split scan --

__device__ __forceinline__ void split_flags(uint8_t r, bool poly, uint32_t& a, uint32_t& b, uint32_t& l) {
    if (poly) {
        a = (r != kToLeaf && r != kBelow) ? 1u : 0u;
        b = (r != kToLeaf && r != kAbove) ? 1u : 0u;
        l = r == kToLeaf ? 1u : 0u;
    } else {
        a = (r == kAbove || r == kStraddle) ? 1u : 0u;
        b = (r == kBelow || r == kStraddle) ? 1u : 0u;
        l = 0;
    }
}

__global__ void __launch_bounds__(kThreads) k_kdb_split_reduce(KdDev D, uint32_t L) {
    __shared__ uint3 s3[kThreads / 64];
    if (D.st->overflow) return;
    const uint32_t np = D.lvl[L].n_polys, nc = D.lvl[L].n_cands;
    const uint32_t tp = (np + kTile - 1) / kTile, tc = (nc + kTile - 1) / kTile;
    const uint32_t cur = L & 1;
    const LNode* nodes = D.nodes + slot_base(L);
    for (uint32_t t = blockIdx.x; t < tp + tc; t += gridDim.x) {
        const bool is_poly = t < tp;
        const uint32_t tt = is_poly ? t : t - tp;
        const uint32_t n = is_poly ? np : nc;
        uint32_t a = 0, b = 0, l = 0;
        const uint32_t i0 = tt * kTile + threadIdx.x * kItems;
        for (uint32_t k = 0; k < kItems; k++) {
            const uint32_t i = i0 + k;
            if (i >= n) break;
            uint8_t r = kToLeaf;
            if (is_poly) {
                const uint32_t slot = D.pslot[cur][i], p = D.poly[cur][i];
                if (in_range(D, slot, 1u << L) && in_range(D, p, D.n_tris)) {
                    const LNode& nd = nodes[slot];
                    if (nd.flags & kSlotSplit) r = classify(D, p, (nd.flags >> 4) & 3u, nd.plane);
                }
                D.rp[i] = r;
            } else {
                const uint32_t slot = D.cseg[cur][i] >> 2, p = D.cand[cur][i].y & kPolyMask;
                if (in_range(D, slot, 1u << L) && in_range(D, p, D.n_tris)) {
                    const LNode& nd = nodes[slot];
                    if (nd.flags & kSlotSplit) r = classify(D, p, (nd.flags >> 4) & 3u, nd.plane);
                }
                D.rc[i] = r;
            }
            uint32_t fa, fb, fl;
            split_flags(r, is_poly, fa, fb, fl);
            a += fa;
            b += fb;
            l += fl;
        }
        uint3 tot;
        block_excl3<kThreads, false>(make_uint3(a, b, l), s3, tot);
        if (threadIdx.x == 0) (is_poly ? D.part_p : D.part_c)[tt] = make_uint4(tot.x, tot.y, tot.z, 0);
    }
}

__global__ void __launch_bounds__(kThreads) k_kdb_split_apply(KdDev D, uint32_t L) {
    __shared__ uint3 s3[kThreads / 64];
    if (D.st->overflow) return;
    const uint32_t np = D.lvl[L].n_polys, nc = D.lvl[L].n_cands;
    const uint32_t tp = (np + kTile - 1) / kTile, tc = (nc + kTile - 1) / kTile;
    for (uint32_t t = blockIdx.x; t < tp + tc; t += gridDim.x) {
        const bool is_poly = t < tp;
        const uint32_t tt = is_poly ? t : t - tp;
        const uint32_t n = is_poly ? np : nc;
        const uint8_t* rr = is_poly ? D.rp : D.rc;
        uint32_t ak[kItems], bk[kItems], lk[kItems];
        uint32_t a = 0, b = 0, l = 0;
        const uint32_t i0 = tt * kTile + threadIdx.x * kItems;
#pragma unroll
        for (uint32_t k = 0; k < kItems; k++) {
            ak[k] = bk[k] = lk[k] = 0;
            if (i0 + k < n) split_flags(rr[i0 + k], is_poly, ak[k], bk[k], lk[k]);
            a += ak[k];
            b += bk[k];
            l += lk[k];
        }
        const uint4 P = (is_poly ? D.part_p : D.part_c)[tt];
        uint3 tot;
        const uint3 ex = block_excl3<kThreads, false>(make_uint3(a, b, l), s3, tot);
        uint32_t A = P.x + ex.x;
        uint32_t B = P.y + ex.y;
        uint32_t C = P.z + ex.z;
#pragma unroll
        for (uint32_t k = 0; k < kItems; k++) {
            const uint32_t i = i0 + k;
            if (i >= n) break;
            if (is_poly) {
                D.pa[i] = A;
                D.pb[i] = B;
                D.pl[i] = C;
            } else {
                D.ca[i] = A;
                D.cb[i] = B;
            }
            A += ak[k];
            B += bk[k];
            C += lk[k];
        }
    }
}

// --------------------------------------------------------------- scatter --

__global__ void __launch_bounds__(kThreads) k_kdb_scatter(KdDev D, uint32_t L) {
    if (D.st->overflow) return;
    const uint32_t np = D.lvl[L].n_polys, nc = D.lvl[L].n_cands;
    const uint32_t cur = L & 1, nxt = cur ^ 1u;
    LNode* nodes = D.nodes + slot_base(L);
    const uint32_t stage_base = D.lvl[L].stage_base;
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < np + nc; g += gridDim.x * blockDim.x) {
        if (g < np) {
            const uint32_t i = g;
            const uint32_t p = D.poly[cur][i], slot = D.pslot[cur][i];
            if (!in_range(D, slot, 1u << L)) continue;
            LNode& nd = nodes[slot];
            const uint8_t r = D.rp[i];
            if (r == kToLeaf) {
                const uint32_t d = stage_base + D.pl[i];
                D.stage_poly[d] = p;
                D.stage_owner[d] = slot_base(L) + slot;
                if (i == nd.pstart) nd.stage = d;
                continue;
            }
            const uint32_t ps = nd.pstart, pe = ps + nd.pcount;
            if (!in_range(D, ps, i + 1) || !in_range(D, pe, np + 1)) continue;
            const uint32_t pas = D.pa[ps], pbs = D.pb[ps];
            const uint32_t na = D.pa[pe] - pas, nb = D.pb[pe] - pbs;
            const uint32_t base = pas + pbs;
            if (r != kBelow) {
                const uint32_t d = base + (D.pa[i] - pas);
                D.poly[nxt][d] = p;
                D.pslot[nxt][d] = 2u * slot;
            }
            if (r != kAbove) {
                const uint32_t d = base + na + (D.pb[i] - pbs);
                D.poly[nxt][d] = p;
                D.pslot[nxt][d] = 2u * slot + 1u;
            }
            if (i == ps) {
                // children's ranges in the next depth's arrays
                LNode& ca = D.nodes[slot_base(L + 1) + 2u * slot];
                LNode& cb = D.nodes[slot_base(L + 1) + 2u * slot + 1u];
                ca.pstart = base;
                ca.pcount = na;
                cb.pstart = base + na;
                cb.pcount = nb;
                const uint32_t cs0 = nd.cstart[0], ce2 = nd.cstart[2] + nd.ccount[2];
                if (!in_range(D, ce2, nc + 1) || !in_range(D, cs0, ce2 + 1)) continue;
                const uint32_t cas = D.ca[cs0], cbs = D.cb[cs0];
                const uint32_t nca = D.ca[ce2] - cas, ncb = D.cb[ce2] - cbs;
                const uint32_t cbase = cas + cbs;
                for (uint32_t j = 0; j < 3; j++) {
                    const uint32_t s = nd.cstart[j], e = s + nd.ccount[j];
                    ca.cstart[j] = cbase + (D.ca[s] - cas);
                    ca.ccount[j] = D.ca[e] - D.ca[s];
                    cb.cstart[j] = cbase + nca + (D.cb[s] - cbs);
                    cb.ccount[j] = D.cb[e] - D.cb[s];
                }
                ca.ev = nca / 6u;
                cb.ev = ncb / 6u;
            }
        } else {
            const uint32_t i = g - np;
            const uint8_t r = D.rc[i];
            if (r == kToLeaf || r == kCoplanar) continue;
            const uint2 c = D.cand[cur][i];
            const uint32_t seg = D.cseg[cur][i], slot = seg >> 2, j = seg & 3u;
            if (!in_range(D, slot, 1u << L)) continue;
            const LNode& nd = nodes[slot];
            const uint32_t axis = (nd.flags >> 4) & 3u;
            const float plane = nd.plane, v = cval(c);
            const uint32_t cs0 = nd.cstart[0], ce2 = nd.cstart[2] + nd.ccount[2];
            if (!in_range(D, ce2, nc + 1) || !in_range(D, cs0, i + 1)) continue;
            const uint32_t cas = D.ca[cs0], cbs = D.cb[cs0];
            const uint32_t nca = D.ca[ce2] - cas;
            const uint32_t cbase = cas + cbs;
            const bool clamp = r == kStraddle && j == axis;
            if (r != kBelow) {
                const uint32_t d = cbase + (D.ca[i] - cas);
                const float w = clamp ? (plane > v ? plane : v) : v;  // MAX, util.hpp:66
                D.cand[nxt][d] = make_uint2(__float_as_uint(w), c.y);
                D.cseg[nxt][d] = ((2u * slot) << 2) | j;
            }
            if (r != kAbove) {
                const uint32_t d = cbase + nca + (D.cb[i] - cbs);
                const float w = clamp ? (plane < v ? plane : v) : v;  // MIN, util.hpp:67
                D.cand[nxt][d] = make_uint2(__float_as_uint(w), c.y);
                D.cseg[nxt][d] = ((2u * slot + 1u) << 2) | j;
            }
        }
    }
}

// -------------------------------------------------------------- assembly --

__global__ void __launch_bounds__(kThreads) k_kdb_bottom_up(KdDev D, uint32_t L) {
    if (D.st->overflow) return;
    KDB_FOR_NODES(D, L, i) {
        LNode& nd = D.nodes[slot_base(L) + i];
        if (nd.flags & kSlotSplit) {
            // a child without polygons is no node (its record holds no counts)
            const LNode& a = D.nodes[slot_base(L + 1) + 2u * i];
            const LNode& b = D.nodes[slot_base(L + 1) + 2u * i + 1u];
            nd.size = 1u + (a.pcount ? a.size : 0u) + (b.pcount ? b.size : 0u);
            nd.nleaf = (a.pcount ? a.nleaf : 0u) + (b.pcount ? b.nleaf : 0u);
            nd.nlp = (a.pcount ? a.nlp : 0u) + (b.pcount ? b.nlp : 0u);
        } else {
            nd.size = 1u;
            nd.nleaf = 1u;
            nd.nlp = nd.pcount;
        }
    }
}

// One block: device-layout indices of depths 0..12 and the level prefix.
// Depths 0..11 are stored breadth first (slot order within a depth); the
// subtrees of depth-12 nodes follow in slot order, each depth first.
__global__ void __launch_bounds__(1024) k_kdb_top_order(KdDev D) {
    __shared__ uint32_t s0[16];
    __shared__ uint8_t exists[1u << kKdTopLevels];
    if (D.st->overflow) return;
    uint32_t prefix = 0;
    for (uint32_t L = 0; L <= kKdTopLevels; L++) {
        if (threadIdx.x == 0) D.st->prefix[L] = prefix;
        if (level_empty(D, L)) continue;
        const uint32_t n = 1u << L;
        for (uint32_t i = threadIdx.x; i < n; i += 1024) exists[i] = 0;
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < D.lvl[L].n_nodes; k += 1024) {
            const uint32_t i = D.nlist[slot_base(L) + k];
            if (is_node(D, L, i)) exists[i] = 1;
        }
        __syncthreads();
        uint32_t running = prefix;
        for (uint32_t c = 0; c < n; c += 1024) {
            const uint32_t i = c + threadIdx.x;
            const bool ex = i < n && exists[i];
            LNode* nd = ex ? &D.nodes[slot_base(L) + i] : nullptr;
            const uint32_t v = ex ? (L == kKdTopLevels ? nd->size : 1u) : 0u;
            uint32_t tot;
            const uint32_t off = block_excl<1024>(v, SumOp(), s0, tot);
            if (ex) nd->dev = running + off;
            running += tot;
        }
        if (L < kKdTopLevels) prefix = running;
    }
    if (threadIdx.x == 0) {
        // a split always leaves a child with polygons, so the deepest depth
        // holding nodes is one below the deepest junction
        uint32_t levels = 1;
        for (uint32_t L = 1; L < kKdLevels; L++)
            if (D.lvl[L].n_polys) levels = L + 1;
        D.st->levels = levels;
        D.st->max_jdepth = levels - 1;
        LNode& r = D.nodes[0];
        r.pre = 0;
        r.lid = 0;
        r.lps = 0;
        D.st->n_nodes = r.size;
        D.st->n_lp = r.nlp;
        D.st->stage_total = D.lvl[kKdLevels].stage_base;
    }
}

// Top down, a depth per launch: the children's reference index, leaf id,
// leaf start and (below the breadth-first top) device index, then the
// node's own records in both layouts (its indices are final: the parent's
// launch, or the top order, wrote them).
__global__ void __launch_bounds__(kThreads) k_kdb_top_down(KdDev D, uint32_t L) {
    if (D.st->overflow) return;
    const uint32_t n_nodes = D.st->n_nodes;
    KDB_FOR_NODES(D, L, i) {
        const LNode& nd = D.nodes[slot_base(L) + i];
        if (!in_range(D, nd.pre, n_nodes) || !in_range(D, nd.dev, n_nodes)) continue;
        if (nd.flags & kSlotSplit) {
            LNode& a = D.nodes[slot_base(L + 1) + 2u * i];
            LNode& b = D.nodes[slot_base(L + 1) + 2u * i + 1u];
            const bool deep = L + 1 > kKdTopLevels;
            const uint32_t a_size = a.pcount ? a.size : 0u;
            if (a.pcount) {
                a.pre = nd.pre + 1u;
                a.lid = nd.lid;
                a.lps = nd.lps;
                if (deep) a.dev = nd.dev + 1u;
            }
            if (b.pcount) {
                b.pre = nd.pre + 1u + a_size;
                b.lid = nd.lid + (a.pcount ? a.nleaf : 0u);
                b.lps = nd.lps + (a.pcount ? a.nlp : 0u);
                if (deep) b.dev = nd.dev + 1u + a_size;
            }
            const uint32_t fl = HULK_NODE_JUNCTION | ((nd.flags >> 4) & 3u);
            const uint32_t pl = __float_as_uint(nd.plane);
            D.ref_nodes[nd.pre] = make_uint4(fl, a.pcount ? a.pre : 0u, b.pcount ? b.pre : 0u, pl);
            D.dev_nodes[nd.dev] = make_uint4(fl, a.pcount ? a.dev : 0u, b.pcount ? b.dev : 0u, pl);
        } else {
            const uint4 leaf = make_uint4(HULK_NODE_LEAF, nd.lps, nd.pcount, nd.lid);
            D.ref_nodes[nd.pre] = leaf;
            D.dev_nodes[nd.dev] = leaf;
        }
    }
}

__global__ void __launch_bounds__(kThreads) k_kdb_leaf_copy(KdDev D) {
    if (D.st->overflow) return;
    const uint32_t n = D.lvl[kKdLevels].stage_base;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t owner = D.stage_owner[k];
        if (!in_range(D, owner, kSlots)) continue;
        const LNode& nd = D.nodes[owner];
        const uint32_t dst = nd.lps + (k - nd.stage);
        if (!in_range(D, k - nd.stage, nd.pcount) || !in_range(D, dst, n)) continue;
        D.lp[dst] = D.stage_poly[k];
    }
}

uint32_t blocks_for(uint64_t n, uint32_t per) {
    return static_cast<uint32_t>(std::max<uint64_t>(1, (n + per - 1) / per));
}

// Blocks for a pass over the nodes of depth L (at most 2^L, counted on the
// device): enough for every node of the shallow depths, a grid-stride loop
// below.
uint32_t node_grid(uint32_t L) { return std::min<uint32_t>(blocks_for(1ull << L, kThreads), 512u); }

}  // namespace

struct KdBuilder {
    void* arena = nullptr;
    size_t arena_cap = 0;
    void* sort_tmp = nullptr;
    size_t sort_cap = 0;
    Status* host_status = nullptr;  // pinned
    KdDev D{};
    uint32_t ns_cap = 0;  // sampled triangles the sort buffers hold
    unsigned long long* keys[2] = {nullptr, nullptr};
    uint32_t* vals[2] = {nullptr, nullptr};
};

KdBuilder* kdb_create() {
    KdBuilder* b = new (std::nothrow) KdBuilder;
    if (!b) return nullptr;
    if (hipHostMalloc(reinterpret_cast<void**>(&b->host_status), sizeof(Status), hipHostMallocDefault) != hipSuccess) {
        delete b;
        return nullptr;
    }
    return b;
}

void kdb_destroy(KdBuilder* b) {
    if (!b) return;
    if (b->arena) (void)hipFree(b->arena);
    if (b->sort_tmp) (void)hipFree(b->sort_tmp);
    if (b->host_status) (void)hipHostFree(b->host_status);
    delete b;
}

const uint4* kdb_dev_nodes(const KdBuilder* b) { return b ? b->D.dev_nodes : nullptr; }
const hulk_kd_node* kdb_ref_nodes(const KdBuilder* b) {
    return b ? reinterpret_cast<const hulk_kd_node*>(b->D.ref_nodes) : nullptr;
}
const uint32_t* kdb_leaf_polys(const KdBuilder* b) { return b ? b->D.lp : nullptr; }

namespace {

// Carve every device array of a build with capacities `cap` (and sort
// buffers for `ns` sampled triangles) out of one allocation.
hipError_t layout(KdBuilder* b, const Caps& cap, uint32_t n_tris, uint32_t ns, bool allocate, size_t* bytes_out) {
    size_t off = 0;
    char* base = static_cast<char*>(b->arena);
    auto take = [&](size_t bytes) -> void* {
        off = (off + 255) & ~size_t(255);
        void* p = allocate ? base + off : nullptr;
        off += bytes;
        return p;
    };
    KdDev& D = b->D;
    const size_t P = cap.polys, C = cap.cands, S = cap.stage;
    D.tbox = static_cast<float2*>(take(sizeof(float2) * 3 * size_t(std::max<uint32_t>(n_tris, 1))));
    D.nodes = static_cast<LNode*>(take(sizeof(LNode) * kSlots));
    D.lvl = static_cast<LevelInfo*>(take(sizeof(LevelInfo) * (kKdLevels + 1)));
    D.st = static_cast<Status*>(take(sizeof(Status)));
    for (int k = 0; k < 2; k++) {
        D.poly[k] = static_cast<uint32_t*>(take(4 * P));
        D.pslot[k] = static_cast<uint32_t*>(take(4 * P));
        D.cand[k] = static_cast<uint2*>(take(8 * C));
        D.cseg[k] = static_cast<uint32_t*>(take(4 * C));
    }
    D.bx = static_cast<uint32_t*>(take(4 * C));
    D.ei = static_cast<uint32_t*>(take(4 * C));
    D.hd = static_cast<uint32_t*>(take(4 * C));
    D.pa = static_cast<uint32_t*>(take(4 * (P + 1)));
    D.pb = static_cast<uint32_t*>(take(4 * (P + 1)));
    D.pl = static_cast<uint32_t*>(take(4 * (P + 1)));
    D.ca = static_cast<uint32_t*>(take(4 * (C + 1)));
    D.cb = static_cast<uint32_t*>(take(4 * (C + 1)));
    D.rp = static_cast<uint8_t*>(take(P));
    D.rc = static_cast<uint8_t*>(take(C));
    D.part_p = static_cast<uint4*>(take(16 * (P / kTile + 2)));
    D.part_c = static_cast<uint4*>(take(16 * (C / kTile + 2)));
    D.nlist = static_cast<uint32_t*>(take(4 * size_t(kSlots)));
    D.stage_poly = static_cast<uint32_t*>(take(4 * S));
    D.stage_owner = static_cast<uint32_t*>(take(4 * S));
    D.ref_nodes = static_cast<uint4*>(take(16 * size_t(kSlots)));
    D.dev_nodes = static_cast<uint4*>(take(16 * size_t(kSlots)));
    D.lp = static_cast<uint32_t*>(take(4 * S));
    for (int k = 0; k < 2; k++) {
        b->keys[k] = static_cast<unsigned long long*>(take(8 * 6 * size_t(ns)));
        b->vals[k] = static_cast<uint32_t*>(take(4 * 6 * size_t(ns)));
    }
    D.cap = cap;
    *bytes_out = off + 256;
    return hipSuccess;
}

}  // namespace

int kdb_build(KdBuilder* b, const float4* verts, uint32_t n_verts, const uint4* tris, uint32_t n_tris,
              const SceneBox& box, uint32_t capacity, hipStream_t s, KdBuildResult* out, std::string& err) {
    (void)n_verts;
    if (!b || !out) {
        err = "null argument";
        return HULK_E_ILLEGAL_INPUT;
    }
    if (n_tris > kPolyMask) {
        err = "more than 2^30 - 1 triangles";
        return HULK_E_ILLEGAL_INPUT;
    }
    const uint32_t ns = (n_tris + kReduction - 1) / kReduction;  // sampled triangles
    uint64_t pcap = capacity ? capacity : std::max<uint64_t>(6ull * n_tris, 4096);
    pcap = std::max<uint64_t>(pcap, std::max<uint64_t>(n_tris, 1));
    KdDev& D = b->D;
    D.verts = verts;
    D.tris = tris;
    D.n_tris = n_tris;
#define KDB_TRY(expr)                                            \
    do {                                                         \
        hipError_t _e = (expr);                                  \
        if (_e != hipSuccess) {                                  \
            err = std::string(#expr) + ": " + hipGetErrorString(_e); \
            return HULK_E_HIP;                                   \
        }                                                        \
    } while (0)
    for (uint32_t retry = 0;; retry++) {
        const uint64_t ccap = std::max<uint64_t>(6ull * ns, 6 * ((pcap + 4) / 5) + 64);
        if (pcap >= (1ull << 31) || ccap >= (1ull << 31)) {
            err = "device KD build: more than 2^31 references in one depth";
            return HULK_E_ALLOC;
        }
        Caps cap{static_cast<uint32_t>(pcap), static_cast<uint32_t>(ccap), static_cast<uint32_t>(pcap)};
        size_t bytes = 0;
        layout(b, cap, n_tris, std::max<uint32_t>(ns, 1), false, &bytes);
        if (bytes > b->arena_cap) {
            if (b->arena) {
                KDB_TRY(hipFree(b->arena));
                b->arena = nullptr;
                b->arena_cap = 0;
            }
            KDB_TRY(hipMalloc(&b->arena, bytes));
            b->arena_cap = bytes;
        }
        layout(b, cap, n_tris, std::max<uint32_t>(ns, 1), true, &bytes);
        D.verts = verts;
        D.tris = tris;
        D.n_tris = n_tris;

        hipLaunchKernelGGL(k_kdb_init, dim3(1), dim3(64), 0, s, D, box, n_tris, ns);
        if (n_tris) {
            hipLaunchKernelGGL(k_kdb_root, dim3(blocks_for(n_tris, kThreads)), dim3(kThreads), 0, s, D, n_tris, ns,
                               b->keys[0], b->vals[0]);
            // one sort of the three axes: 34 key bits (axis, ordered value)
            size_t need = 0;
            KDB_TRY(rocprim::radix_sort_pairs(nullptr, need, b->keys[0], b->keys[1], b->vals[0], b->vals[1],
                                              6u * ns, 0, 34, s));
            if (need > b->sort_cap) {
                if (b->sort_tmp) KDB_TRY(hipFree(b->sort_tmp));
                b->sort_tmp = nullptr;
                b->sort_cap = 0;
                KDB_TRY(hipMalloc(&b->sort_tmp, need));
                b->sort_cap = need;
            }
            size_t sz = b->sort_cap;
            KDB_TRY(rocprim::radix_sort_pairs(b->sort_tmp, sz, b->keys[0], b->keys[1], b->vals[0], b->vals[1], 6u * ns,
                                              0, 34, s));
            hipLaunchKernelGGL(k_kdb_root_cands, dim3(blocks_for(6ull * ns, kThreads)), dim3(kThreads), 0, s, D, ns,
                               b->keys[1], b->vals[1]);
        }
        const uint32_t grid_c = std::min<uint32_t>(kMaxGrid, blocks_for(cap.cands, kTile));
        const uint32_t grid_pc = std::min<uint32_t>(kMaxGrid, blocks_for(uint64_t(cap.polys) + cap.cands, kTile));
        for (uint32_t L = 0; L < kKdLevels; L++) {
            hipLaunchKernelGGL(k_kdb_group_reduce, dim3(grid_c), dim3(kThreads), 0, s, D, L);
            hipLaunchKernelGGL(k_kdb_scan_parts, dim3(1), dim3(1024), 0, s, D, L, 0u);
            hipLaunchKernelGGL(k_kdb_group_apply, dim3(grid_c), dim3(kThreads), 0, s, D, L);
            hipLaunchKernelGGL(k_kdb_sah, dim3(L <= kSahLdsDepth ? std::min<uint32_t>(grid_c, 1024) : grid_c),
                               dim3(kThreads), 0, s, D, L);
            hipLaunchKernelGGL(k_kdb_decide, dim3(node_grid(L)), dim3(kThreads), 0, s, D, L);
            hipLaunchKernelGGL(k_kdb_split_reduce, dim3(grid_pc), dim3(kThreads), 0, s, D, L);
            hipLaunchKernelGGL(k_kdb_scan_parts, dim3(2), dim3(1024), 0, s, D, L, 1u);
            hipLaunchKernelGGL(k_kdb_split_apply, dim3(grid_pc), dim3(kThreads), 0, s, D, L);
            hipLaunchKernelGGL(k_kdb_scatter, dim3(grid_pc * 4), dim3(kThreads), 0, s, D, L);
        }
        for (uint32_t L = kKdLevels; L-- > 0;)
            hipLaunchKernelGGL(k_kdb_bottom_up, dim3(node_grid(L)), dim3(kThreads), 0, s, D, L);
        hipLaunchKernelGGL(k_kdb_top_order, dim3(1), dim3(1024), 0, s, D);
        for (uint32_t L = 0; L < kKdLevels; L++)
            hipLaunchKernelGGL(k_kdb_top_down, dim3(node_grid(L)), dim3(kThreads), 0, s, D, L);
        hipLaunchKernelGGL(k_kdb_leaf_copy, dim3(std::min<uint32_t>(kMaxGrid * 4, blocks_for(cap.stage, kThreads))),
                           dim3(kThreads), 0, s, D);
        KDB_TRY(hipGetLastError());
        KDB_TRY(hipMemcpyAsync(b->host_status, D.st, sizeof(Status), hipMemcpyDeviceToHost, s));
        KDB_TRY(hipStreamSynchronize(s));
        const Status& st = *b->host_status;
        if (st.bad_input) {
            err = "device KD build: a triangle has a NaN vertex coordinate";
            return HULK_E_ILLEGAL_INPUT;
        }
        if (st.internal) {
            err = "device KD build: an index failed a consistency check";
            return HULK_E_HIP;
        }
        if (st.overflow) {
            pcap *= 2;
            continue;
        }
        out->n_nodes = st.n_nodes;
        out->n_lp = st.n_lp;
        out->max_jdepth = st.max_jdepth;
        for (uint32_t d = 0; d <= kKdTopLevels; d++) out->level_prefix[d] = st.prefix[d];
        out->levels = st.levels;
        out->capacity = cap.polys;
        out->retries = retry;
        if (st.stage_total != st.n_lp) {
            err = "device KD build: leaf staging and subtree counts disagree";
            return HULK_E_HIP;
        }
        return HULK_OK;
    }
#undef KDB_TRY
}

}  // namespace hulk
