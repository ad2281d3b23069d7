// hulk_kernels.hip - gfx950 traversal, shading and compaction kernels.
//
// One lane = one ray.  Every traversal variant is a restatement of one
// reference kernel (cited per function) and must produce the same nearest
// hit, t and barycentrics as the CPU oracle of that variant, bit for bit.
//
// Layout in HBM (built by hulk_api.hip on upload):
//   nodes  : uint4 per node {flags, above|start, below|count, plane|leaf id};
//            re-ordered so that nodes of depth < 12 come first in
//            breadth-first order (the LDS top tree is a prefix), the rest
//            in depth-first order per subtree.  Root = 0, empty child = 0.
//   recs   : TriRec (64 B) per leaf reference, leaf-contiguous.
//   ropes  : rope node (16 B) + RopeSlot (64 B) per rope slot, built on the
//            device (3 slots per node: the node and two empty-leaf slots).
// Work distribution: chunks of 64 rays per wave, either statically mapped
// with an XCD-aware block order or fetched from 8 per-XCD slices by
// persistent waves; the refill kernel fetches rays per idle lane.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hulk_abi.h"
#include "hulk_device.hpp"
#include "hulk_kernels.hpp"

namespace hulk {

// ------------------------------------------------------------ utilities --

// A 16-B node from HBM / L2 by one raw buffer load (word 3 of the
// resource: 32-bit data format, the gfx9 value).  One instruction for the
// whole node: a plain load of the select between the LDS and the global
// copy compiles to flat loads, with the plane fetched by a second, dependent
// load after the junction test.  The buffer's range check (num_bytes)
// returns zeros past the end: flags 0, neither junction nor leaf.
__device__ __forceinline__ uint4 load_node(const uint4* g, uint32_t num_bytes, uint32_t i) {
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint4*>(g), (short)0, (int)num_bytes, 0x00020000);
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, i * 16u, 0, 0);
    return make_uint4(v[0], v[1], v[2], v[3]);
}

__device__ __forceinline__ float4 load_f4(__amdgpu_buffer_rsrc_t rs, uint32_t off) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
    return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]), __uint_as_float(v[3]));
}

__device__ __forceinline__ uint4 fetch_node(const TraceParams& P, const uint4* lds, uint32_t m, uint32_t i) {
    if (i < m) return lds[i];
    return load_node(P.nodes, P.node_bytes, i);
}

__device__ __forceinline__ void junction(const uint4& n, const Ray& r, float& split, uint32_t& first,
                                         uint32_t& second) {
    const uint32_t axis = n.x & 3u;
    const float plane = __uint_as_float(n.w);
    const float ro = lane(r.o, axis), iv = lane(r.inv, axis), rd = lane(r.d, axis);
    split = (plane - ro) * iv;
    // below child first when ro < plane, or ro == plane and rd <= 0
    // (bitwise, so that the lane masks combine without a branch)
    const bool below = (ro < plane) | ((ro == plane) & (rd <= 0.0f));
    first = below ? n.z : n.y;
    second = below ? n.y : n.z;
}

struct Cnt {
    uint32_t nodes, leaves, tris;
};

// Frame of global ray / pixel index i in a batch (offsets ascending).
__device__ __forceinline__ uint32_t find_frame(const FrameDev* F, uint32_t n, uint32_t i) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (F[mid].offset <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Ray i of a camera launch -> pixel (x, y).  Images whose sides are
// multiples of 8 are walked in 8x8 tiles, so each wave traces one compact
// tile (coherent rays, as the reference's 8x8 work-groups,
// GLVisitorGPUTrace.cpp:18); other sizes go row by row.
__device__ __forceinline__ void tile_pixel(uint32_t i, uint32_t W, uint32_t H, uint32_t& x, uint32_t& y) {
    if (((W | H) & 7u) == 0) {
        const uint32_t tile = i >> 6, k = i & 63u, tiles_x = W >> 3;
        const uint32_t ty = tile / tiles_x, tx = tile - ty * tiles_x;
        x = (tx << 3) | (k & 7u);
        y = (ty << 3) | (k >> 3);
    } else {
        y = i / W;
        x = i - y * W;
    }
}

__device__ __forceinline__ uint32_t udiv(uint32_t a, uint32_t d, uint32_t m) {
    uint32_t q = __umulhi(a, m);
    while (a - q * d >= d) q++;
    return q;
}

// tile_pixel with the divisions by W and W / 8 done by host multipliers
// (camera launches: W is the same for every ray, and the multipliers stay
// kernel arguments instead of per-wave reciprocals held in registers).
// STRIPS: the tiles in vertical strips of 16 tiles (128 pixels), row after
// row within a strip, strip after strip, where the tile columns fill whole
// strips (ty_m: the multiplier of H / 8).  The rays in flight on one XCD
// (consecutive ray indices) then cover a region about as tall as wide
// instead of a band across the image, so that they share more of the tree
// and the triangle records in that XCD's L2.
template <bool STRIPS = false>
__device__ __forceinline__ void tile_pixel_m(uint32_t i, uint32_t W, uint32_t H, uint32_t w_m, uint32_t tx_m,
                                             uint32_t ty_m, uint32_t& x, uint32_t& y) {
    if (((W | H) & 7u) == 0) {
        // the lane's place in its tile, recomputed per ray: hoisted out of
        // the ray loop it would be one more register held (and spilled)
        // across every traversal
        uint32_t k = i & 63u;
        asm volatile("" : "+v"(k));
        const uint32_t tile = i >> 6, tiles_x = W >> 3;
        uint32_t tx, ty;
        if (STRIPS && (tiles_x & 15u) == 0u) {
            const uint32_t tiles_y = H >> 3;
            const uint32_t strip = udiv(tile >> 4, tiles_y, ty_m);
            const uint32_t r = tile - strip * (tiles_y << 4);
            ty = r >> 4;
            tx = (strip << 4) | (r & 15u);
        } else {
            ty = udiv(tile, tiles_x, tx_m);
            tx = tile - ty * tiles_x;
        }
        x = (tx << 3) | (k & 7u);
        y = (ty << 3) | (k >> 3);
    } else {
        y = udiv(i, W, w_m);
        x = i - y * W;
    }
}

// Ray i of frame fr: pixel index (frame offset + y*W + x) returned, 8x8
// tiles inside the frame when its sides are multiples of 8, the divisions
// by W and W / 8 done by the frame's multipliers.
template <bool STRIPS = false>
__device__ __forceinline__ uint32_t frame_pixel_ray(const FrameDev& fr, uint32_t i, Ray& r) {
    uint32_t x, y;
    tile_pixel_m<STRIPS>(i - fr.offset, fr.W, fr.H, fr.w_m, fr.tx_m, fr.ty_m, x, y);
    cast_ray_f(fr.cam, x, y, fr.fW, fr.fH, r);
    return fr.offset + y * fr.W + x;
}

// Batched primary ray i (f: its frame).  The active lanes take consecutive
// rays, nearly always of one frame: the frame of the first active lane's ray
// is searched once for the wave (wave-uniform, so scalar loads), and the
// lanes search their own frames only when some ray lies outside it.
template <bool STRIPS = false>
__device__ __forceinline__ uint32_t frame_ray_tiled(const FrameDev* F, uint32_t n, uint32_t i, Ray& r,
                                                    uint32_t& f) {
    const uint32_t f0 = find_frame(F, n, __builtin_amdgcn_readfirstlane(i));
    const uint32_t lo = F[f0].offset, hi = f0 + 1u < n ? F[f0 + 1u].offset : 0xFFFFFFFFu;
    if (__ballot(i < lo || i >= hi) == 0ull) {
        f = f0;
        return frame_pixel_ray<STRIPS>(F[f0], i, r);
    }
    f = find_frame(F, n, i);
    return frame_pixel_ray<STRIPS>(F[f], i, r);
}

// Batched primary ray i of frames of one size (P.xs_rps rays per strip of
// 16 tile columns): strip-major across the frames, strip s of every frame
// before strip s + 1, tile rows within a strip.  An XCD's slice of the ray
// indices then holds the same strips of all frames, so its L2 keeps serving
// the geometry those strips see.  The frame of the wave's first active lane
// is read with scalar loads when every lane shares it.
__device__ __forceinline__ uint32_t frame_ray_xstrips(const TraceParams& P, uint32_t i, Ray& r, uint32_t& f) {
    const uint32_t rps = P.xs_rps, nf = P.n_frames;
    const uint32_t sg = udiv(i, rps, P.xs_rps_m);
    const uint32_t strip = udiv(sg, nf, P.xs_nf_m);
    f = sg - strip * nf;
    uint32_t k = i - sg * rps;  // ray within the strip
    const uint32_t t = k >> 6;
    k &= 63u;
    const uint32_t x = (((strip << 4) | (t & 15u)) << 3) | (k & 7u), y = ((t >> 4) << 3) | (k >> 3);
    const uint32_t f0 = __builtin_amdgcn_readfirstlane(f);
    if (__ballot(f != f0) == 0ull) {
        const FrameDev& fr = P.frames[f0];
        cast_ray_f(fr.cam, x, y, fr.fW, fr.fH, r);
        return fr.offset + y * fr.W + x;
    }
    const FrameDev& fr = P.frames[f];
    cast_ray_f(fr.cam, x, y, fr.fW, fr.fH, r);
    return fr.offset + y * fr.W + x;
}

__device__ __forceinline__ uint32_t frame_ray(const FrameDev* F, uint32_t n, uint32_t i, Ray& r) {
    const uint32_t f = find_frame(F, n, i);
    const FrameDev& fr = F[f];
    const uint32_t local = i - fr.offset;
    const uint32_t y = local / fr.W, x = local - y * fr.W;
    cast_ray(fr.cam, x, y, fr.W, fr.H, r);
    return f;
}

constexpr float kNoStop = -__builtin_inff();
constexpr uint32_t kRecBytes = sizeof(TriRec);
static_assert(kRecBytes == 64, "triangle records are 64 B");

// Leaf modes (template parameter LB of the traversal kernels): one record
// in flight, two records in flight, or one record with the Moller-Trumbore
// test instead of the reference's.
constexpr int kLeafMt = 3;

// Leaf loop.  tstop: a shadow ray only needs to know whether its nearest
// hit lies below the light distance; once best.t < tstop that answer is
// fixed (best.t only decreases), so the ray stops.  Primary and mirror rays
// pass kNoStop = -inf, which never stops them: no float is below it, so where
// it is a constant after inlining the compiler drops the compares.
template <bool COUNT, int LB>
__device__ __forceinline__ void test_leaf(const TriRec* __restrict__ recs, uint32_t rec_bytes, const uint4& n,
                                          const Ray& r, Best& best, Cnt& c, float tstop) {
    const uint32_t start = n.y, end = n.y + n.z;
    if (COUNT) {
        c.leaves++;
        c.tris += n.z;
    }
    if (LB == 2) {
        // two records in flight per lane: the loads of a pair are issued
        // together (pays off where records come from L2 / MALL, not L1)
        uint32_t i = start;
        for (; i + 1 < end; i += 2) {
            const TriRec r0 = recs[i], r1 = recs[i + 1];
            if (test_record(r, r0, best) && best.t < tstop) return;
            if (test_record(r, r1, best) && best.t < tstop) return;
        }
        if (i < end) {
            const TriRec r0 = recs[i];
            test_record(r, r0, best);
        }
        return;
    }
    // A record array holds fewer than 2^26 records (the upload, the device
    // build and the leaf pool refuse more), so buffer loads at a 32-bit byte
    // offset reach every record: one add per record, the four 16-B parts at
    // immediate offsets.  Past rec_bytes the loads return zeros (normal 0:
    // no hit).
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<TriRec*>(recs), (short)0, (int)rec_bytes, 0x00020000);
    for (uint32_t off = start * kRecBytes, off_end = end * kRecBytes; off < off_end; off += kRecBytes) {
        TriRec rec;
        rec.a = load_f4(rs, off);
        rec.b = load_f4(rs, off + 16u);
        rec.c = load_f4(rs, off + 32u);
        rec.d = load_f4(rs, off + 48u);
        if (test_any<LB == kLeafMt>(r, rec, best) && best.t < tstop) break;
    }
}

// ------------------------------------------------------- traversals -----

// Full stack: trace_ray of GPUTrace.cl:6-90 (CPU twin KDTreeCPU::
// searchKDTree, KDTreeCPU.cpp:285-361).  The stack lives in LDS, one
// column per lane.  Entries keep (node, tStart) only: the tEnd of entry k
// equals the tStart of entry k-1 (tEnd0 for the bottom entry), because tEnd
// changes only by a push (to the pushed split) or a pop.
template <bool COUNT, int LB>
__device__ void trav_fullstack(const TraceParams& P, const uint4* lds_nodes, uint32_t* stk, uint32_t stride,
                               const Ray& r, float tclip, float tstop, Best& best, Cnt& c, uint32_t& flags) {
    float ts = 0.0f, te = kFltMax;
    if (!ray_box(P.box, r, ts, te)) return;
    if (tclip < te) te = tclip;
    const float te0 = te;
    uint32_t sp = 0, node = 0, guard = 0;
    do {
        if (best.t < ts) break;
        const uint4 n = fetch_node(P, lds_nodes, P.lds_node_count, node);
        if (COUNT) c.nodes++;
        if (n.x & kNodeLeaf) {
            test_leaf<COUNT, LB>(P.recs, P.rec_bytes, n, r, best, c, tstop);
            if (best.t < tstop) break;
            node = 0;
        } else {
            float split;
            uint32_t first, second;
            junction(n, r, split, first, second);
            if (split > te || split <= 0.0f) {
                node = first;
            } else if (split < ts) {
                node = second;
            } else {
                if (second != 0) {
                    if (sp >= P.stack_depth) {  // deeper than the uploaded tree allows
                        flags |= kHitFlagOverflow;
                        break;
                    }
                    stk[(2 * sp) * stride] = second;
                    stk[(2 * sp + 1) * stride] = __float_as_uint(split);
                    sp++;
                }
                node = first;
                te = split;
            }
        }
        if (node == 0 && sp > 0) {
            sp--;
            node = stk[(2 * sp) * stride];
            ts = __uint_as_float(stk[(2 * sp + 1) * stride]);
            te = sp > 0 ? __uint_as_float(stk[(2 * sp - 1) * stride]) : te0;
        }
        if (++guard > P.guard) {
            flags |= kHitFlagGuard;
            break;
        }
    } while (node != 0);
}

// Short stack with restart: trace_ray of GPUTrace_shortStack.cl:6-105 and,
// with the stack in LDS, GPUTrace_LDSshortStack.cl:7-105.  Circular stack of
// `depth` entries that drops the oldest on overflow; an empty stack
// restarts from the root with tmin = tmax.  Frame i of lane L lives at
// stk[(3 i + k) * stride] (k = tStart, tEnd, node).
template <bool COUNT, int LB>
__device__ void trav_shortstack(const TraceParams& P, const uint4* lds_nodes, uint32_t* stk, uint32_t stride,
                                const Ray& r, float tclip, float tstop, Best& best, Cnt& c, uint32_t& flags) {
    float ts = 0.0f, te = kFltMax;
    if (!ray_box(P.box, r, ts, te)) return;
    if (tclip < te) te = tclip;
    const int depth = (int)P.stack_depth;
    // entry idx starts at word off = 3 idx stride: kept premultiplied, so
    // that a push or pop is an add and a compare
    const int step = 3 * (int)stride, wrap = depth * step;
    int size = 0, off = 0;
    uint32_t node = 0, guard = 0;
    float tmax = ts, tmin = ts;
    while (tmax < te) {
        if (size == 0) {
            tmin = tmax;
            tmax = te;
            node = 0;
        } else {
            --size;
            off -= step;
            if (off < 0) off += wrap;
            tmin = __uint_as_float(stk[off]);
            tmax = __uint_as_float(stk[off + stride]);
            node = stk[off + 2 * stride];
        }
        uint4 n = fetch_node(P, lds_nodes, P.lds_node_count, node);
        if (COUNT) c.nodes++;
        while (is_junction(n.x)) {
            float split;
            uint32_t first, second;
            junction(n, r, split, first, second);
            if (split >= tmax || split <= 0.0f) {
                node = first;
            } else if (split <= tmin) {
                node = second;
            } else {
                stk[off] = __float_as_uint(split);
                stk[off + stride] = __float_as_uint(tmax);
                stk[off + 2 * stride] = second;
                off = off + step == wrap ? 0 : off + step;
                size = size + 1 > depth ? depth : size + 1;
                node = first;
                tmax = split;
            }
            n = fetch_node(P, lds_nodes, P.lds_node_count, node);
            if (COUNT) c.nodes++;
            if (node == 0) break;
        }
        if (n.x & kNodeLeaf) {
            test_leaf<COUNT, LB>(P.recs, P.rec_bytes, n, r, best, c, tstop);
            if (best.t < tmax || best.t < tstop) break;
        }
        if (++guard > P.guard) {
            flags |= kHitFlagGuard;
            break;
        }
    }
}

// The same short stack kept in a private array (registers / scratch).
template <bool COUNT, int LB>
__device__ void trav_shortstack_private(const TraceParams& P, const uint4* lds_nodes, const Ray& r, float tclip,
                                        float tstop, Best& best, Cnt& c, uint32_t& flags) {
    float ts = 0.0f, te = kFltMax;
    if (!ray_box(P.box, r, ts, te)) return;
    if (tclip < te) te = tclip;
    const int depth = (int)P.stack_depth;
    float f_ts[kMaxShortStack], f_te[kMaxShortStack];
    uint32_t f_node[kMaxShortStack];
    int size = 0, idx = 0;
    uint32_t node = 0, guard = 0;
    float tmax = ts, tmin = ts;
    while (tmax < te) {
        if (size == 0) {
            tmin = tmax;
            tmax = te;
            node = 0;
        } else {
            --size;
            --idx;
            if (idx == -1) idx += depth;
            tmin = f_ts[idx];
            tmax = f_te[idx];
            node = f_node[idx];
        }
        uint4 n = fetch_node(P, lds_nodes, P.lds_node_count, node);
        if (COUNT) c.nodes++;
        while (is_junction(n.x)) {
            float split;
            uint32_t first, second;
            junction(n, r, split, first, second);
            if (split >= tmax || split <= 0.0f) {
                node = first;
            } else if (split <= tmin) {
                node = second;
            } else {
                f_ts[idx] = split;
                f_te[idx] = tmax;
                f_node[idx] = second;
                idx = idx + 1 == depth ? 0 : idx + 1;
                size = size + 1 > depth ? depth : size + 1;
                node = first;
                tmax = split;
            }
            n = fetch_node(P, lds_nodes, P.lds_node_count, node);
            if (COUNT) c.nodes++;
            if (node == 0) break;
        }
        if (n.x & kNodeLeaf) {
            test_leaf<COUNT, LB>(P.recs, P.rec_bytes, n, r, best, c, tstop);
            if (best.t < tmax || best.t < tstop) break;
        }
        if (++guard > P.guard) {
            flags |= kHitFlagGuard;
            break;
        }
    }
}

// kd-restart: trace_ray of GPUTrace_stackless.cl:6-120 with every leaf
// resident.  ThreadContext starts zeroed, so tmin = tmax = 0 (the box entry
// is not used); each leaf is followed by a restart from the root; there is
// no early exit after a hit.
template <bool COUNT, int LB>
__device__ void trav_restart(const TraceParams& P, const uint4* lds_nodes, const Ray& r, float tclip, float tstop,
                             Best& best, Cnt& c, uint32_t& flags) {
    float ts = 0.0f, te = kFltMax;
    if (!ray_box(P.box, r, ts, te)) return;
    if (tclip < te) te = tclip;
    float tmax = 0.0f, tmin = 0.0f;
    if (best.t < tmax) return;
    uint32_t guard = 0;
    while (tmax < te) {
        tmin = tmax;
        tmax = te;
        uint32_t node = 0;
        uint4 n = fetch_node(P, lds_nodes, P.lds_node_count, 0);
        if (COUNT) c.nodes++;
        while (is_junction(n.x)) {
            float split;
            uint32_t first, second;
            junction(n, r, split, first, second);
            if (split >= tmax || split <= 0.0f) {
                node = first;
            } else if (split <= tmin) {
                node = second;
            } else {
                node = first;
                tmax = split;
            }
            n = fetch_node(P, lds_nodes, P.lds_node_count, node);
            if (COUNT) c.nodes++;
            if (node == 0) break;
        }
        if (n.x & kNodeLeaf) {
            test_leaf<COUNT, LB>(P.recs, P.rec_bytes, n, r, best, c, tstop);
            if (best.t < tstop) break;
        }
        if (++guard > P.guard) {
            flags |= kHitFlagGuard;
            break;
        }
    }
}

// Stackless rope traversal on the unchanged tree (no reference kernel; the
// reference's "stackless" kernel is kd-restart).  Rope nodes: junction
// {flags, above slot, below slot, plane}, leaf {flags, start, count, id};
// empty children are empty leaves with their own cell and ropes.
// Rules: descend from a slot at time t taking first if split <= 0 or
// split > t, else second; test the leaf; leave through the face of minimum
// slab time (lowest axis on ties); stop when best.t <= t_out, t_out >= te,
// or there is no rope; else t = max(t, t_out) and follow the rope.
struct RopeState {
    uint32_t slot, guard;
    float t, te;
};

// Start of a rope traversal from the ray's scene-box interval [ts, te]
// (inbox false: the ray misses the box, and so does the traversal).
__device__ __forceinline__ bool ropes_begin_box(bool inbox, float ts, float te, float tclip, RopeState& s) {
    if (!inbox) return false;
    if (tclip < te) te = tclip;
    s.t = ts;
    s.te = te;
    s.slot = 0;
    s.guard = 0;
    return true;
}

// Start of a rope traversal; false when the ray misses the scene box.
__device__ __forceinline__ bool ropes_begin(const TraceParams& P, const Ray& r, float tclip, RopeState& s) {
    float ts = 0.0f, te = kFltMax;
    const bool inbox = ray_box(P.box, r, ts, te);
    return ropes_begin_box(inbox, ts, te, tclip, s);
}

// Descent of a rope traversal from the current slot to a leaf; returns the
// leaf's rope node (s.slot is its slot).
template <bool COUNT>
__device__ __forceinline__ uint4 ropes_descend(const TraceParams& P, const Ray& r, RopeState& s, Cnt& c) {
    uint4 n = load_node(P.rope_nodes, P.rope_node_bytes, s.slot);
    if (COUNT) c.nodes++;
    while (is_junction(n.x)) {
        float split;
        uint32_t first, second;
        junction(n, r, split, first, second);
        s.slot = (split <= 0.0f || split > s.t) ? first : second;
        n = load_node(P.rope_nodes, P.rope_node_bytes, s.slot);
        if (COUNT) c.nodes++;
    }
    return n;
}

// The exit of the leaf at s.slot from its rope slot: the face the ray
// leaves through (d[a] > 0: max face, d[a] < 0: min face, ties to the lowest
// axis), its slab time and the rope across it.  Independent of the leaf's
// hits, so the lane-refill loop computes it before the leaf is tested.  It
// keeps the leaf's cell, which the start-leaf check of the next shadow or
// mirror ray reads when that ray leaves the same leaf.
struct RopeExit {
    float tout;
    uint32_t nxt;        // kRopeNone when the ray leaves through no face
    uint32_t slot;       // the leaf slot of this exit, and its cell
    float lo[3], hi[3];
};

// v_i of (v_0, v_1, v_2), from values (no address select).
template <typename T>
__device__ __forceinline__ T sel3(int i, T v0, T v1, T v2) {
    return i == 0 ? v0 : (i == 1 ? v1 : v2);
}

// Cell of rope slot s (words 0-5 of its RopeSlot).
__device__ __forceinline__ void slot_cell(const RopeSlot* slots, uint32_t s, float lo[3], float hi[3]) {
    const uint4* p = reinterpret_cast<const uint4*>(slots + s);
    const uint4 a = p[0];
    const uint2 b = reinterpret_cast<const uint2*>(p + 1)[0];
    lo[0] = __uint_as_float(a.x);
    lo[1] = __uint_as_float(a.y);
    lo[2] = __uint_as_float(a.z);
    hi[0] = __uint_as_float(a.w);
    hi[1] = __uint_as_float(b.x);
    hi[2] = __uint_as_float(b.y);
}

// Bit 31 of RopeState::guard: the ray may take pushed ropes.  |o| and |inv|
// within [2^-64, FLT_MAX] and [2^-32, FLT_MAX] (every comparison with NaN is
// false): then split(v) = (v - o) * inv is finite and nonzero for every
// plane v != o, which the proof of the pushed ropes (rope_exit_of) needs.
constexpr uint32_t kRopePushOk = 0x80000000u;
__device__ __forceinline__ uint32_t rope_push_ok(const Ray& r) {
    const float o[3] = {r.o.x, r.o.y, r.o.z}, iv[3] = {r.inv.x, r.inv.y, r.inv.z};
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float ao = __builtin_fabsf(o[a]), ai = __builtin_fabsf(iv[a]);
        ok &= (ao >= 0x1p-64f) & (ao <= kFltMax) & (ai >= 0x1p-32f) & (ai <= kFltMax);
    }
    return ok ? kRopePushOk : 0u;
}

// The exit of the leaf at `slot`.  PUSHED (the lane-refill kernel): the
// rope taken is the face's pushed rope when the descent from the plain rope
// at the next time t = max(t_prev, t_out) provably passes through it, else
// the plain rope; both descents end in the same leaf.  The push leaves the
// plain rope target N through junctions split across the face (the child
// that holds the face's extent) and junctions split on the face's axis
// (the child toward the face).  Across the face, on an axis b whose bound
// the ray leaves through after the face (d_b > 0: the max bound), the planes
// pushed through lie beyond that bound, so their split exceeds the leaf's
// slab time on b, and the descent keeps to the face's side when that time
// exceeds t: the exit is strict, t2 > t with t2 the second of the three
// slab times.  On the other side (the bound the ray entered through) the
// leaf's own descent, and every earlier one of the traversal, took that
// side at an earlier t, and split is monotone in the plane, so the planes
// beyond it decide the same way (the rope builder never pushes through a
// plane on the scene box, which no descent decided); kRopePushOk keeps
// these splits nonzero.  On the face's axis the planes pushed through lie
// beyond the face, and the nearest (RopeFace::far) decides for all: its
// split must exceed t.  Sufficient, not necessary: ties take the plain rope.
template <bool PUSHED = false>
__device__ __forceinline__ RopeExit rope_exit_of(const Ray& r, const RopeSlot* slots, uint32_t slot,
                                                 float t_prev = 0.0f, uint32_t push_ok = 0u) {
    float mn[3], mx[3];
    slot_cell(slots, slot, mn, mx);
    const float o[3] = {r.o.x, r.o.y, r.o.z};
    const float d[3] = {r.d.x, r.d.y, r.d.z};
    const float iv[3] = {r.inv.x, r.inv.y, r.inv.z};
    // PUSHED: the face record of each axis the ray may leave through (the
    // direction's sign picks it), read with the cell from the same line
    uint32_t frope[3] = {0u, 0u, 0u}, fpush[3] = {0u, 0u, 0u};
    float ffar[3] = {0.0f, 0.0f, 0.0f};
    if (PUSHED) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const uint4 q = reinterpret_cast<const uint4*>(slots[slot].face)[2 * a + (d[a] > 0.0f ? 1 : 0)];
            frope[a] = q.x;
            fpush[a] = q.y;
            ffar[a] = __uint_as_float(q.z);
        }
    }
    RopeExit e;
    e.tout = kFltMax;
    int face = -1;
    e.slot = slot;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        e.lo[a] = mn[a];
        e.hi[a] = mx[a];
    }
    float ts[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float ta;
        int f;
        ts[a] = __builtin_inff();
        if (d[a] > 0.0f) {
            ta = (mx[a] - o[a]) * iv[a];
            f = 2 * a + 1;
        } else if (d[a] < 0.0f) {
            ta = (mn[a] - o[a]) * iv[a];
            f = 2 * a;
        } else {
            continue;
        }
        ts[a] = ta;
        if (ta < e.tout) {
            e.tout = ta;
            face = f;
        }
    }
    if (!PUSHED) {
        e.nxt = face < 0 ? kRopeNone : slots[slot].face[face].rope;
        return e;
    }
    const int ax = face >> 1;  // face < 0: unused
    const float t = e.tout > t_prev ? e.tout : t_prev;
    const float t2 = __builtin_amdgcn_fmed3f(ts[0], ts[1], ts[2]);
    const float sfar = (sel3(ax, ffar[0], ffar[1], ffar[2]) - sel3(ax, o[0], o[1], o[2])) * sel3(ax, iv[0], iv[1], iv[2]);
    const bool take = (push_ok & kRopePushOk) && t2 > t && sfar > t;
    e.nxt = face < 0 ? kRopeNone
                     : (take ? sel3(ax, fpush[0], fpush[1], fpush[2]) : sel3(ax, frope[0], frope[1], frope[2]));
    return e;
}

// After the leaf of s.slot was tested: follow the rope out of it.  False
// once the traversal is over.
__device__ __forceinline__ bool ropes_exit(const TraceParams& P, RopeState& s, const RopeExit& e, float tstop,
                                           const Best& best, uint32_t& flags) {
    if (best.t < tstop) return false;
    if (best.t <= e.tout || e.tout >= s.te || e.nxt == kRopeNone) return false;
    if (e.tout > s.t) s.t = e.tout;
    s.slot = e.nxt;
    if ((++s.guard & ~kRopePushOk) > P.guard) {
        flags |= kHitFlagGuard;
        return false;
    }
    return true;
}

// One leaf of a rope traversal: descend from the current slot to a leaf,
// test it, and follow the rope out of it.  False once the traversal is over.
template <bool COUNT, int LB>
__device__ __forceinline__ bool ropes_step(const TraceParams& P, const Ray& r, RopeState& s, float tstop, Best& best,
                                           Cnt& c, uint32_t& flags) {
    const uint4 n = ropes_descend<COUNT>(P, r, s, c);
    test_leaf<COUNT, LB>(P.recs, P.rec_bytes, n, r, best, c, tstop);
    if (best.t < tstop) return false;
    return ropes_exit(P, s, rope_exit_of(r, P.rope_slots, s.slot), tstop, best, flags);
}

// The LDS short stack of trav_shortstack split at its leaves, so that a
// lane-refill loop can interleave it with shading: ss_begin enters the
// scene box, ss_descend runs one pass of the outer loop up to its leaf (a
// zero node when the pass ends without one), ss_exit applies the loop's
// tests after the leaf.  The state below is the loop's locals.
struct SsState {
    float tmin, tmax, te;
    int size, off;
    uint32_t guard;
};

__device__ __forceinline__ bool ss_begin_box(bool inbox, float ts, float te, float tclip, SsState& s) {
    if (!inbox) return false;
    if (tclip < te) te = tclip;
    s.tmin = s.tmax = ts;
    s.te = te;
    s.size = s.off = 0;
    s.guard = 0;
    return s.tmax < s.te;
}

__device__ __forceinline__ bool ss_begin(const TraceParams& P, const Ray& r, float tclip, SsState& s) {
    float ts = 0.0f, te = kFltMax;
    const bool inbox = ray_box(P.box, r, ts, te);
    return ss_begin_box(inbox, ts, te, tclip, s);
}

template <bool COUNT>
__device__ __forceinline__ uint4 ss_descend(const TraceParams& P, const uint4* lds_nodes, uint32_t* stk,
                                            uint32_t stride, const Ray& r, SsState& s, Cnt& c) {
    const int step = 3 * (int)stride, wrap = (int)P.stack_depth * step;
    uint32_t node;
    if (s.size == 0) {
        s.tmin = s.tmax;
        s.tmax = s.te;
        node = 0;
    } else {
        --s.size;
        s.off -= step;
        if (s.off < 0) s.off += wrap;
        s.tmin = __uint_as_float(stk[s.off]);
        s.tmax = __uint_as_float(stk[s.off + stride]);
        node = stk[s.off + 2 * stride];
    }
    uint4 n = fetch_node(P, lds_nodes, P.lds_node_count, node);
    if (COUNT) c.nodes++;
    while (is_junction(n.x)) {
        float split;
        uint32_t first, second;
        junction(n, r, split, first, second);
        if (split >= s.tmax || split <= 0.0f) {
            node = first;
        } else if (split <= s.tmin) {
            node = second;
        } else {
            stk[s.off] = __float_as_uint(split);
            stk[s.off + stride] = __float_as_uint(s.tmax);
            stk[s.off + 2 * stride] = second;
            s.off = s.off + step == wrap ? 0 : s.off + step;
            s.size = s.size + 1 > (int)P.stack_depth ? (int)P.stack_depth : s.size + 1;
            node = first;
            s.tmax = split;
        }
        n = fetch_node(P, lds_nodes, P.lds_node_count, node);
        if (COUNT) c.nodes++;
        if (node == 0) break;
    }
    return (n.x & kNodeLeaf) ? n : make_uint4(0u, 0u, 0u, 0u);
}

__device__ __forceinline__ bool ss_exit(const TraceParams& P, SsState& s, const uint4& leaf, float tstop,
                                        const Best& best, uint32_t& flags) {
    if ((leaf.x & kNodeLeaf) && (best.t < s.tmax || best.t < tstop)) return false;
    if (++s.guard > P.guard) {
        flags |= kHitFlagGuard;
        return false;
    }
    return s.tmax < s.te;
}

// Traversal of the refill kernel: ropes (TV = HULK_VARIANT_ROPES) or the
// LDS short stack (TV = HULK_VARIANT_LDSSTACK).
struct TravState {
    RopeState rope;
    RopeExit ex;  // ropes: the exit of the leaf the last descent reached
    SsState ss;
};

// trav_begin from the ray's scene-box interval, computed by the caller.
template <int TV>
__device__ __forceinline__ bool trav_begin_box(bool inbox, float ts, float te, float tclip, TravState& s) {
    if (TV == HULK_VARIANT_LDSSTACK) return ss_begin_box(inbox, ts, te, tclip, s.ss);
    return ropes_begin_box(inbox, ts, te, tclip, s.rope);
}

// Start of the lane-refill kernel's traversal (branch-free box test).
template <int TV>
__device__ __forceinline__ bool trav_begin(const TraceParams& P, const Ray& r, float tclip, TravState& s) {
    float ts = 0.0f, te = kFltMax;
    const bool inbox = ray_box_nb(P.box, r, ts, te);
    return trav_begin_box<TV>(inbox, ts, te, tclip, s);
}

// Rope traversal of a ray whose origin lies in the scene box (entry t 0)
// started at leaf slot `hint` instead of the root when the origin lies
// strictly inside the hint's cell: with t 0 every descent decision takes the
// origin's side of the plane (o != plane), so the root descent ends there.
// The cell comes from the last rope exit when that was the hint's leaf.
__device__ __forceinline__ void ropes_hint(const TraceParams& P, const Ray& r, RopeState& s, uint32_t hint,
                                           const RopeExit& ex) {
    if (s.t != 0.0f || hint == 0u) return;
    float lo[3], hi[3];
    if (hint == ex.slot) {  // the cell the last rope exit read
#pragma unroll
        for (int a = 0; a < 3; a++) {
            lo[a] = ex.lo[a];
            hi[a] = ex.hi[a];
        }
    } else {
        slot_cell(P.rope_slots, hint, lo, hi);
    }
    const bool in = lo[0] < r.o.x && r.o.x < hi[0] && lo[1] < r.o.y && r.o.y < hi[1] && lo[2] < r.o.z &&
                    r.o.z < hi[2];
    if (in) s.slot = hint;
}

template <int TV, bool COUNT>
__device__ __forceinline__ uint4 trav_descend(const TraceParams& P, const uint4* lds_nodes, uint32_t* stk,
                                              uint32_t stride, const Ray& r, TravState& s, Cnt& c) {
    if (TV == HULK_VARIANT_LDSSTACK) return ss_descend<COUNT>(P, lds_nodes, stk, stride, r, s.ss, c);
    const uint4 n = ropes_descend<COUNT>(P, r, s.rope, c);
    s.ex = rope_exit_of<!COUNT>(r, P.rope_slots, s.rope.slot, s.rope.t, s.rope.guard);
    return n;
}

template <int TV>
__device__ __forceinline__ bool trav_exit(const TraceParams& P, const Ray& r, TravState& s, const uint4& leaf,
                                          float tstop, const Best& best, uint32_t& flags) {
    if (TV == HULK_VARIANT_LDSSTACK) return ss_exit(P, s.ss, leaf, tstop, best, flags);
    return ropes_exit(P, s.rope, s.ex, tstop, best, flags);
}

// Inclusive prefix sum over the 64 lanes of a wave with DPP: row shifts by
// 1, 2, 4 and 8 (zero from outside the row of 16), then the last lane of
// row 0 added to row 1 and of row 2 to row 3, then the last lane of row 1 to
// rows 2 and 3.
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true);  // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return v;
}

// Per-wave LDS of the cooperative leaf tests (kCoopWaveWords words): 64
// owner entries of two 16-B words (first record minus first packed index,
// best t before the leaf, ray origin x, y; origin z, direction), the owners'
// 64-bit keys, then one 64-bit word of segment-start bits.
constexpr uint32_t kCoopWaveWords = 11u * 64u;
constexpr uint32_t kCoopKeys = 128u;    // uint4 index of the keys (word 512)
constexpr uint32_t kCoopStarts = 160u;  // uint4 index of the segment-start word (word 640)

// The leaves of the lanes in `want` (a wave-uniform ballot; every lane of
// the wave executes this call) tested by the whole wave, 64 records per
// pass, the records of all owner lanes packed back to back in owner lane
// order: lane k of a pass tests packed record k against its owner's ray and
// that owner's best t before the leaf.  The owner's key keeps its minimum
// of (t bits, packed index) over all passes: the smallest t, and of equal t the lowest leaf position, which is
// the hit the owner's own loop keeps (strict '<' in leaf order: the first of
// equal t wins, GPU_common.cl:62-68), with the same per-test arithmetic
// (test_record).  After a pass in which some lane hit, each owner whose
// minimum lies in that pass takes the hit fields of the lane that tested
// it; a pass without hits ends at its tests.  The records of the next pass
// load while a pass is tested (two record buffers in turn).  A shadow ray (best.t <
// tstop ends it) may test more records after its first hit closer than the
// light, where its own loop would stop: its hit may then be another triangle
// closer than the light, which leaves the shading decision (light distance
// > t) unchanged.  A wave's leaves then cost one memory latency per 64
// records of all of them instead of one per record or two, and the tests
// run on dense lanes; with coop_lanes 64 every leaf step is tested so.
template <bool MT>
__device__ __forceinline__ void coop_leaves(const TriRec* __restrict__ recs, const uint4& n, const Ray& r,
                                            Best& best, uint64_t want, uint32_t lane_id, uint4* otab) {
    const bool own = (want >> lane_id) & 1u;
    const uint32_t cnt = own ? n.z : 0u;
    const uint32_t pre = wave_inclusive_sum(cnt);  // the owners' record counts, scanned in lane order
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)pre, 63);
    const uint32_t first = pre - cnt;  // this owner's first packed index
    const uint32_t my_rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(want >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)want, 0u));
    unsigned long long* keys = reinterpret_cast<unsigned long long*>(otab + kCoopKeys);
    if (own) {
        otab[2u * my_rank] = make_uint4(n.y - first, __float_as_uint(best.t), __float_as_uint(r.o.x),
                                        __float_as_uint(r.o.y));
        otab[2u * my_rank + 1u] = make_uint4(__float_as_uint(r.o.z), __float_as_uint(r.d.x), __float_as_uint(r.d.y),
                                             __float_as_uint(r.d.z));
        keys[my_rank] = ~0ull;
    }
    unsigned long long* starts_word = reinterpret_cast<unsigned long long*>(otab + kCoopStarts);
    // owner rank of packed record base + lane: the owners that began before
    // the pass, plus those whose first index lies in the pass at or below it
    auto owner = [&](uint32_t base) -> uint32_t {
        const uint32_t before = (uint32_t)__builtin_popcountll(__ballot(own && first < base));
        if (lane_id == 0u) *starts_word = 0ull;
        if (own && first >= base && first < base + 64u) atomicOr(starts_word, 1ull << (first - base));
        const unsigned long long starts = *starts_word;
        const unsigned long long upto = lane_id == 63u ? ~0ull : (2ull << lane_id) - 1ull;
        const uint32_t cnt_before = before + (uint32_t)__builtin_popcountll(starts & upto);
        return cnt_before ? cnt_before - 1u : 0u;
    };
    // lanes past the last record load the last one (their owner is the last
    // owner), so that every path consumes the loads
    // The entry's base (first record minus first packed index) may wrap
    // below zero: the record index is its 32-bit sum with the packed index.
    auto fetch = [&](uint32_t rank, uint32_t g, TriRec& rec) {
        const uint32_t idx = otab[2u * rank].x + (g < total ? g : total - 1u);
        rec = recs[idx];
    };
    // the tests of the pass at `base` on its records, loaded one pass ahead
    auto pass = [&](uint32_t base, uint32_t rank, const TriRec& rec) {
        const uint32_t g = base + lane_id;
        Best b;
        const uint4 e0 = otab[2u * rank], e1 = otab[2u * rank + 1u];
        Ray q;
        q.o = v3(__uint_as_float(e0.z), __uint_as_float(e0.w), __uint_as_float(e1.x));
        q.d = v3(__uint_as_float(e1.y), __uint_as_float(e1.z), __uint_as_float(e1.w));
        b.t = __uint_as_float(e0.y);  // the owner's best t before the leaf
        const bool hit = test_any<MT>(q, rec, b) && g < total;
        // every part of the record is consumed on every path (the test may
        // stop before it reads the edges), so that a load in flight of the
        // other buffer is never waited for here
        asm volatile("" ::"v"(rec.a.x), "v"(rec.b.x), "v"(rec.c.x), "v"(rec.d.x));
        // t >= 0 here (or -0): the bits without the sign order like the value
        if (hit) atomicMin(keys + rank, ((unsigned long long)(__float_as_uint(b.t) & 0x7FFFFFFFu) << 32) | g);
        if (__ballot(hit) == 0ull) return;
        // an owner whose minimum now lies in this pass takes the hit fields
        // of the lane of the pass that tested it
        const unsigned long long k = own ? keys[my_rank] : ~0ull;
        const uint32_t w = ((uint32_t)k - base) & 63u;
        const float t = __shfl(b.t, w, 64), s1 = __shfl(b.s1, w, 64), s2 = __shfl(b.s2, w, 64),
                    det = __shfl(b.det, w, 64), nx = __shfl(b.nx, w, 64), ny = __shfl(b.ny, w, 64),
                    nz = __shfl(b.nz, w, 64);
        const uint32_t tri = (uint32_t)__shfl((int)b.tri, w, 64);
        if (k != ~0ull && (uint32_t)k >= base) {
            best.t = t;
            best.s1 = s1;
            best.s2 = s2;
            best.det = det;
            best.tri = tri;
            best.nx = nx;
            best.ny = ny;
            best.nz = nz;
        }
    };
    // two record buffers in turn, so that no copy waits for a load in flight;
    // the next pass's records are loaded unconditionally (past the last
    // pass, the last record again), so that every path has the same loads
    // in flight and the test of a pass waits only for its own
    TriRec ra, rb;
    uint32_t rank_a = owner(0u), rank_b = 0u;
    fetch(rank_a, lane_id, ra);
    for (uint32_t base = 0u;;) {
        rank_b = owner(base + 64u);
        fetch(rank_b, base + 64u + lane_id, rb);
        pass(base, rank_a, ra);
        base += 64u;
        if (base >= total) break;
        rank_a = owner(base + 64u);
        fetch(rank_a, base + 64u + lane_id, ra);
        pass(base, rank_b, rb);
        base += 64u;
        if (base >= total) break;
    }
}

template <bool COUNT, int LB>
__device__ void trav_ropes(const TraceParams& P, const Ray& r, float tclip, float tstop, Best& best, Cnt& c,
                           uint32_t& flags) {
    RopeState s;
    if (!ropes_begin(P, r, tclip, s)) return;
    while (ropes_step<COUNT, LB>(P, r, s, tstop, best, c, flags)) {
    }
}

template <bool COUNT, int LB>
__device__ void trav_brute(const TraceParams& P, const Ray& r, float tstop, Best& best, Cnt& c) {
    for (uint32_t i = 0; i < P.n_tris; i++) {
        const uint4 t = P.tris[i];
        const float4 a = P.verts[t.x], b = P.verts[t.y], cc = P.verts[t.z];
        test_vertices<LB == kLeafMt>(r, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(cc.x, cc.y, cc.z), i, best);
        if (best.t < tstop) break;
    }
    if (COUNT) c.tris += P.n_tris;
}

template <int VARIANT, bool COUNT, int LB>
__device__ __forceinline__ void trace_one(const TraceParams& P, const uint4* lds_nodes, uint32_t* stk,
                                          uint32_t stride, const Ray& r, float tclip, float tstop, Best& best,
                                          Cnt& c, uint32_t& flags) {
    if (VARIANT == HULK_VARIANT_FULLSTACK)
        trav_fullstack<COUNT, LB>(P, lds_nodes, stk, stride, r, tclip, tstop, best, c, flags);
    else if (VARIANT == HULK_VARIANT_SHORTSTACK)
        trav_shortstack_private<COUNT, LB>(P, lds_nodes, r, tclip, tstop, best, c, flags);
    else if (VARIANT == HULK_VARIANT_LDSSTACK)
        trav_shortstack<COUNT, LB>(P, lds_nodes, stk, stride, r, tclip, tstop, best, c, flags);
    else if (VARIANT == HULK_VARIANT_RESTART) trav_restart<COUNT, LB>(P, lds_nodes, r, tclip, tstop, best, c, flags);
    else if (VARIANT == HULK_VARIANT_ROPES) trav_ropes<COUNT, LB>(P, r, tclip, tstop, best, c, flags);
    else trav_brute<COUNT, LB>(P, r, tstop, best, c);
}

// ---------------------------------------------------------- trace kernel --

// FrameBuffer::replacePixel (FrameBuffer.hpp:196-217): min(1, c) * 255
// rounded to nearest even, packed R | G << 8 | B << 16.
__device__ __forceinline__ uint32_t to_u8(float c) {
    c = c > 1.0f ? 1.0f : c;
    c = c > 0.0f ? c : 0.0f;
    return (uint32_t)__builtin_rintf(c * 255.0f);
}

// Lambert term of one light at hit point Pp with normal nn (fused
// semantics, GPUTrace.cl:122-160); for a point light, L and its distance
// are returned for the shadow ray.
__device__ __forceinline__ float light_intensity(const float4& col, const float4& dat, const V3& Pp, const V3& nn,
                                                 V3& L, float& dist) {
    float intensity = 0.0f;
    if (col.w == 0.0f) intensity = dot(nn, v3(dat.x, dat.y, dat.z));
    if (col.w == 1.0f) {
        L = sub(v3(dat.x, dat.y, dat.z), Pp);
        const float dd = dot(L, L);
        dist = __builtin_sqrtf(dd);
        L = scale(L, rcp_exact(__builtin_sqrtf(dd)));
        intensity = dot(nn, L);
    }
    return intensity;
}

// Direction and distance of a point light at pos (w = 1 light) from P, with
// light_intensity's operations; true when the surface faces it (n.L > 0),
// the only case that needs a shadow ray.
__device__ __forceinline__ bool point_light_dir(const float4& pos, const V3& Pp, const V3& nn, V3& L, float& dist) {
    L = sub(v3(pos.x, pos.y, pos.z), Pp);
    const float dd = dot(L, L);
    dist = __builtin_sqrtf(dd);
    L = scale(L, rcp_exact(__builtin_sqrtf(dd)));
    return dot(nn, L) > 0.0f;
}

// Whole-frame path of one ray (render kernel): the same values, in the same
// order, as the queued pipeline (k_gen_secondary, the shadow trace,
// k_shade_accumulate and the mirror trace per generation, then k_pack),
// with every intermediate kept in registers.  Generation g shades its hit:
// c = 0.1 + sum_k color_k * max(I_k, 0), a point light's I halved when its
// shadow ray (origin P + 0.001 L, nearest hit, stopped once t < distance)
// is blocked; no shadow ray where I <= 0 (the term is 0 either way, NaN
// included); acc += c * keep, keep = 0.2 w for a mirror hit that spawns a
// reflection (r = d - 2 (d.n) n, origin P + 0.001 r, weight 0.8 w), else w.
template <int VARIANT, bool COUNT, int LB>
__device__ void render_ray(const TraceParams& P, const uint4* lds_nodes, uint32_t* stk, uint32_t stride, Ray r,
                           Best best, uint32_t max_bounce, Cnt& c, uint64_t& crays, uint64_t& cshadow,
                           uint64_t& chits, float& ar, float& ag, float& ab) {
    float w = 1.0f;
    for (uint32_t g = 0;; g++) {
        if (best.tri == kNoHit) break;
        const V3 Pp = add(r.o, scale(r.d, best.t));
        const V3 nn = v3(best.nx, best.ny, best.nz);
        float cr = 0.1f, cg = 0.1f, cb = 0.1f;
        for (uint32_t k = 0; k < P.n_lights; k++) {
            const float4 col = P.lights[2 * k];
            const float4 dat = P.lights[2 * k + 1];
            V3 L = v3(0.0f, 0.0f, 0.0f);
            float dist = 0.0f;
            float intensity = light_intensity(col, dat, Pp, nn, L, dist);
            // a point light the surface faces away from adds max(I, 0) = 0
            // whether or not it is blocked: no shadow ray
            if (col.w == 1.0f && intensity > 0.0f) {
                Ray sr;
                sr.o = add(Pp, scale(L, 0.001f));
                sr.d = L;
                set_inv(sr);
                Best sb;
                best_init(sb);
                uint32_t sf = 0;
                trace_one<VARIANT, COUNT, LB>(P, lds_nodes, stk, stride, sr, kFltMax, dist, sb, c, sf);
                if (COUNT) {
                    cshadow++;
                    chits += sb.tri != kNoHit ? 1 : 0;
                }
                if (dist > sb.t) intensity *= 0.5f;
            }
            const float m = intensity > 0.0f ? intensity : 0.0f;
            cr += col.x * m;
            cg += col.y * m;
            cb += col.z * m;
        }
        const bool spawn = P.mirror && g < max_bounce && P.mirror[best.tri];
        const float keep = spawn ? w * 0.2f : w;
        ar += cr * keep;
        ag += cg * keep;
        ab += cb * keep;
        if (!spawn) break;
        const V3 dn = normalize(r.d);
        const float kk = 2.0f * dot(dn, nn);
        const V3 rr = sub(dn, scale(nn, kk));
        r.o = add(Pp, scale(rr, 0.001f));
        r.d = rr;
        set_inv(r);
        w = w * 0.8f;
        best_init(best);
        uint32_t mf = 0;
        trace_one<VARIANT, COUNT, LB>(P, lds_nodes, stk, stride, r, kFltMax, kNoStop, best, c, mf);
        if (COUNT) {
            crays++;
            chits += best.tri != kNoHit ? 1 : 0;
        }
    }
}

__device__ __forceinline__ uint64_t wave_sum(uint64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Generic trace kernel.  Ray source: the camera (one ray per pixel) or a
// ray buffer (optionally a compacted queue whose length is read on the
// device).  Output: a full hit record, or only t scattered to t_slot[i].
// RENDER: camera / frame rays only; after the primary hit record each ray
// runs its whole shading path (render_ray) and writes its packed pixel.
// Two records in flight (LB = 2) want about 100 VGPRs; a cap of 5 waves per
// SIMD (96 VGPRs, a few bytes of spill) measured fastest for this kernel.
// The LDS short stack with one record in flight takes 86 VGPRs (5 waves);
// at 6 waves (80 VGPRs, a few spills outside the traversal loops) its
// 10-frame terrain batch runs 0.6-1.4 % faster.
// Scene of the chunk of 64 rays starting at ray i0 of a batch over several
// scenes, bound into P (every lane of the wave loads the same entry, so the
// fields stay scalar).
__device__ __forceinline__ void bind_scene(TraceParams& P, uint32_t i0) {
    const uint32_t f = __builtin_amdgcn_readfirstlane(find_frame(P.frames, P.n_frames, i0));
    const SceneDev& S = P.scenes[__builtin_amdgcn_readfirstlane(P.frames[f].scene)];
    P.nodes = S.nodes;
    P.recs = S.recs;
    P.rope_nodes = S.rope_nodes;
    P.rope_slots = S.rope_slots;
    P.lights = S.lights;
    P.mirror = S.mirror;
    P.verts = S.verts;
    P.tris = S.tris;
    P.box = S.box;
    P.node_bytes = S.node_bytes;
    P.rec_bytes = S.rec_bytes;
    P.rope_node_bytes = S.rope_node_bytes;
    P.guard = S.guard;
    P.n_lights = S.n_lights;
    P.n_tris = S.n_tris;
    P.stack_depth = S.stack_depth;
    P.lds_node_count = S.lds_node_count;
}

// Chunk range owner of a static (non-persistent) launch: XCD-aware order
// (blocks are dispatched round-robin over the 8 XCDs, so consecutive chunks
// go to blocks on the same XCD).  When the block count is not a multiple of
// 8 some mapped ids would overflow, and a queue whose length is only known
// on the device fills its low chunks, which round-robin dispatch already
// spreads over the XCDs: plain block order then.
__device__ __forceinline__ uint32_t static_block(const TraceParams& P) {
    const uint32_t nblocks = gridDim.x, b = blockIdx.x;
    if ((nblocks & 7u) != 0 || P.n_dev) return b;
    const uint32_t per_xcd = (nblocks + 7u) >> 3;
    return (b & 7u) * per_xcd + (b >> 3);
}

template <int VARIANT, bool COUNT, bool RENDER, int LB, bool MULTI = false>
__global__ void __launch_bounds__(1024)
    __attribute__((amdgpu_waves_per_eu(LB == 2 ? 5 : (VARIANT == HULK_VARIANT_LDSSTACK ? 6 : 1))))
k_trace(TraceParams P) {
    extern __shared__ uint4 smem[];
    const uint32_t tid = threadIdx.x;
    const uint32_t block = blockDim.x;
    uint4* lds_nodes = smem;
    uint32_t* stk = reinterpret_cast<uint32_t*>(smem + P.lds_node_count) + tid;
    const uint32_t stride = block;

    // Several scenes with a top tree in LDS (static launches whose blocks
    // each lie in one frame): the block's scene first.  P.lds_node_count is
    // the largest scene's count here, which placed the stack columns above.
    if (MULTI && P.lds_node_count) bind_scene(P, static_block(P) * block);
    // Stage the top tree levels in LDS.
    for (uint32_t i = tid; i < P.lds_node_count; i += block) lds_nodes[i] = P.nodes[i];
    if (P.lds_node_count) __syncthreads();

    const uint32_t n_rays = P.n_dev ? *P.n_dev : P.n;
    const uint32_t wave = tid >> 6, lane_id = tid & 63u;
    const uint32_t waves_per_block = block >> 6;
    uint64_t cn = 0, cl = 0, ct = 0, cfb = 0, chits = 0, crays = 0, cshadow = 0;

    // Static: XCD-aware chunk order (blocks are dispatched round-robin over
    // the 8 XCDs, so consecutive chunks go to blocks on the same XCD).
    // Persistent: the chunk space is cut into 8 slices, one per XCD, each
    // with its own counter; a wave takes P.fetch chunks per atomic from the
    // slice of its block's XCD and moves on to the next slice once its own
    // is drained, so atomics stay spread over 8 addresses and the last
    // chunks are shared by every XCD.
    uint32_t chunk = 0, chunk_end = 0, slice = blockIdx.x & 7u, tries = 0;
    const uint32_t nchunks = (n_rays + 63u) >> 6;
    auto next_chunks = [&]() -> bool {
        while (tries < 8) {
            const uint32_t lo = slice * nchunks / 8u, hi = (slice + 1u) * nchunks / 8u;
            uint32_t got = 0;
            if (lane_id == 0) got = atomicAdd(P.work_counter + slice * kWorkCounterStride, P.fetch);
            got = __shfl(got, 0, 64);
            if (got < hi - lo) {
                chunk = lo + got;
                chunk_end = min(chunk + P.fetch, hi);
                return true;
            }
            slice = (slice + 1u) & 7u;
            tries++;
        }
        return false;
    };
    bool have;
    if (P.persistent) {
        have = next_chunks();
    } else {
        chunk = static_block(P) * waves_per_block + wave;
        have = chunk < nchunks;
    }
    while (have) {
        const uint32_t i = chunk * 64u + lane_id;
        if (MULTI) bind_scene(P, chunk * 64u);  // frames hold multiples of 64 rays
        if (i < n_rays) {
            Ray r;
            uint32_t o = i;  // output / bound index: the pixel index for camera rays
            uint32_t max_bounce = P.max_bounce;
            float tstop = kNoStop;
            if (P.source == kSourceCamera) {
                uint32_t x, y;
                tile_pixel_m(i, P.W, P.H, P.w_m, P.tx_m, P.ty_m, x, y);
                o = y * P.W + x;
                cast_ray_f(P.cam, x, y, P.fW, P.fH, r);
            } else if (P.source == kSourceFrames) {
                uint32_t f;
                o = frame_ray_tiled(P.frames, P.n_frames, i, r, f);
                if (RENDER) max_bounce = P.frames[f].max_bounce;
            } else {
                const hulk_ray ray = P.rays[i];
                r.o = v3(ray.start.x, ray.start.y, ray.start.z);
                r.d = v3(ray.dir.x, ray.dir.y, ray.dir.z);
                // shadow queue: start.w carries the light distance
                if (!RENDER && P.stop_at_w) tstop = ray.start.w;
            }
            set_inv(r);
            Best best;
            best_init(best);
            Cnt c = {0, 0, 0};
            uint32_t flags = 0;
            const float tb = P.tbound ? P.tbound[o] : kFltMax;
            // Bounded pass: box exit clipped to tb * (1 + 2^-10).  When the
            // clip does not shorten the interval the pass is the unbounded
            // traversal; otherwise only a hit below tb is kept and every other
            // ray runs again unbounded, so the result equals the plain one.
            float bts = 0.0f, bte = 0.0f;
            const float tclip = tb * (1.0f + 1.0f / 1024.0f);
            const bool clipped = tb < kFltMax && ray_box(P.box, r, bts, bte) && tclip < bte;
            if (clipped) {
                trace_one<VARIANT, COUNT, LB>(P, lds_nodes, stk, stride, r, tclip, tstop, best, c, flags);
                if (!(best.tri != kNoHit && best.t < tb)) {
                    best_init(best);
                    flags |= HULK_HITF_BOUNDED_FALLBACK;
                    trace_one<VARIANT, COUNT, LB>(P, lds_nodes, stk, stride, r, kFltMax, tstop, best, c, flags);
                    if (COUNT) cfb++;
                }
            } else {
                trace_one<VARIANT, COUNT, LB>(P, lds_nodes, stk, stride, r, kFltMax, tstop, best, c, flags);
            }
            if (P.t_out) {
                P.t_out[P.t_slot ? P.t_slot[i] : o] = best.t;
            } else {
                float u = 0.0f, v = 0.0f;
                if (best.tri != kNoHit) {
                    u = best.s1 / best.det;
                    v = best.s2 / best.det;
                }
                float4* h = reinterpret_cast<float4*>(P.hits + o);
                h[0] = make_float4(best.t, u, v, __uint_as_float(best.tri));
                h[1] = make_float4(best.nx, best.ny, best.nz, __uint_as_float(flags));
            }
            if (COUNT) {
                chits += best.tri != kNoHit ? 1 : 0;
                crays += 1;
            }
            if (RENDER) {
                float ar = 0.0f, ag = 0.0f, ab = 0.0f;
                render_ray<VARIANT, COUNT, LB>(P, lds_nodes, stk, stride, r, best, max_bounce, c, crays, cshadow, chits,
                                           ar, ag, ab);
                P.rgba[o] = to_u8(ar) | (to_u8(ag) << 8) | (to_u8(ab) << 16);
            }
            if (COUNT) {
                cn += c.nodes;
                cl += c.leaves;
                ct += c.tris;
            }
        }
        if (P.persistent) {
            if (++chunk == chunk_end) have = next_chunks();
        } else {
            have = false;
        }
    }
    if (COUNT) {
        cn = wave_sum(cn);
        cl = wave_sum(cl);
        ct = wave_sum(ct);
        cfb = wave_sum(cfb);
        chits = wave_sum(chits);
        crays = wave_sum(crays);
        cshadow = wave_sum(cshadow);
        if (lane_id == 0) {
            // rays with a full hit record (primary / mirror) vs shadow rays
            atomicAdd(reinterpret_cast<unsigned long long*>(&P.counters[P.t_out ? 6 : 0]), (unsigned long long)crays);
            if (RENDER) atomicAdd(reinterpret_cast<unsigned long long*>(&P.counters[6]), (unsigned long long)cshadow);
            atomicAdd(reinterpret_cast<unsigned long long*>(&P.counters[1]), (unsigned long long)cn);
            atomicAdd(reinterpret_cast<unsigned long long*>(&P.counters[2]), (unsigned long long)cl);
            atomicAdd(reinterpret_cast<unsigned long long*>(&P.counters[3]), (unsigned long long)ct);
            atomicAdd(reinterpret_cast<unsigned long long*>(&P.counters[4]), (unsigned long long)chits);
            atomicAdd(reinterpret_cast<unsigned long long*>(&P.counters[5]), (unsigned long long)cfb);
        }
    }
}

// ------------------------------------------------- lane-refill frames --

// Whole frames with the rope traversal, lane by lane: a lane whose ray path
// (primary hit, each point light's shadow ray, mirror generations) is over
// takes the next ray of the frame queue while the other lanes of its wave go
// on, so a wave never waits for its slowest lane.  Every lane runs the same
// arithmetic in the same order as render_ray + k_trace (bounded pass with
// fallback, hit record, shading, pixel), so pixels, hit records and counters
// are identical; only the assignment of rays to lanes differs.  Waves take
// rays from 8 per-XCD slices of the ray range, a count of idle lanes per
// atomic, refilling once at least kRefillMin lanes are idle (28 of 64:
// with the strip ray orders, fewer and larger refills ran the 1M batch 2 %
// faster than 16 at the same config-3 frame time; 8 and 32 slower).
constexpr uint32_t kRefillMin = 28;
// Lights staged in the refill kernel's LDS (2 float4 each) when they fit.
constexpr uint32_t kRefillLdsLights = 64;

enum LanePhase : uint32_t { kLaneIdle = 0, kLanePrimary, kLaneFallback, kLaneShadow, kLaneMirror };

// Path state a lane keeps between traversals lives in per-lane LDS columns
// (field f of thread t at word f * blockDim + t), so that registers hold
// only what the traversal step uses.
enum LaneField : uint32_t {
    kFPx, kFPy, kFPz, kFNx, kFNy, kFNz, kFDx, kFDy, kFDz,  // generation hit point, normal, ray direction
    kFW, kFAr, kFAg, kFAb, kFCr, kFCg, kFCb,              // weight, pixel sums, generation sums
    kFIntensity,                                          // current light's term
    kUO, kURem, kUK, kUFlags,                             // pixel, mirror bounces left (bit 31: the hit surface
                                                          // is a mirror), light, flags
                                                          // (primary hit record; then the rope slot of the
                                                          // generation's hit leaf)
    kLaneFields
};
// per lane: the lane fields and the lane's share of the kCoopWaveWords of
// the cooperative leaf tests
constexpr uint32_t kRefillLdsWords = kLaneFields + kCoopWaveWords / 64u;
static_assert(kRefillLdsWords * 64u * 4u % 16u == 0u && kLaneFields * 64u * 4u % 16u == 0u,
              "the owner table of each wave is 16-B aligned");

// Waves per SIMD: the cooperative instances (no per-lane leaf loop) run 4
// waves per SIMD without spilling (5, at 96 VGPRs, spill and measured
// slower); the others measured fastest uncapped.  A wave's LDS region is
// 8 KiB (32 words per lane).
// LDS: the lights (up to kRefillLdsLights; more are read from memory), the
// top tree levels (LDS short stack), one region per wave (lane fields, owner
// table), then the short-stack columns.  COOP: every
// leaf step is a cooperative one (coop_lanes 64; no per-lane leaf loop in
// the instance).
template <int TV, bool COUNT, int LB, bool COOP = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(COOP ? 4 : 1)))
k_render_refill(TraceParams P) {
    extern __shared__ uint4 smem[];
    const uint32_t tid = threadIdx.x, block = blockDim.x;
    const uint32_t lane_id = tid & 63u;
    float4* lds_lights = reinterpret_cast<float4*>(smem);
    uint4* lds_nodes = smem + (P.n_lights <= kRefillLdsLights ? 2u * P.n_lights : 0u);
    uint32_t* lane_mem = reinterpret_cast<uint32_t*>(lds_nodes + P.lds_node_count);
    const bool ll_lds = P.n_lights <= kRefillLdsLights;
    if (ll_lds) for (uint32_t i = tid; i < 2u * P.n_lights; i += block) lds_lights[i] = P.lights[i];
    if (P.n_lights) __syncthreads();
    auto LL = [&](uint32_t i) -> float4 {
        if (ll_lds) return lds_lights[i];
        return P.lights[i];
    };
    // one region of kRefillLdsWords * 64 words per wave: lane field f of
    // lane l at word 64 f + l, so that every field is one base register and
    // an immediate offset; then the owner table
    uint32_t* wave_mem = lane_mem + (tid >> 6) * (kRefillLdsWords * 64u);
    uint32_t* fields = wave_mem + lane_id;
    uint4* otab = reinterpret_cast<uint4*>(wave_mem + kLaneFields * 64u);
    uint32_t* stk = lane_mem + kRefillLdsWords * block + tid;
    for (uint32_t i = tid; i < P.lds_node_count; i += block) lds_nodes[i] = P.nodes[i];
    if (P.lds_node_count) __syncthreads();
    auto F = [&](uint32_t f) -> float& { return reinterpret_cast<float*>(fields)[f * 64u]; };
    auto U = [&](uint32_t f) -> uint32_t& { return fields[f * 64u]; };
    const uint32_t n_rays = P.n;
    uint64_t cfb = 0, chits = 0, crays = 0, cshadow = 0;

    // work queue: 8 slices, one counter each, a wave starts at its block's XCD
    uint32_t slice = blockIdx.x & 7u, tries = 0;
    bool more = n_rays > 0;

    // traversal state (registers)
    uint32_t phase = kLaneIdle, tfl = 0;
    float tstop = -kFltMax;
    bool clipped = false, active = false;  // active: a traversal is in progress
    Ray tr;
    Best bt;
    TravState st;
    Cnt c = {0, 0, 0};
    best_init(bt);
    st.rope.slot = st.rope.guard = 0;
    st.rope.t = st.rope.te = 0.0f;
    st.ss.tmin = st.ss.tmax = st.ss.te = 0.0f;
    st.ss.size = st.ss.off = 0;
    st.ss.guard = 0;
    tr.o = tr.d = tr.inv = v3(0.0f, 0.0f, 0.0f);

    for (;;) {
        // ---- refill (wave-uniform decision)
        const uint64_t idle = __ballot(phase == kLaneIdle);
        const uint32_t n_idle = (uint32_t)__builtin_popcountll(idle);
        if (more && (n_idle >= kRefillMin || n_idle == 64u)) {
            const uint32_t lo = (uint32_t)((uint64_t)slice * n_rays / 8u);
            const uint32_t hi = (uint32_t)((uint64_t)(slice + 1u) * n_rays / 8u);
            uint32_t got = 0;
            if (lane_id == 0) got = atomicAdd(P.work_counter + slice * kWorkCounterStride, n_idle);
            got = __shfl(got, 0, 64);
            const uint32_t avail = got < hi - lo ? min(n_idle, hi - lo - got) : 0u;
            if (avail < n_idle) {  // this slice is drained: the next fetch tries the next one
                slice = (slice + 1u) & 7u;
                if (++tries == 8u) more = false;
            }
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
            if (phase == kLaneIdle && rank < avail) {
                const uint32_t i = lo + got + rank;
                uint32_t o, mb = P.max_bounce;
                if (P.source == kSourceCamera) {
                    uint32_t x, y;
                    tile_pixel_m<true>(i, P.W, P.H, P.w_m, P.tx_m, P.ty_m, x, y);
                    o = y * P.W + x;
                    cast_ray_f(P.cam, x, y, P.fW, P.fH, tr);
                } else {
                    uint32_t f;
                    o = P.xs_rps ? frame_ray_xstrips(P, i, tr, f)
                                 : frame_ray_tiled<true>(P.frames, P.n_frames, i, tr, f);
                    mb = P.frames[f].max_bounce;
                }
                set_inv(tr);
                best_init(bt);
                U(kUO) = o;
                U(kURem) = mb < 0x7FFFFFFFu ? mb : 0x7FFFFFFFu;
                U(kUFlags) = 0;
                F(kFW) = 1.0f;
                F(kFAr) = F(kFAg) = F(kFAb) = 0.0f;
                const float tb = P.tbound ? P.tbound[o] : kFltMax;
                // one box test for the bounded pass's clip and the traversal
                float bts = 0.0f, bte = kFltMax;
                const bool inbox = ray_box_nb(P.box, tr, bts, bte);
                const float tclip = tb * (1.0f + 1.0f / 1024.0f);
                clipped = tb < kFltMax && inbox && tclip < bte;
                phase = kLanePrimary;
                tstop = -kFltMax;
                tfl = 0;
                active = trav_begin_box<TV>(inbox, bts, bte, clipped ? tclip : kFltMax, st);
                if (TV == HULK_VARIANT_ROPES && !COUNT) st.rope.guard = rope_push_ok(tr);
            }
        }
        if (__ballot(phase != kLaneIdle) == 0) {
            if (!more) break;
            continue;
        }
        // ---- one leaf of this lane's traversal.  Idle lanes stay in step
        // (no early continue), so that the wave can test the leaves of its
        // few busy lanes together when at most P.coop_lanes lanes are at one
        // (the counted instance keeps every test per lane, so its counters
        // equal the other kernels').
        const bool busy = phase != kLaneIdle;
        bool done = false;
        if (COUNT || (!COOP && P.coop_lanes == 0)) {
            if (busy) {
                done = !active;
                if (active) {
                    const uint4 ln = trav_descend<TV, COUNT>(P, lds_nodes, stk, block, tr, st, c);
                    if (ln.x & kNodeLeaf) test_leaf<COUNT, LB>(P.recs, P.rec_bytes, ln, tr, bt, c, tstop);
                    done = !trav_exit<TV>(P, tr, st, ln, tstop, bt, tfl);
                }
            }
        } else {
            const bool walk = busy && active;
            uint4 ln = make_uint4(0u, 0u, 0u, 0u);
            if (walk) ln = trav_descend<TV, COUNT>(P, lds_nodes, stk, block, tr, st, c);
            const uint64_t at_leaf = __ballot(walk && ln.z != 0u);
            if (at_leaf && (COOP || (uint32_t)__builtin_popcountll(at_leaf) <= P.coop_lanes))
                coop_leaves<LB == kLeafMt>(P.recs, ln, tr, bt, at_leaf, lane_id, otab);
            else if (!COOP && walk && (ln.x & kNodeLeaf))
                test_leaf<COUNT, LB>(P.recs, P.rec_bytes, ln, tr, bt, c, tstop);
            if (busy) done = !walk || !trav_exit<TV>(P, tr, st, ln, tstop, bt, tfl);
        }
        // ---- state transitions of lanes whose traversal is over (a new
        // traversal may end at once when it misses the box: loop)
        while (done) {
            done = false;
            // a lane that starts a new traversal of tr (the unbounded re-run
            // of a primary ray, a shadow ray or a mirror ray) begins it after
            // this block, so that the wave runs one box test for all of them
            bool begin = false;
            do {
                bool shade = false;  // start shading the generation's hit in bt
                // clipped: the lane's bound (P.tbound, below FLT_MAX) clipped its pass
                if (phase == kLanePrimary && clipped && !(bt.tri != kNoHit && bt.t < P.tbound[U(kUO)])) {
                    best_init(bt);
                    U(kUFlags) |= tfl | HULK_HITF_BOUNDED_FALLBACK;
                    if (COUNT) cfb++;
                    phase = kLaneFallback;
                    tfl = 0;
                    begin = true;
                    break;
                }
                if (phase == kLanePrimary || phase == kLaneFallback) {
                    const uint32_t flags = U(kUFlags) | tfl;
                    float u = 0.0f, v = 0.0f;
                    if (bt.tri != kNoHit) {
                        u = bt.s1 / bt.det;
                        v = bt.s2 / bt.det;
                    }
                    float4* h = reinterpret_cast<float4*>(P.hits + U(kUO));
                    h[0] = make_float4(bt.t, u, v, __uint_as_float(bt.tri));
                    h[1] = make_float4(bt.nx, bt.ny, bt.nz, __uint_as_float(flags));
                    if (COUNT) {
                        crays++;
                        chits += bt.tri != kNoHit ? 1 : 0;
                    }
                    shade = true;
                } else if (phase == kLaneMirror) {
                    if (COUNT) {
                        crays++;
                        chits += bt.tri != kNoHit ? 1 : 0;
                    }
                    shade = true;
                } else {  // kLaneShadow: point light k of the generation
                    if (COUNT) {
                        cshadow++;
                        chits += bt.tri != kNoHit ? 1 : 0;
                    }
                    float intensity = F(kFIntensity);
                    if (tstop > bt.t) intensity *= 0.5f;  // tstop holds the light distance
                    const float4 col = LL(2 * U(kUK));
                    const float m = intensity > 0.0f ? intensity : 0.0f;
                    F(kFCr) += col.x * m;
                    F(kFCg) += col.y * m;
                    F(kFCb) += col.z * m;
                    U(kUK) += 1;
                }
                if (shade) {
                    if (bt.tri == kNoHit) {  // the path ends: pixel
                        P.rgba[U(kUO)] = to_u8(F(kFAr)) | (to_u8(F(kFAg)) << 8) | (to_u8(F(kFAb)) << 16);
                        phase = kLaneIdle;
                        active = false;
                        break;
                    }
                    const V3 hp = add(tr.o, scale(tr.d, bt.t));
                    F(kFPx) = hp.x;
                    F(kFPy) = hp.y;
                    F(kFPz) = hp.z;
                    F(kFNx) = bt.nx;
                    F(kFNy) = bt.ny;
                    F(kFNz) = bt.nz;
                    F(kFDx) = tr.d.x;
                    F(kFDy) = tr.d.y;
                    F(kFDz) = tr.d.z;
                    const uint32_t rem = U(kURem) & 0x7FFFFFFFu;
                    U(kURem) = rem | ((P.mirror && rem && P.mirror[bt.tri]) ? 0x80000000u : 0u);
                    if (TV == HULK_VARIANT_ROPES) U(kUFlags) = st.rope.slot;
                    F(kFCr) = F(kFCg) = F(kFCb) = 0.1f;
                    U(kUK) = 0;
                }
                // lights from k on: directional ones inline, the next point light as a shadow ray
                const V3 Pp = v3(F(kFPx), F(kFPy), F(kFPz)), nn = v3(F(kFNx), F(kFNy), F(kFNz));
                for (uint32_t k = U(kUK); k < P.n_lights; k++) {
                    const float4 col = LL(2 * k);
                    const float4 dat = LL(2 * k + 1);
                    V3 L = v3(0.0f, 0.0f, 0.0f);
                    float dist = 0.0f;
                    const float intensity = light_intensity(col, dat, Pp, nn, L, dist);
                    if (col.w == 1.0f && intensity > 0.0f) {  // facing away: no shadow ray, term 0
                        U(kUK) = k;
                        F(kFIntensity) = intensity;
                        tr.o = add(Pp, scale(L, 0.001f));
                        tr.d = L;
                        tstop = dist;
                        phase = kLaneShadow;
                        begin = true;
                        break;
                    }
                    const float m = intensity > 0.0f ? intensity : 0.0f;
                    F(kFCr) += col.x * m;
                    F(kFCg) += col.y * m;
                    F(kFCb) += col.z * m;
                }
                if (begin) break;
                // generation complete: accumulate, then a mirror ray or the pixel
                const uint32_t rem = U(kURem);
                const bool spawn = rem > 0x80000000u;  // a mirror, and bounces left
                const float w = F(kFW);
                const float keep = spawn ? w * 0.2f : w;
                F(kFAr) += F(kFCr) * keep;
                F(kFAg) += F(kFCg) * keep;
                F(kFAb) += F(kFCb) * keep;
                if (!spawn) {
                    P.rgba[U(kUO)] = to_u8(F(kFAr)) | (to_u8(F(kFAg)) << 8) | (to_u8(F(kFAb)) << 16);
                    phase = kLaneIdle;
                    active = false;
                    break;
                }
                const V3 dn = normalize(v3(F(kFDx), F(kFDy), F(kFDz)));
                const float kk = 2.0f * dot(dn, nn);
                const V3 rr = sub(dn, scale(nn, kk));
                tr.o = add(Pp, scale(rr, 0.001f));
                tr.d = rr;
                F(kFW) = w * 0.8f;
                U(kURem) = (rem & 0x7FFFFFFFu) - 1u;
                tstop = -kFltMax;
                phase = kLaneMirror;
                begin = true;
            } while (false);
            if (begin) {
                if (phase != kLaneFallback) set_inv(tr);
                best_init(bt);
                tfl = 0;
                active = trav_begin<TV>(P, tr, kFltMax, st);
                if (TV == HULK_VARIANT_ROPES && !COUNT) st.rope.guard = rope_push_ok(tr);
                // shadow and mirror rays leave the generation's hit point
                if (TV == HULK_VARIANT_ROPES && !COUNT && active && phase != kLaneFallback)
                    ropes_hint(P, tr, st.rope, U(kUFlags), st.ex);
                done = !active;
            }
        }
    }
    if (COUNT) {
        const uint64_t cn = wave_sum((uint64_t)c.nodes), cl = wave_sum((uint64_t)c.leaves),
                       ct = wave_sum((uint64_t)c.tris);
        cfb = wave_sum(cfb);
        chits = wave_sum(chits);
        crays = wave_sum(crays);
        cshadow = wave_sum(cshadow);
        if (lane_id == 0) {
            atomicAdd(reinterpret_cast<unsigned long long*>(&P.counters[0]), (unsigned long long)crays);
            atomicAdd(reinterpret_cast<unsigned long long*>(&P.counters[6]), (unsigned long long)cshadow);
            atomicAdd(reinterpret_cast<unsigned long long*>(&P.counters[1]), (unsigned long long)cn);
            atomicAdd(reinterpret_cast<unsigned long long*>(&P.counters[2]), (unsigned long long)cl);
            atomicAdd(reinterpret_cast<unsigned long long*>(&P.counters[3]), (unsigned long long)ct);
            atomicAdd(reinterpret_cast<unsigned long long*>(&P.counters[4]), (unsigned long long)chits);
            atomicAdd(reinterpret_cast<unsigned long long*>(&P.counters[5]), (unsigned long long)cfb);
        }
    }
}

template <int TV>
static void launch_refill_tv(const TraceParams& P, bool count, uint32_t grid, uint32_t block, size_t lds,
                             hipStream_t s) {
    if (!count && P.coop_lanes >= 64) {  // every leaf step cooperative: the records in flight do not matter
        if (P.leaf_batch == kLeafMt)
            hipLaunchKernelGGL((k_render_refill<TV, false, kLeafMt, true>), dim3(grid), dim3(block), lds, s, P);
        else hipLaunchKernelGGL((k_render_refill<TV, false, 1, true>), dim3(grid), dim3(block), lds, s, P);
    } else if (P.leaf_batch == 2) {
        if (count) hipLaunchKernelGGL((k_render_refill<TV, true, 2>), dim3(grid), dim3(block), lds, s, P);
        else hipLaunchKernelGGL((k_render_refill<TV, false, 2>), dim3(grid), dim3(block), lds, s, P);
    } else if (P.leaf_batch == kLeafMt) {
        if (count) hipLaunchKernelGGL((k_render_refill<TV, true, kLeafMt>), dim3(grid), dim3(block), lds, s, P);
        else hipLaunchKernelGGL((k_render_refill<TV, false, kLeafMt>), dim3(grid), dim3(block), lds, s, P);
    } else {
        if (count) hipLaunchKernelGGL((k_render_refill<TV, true, 1>), dim3(grid), dim3(block), lds, s, P);
        else hipLaunchKernelGGL((k_render_refill<TV, false, 1>), dim3(grid), dim3(block), lds, s, P);
    }
}

size_t refill_lds_bytes(size_t trace_lds, uint32_t block, uint32_t n_lights) {
    return trace_lds + (size_t)kRefillLdsWords * block * 4 + (n_lights <= kRefillLdsLights ? 32u * n_lights : 0u);
}

hipError_t launch_render_refill(int variant, const TraceParams& P, bool count, uint32_t grid, uint32_t block,
                                size_t trace_lds, hipStream_t s) {
    const size_t lds = refill_lds_bytes(trace_lds, block, P.n_lights);
    if (variant == HULK_VARIANT_LDSSTACK) launch_refill_tv<HULK_VARIANT_LDSSTACK>(P, count, grid, block, lds, s);
    else launch_refill_tv<HULK_VARIANT_ROPES>(P, count, grid, block, lds, s);
    return hipGetLastError();
}

// ------------------------------------------------------ record builder --

// Per-leaf-reference records: the normal, edges and Gram terms of
// checkRayToPoly (SharedFunctions.hpp:53-85), same operations and order.
__global__ void k_build_records(const float4* __restrict__ verts, const uint4* __restrict__ tris,
                                const uint32_t* __restrict__ lp, uint32_t n, TriRec* __restrict__ recs) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = lp[i];
    const uint4 t = tris[id];
    const float4 a = verts[t.x], b = verts[t.y], c = verts[t.z];
    const V3 p1 = v3(a.x, a.y, a.z), p2 = v3(b.x, b.y, b.z), p3 = v3(c.x, c.y, c.z);
    const V3 e0 = sub(p2, p1), e1 = sub(p3, p1), e2 = sub(p3, p2);
    const V3 nn = normalize(add(add(cross(e0, e1), cross(e1, e2)), cross(e2, e0)));
    TriRec r;
    r.a = make_float4(nn.x, nn.y, nn.z, __uint_as_float(id));
    r.b = make_float4(p1.x, p1.y, p1.z, dot(e0, e0));
    r.c = make_float4(e0.x, e0.y, e0.z, dot(e0, e1));
    r.d = make_float4(e1.x, e1.y, e1.z, dot(e1, e1));
    recs[i] = r;
}

// ------------------------------------------------- secondary generation --

// Secondary rays of generation g, compacted without atomics in three
// passes: k_gen_secondary<false> counts per 256-ray block (wave ballots),
// k_scan_pairs turns the block counts into offsets (and the queue lengths),
// k_gen_secondary<true> writes each ray at its block offset + the counts of
// the earlier waves of its block + its lane rank.  Queue order is therefore
// ray order, deterministic and coherent.
//
// For every ray of generation g with a hit: one shadow ray per point light
// (origin P + 0.001*L, direction L = normalize(pos - P), start.w = the light
// distance), slot = ray * n_point + k; and, for a mirror triangle while the
// ray's frame allows generation g + 1, the reflected ray
// d' = dn - (2 (dn.n)) n (dn = normalize(d)), origin P + 0.001*d', appended
// to the next generation with weight w*0.8.
template <bool WRITE>
__global__ void __launch_bounds__(256) k_gen_secondary(GenParams G) {
    __shared__ uint32_t wsum[2][4];
    const uint32_t n = G.n_dev ? *G.n_dev : G.n;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t wave = threadIdx.x >> 6, lane_id = threadIdx.x & 63u;
    const bool active = i < n;
    hulk_hit h;
    Ray r;
    uint32_t pixel = 0;
    float w = 1.0f;
    bool may_spawn = true;
    if (active) {
        h = G.hits[i];
        if (G.camera_rays && G.frames) {
            const uint32_t f = frame_ray(G.frames, G.n_frames, i, r);
            pixel = i;
            may_spawn = G.frames[f].max_bounce > G.gen;
        } else if (G.camera_rays) {
            const uint32_t y = i / G.W, x = i - y * G.W;
            cast_ray(G.cam, x, y, G.W, G.H, r);
            pixel = i;
        } else {
            const hulk_ray ray = G.rays[i];
            r.o = v3(ray.start.x, ray.start.y, ray.start.z);
            r.d = v3(ray.dir.x, ray.dir.y, ray.dir.z);
            pixel = G.pixel_in[i];
            w = G.weight_in[i];
            if (G.frames) may_spawn = G.frames[find_frame(G.frames, G.n_frames, pixel)].max_bounce > G.gen;
        }
    }
    const bool hit = active && h.tri != kNoHit;
    const bool mirror = hit && G.next_rays && may_spawn && G.mirror && G.mirror[h.tri];
    const uint64_t mir_mask = __ballot(mirror);
    // shadow rays: point light k of a hit whose surface faces it (I > 0;
    // elsewhere the light adds 0 whether blocked or not); light-major
    // within the wave
    const V3 P = hit ? add(r.o, scale(r.d, h.t)) : v3(0.0f, 0.0f, 0.0f);
    const V3 nn = hit ? v3(h.nx, h.ny, h.nz) : v3(0.0f, 0.0f, 0.0f);
    uint32_t n_shadow = 0;
    for (uint32_t k = 0; k < G.n_point; k++) {
        V3 L;
        float dist;
        const bool lit = hit && point_light_dir(G.point_pos[k], P, nn, L, dist);
        n_shadow += (uint32_t)__builtin_popcountll(__ballot(lit));
    }
    if (lane_id == 0) {
        wsum[0][wave] = n_shadow;
        wsum[1][wave] = (uint32_t)__builtin_popcountll(mir_mask);
    }
    __syncthreads();
    if (!WRITE) {
        if (threadIdx.x == 0) {
            G.block_counts[2 * blockIdx.x] = wsum[0][0] + wsum[0][1] + wsum[0][2] + wsum[0][3];
            G.block_counts[2 * blockIdx.x + 1] = wsum[1][0] + wsum[1][1] + wsum[1][2] + wsum[1][3];
        }
        return;
    }
    // this wave's offsets: the block's scanned offset plus the earlier waves
    uint32_t shadow_base = G.block_offsets[2 * blockIdx.x], mirror_base = G.block_offsets[2 * blockIdx.x + 1];
    for (uint32_t k = 0; k < wave; k++) {
        shadow_base += wsum[0][k];
        mirror_base += wsum[1][k];
    }
    const uint32_t mir_rank =
        __builtin_amdgcn_mbcnt_hi((uint32_t)(mir_mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mir_mask, 0u));
    uint32_t j = shadow_base;
    for (uint32_t k = 0; k < G.n_point; k++) {
        V3 L;
        float dist = 0.0f;
        const bool lit = hit && point_light_dir(G.point_pos[k], P, nn, L, dist);
        const uint64_t m = __ballot(lit);
        if (lit) {
            const uint32_t rank =
                __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            hulk_ray sr;
            const V3 so = add(P, scale(L, 0.001f));
            sr.start = hulk_float4{so.x, so.y, so.z, dist};
            sr.dir = hulk_float4{L.x, L.y, L.z, 0.0f};
            G.shadow_rays[j + rank] = sr;
            G.shadow_slot[j + rank] = i * G.n_point + k;
        }
        j += (uint32_t)__builtin_popcountll(m);
    }
    if (!hit) return;
    if (mirror) {
        const uint32_t j = mirror_base + mir_rank;
        const V3 nn = v3(h.nx, h.ny, h.nz);
        const V3 dn = normalize(r.d);
        const float kk = 2.0f * dot(dn, nn);
        const V3 rr = sub(dn, scale(nn, kk));
        const V3 o = add(P, scale(rr, 0.001f));
        hulk_ray nr;
        nr.start = hulk_float4{o.x, o.y, o.z, 0.0f};
        nr.dir = hulk_float4{rr.x, rr.y, rr.z, 0.0f};
        G.next_rays[j] = nr;
        G.pixel_out[j] = pixel;
        G.weight_out[j] = w * 0.8f;
    }
}

// Exclusive scan of n pairs (shadow, mirror) of block counts by one block:
// each thread sums a contiguous run (unrolled, so its loads are in flight
// together), the block scans the run sums in LDS, then each thread writes
// its run's offsets.  Totals go to the queue counters.
__global__ void __launch_bounds__(1024) k_scan_pairs(const uint2* counts, uint32_t n, uint2* offsets,
                                                     uint32_t* total0, uint32_t* total1) {
    __shared__ uint32_t s0[1024], s1[1024];
    const uint32_t t = threadIdx.x, T = blockDim.x;
    const uint32_t per = (n + T - 1) / T;
    const uint32_t a = min(n, t * per), b = min(n, a + per);
    uint32_t c0 = 0, c1 = 0;
#pragma unroll 8
    for (uint32_t k = a; k < b; k++) {
        const uint2 v = counts[k];
        c0 += v.x;
        c1 += v.y;
    }
    s0[t] = c0;
    s1[t] = c1;
    __syncthreads();
    for (uint32_t off = 1; off < T; off <<= 1) {  // Hillis-Steele inclusive scan
        const uint32_t v0 = t >= off ? s0[t - off] : 0u, v1 = t >= off ? s1[t - off] : 0u;
        __syncthreads();
        s0[t] += v0;
        s1[t] += v1;
        __syncthreads();
    }
    uint32_t o0 = s0[t] - c0, o1 = s1[t] - c1;
#pragma unroll 8
    for (uint32_t k = a; k < b; k++) {
        const uint2 v = counts[k];
        offsets[k] = make_uint2(o0, o1);
        o0 += v.x;
        o1 += v.y;
    }
    if (t == T - 1) {
        if (total0) *total0 = s0[t];
        if (total1) *total1 = s1[t];
    }
}

// ------------------------------------------------------------ shading --

// Lambert shading of generation g (fused semantics, GPUTrace.cl:122-160):
// c = 0.1 + sum_k color_k * max(I_k, 0), I = n.dir (directional) or n.L
// halved when dist > t_shadow (point).  acc[pixel] += c * keep, keep = w*0.2
// for a mirror hit that spawned a reflection, else w.
__global__ void k_shade_accumulate(ShadeParams S) {
    const uint32_t n = S.n_dev ? *S.n_dev : S.n;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const hulk_hit h = S.hits[i];
    if (h.tri == kNoHit) return;
    Ray r;
    uint32_t pixel;
    float w;
    if (S.camera_rays && S.frames) {
        frame_ray(S.frames, S.n_frames, i, r);
        pixel = i;
        w = 1.0f;
    } else if (S.camera_rays) {
        const uint32_t y = i / S.W, x = i - y * S.W;
        cast_ray(S.cam, x, y, S.W, S.H, r);
        pixel = i;
        w = 1.0f;
    } else {
        const hulk_ray ray = S.rays[i];
        r.o = v3(ray.start.x, ray.start.y, ray.start.z);
        r.d = v3(ray.dir.x, ray.dir.y, ray.dir.z);
        pixel = S.pixel_in[i];
        w = S.weight_in[i];
    }
    const V3 P = add(r.o, scale(r.d, h.t));
    const V3 nn = v3(h.nx, h.ny, h.nz);
    float cr = 0.1f, cg = 0.1f, cb = 0.1f;
    uint32_t kp = 0;
    for (uint32_t k = 0; k < S.n_lights; k++) {
        const float4 col = S.lights[2 * k];
        const float4 dat = S.lights[2 * k + 1];
        V3 L = v3(0.0f, 0.0f, 0.0f);
        float dist = 0.0f;
        float intensity = light_intensity(col, dat, P, nn, L, dist);
        if (col.w == 1.0f) {
            if (intensity > 0.0f) {  // only these had a shadow ray
                const float tsh = S.shadow_t[i * S.n_point + kp];
                if (dist > tsh) intensity *= 0.5f;
            }
            kp++;
        }
        const float m = intensity > 0.0f ? intensity : 0.0f;
        cr += col.x * m;
        cg += col.y * m;
        cb += col.z * m;
    }
    const bool spawned = S.frames ? S.frames[find_frame(S.frames, S.n_frames, pixel)].max_bounce > S.gen
                                  : S.spawned != 0;
    const bool mirror = S.mirror && spawned && S.mirror[h.tri];
    const float keep = mirror ? w * 0.2f : w;
    float* acc = S.acc + 3 * (size_t)pixel;
    acc[0] += cr * keep;
    acc[1] += cg * keep;
    acc[2] += cb * keep;
}

__global__ void k_pack(const float* __restrict__ acc, uint32_t n, uint32_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = to_u8(acc[3 * i]) | (to_u8(acc[3 * i + 1]) << 8) | (to_u8(acc[3 * i + 2]) << 16);
}

__global__ void k_fill_f32(float* p, size_t n, float v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ShadePicture, shadePicture.cl:7-55: base 0.1 for every pixel; a hit
// (t != FLT_MAX and t > 0) adds, per light, n.dir (directional) or
// (n.L)/2 (point, no shadow ray).
__global__ void k_shade_picture(ShadeParams S, uint32_t* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.n) return;
    const hulk_hit h = S.hits[i];
    float cr = 0.1f, cg = 0.1f, cb = 0.1f;
    if (h.t != kFltMax && h.t > 0.0f) {
        Ray r;
        const uint32_t y = i / S.W, x = i - y * S.W;
        cast_ray(S.cam, x, y, S.W, S.H, r);
        const V3 P = add(r.o, scale(r.d, h.t));
        const V3 nn = v3(h.nx, h.ny, h.nz);
        for (uint32_t k = 0; k < S.n_lights; k++) {
            const float4 col = S.lights[2 * k];
            const float4 dat = S.lights[2 * k + 1];
            float intensity = 0.0f;
            if (col.w == 1.0f) {
                V3 L = sub(v3(dat.x, dat.y, dat.z), P);
                L = scale(L, rcp_exact(__builtin_sqrtf(dot(L, L))));
                intensity = dot(nn, L) / 2.0f;
            }
            if (col.w == 0.0f) intensity = dot(nn, v3(dat.x, dat.y, dat.z));
            const float m = intensity > 0.0f ? intensity : 0.0f;
            cr += col.x * m;
            cg += col.y * m;
            cb += col.z * m;
        }
    }
    out[i] = to_u8(cr) | (to_u8(cg) << 8) | (to_u8(cb) << 16);
}

// Primary rays of a camera as ray records, plus the scene-box exit time
// (0 when the ray misses the box): training input of the learned bound.
__global__ void k_camera_rays(Camera cam, uint32_t W, uint32_t H, SceneBox box, hulk_ray* rays, float* texit) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= W * H) return;
    const uint32_t y = i / W, x = i - y * W;
    Ray r;
    cast_ray(cam, x, y, W, H, r);
    hulk_ray q;
    q.start = hulk_float4{r.o.x, r.o.y, r.o.z, 0.0f};
    q.dir = hulk_float4{r.d.x, r.d.y, r.d.z, 0.0f};
    rays[i] = q;
    if (texit) {
        float ts = 0.0f, te = 0.0f;
        if (!ray_box(box, r, ts, te)) te = 0.0f;
        texit[i] = te;
    }
}

// -------------------------------------------------------------- selftest --

// Every finite nonzero float x: out[0] counts those with
// 2^-126 <= |x| < 2^126, out[1] those of them where v_rcp_f32 + one Newton
// step differs from the IEEE 1.0f / x, out[2] the differences of that
// formula outside the range, out[3] the differences of rcp_exact (range
// test + fallback) over every x.
__global__ void k_selftest_rcp(unsigned long long* out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long in_range = 0, d_in = 0, d_out = 0, d_exact = 0;
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < (1ull << 32); b += stride) {
        const uint32_t bits = (uint32_t)b;
        if ((bits & 0x7F800000u) == 0x7F800000u || (bits & 0x7FFFFFFFu) == 0) continue;
        const float x = __uint_as_float(bits);
        const float y0 = __builtin_amdgcn_rcpf(x);
        const uint32_t nr = __float_as_uint(__builtin_fmaf(__builtin_fmaf(-x, y0, 1.0f), y0, y0));
        const uint32_t ieee = __float_as_uint(1.0f / x);
        const float ax = __builtin_fabsf(x);
        if (ax >= 0x1p-126f && ax < 0x1p126f) {
            in_range++;
            d_in += nr != ieee;
        } else {
            d_out += nr != ieee;
        }
        d_exact += __float_as_uint(rcp_exact(x)) != ieee;
    }
    atomicAdd(&out[0], in_range);
    atomicAdd(&out[1], d_in);
    atomicAdd(&out[2], d_out);
    atomicAdd(&out[3], d_exact);
}

hipError_t launch_selftest_rcp(unsigned long long* out, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, reinterpret_cast<uint32_t*>(out), (size_t)8, 0u);
    hipLaunchKernelGGL(k_selftest_rcp, dim3(4096), dim3(256), 0, s, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------- ropes --

// Rope slots of the rope variant, built on the device from the uploaded
// node array.  Slot s < n is node s; the empty above / below child of
// junction i becomes the empty leaf slot n + 2i / n + 2i + 1.  A child's
// cell is its parent's cell with the split plane as its min (above) or max
// (below) on the split axis, and its rope on that face points at its
// sibling; every other face is inherited.  So for a slot the deepest
// ancestor split on each (axis, side) fixes that bound and rope, found by
// walking up the parent links; faces no ancestor fixes keep the scene box
// and kRopeNone.
constexpr uint32_t kLinkRoot = 0xFFFFFFFFu, kLinkUnused = 0xFFFFFFFEu;
constexpr uint32_t kRopePushLevels = 6;

__global__ void k_rope_links(const uint4* __restrict__ nodes, uint32_t n, uint32_t* __restrict__ up,
                             uint4* __restrict__ rn) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint4 nd = nodes[i];
    if (i == 0) up[0] = kLinkRoot;
    if (nd.x & kNodeJunction) {
        const uint32_t a = nd.y ? nd.y : n + 2 * i, b = nd.z ? nd.z : n + 2 * i + 1;
        rn[i] = make_uint4(nd.x, a, b, nd.w);
        up[a] = i;
        up[b] = i | 0x80000000u;
        if (!nd.y) rn[a] = make_uint4(kNodeLeaf, 0, 0, 0);
        if (!nd.z) rn[b] = make_uint4(kNodeLeaf, 0, 0, 0);
    } else {
        rn[i] = nd;
    }
}

__global__ void k_rope_slots(const uint4* __restrict__ rn, const uint32_t* __restrict__ up, uint32_t n_slots,
                             SceneBox box, RopeSlot* __restrict__ rs) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_slots) return;
    uint32_t link = up[s];
    // traversals read the slots of leaves only (rope exits, start leaves)
    if (link == kLinkUnused || is_junction(rn[s].x)) return;
    float lo[3] = {box.mn[0], box.mn[1], box.mn[2]}, hi[3] = {box.mx[0], box.mx[1], box.mx[2]};
    uint32_t rope[6] = {kRopeNone, kRopeNone, kRopeNone, kRopeNone, kRopeNone, kRopeNone};
    uint32_t fixed = 0;  // bit f: face f (2a = min, 2a + 1 = max on axis a) already set by a deeper split
    for (uint32_t step = 0; link != kLinkRoot && step < 2 * kMaxFullStack; step++) {
        const uint32_t p = link & 0x7FFFFFFFu;
        const uint4 pn = rn[p];
        const uint32_t a = pn.x & 3u;
        const uint32_t f = (link >> 31) ? 2 * a + 1 : 2 * a;  // below child: max face; above child: min face
        if (!(fixed & (1u << f))) {
            fixed |= 1u << f;
            const float v = __uint_as_float(pn.w);
            if (f & 1u) {
                hi[a] = v;
                rope[f] = pn.y;  // the above sibling
            } else {
                lo[a] = v;
                rope[f] = pn.z;  // the below sibling
            }
        }
        link = up[p];
    }
    RopeSlot r;
    for (int a = 0; a < 3; a++) {
        r.lo[a] = lo[a];
        r.hi[a] = hi[a];
    }
    r.pad[0] = r.pad[1] = 0;
    for (int k = 0; k < 6; k++) {  // pushed ropes: the ropes themselves until k_rope_push runs
        r.face[k].rope = r.face[k].push = rope[k];
        r.face[k].far = (k & 1) ? __builtin_inff() : -__builtin_inff();
        r.face[k].pad = 0;
    }
    rs[s] = r;
}

// Pushed ropes of the leaf slots: from the rope of face f (2a: the min face,
// 2a + 1: the max face on axis a), down through every junction split across
// a whose plane leaves the leaf's extent on one side (the child holding the
// whole face; never a plane on the scene box, which no descent decides) and
// every junction split on a (the child toward the face; the last such plane,
// the nearest to the face, is RopeFace::far), to the first junction that
// cuts the face.  rope_exit_of takes a push only when the descent from the
// rope provably passes through it.
// One thread per (slot, face): thread i handles face i % 6 of slot i / 6.
// A push goes at most kRopePushLevels levels: the build is a chain of
// dependent node loads, and on the 1M-triangle frames 6 levels keep 98 % of
// the descent loads an unbounded push saves (a CPU walk of the rope rules).
__global__ void k_rope_push(const uint4* __restrict__ rn, const uint32_t* __restrict__ up, uint32_t n_slots,
                            SceneBox box, RopeSlot* __restrict__ rs) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t s = (uint32_t)(i / 6u), f = (uint32_t)(i % 6u);
    if (s >= n_slots || up[s] == kLinkUnused || is_junction(rn[s].x)) return;
    float lo[3], hi[3];
    slot_cell(rs, s, lo, hi);
    {
        const uint32_t a = f >> 1;
        uint32_t d = rs[s].face[f].rope;
        float far = rs[s].face[f].far;
        for (uint32_t step = 0; d != kRopeNone && step < kRopePushLevels; step++) {
            const uint4 m = rn[d];
            if (!is_junction(m.x)) break;
            const uint32_t b = m.x & 3u;
            const float v = __uint_as_float(m.w);
            if (b == a) {  // max face: the ray enters the region above it, whose side toward the face is below
                d = (f & 1u) ? m.z : m.y;
                far = v;
            } else if (v == box.mn[b] || v == box.mx[b]) {
                break;
            } else if (v >= hi[b]) {
                d = m.z;
            } else if (v <= lo[b]) {
                d = m.y;
            } else {
                break;
            }
        }
        rs[s].face[f].push = d;
        rs[s].face[f].far = far;
    }
}

hipError_t launch_build_ropes(const uint4* nodes, uint32_t n, uint32_t* up, uint4* rn, RopeSlot* rs,
                              const SceneBox& box, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const uint32_t n_slots = 3 * n;
    hipLaunchKernelGGL(k_fill_u32, dim3((n_slots + 255) / 256), dim3(256), 0, s, up, (size_t)n_slots, kLinkUnused);
    hipLaunchKernelGGL(k_rope_links, dim3((n + 255) / 256), dim3(256), 0, s, nodes, n, up, rn);
    hipLaunchKernelGGL(k_rope_slots, dim3((n_slots + 255) / 256), dim3(256), 0, s, rn, up, n_slots, box, rs);
    return hipGetLastError();
}

hipError_t launch_push_ropes(const uint4* rn, const uint32_t* up, uint32_t n, RopeSlot* rs, const SceneBox& box,
                             hipStream_t s) {
    if (n == 0) return hipSuccess;
    const uint32_t n_slots = 3 * n;
    hipLaunchKernelGGL(k_rope_push, dim3((6 * n_slots + 255) / 256), dim3(256), 0, s, rn, up, n_slots, box, rs);
    return hipGetLastError();
}

// ---------------------------------------------------------- launchers --

template <int V, bool C>
static hipError_t launch_trace_t(const TraceParams& P, bool render, uint32_t leaf_batch, uint32_t grid,
                                 uint32_t block, size_t lds, hipStream_t s) {
    if (leaf_batch == 2) {
        if (render) hipLaunchKernelGGL((k_trace<V, C, true, 2>), dim3(grid), dim3(block), lds, s, P);
        else hipLaunchKernelGGL((k_trace<V, C, false, 2>), dim3(grid), dim3(block), lds, s, P);
    } else if (leaf_batch == kLeafMt) {
        if (render) hipLaunchKernelGGL((k_trace<V, C, true, kLeafMt>), dim3(grid), dim3(block), lds, s, P);
        else hipLaunchKernelGGL((k_trace<V, C, false, kLeafMt>), dim3(grid), dim3(block), lds, s, P);
    } else {
        if (render) hipLaunchKernelGGL((k_trace<V, C, true, 1>), dim3(grid), dim3(block), lds, s, P);
        else hipLaunchKernelGGL((k_trace<V, C, false, 1>), dim3(grid), dim3(block), lds, s, P);
    }
    return hipGetLastError();
}

template <int V>
static hipError_t launch_trace_v(const TraceParams& P, bool count, bool render, uint32_t leaf_batch, uint32_t grid,
                                 uint32_t block, size_t lds, hipStream_t s) {
    return count ? launch_trace_t<V, true>(P, render, leaf_batch, grid, block, lds, s)
                 : launch_trace_t<V, false>(P, render, leaf_batch, grid, block, lds, s);
}

hipError_t launch_trace(int variant, const TraceParams& P, bool count, bool render, uint32_t leaf_batch,
                        uint32_t grid, uint32_t block, size_t lds, hipStream_t s) {
    switch (variant) {
        case HULK_VARIANT_BRUTEFORCE:
            return launch_trace_v<HULK_VARIANT_BRUTEFORCE>(P, count, render, leaf_batch, grid, block, lds, s);
        case HULK_VARIANT_FULLSTACK:
            return launch_trace_v<HULK_VARIANT_FULLSTACK>(P, count, render, leaf_batch, grid, block, lds, s);
        case HULK_VARIANT_SHORTSTACK:
            return launch_trace_v<HULK_VARIANT_SHORTSTACK>(P, count, render, leaf_batch, grid, block, lds, s);
        case HULK_VARIANT_LDSSTACK:
            return launch_trace_v<HULK_VARIANT_LDSSTACK>(P, count, render, leaf_batch, grid, block, lds, s);
        case HULK_VARIANT_RESTART:
            return launch_trace_v<HULK_VARIANT_RESTART>(P, count, render, leaf_batch, grid, block, lds, s);
        case HULK_VARIANT_ROPES:
            return launch_trace_v<HULK_VARIANT_ROPES>(P, count, render, leaf_batch, grid, block, lds, s);
        default: return hipErrorInvalidValue;
    }
}

template <int V>
static hipError_t launch_trace_scenes_v(const TraceParams& P, bool count, uint32_t leaf_batch, uint32_t grid,
                                        uint32_t block, size_t lds, hipStream_t s) {
    if (count) {
        if (leaf_batch == 2) hipLaunchKernelGGL((k_trace<V, true, true, 2, true>), dim3(grid), dim3(block), lds, s, P);
        else hipLaunchKernelGGL((k_trace<V, true, true, 1, true>), dim3(grid), dim3(block), lds, s, P);
    } else {
        if (leaf_batch == 2) hipLaunchKernelGGL((k_trace<V, false, true, 2, true>), dim3(grid), dim3(block), lds, s, P);
        else hipLaunchKernelGGL((k_trace<V, false, true, 1, true>), dim3(grid), dim3(block), lds, s, P);
    }
    return hipGetLastError();
}

hipError_t launch_trace_scenes(int variant, const TraceParams& P, bool count, uint32_t leaf_batch, uint32_t grid,
                               uint32_t block, size_t lds, hipStream_t s) {
    if (leaf_batch == kLeafMt || !P.scenes || !P.frames) return hipErrorInvalidValue;
    switch (variant) {
        case HULK_VARIANT_BRUTEFORCE:
            return launch_trace_scenes_v<HULK_VARIANT_BRUTEFORCE>(P, count, leaf_batch, grid, block, lds, s);
        case HULK_VARIANT_FULLSTACK:
            return launch_trace_scenes_v<HULK_VARIANT_FULLSTACK>(P, count, leaf_batch, grid, block, lds, s);
        case HULK_VARIANT_SHORTSTACK:
            return launch_trace_scenes_v<HULK_VARIANT_SHORTSTACK>(P, count, leaf_batch, grid, block, lds, s);
        case HULK_VARIANT_LDSSTACK:
            return launch_trace_scenes_v<HULK_VARIANT_LDSSTACK>(P, count, leaf_batch, grid, block, lds, s);
        case HULK_VARIANT_RESTART:
            return launch_trace_scenes_v<HULK_VARIANT_RESTART>(P, count, leaf_batch, grid, block, lds, s);
        case HULK_VARIANT_ROPES:
            return launch_trace_scenes_v<HULK_VARIANT_ROPES>(P, count, leaf_batch, grid, block, lds, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_build_records(const float4* verts, const uint4* tris, const uint32_t* lp, uint32_t n,
                                TriRec* recs, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_build_records, dim3((n + 255) / 256), dim3(256), 0, s, verts, tris, lp, n, recs);
    return hipGetLastError();
}

hipError_t launch_gen_secondary(const GenParams& G, uint32_t n_max, hipStream_t s) {
    if (n_max == 0) return hipSuccess;
    const uint32_t blocks = (n_max + 255) / 256;
    hipLaunchKernelGGL(k_gen_secondary<false>, dim3(blocks), dim3(256), 0, s, G);
    hipLaunchKernelGGL(k_scan_pairs, dim3(1), dim3(1024), 0, s, reinterpret_cast<const uint2*>(G.block_counts),
                       blocks, reinterpret_cast<uint2*>(G.block_offsets), G.shadow_count, G.next_count);
    hipLaunchKernelGGL(k_gen_secondary<true>, dim3(blocks), dim3(256), 0, s, G);
    return hipGetLastError();
}

hipError_t launch_shade_accumulate(const ShadeParams& S, uint32_t n_max, hipStream_t s) {
    if (n_max == 0) return hipSuccess;
    hipLaunchKernelGGL(k_shade_accumulate, dim3((n_max + 255) / 256), dim3(256), 0, s, S);
    return hipGetLastError();
}

hipError_t launch_shade_picture(const ShadeParams& S, uint32_t* out, hipStream_t s) {
    if (S.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_shade_picture, dim3((S.n + 255) / 256), dim3(256), 0, s, S, out);
    return hipGetLastError();
}

hipError_t launch_pack(const float* acc, uint32_t n, uint32_t* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack, dim3((n + 255) / 256), dim3(256), 0, s, acc, n, out);
    return hipGetLastError();
}

hipError_t launch_camera_rays(const Camera& cam, uint32_t W, uint32_t H, const SceneBox& box, hulk_ray* rays,
                              float* texit, hipStream_t s) {
    const uint32_t n = W * H;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_camera_rays, dim3((n + 255) / 256), dim3(256), 0, s, cam, W, H, box, rays, texit);
    return hipGetLastError();
}

hipError_t launch_fill_f32(float* p, size_t n, float v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}
hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

// ------------------------------------------------ out-of-core streaming --

// kd-restart with leaves paged into a record pool (the residency protocol of
// GPUTrace_stackless.cl:27-115 and the relaunch loop of GLVisitorGPUTrace::
// launchTraceIterations, GLVisitorGPUTrace.cpp:412-489).  A lane traces its
// pixel as trav_restart does, but a leaf's records are read from the pool at
// the leaf's base, and only while the leaf is resident.  At a leaf that is
// not, the lane flags the leaf as requested, saves (tmin, tmax, the leaf's
// node) and its best hit so far, and queues its pixel for the next launch,
// which tests that leaf with the same interval and goes on.  The sequence of
// leaf tests is therefore the in-core kernel's, and so is the hit.  (The
// reference's relaunch returns early once its hit lies below tmax; the
// in-core kernel has no such exit, and neither has this one, so the two
// agree bit for bit.)
__global__ void __launch_bounds__(256) k_trace_streamed(StreamParams S) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane_id = threadIdx.x & 63u;
    bool queued = false;
    uint32_t o = 0;
    if (i < S.n_in) {
        uint32_t x, y;
        if (S.queue_in) {
            o = S.queue_in[i];
            y = o / S.W;
            x = o - y * S.W;
        } else {
            tile_pixel(i, S.W, S.H, x, y);
            o = y * S.W + x;
        }
        Ray r;
        cast_ray(S.cam, x, y, S.W, S.H, r);
        set_inv(r);
        Best best;
        float tmin = 0.0f, tmax = 0.0f;
        uint32_t pending = 0, flags = 0;
        if (S.queue_in) {
            best = S.best[o];
            const float4 st = S.state[o];
            tmin = st.x;
            tmax = st.y;
            pending = __float_as_uint(st.z);
        } else {
            best_init(best);
        }
        Cnt c = {0, 0, 0};
        float ts = 0.0f, te = kFltMax;
        if (ray_box(S.box, r, ts, te)) {
            uint32_t guard = 0;
            if (pending) {
                const uint4 n = S.nodes[pending];
                const uint2 lt = S.leaf_table[n.w];
                if (lt.y) {
                    test_leaf<false, 1>(S.pool, S.pool_bytes, make_uint4(n.x, lt.x, n.z, n.w), r, best, c, kNoStop);
                    pending = 0;
                } else {
                    S.request[n.w] = 1u;
                    queued = true;
                }
            }
            while (!queued && tmax < te) {
                tmin = tmax;
                tmax = te;
                uint32_t node = 0;
                uint4 n = S.nodes[0];
                while (is_junction(n.x)) {
                    float split;
                    uint32_t first, second;
                    junction(n, r, split, first, second);
                    if (split >= tmax || split <= 0.0f) {
                        node = first;
                    } else if (split <= tmin) {
                        node = second;
                    } else {
                        node = first;
                        tmax = split;
                    }
                    n = S.nodes[node];
                    if (node == 0) break;
                }
                if (n.x & kNodeLeaf) {
                    const uint2 lt = S.leaf_table[n.w];
                    if (lt.y) {
                        test_leaf<false, 1>(S.pool, S.pool_bytes, make_uint4(n.x, lt.x, n.z, n.w), r, best, c, kNoStop);
                    } else {
                        S.request[n.w] = 1u;
                        pending = node;
                        queued = true;
                    }
                }
                if (++guard > S.guard) {
                    flags |= kHitFlagGuard;
                    break;
                }
            }
        }
        if (queued) {
            S.state[o] = make_float4(tmin, tmax, __uint_as_float(pending), 0.0f);
            S.best[o] = best;
        } else {
            float u = 0.0f, v = 0.0f;
            if (best.tri != kNoHit) {
                u = best.s1 / best.det;
                v = best.s2 / best.det;
            }
            float4* h = reinterpret_cast<float4*>(S.hits + o);
            h[0] = make_float4(best.t, u, v, __uint_as_float(best.tri));
            h[1] = make_float4(best.nx, best.ny, best.nz, __uint_as_float(flags));
        }
    }
    // queue the stopped pixels: one atomic per wave
    const unsigned long long m = __ballot(queued);
    if (m) {
        const int leader = __ffsll(static_cast<long long>(m)) - 1;
        uint32_t base = 0;
        if (static_cast<int>(lane_id) == leader) base = atomicAdd(S.n_out, static_cast<uint32_t>(__popcll(m)));
        base = __shfl(base, leader, 64);
        if (queued) S.queue_out[base + __popcll(m & ((1ull << lane_id) - 1ull))] = o;
    }
}

// Requested leaves (flag set by a lane this launch) -> list; flags cleared.
__global__ void __launch_bounds__(256) k_collect_requests(uint32_t* request, uint32_t n, uint32_t* list,
                                                          uint32_t* n_list) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane_id = threadIdx.x & 63u;
    const bool req = i < n && request[i] != 0u;
    if (req) request[i] = 0u;
    const unsigned long long m = __ballot(req);
    if (m) {
        const int leader = __ffsll(static_cast<long long>(m)) - 1;
        uint32_t base = 0;
        if (static_cast<int>(lane_id) == leader) base = atomicAdd(n_list, static_cast<uint32_t>(__popcll(m)));
        base = __shfl(base, leader, 64);
        if (req) list[base + __popcll(m & ((1ull << lane_id) - 1ull))] = i;
    }
}

hipError_t launch_trace_streamed(const StreamParams& S, hipStream_t s) {
    if (S.n_in == 0) return hipSuccess;
    hipLaunchKernelGGL(k_trace_streamed, dim3((S.n_in + 255) / 256), dim3(256), 0, s, S);
    return hipGetLastError();
}

hipError_t launch_collect_requests(uint32_t* request, uint32_t n, uint32_t* list, uint32_t* n_list, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_collect_requests, dim3((n + 255) / 256), dim3(256), 0, s, request, n, list, n_list);
    return hipGetLastError();
}

}  // namespace hulk
