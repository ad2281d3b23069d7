// hulk_kernels.hpp - parameter blocks and launchers shared by the kernels
// (hulk_kernels.hip) and the C-ABI host code (hulk_api.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hulk_abi.h"
#include "hulk_device.hpp"

namespace hulk {

constexpr uint32_t kMaxShortStack = 16;
constexpr uint32_t kMaxFullStack = 64;
// Leaf references (64-B triangle records) per tree or leaf pool: records are
// addressed by 32-bit byte offsets.
constexpr uint32_t kMaxLeafRefs = 1u << 26;
constexpr uint32_t kSourceCamera = 0;
constexpr uint32_t kSourceRays = 1;
constexpr uint32_t kSourceFrames = 2;  // batched primary rays of a device frame table
constexpr uint32_t kHitFlagOverflow = 1u << 30;  // full stack deeper than the uploaded tree depth
constexpr uint32_t kHitFlagGuard = 1u << 31;     // iteration guard reached (malformed tree)

// Rope slot: cell box + six ropes (face 2*axis + 0 = min, + 1 = max).
// One face of a rope slot: the rope (the slot across the face), the pushed
// rope (k_rope_push: a node below the rope whose cell holds the whole face)
// and the plane on the face's axis that bounds the push (+-inf: none).
struct RopeFace {
    uint32_t rope;
    uint32_t push;
    float far;
    uint32_t pad;
};

// 128 B, one cache line: the cell, then face f (2 axis + 0: min face, + 1:
// max face), so that a rope exit reads its cell and its faces from one line.
// Written for leaf slots only (the slots traversals read).
struct RopeSlot {
    float lo[3];  // cell min
    float hi[3];  // cell max
    uint32_t pad[2];
    RopeFace face[6];
};
static_assert(sizeof(RopeSlot) == 128, "rope slots are one 128-B line");

// One frame of a batch: rays offset .. offset + W*H - 1 of the batch, pixel
// index y*W + x inside the frame; scene: its entry of the batch's scene
// table (batches over several scenes).
struct FrameDev {
    Camera cam;
    uint32_t W, H, offset, max_bounce;
    uint32_t scene;
    uint32_t w_m, tx_m, ty_m;  // udiv_multiplier(W), udiv_multiplier(W / 8), udiv_multiplier(H / 8)
    float fW, fH;        // W and H as floats
};

// Everything a traversal and its shading read from one scene, for batches
// whose frames come from several scenes (TraceParams fields of the same
// names).
struct SceneDev {
    const uint4* nodes;
    const TriRec* recs;
    const uint4* rope_nodes;
    const RopeSlot* rope_slots;
    const float4* lights;
    const uint8_t* mirror;
    const float4* verts;
    const uint4* tris;
    SceneBox box;
    uint32_t node_bytes, rec_bytes, rope_node_bytes, guard, n_lights, n_tris, stack_depth;
    uint32_t lds_node_count;  // top-tree nodes a block stages in LDS (static launches only)
};

struct TraceParams {
    const uint4* nodes;
    uint32_t node_bytes;       // 16 * node count (range of the node buffer loads)
    uint32_t rope_node_bytes;  // 16 * rope slot count
    uint32_t lds_node_count;
    uint32_t stack_depth;
    const TriRec* recs;
    uint32_t rec_bytes;        // 64 * leaf reference count (range of the record buffer loads)
    const float4* verts;
    const uint4* tris;
    uint32_t n_tris;
    const uint4* rope_nodes;
    const RopeSlot* rope_slots;
    SceneBox box;
    uint32_t guard;
    // ray source
    uint32_t source;
    Camera cam;
    uint32_t W, H;
    uint32_t w_m, tx_m, ty_m;  // camera source: udiv multipliers of W, W / 8 and H / 8 (udiv_multiplier)
    float fW, fH;        // camera source: (float)W, (float)H
    const hulk_ray* rays;
    const FrameDev* frames;
    uint32_t n_frames;
    // frames of one size (lane-refill kernel): rays per strip of 16 tile
    // columns (0: the frames' own order) and the multipliers of it and of
    // n_frames
    uint32_t xs_rps, xs_rps_m, xs_nf_m;
    const uint32_t* n_dev;
    uint32_t n;
    const float* tbound;
    // outputs
    hulk_hit* hits;
    float* t_out;
    const uint32_t* t_slot;
    unsigned long long* counters;  // rays, nodes, leaves, tris, hits, fallbacks
    // persistent mode: 8 counters, one per XCD slice of the chunk space,
    // kWorkCounterStride words apart; `fetch` chunks of 64 rays per atomic
    uint32_t* work_counter;
    uint32_t persistent;
    uint32_t fetch;
    // whole-frame launch (render kernel): shading with point-light shadow
    // rays and mirror generations per ray, packed pixels to rgba[o]
    const float4* lights;  // 2 float4 per light
    uint32_t n_lights;
    const uint8_t* mirror;  // mirror flag per triangle, or null (no bounces)
    uint32_t max_bounce;    // camera source; frames carry their own
    uint32_t* rgba;
    uint32_t leaf_batch;    // host side: picks the kernel instance (1 or 2 records in flight, 3 = one
                            // record with the Moller-Trumbore test)
    uint32_t coop_lanes;    // refill kernel: wave-cooperative leaf tests at <= this many busy lanes
    uint32_t stop_at_w;  // shadow queue: stop once best.t < ray.start.w (the light distance)
    // several scenes (frame source, whole-frame launch): the frame's scene
    // entry replaces the scene fields above for each chunk of 64 rays
    const SceneDev* scenes;
};

constexpr uint32_t kWorkCounterStride = 64;  // 256 B between the per-XCD counters
constexpr uint32_t kWorkCounterWords = 8 * kWorkCounterStride;

// a / d for a divisor d >= 1 shared by every lane: m = udiv_multiplier(d)
// from the host, umulhi(a, m) is at most two below the quotient.
inline uint32_t udiv_multiplier(uint32_t d) { return d ? 0xFFFFFFFFu / d : 0u; }

struct GenParams {
    const hulk_hit* hits;
    uint32_t camera_rays;
    Camera cam;
    uint32_t W, H;
    const FrameDev* frames;  // batch: camera per frame, spawn if max_bounce > gen
    uint32_t n_frames, gen;
    const hulk_ray* rays;
    const uint32_t* pixel_in;
    const float* weight_in;
    const uint32_t* n_dev;
    uint32_t n;
    const float4* point_pos;
    uint32_t n_point;
    hulk_ray* shadow_rays;
    uint32_t* shadow_slot;
    uint32_t* shadow_count;
    const uint8_t* mirror;
    hulk_ray* next_rays;
    uint32_t* pixel_out;
    float* weight_out;
    uint32_t* next_count;
    uint32_t* block_counts;   // 2 per 256-ray block of the count pass
    uint32_t* block_offsets;  // 2 per block, exclusive scan of block_counts
};

struct ShadeParams {
    const hulk_hit* hits;
    uint32_t camera_rays;
    Camera cam;
    uint32_t W, H;
    const FrameDev* frames;
    uint32_t n_frames, gen;
    const hulk_ray* rays;
    const uint32_t* pixel_in;
    const float* weight_in;
    const uint32_t* n_dev;
    uint32_t n;
    const float4* lights;  // 2 float4 per light
    uint32_t n_lights;
    uint32_t n_point;
    const float* shadow_t;
    const uint8_t* mirror;
    uint32_t spawned;
    float* acc;
};

// Out-of-core trace (kd-restart over a paged record pool).  Per pixel:
// state = (tmin, tmax, pending leaf node, -) and the best hit so far, kept
// between launches; queue_in = pixels to resume (null: every pixel, first
// launch); pixels that stop at a non-resident leaf go to queue_out.
struct StreamParams {
    const uint4* nodes;
    const TriRec* pool;
    uint32_t pool_bytes;      // 64 * pool records
    const uint2* leaf_table;  // per leaf id: (pool base, resident)
    uint32_t* request;        // per leaf id: set to 1 by a lane that needs the leaf
    SceneBox box;
    Camera cam;
    uint32_t W, H;
    const uint32_t* queue_in;
    uint32_t n_in;
    uint32_t* queue_out;
    uint32_t* n_out;
    float4* state;
    Best* best;
    hulk_hit* hits;
    uint32_t guard;
};

// leaf_batch: triangle records in flight per lane in the leaf loop (1 or 2)
hipError_t launch_trace(int variant, const TraceParams& P, bool count, bool render, uint32_t leaf_batch,
                        uint32_t grid, uint32_t block, size_t lds, hipStream_t s);
// Whole frames of a batch over several scenes (P.scenes, P.frames): the
// fused frame kernel with each chunk's scene bound from the table.
hipError_t launch_trace_scenes(int variant, const TraceParams& P, bool count, uint32_t leaf_batch, uint32_t grid,
                               uint32_t block, size_t lds, hipStream_t s);
hipError_t launch_build_records(const float4* verts, const uint4* tris, const uint32_t* lp, uint32_t n,
                                TriRec* recs, hipStream_t s);
hipError_t launch_gen_secondary(const GenParams& G, uint32_t n_max, hipStream_t s);
hipError_t launch_shade_accumulate(const ShadeParams& S, uint32_t n_max, hipStream_t s);
hipError_t launch_shade_picture(const ShadeParams& S, uint32_t* out, hipStream_t s);
hipError_t launch_pack(const float* acc, uint32_t n, uint32_t* out, hipStream_t s);
hipError_t launch_fill_f32(float* p, size_t n, float v, hipStream_t s);
hipError_t launch_camera_rays(const Camera& cam, uint32_t W, uint32_t H, const SceneBox& box, hulk_ray* rays,
                              float* texit, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s);
hipError_t launch_selftest_rcp(unsigned long long* out, hipStream_t s);  // out: 4 counters
// whole frames, rope traversal, lanes refilled from the ray queue (work_counter: 8 slices)
// variant: HULK_VARIANT_ROPES or HULK_VARIANT_LDSSTACK; trace_lds: the LDS of
// the traversal (top tree levels, short-stack columns) before the lane fields
size_t refill_lds_bytes(size_t trace_lds, uint32_t block, uint32_t n_lights);
hipError_t launch_render_refill(int variant, const TraceParams& P, bool count, uint32_t grid, uint32_t block,
                                size_t trace_lds, hipStream_t s);
// rope slots (3 n slots: nodes, then two empty-leaf slots per junction)
hipError_t launch_trace_streamed(const StreamParams& S, hipStream_t s);
hipError_t launch_collect_requests(uint32_t* request, uint32_t n, uint32_t* list, uint32_t* n_list, hipStream_t s);
hipError_t launch_build_ropes(const uint4* nodes, uint32_t n, uint32_t* up, uint4* rn, RopeSlot* rs,
                              const SceneBox& box, hipStream_t s);
// Pushed ropes of the slots launch_build_ropes built (k_rope_push).
hipError_t launch_push_ropes(const uint4* rn, const uint32_t* up, uint32_t n, RopeSlot* rs, const SceneBox& box,
                             hipStream_t s);

}  // namespace hulk
