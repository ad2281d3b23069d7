// hulk_api.hip - C ABI implementation: contexts, uploads, frame pipelines.
//
// Replaces GLVisitorGPUTrace's OpenCL host code (HulkEngine/src/
// GLVisitorGPUTrace.cpp): initializeVisitor (:94-161) -> hulk_create,
// buildTree / buildPerLeafData (:163-294) -> hulk_upload_scene +
// hulk_upload_kdtree, traceScene / launchTraceIterations (:296-489) ->
// hulk_trace_camera + hulk_shade / hulk_render_frame.  Every leaf is
// resident in HBM, so the residency relaunch loop has no counterpart.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "hulk_abi.h"
#include "hulk_kdbuild.hpp"
#include "hulk_kernels.hpp"

using namespace hulk;

// ABI layouts the bindings rely on (SharedStructs.hpp sizes for the first five).
static_assert(sizeof(hulk_float4) == 16 && sizeof(hulk_tri) == 16 && sizeof(hulk_kd_node) == 16, "layout");
static_assert(sizeof(hulk_box) == 32 && sizeof(hulk_camera) == 64 && sizeof(hulk_light) == 32, "layout");
static_assert(sizeof(hulk_ray) == 32 && sizeof(hulk_hit) == 32, "layout");
static_assert(sizeof(hulk_opts) == 48 && sizeof(hulk_tune_entry) == 36 && sizeof(hulk_tune_request) == 96, "layout");

namespace {

constexpr uint32_t kTopLevelsMax = 12;

// Device scratch of the allocation-free frame call (hulk_render_frame_static):
// the persistent waves' work counters, then the primary hit records of
// frames whose caller passes no hit buffer, up to kStaticMaxRays of them.
constexpr uint32_t kStaticMaxRays = 1u << 18;
constexpr size_t kScratchCounterBytes = kWorkCounterWords * 4u;
constexpr size_t kScratchBytes = kScratchCounterBytes + size_t(kStaticMaxRays) * sizeof(hulk_hit);
constexpr int kMaxDevices = 64;

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

}  // namespace

// One copy per GPU, held until the process ends.  Host code never uses this
// name as an address: each device's address comes from hipGetSymbolAddress
// (scratch_address below).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Buffers of one frame pipeline (primary hits, shadow queue, mirror queues,
// accumulation).  The context owns one set; every batch owns its own, so a
// captured graph never sees a buffer that a later call reallocates.
struct Work {
    DevBuf hits, hits2, hits3, shadow_rays, shadow_slot, shadow_t, qrays[2], qpix[2], qw[2], acc, ctr, wq, bcount,
        boffset;
};

struct hulk_ctx {
    hulk_opts opts{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;
    int num_cus = 256;
    size_t max_lds = 65536;

    // scene
    DevBuf verts, tris;
    uint32_t n_verts = 0, n_tris = 0;
    SceneBox box{};
    bool have_scene = false;

    // tree: device layout (nodes), reference layout (ref_nodes), leaf
    // polygon references (lp) and their triangle records (recs)
    DevBuf nodes, ref_nodes, lp, recs;
    uint32_t n_nodes = 0, n_lp = 0, max_jdepth = 0;
    KdBuilder* kdb = nullptr;
    KdBuildResult kdb_last{};
    uint32_t tree_gen = 0;  // bumped whenever the tree or the vertices change
    uint32_t lights_gen = 0;  // bumped whenever the lights change

    // out-of-core streaming: per-leaf extents of the records, their pinned
    // host copy (the backing store) and the device pool / tables
    uint32_t stream_gen = ~0u;
    std::vector<uint2> leaf_extent;  // per leaf id: (first record, count)
    uint32_t max_leaf = 0;
    TriRec* backing = nullptr;  // pinned
    size_t backing_cap = 0;
    TriRec* staging = nullptr;  // pinned mirror of the device pool
    size_t staging_cap = 0;
    uint32_t* stream_host = nullptr;  // pinned: (stopped pixels, requested leaves)
    DevBuf s_pool, s_table, s_request, s_list, s_queue[2], s_state, s_best, s_hits, s_count;
    uint32_t level_prefix[kTopLevelsMax + 1] = {0};
    bool have_tree = false;
    bool ropes_ready = false;
    bool ropes_pushed = false;  // k_rope_push has run on the current rope slots
    DevBuf rope_nodes, rope_slots, rope_links;
    DevBuf selftest;
    uint32_t n_slots = 0;

    // lights
    std::vector<hulk_light> lights;
    DevBuf d_lights, d_point;
    uint32_t n_point = 0;

    // staging + work buffers of single-frame calls
    DevBuf rays, tbound, rgba, mirror, stats, hits_io;
    Work work;
    unsigned long long* host_stats = nullptr;  // pinned
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    hulk_counters cnt{};
    void* scratch = nullptr;  // this device's address of `scratch` (null: lookup failed, see scratch_err)
    hipError_t scratch_err = hipSuccess;
};

namespace {

int fail(hulk_ctx* c, int code, const char* msg) {
    if (c) c->err = msg;
    return code;
}

int hip_fail(hulk_ctx* c, hipError_t e, const char* where) {
    if (c) {
        c->err = std::string(where) + ": " + hipGetErrorString(e);
    }
    return HULK_E_HIP;
}

#define HIP_TRY(ctx, expr)                                    \
    do {                                                      \
        hipError_t _e = (expr);                               \
        if (_e != hipSuccess) return hip_fail(ctx, _e, #expr); \
    } while (0)

int ensure(hulk_ctx* c, DevBuf& b, size_t bytes) {
    if (bytes <= b.cap && b.p) return HULK_OK;
    if (b.p) {
        hipError_t e = hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
        if (e != hipSuccess) return hip_fail(c, e, "hipFree");
    }
    size_t want = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        b.p = nullptr;
        return hip_fail(c, e, "hipMalloc");
    }
    b.cap = want;
    return HULK_OK;
}

void release(DevBuf& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

// Address of `scratch` on `device`, which must be the current device: looked
// up once per device and kept for the life of the process.
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};

hipError_t scratch_address(int device, void** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_done[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_err[device] = e;
        g_scratch_addr[device] = e == hipSuccess ? p : nullptr;
        g_scratch_done[device] = true;
    }
    *out = g_scratch_addr[device];
    return g_scratch_err[device];
}

int check_opts(hulk_ctx* c, const hulk_opts& o) {
    if (o.variant < 0 || o.variant >= HULK_VARIANT_COUNT) return fail(c, HULK_E_ILLEGAL_INPUT, "unknown variant");
    if (o.block_size < 64 || o.block_size > 1024 || (o.block_size % 64) != 0)
        return fail(c, HULK_E_ILLEGAL_INPUT, "block_size must be a multiple of 64 in [64, 1024]");
    if (o.short_stack_depth < 1 || o.short_stack_depth > kMaxShortStack)
        return fail(c, HULK_E_ILLEGAL_INPUT, "short_stack_depth must be in [1, 16]");
    if (o.lds_top_levels > kTopLevelsMax) return fail(c, HULK_E_ILLEGAL_INPUT, "lds_top_levels must be <= 12");
    if (o.pipeline > HULK_PIPELINE_REFILL) return fail(c, HULK_E_ILLEGAL_INPUT, "unknown pipeline");
    if (o.pipeline == HULK_PIPELINE_REFILL &&
        ((o.variant != HULK_VARIANT_ROPES && o.variant != HULK_VARIANT_LDSSTACK) || o.block_size > 256))
        return fail(c, HULK_E_ILLEGAL_INPUT,
                    "the refill pipeline needs the ropes or LDS short-stack variant and blocks of <= 256");
    if (o.leaf_batch > 2) return fail(c, HULK_E_ILLEGAL_INPUT, "leaf_batch must be 0, 1 or 2");
    if (o.coop_lanes > 64) return fail(c, HULK_E_ILLEGAL_INPUT, "coop_lanes must be in [0, 64]");
    if (o.tri_test > HULK_TRI_MOLLER_TRUMBORE) return fail(c, HULK_E_ILLEGAL_INPUT, "unknown tri_test");
    if (o.tri_test == HULK_TRI_MOLLER_TRUMBORE && o.leaf_batch == 2)
        return fail(c, HULK_E_ILLEGAL_INPUT, "the Moller-Trumbore test runs with one record in flight");
    return HULK_OK;
}

// Frame launch model of hulk_pipeline (AUTO: fused; REFILL is a fused launch too).
bool use_fused(const hulk_opts& o, bool /*bounces*/) { return o.pipeline != HULK_PIPELINE_QUEUED; }

Camera to_cam(const hulk_camera& c) {
    Camera k;
    k.from[0] = c.from.x; k.from[1] = c.from.y; k.from[2] = c.from.z;
    k.vf[0] = c.vertical_factor.x; k.vf[1] = c.vertical_factor.y; k.vf[2] = c.vertical_factor.z;
    k.hf[0] = c.horizontal_factor.x; k.hf[1] = c.horizontal_factor.y; k.hf[2] = c.horizontal_factor.z;
    k.ll[0] = c.lower_left.x; k.ll[1] = c.lower_left.y; k.ll[2] = c.lower_left.z;
    return k;
}

// Rope slots for the rope variant, built on the device (launch_build_ropes)
// the first time a trace needs them after a tree upload; their pushed
// ropes (launch_push_ropes) on the second trace of the same tree, so that a
// tree rebuilt for every frame does not pay the push build (tens of us) for
// the few us per frame it saves.  The pushed targets are the plain ropes
// until then.  Not inside a graph capture (the build would replay with the
// graph): a later eager trace builds them in place, where the graph reads.
int ensure_ropes(hulk_ctx* c) {
    if (c->ropes_ready) {
        if (!c->ropes_pushed) {
            hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
            HIP_TRY(c, hipStreamIsCapturing(c->stream, &st));
            if (st == hipStreamCaptureStatusNone) {
                HIP_TRY(c, launch_push_ropes(static_cast<const uint4*>(c->rope_nodes.p),
                                             static_cast<const uint32_t*>(c->rope_links.p), c->n_nodes,
                                             static_cast<RopeSlot*>(c->rope_slots.p), c->box, c->stream));
                c->ropes_pushed = true;
            }
        }
        return HULK_OK;
    }
    const uint32_t n_slots = 3 * c->n_nodes;
    int r;
    if ((r = ensure(c, c->rope_nodes, (size_t)n_slots * sizeof(uint4)))) return r;
    if ((r = ensure(c, c->rope_slots, (size_t)n_slots * sizeof(RopeSlot)))) return r;
    if ((r = ensure(c, c->rope_links, (size_t)n_slots * 4))) return r;
    HIP_TRY(c, launch_build_ropes(static_cast<const uint4*>(c->nodes.p), c->n_nodes,
                                  static_cast<uint32_t*>(c->rope_links.p), static_cast<uint4*>(c->rope_nodes.p),
                                  static_cast<RopeSlot*>(c->rope_slots.p), c->box, c->stream));
    c->n_slots = n_slots;
    c->ropes_ready = true;
    c->ropes_pushed = false;
    return HULK_OK;
}

// Fill the tree / geometry part of a TraceParams and the LDS budget for the
// context's options.
int trace_setup(hulk_ctx* c, const hulk_opts& o, TraceParams& P, uint32_t& block, size_t& lds) {
    if (!c->have_scene) return fail(c, HULK_E_NO_SCENE, "no scene uploaded");
    const int v = o.variant;
    if (v != HULK_VARIANT_BRUTEFORCE && !c->have_tree) return fail(c, HULK_E_NO_TREE, "no KD tree uploaded");
    if (v == HULK_VARIANT_ROPES) {
        int r = ensure_ropes(c);
        if (r) return r;
    }
    std::memset(&P, 0, sizeof(P));
    P.leaf_batch = o.tri_test == HULK_TRI_MOLLER_TRUMBORE ? 3u : (o.leaf_batch == 2 ? 2u : 1u);
    P.coop_lanes = o.coop_lanes;
    P.nodes = static_cast<const uint4*>(c->nodes.p);
    P.recs = static_cast<const TriRec*>(c->recs.p);
    P.rec_bytes = c->n_lp * static_cast<uint32_t>(sizeof(TriRec));
    P.verts = static_cast<const float4*>(c->verts.p);
    P.tris = static_cast<const uint4*>(c->tris.p);
    P.n_tris = c->n_tris;
    P.rope_nodes = static_cast<const uint4*>(c->rope_nodes.p);
    P.node_bytes = c->n_nodes * 16u;
    P.rope_node_bytes = c->n_slots * 16u;
    P.rope_slots = static_cast<const RopeSlot*>(c->rope_slots.p);
    P.box = c->box;
    P.guard = static_cast<uint32_t>(std::min<uint64_t>(4ull * (c->n_nodes + c->n_slots) + 4096u, 0x7FFFFFFFull));
    block = o.block_size;
    uint32_t lds_nodes = 0;
    if (v != HULK_VARIANT_BRUTEFORCE && v != HULK_VARIANT_ROPES)
        lds_nodes = c->level_prefix[std::min(o.lds_top_levels, kTopLevelsMax)];
    P.lds_node_count = lds_nodes;
    size_t stack_words = 0;
    if (v == HULK_VARIANT_FULLSTACK) {
        P.stack_depth = std::max<uint32_t>(c->max_jdepth, 1u);
        stack_words = 2u * P.stack_depth;
    } else if (v == HULK_VARIANT_SHORTSTACK || v == HULK_VARIANT_LDSSTACK) {
        P.stack_depth = o.short_stack_depth;
        if (v == HULK_VARIANT_LDSSTACK) stack_words = 3u * P.stack_depth;
    }
    lds = (size_t)lds_nodes * 16u + stack_words * 4u * block;
    if (lds > c->max_lds) {
        char buf[160];
        std::snprintf(buf, sizeof(buf),
                      "LDS need %zu B exceeds the per-block limit %zu B (lower block_size or lds_top_levels)", lds,
                      c->max_lds);
        return fail(c, HULK_E_ILLEGAL_INPUT, buf);
    }
    P.counters = static_cast<unsigned long long*>(c->stats.p);
    return HULK_OK;
}

// Launch one trace.  Persistent launches zero their work counter with a
// fill kernel first (graph-capture safe: no memset nodes).
int run_trace(hulk_ctx* c, int variant, TraceParams& P, uint32_t n_max, uint32_t block, size_t lds, bool count,
              bool persistent, uint32_t* work_counter, bool timed, bool render = false) {
    if (n_max == 0) return HULK_OK;
    P.persistent = persistent ? 1u : 0u;
    P.work_counter = work_counter;
    uint32_t grid;
    if (persistent) {
        HIP_TRY(c, launch_fill_u32(work_counter, kWorkCounterWords, 0u, c->stream));
        // one resident wave set (2048 threads per CU); each wave takes
        // `fetch` chunks per atomic, so that every wave still gets about
        // four fetches of its slice
        const uint32_t per_cu = std::max<uint32_t>(1u, 2048u / block);
        grid = std::max<uint32_t>(1u, std::min<uint32_t>((n_max + block - 1) / block, c->num_cus * per_cu));
        const uint64_t waves = uint64_t(grid) * (block / 64);
        const uint64_t chunks = (uint64_t(n_max) + 63) / 64;
        P.fetch = static_cast<uint32_t>(std::min<uint64_t>(8, std::max<uint64_t>(1, chunks / (waves * 4))));
    } else {
        grid = std::max<uint32_t>(1u, (n_max + block - 1) / block);
    }
    if (timed) HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
    HIP_TRY(c, launch_trace(variant, P, count, render, P.leaf_batch, grid, block, lds, c->stream));
    if (timed) HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
    return HULK_OK;
}

int zero_stats(hulk_ctx* c) {
    HIP_TRY(c, launch_fill_u32(static_cast<uint32_t*>(c->stats.p), 16, 0u, c->stream));
    return HULK_OK;
}

int read_stats(hulk_ctx* c) {
    HIP_TRY(c, hipMemcpyAsync(c->host_stats, c->stats.p, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->cnt.rays += c->host_stats[0];
    c->cnt.node_visits += c->host_stats[1];
    c->cnt.leaf_visits += c->host_stats[2];
    c->cnt.tri_tests += c->host_stats[3];
    c->cnt.hits += c->host_stats[4];
    c->cnt.fallbacks += c->host_stats[5];
    c->cnt.shadow_rays += c->host_stats[6];
    return HULK_OK;
}

int finish(hulk_ctx* c, uint32_t flags, bool timed) {
    if (flags & HULK_FLAG_NO_SYNC) return HULK_OK;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (timed) {
        float ms = 0.0f;
        HIP_TRY(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->cnt.last_kernel_ms = ms;
    }
    return HULK_OK;
}

void release_work(Work& w) {
    DevBuf* bufs[] = {&w.hits, &w.hits2, &w.hits3, &w.shadow_rays, &w.shadow_slot, &w.shadow_t, &w.qrays[0],
                      &w.qrays[1], &w.qpix[0], &w.qpix[1], &w.qw[0], &w.qw[1], &w.acc, &w.ctr, &w.wq, &w.bcount,
                      &w.boffset};
    for (DevBuf* b : bufs) release(*b);
}

// Size a work set for n primary rays, np point lights, with or without
// mirror generations.
int ensure_work(hulk_ctx* c, Work& w, uint32_t n, uint32_t np, bool bounces, bool queued = true) {
    int r;
    const size_t m = std::max<uint32_t>(np, 1);
    if ((r = ensure(c, w.ctr, 64))) return r;
    if ((r = ensure(c, w.wq, kWorkCounterWords * 4))) return r;
    if ((r = ensure(c, w.hits, (size_t)n * sizeof(hulk_hit)))) return r;
    if (!queued) return HULK_OK;  // the fused pipeline keeps everything else in registers
    if ((r = ensure(c, w.acc, (size_t)n * 12))) return r;
    if ((r = ensure(c, w.shadow_rays, (size_t)n * m * sizeof(hulk_ray)))) return r;
    if ((r = ensure(c, w.shadow_slot, (size_t)n * m * 4))) return r;
    if ((r = ensure(c, w.shadow_t, (size_t)n * m * 4))) return r;
    const size_t blocks = ((size_t)n + 255) / 256;  // 2 counts per 256-ray block of the compaction
    if ((r = ensure(c, w.bcount, blocks * 8))) return r;
    if ((r = ensure(c, w.boffset, blocks * 8))) return r;
    if (bounces) {
        if ((r = ensure(c, w.hits2, (size_t)n * sizeof(hulk_hit)))) return r;
        if ((r = ensure(c, w.hits3, (size_t)n * sizeof(hulk_hit)))) return r;
        for (int k = 0; k < 2; k++) {
            if ((r = ensure(c, w.qrays[k], (size_t)n * sizeof(hulk_ray)))) return r;
            if ((r = ensure(c, w.qpix[k], (size_t)n * 4))) return r;
            if ((r = ensure(c, w.qw[k], (size_t)n * 4))) return r;
        }
    }
    return HULK_OK;
}

// Primary rays of one camera (cam / W / H) or of a device frame table.
struct FrameSet {
    const FrameDev* frames = nullptr;
    uint32_t n_frames = 0;
    uint32_t frame_w = 0, frame_h = 0;  // frames: their size when every frame has the same, else 0
    Camera cam{};
    uint32_t W = 0, H = 0;
    uint32_t n = 0;
};

int trace_primary(hulk_ctx* c, const hulk_opts& o, const FrameSet& fs, Work& w, const float* tbound, hulk_hit* out,
                  bool count, bool timed) {
    TraceParams P;
    uint32_t block;
    size_t lds;
    int r = trace_setup(c, o, P, block, lds);
    if (r) return r;
    if (fs.frames) {
        P.source = kSourceFrames;
        P.frames = fs.frames;
        P.n_frames = fs.n_frames;
    } else {
        P.source = kSourceCamera;
        P.cam = fs.cam;
        P.W = fs.W;
        P.w_m = udiv_multiplier(fs.W);
        P.tx_m = udiv_multiplier(fs.W >> 3);
        P.ty_m = udiv_multiplier(fs.H >> 3);
        P.fW = static_cast<float>(fs.W);
        P.fH = static_cast<float>(fs.H);
        P.H = fs.H;
    }
    P.n = fs.n;
    P.tbound = tbound;
    P.hits = out;
    return run_trace(c, o.variant, P, fs.n, block, lds, count, o.persistent != 0,
                     static_cast<uint32_t*>(w.wq.p), timed);
}

// Whole frame in one launch with lane refill (REFILL pipeline): the same
// outputs as render_fused.
int render_refill(hulk_ctx* c, const hulk_opts& o, const FrameSet& fs, Work& w, const float* tbound, hulk_hit* out,
                  uint32_t max_bounce, const uint8_t* mirror_dev, uint32_t* rgba, bool count, bool timed) {
    TraceParams P;
    uint32_t block;
    size_t lds;
    int r = trace_setup(c, o, P, block, lds);
    if (r) return r;
    if (refill_lds_bytes(lds, block, static_cast<uint32_t>(c->lights.size())) > c->max_lds)
        return fail(c, HULK_E_ILLEGAL_INPUT, "refill LDS (lane fields + short stack + top levels) exceeds the block limit");
    if (fs.n == 0) return HULK_OK;
    if (fs.frames) {
        P.source = kSourceFrames;
        P.frames = fs.frames;
        P.n_frames = fs.n_frames;
        // frames of one size whose tile columns fill whole strips: ray order
        // strip-major across the frames (frame_ray_xstrips)
        const uint32_t fw = fs.frame_w, fh = fs.frame_h;
        if (fw && fh && ((fw | fh) & 7u) == 0 && ((fw >> 3) & 15u) == 0) {
            P.xs_rps = (fw * fh) / ((fw >> 3) >> 4);
            P.xs_rps_m = udiv_multiplier(P.xs_rps);
            P.xs_nf_m = udiv_multiplier(fs.n_frames);
        }
    } else {
        P.source = kSourceCamera;
        P.cam = fs.cam;
        P.W = fs.W;
        P.w_m = udiv_multiplier(fs.W);
        P.tx_m = udiv_multiplier(fs.W >> 3);
        P.ty_m = udiv_multiplier(fs.H >> 3);
        P.fW = static_cast<float>(fs.W);
        P.fH = static_cast<float>(fs.H);
        P.H = fs.H;
    }
    P.n = fs.n;
    P.tbound = tbound;
    P.hits = out;
    P.lights = static_cast<const float4*>(c->d_lights.p);
    P.n_lights = static_cast<uint32_t>(c->lights.size());
    P.mirror = mirror_dev;
    P.max_bounce = mirror_dev ? max_bounce : 0;
    P.rgba = rgba;
    P.work_counter = static_cast<uint32_t*>(w.wq.p);
    P.persistent = 1;
    HIP_TRY(c, launch_fill_u32(P.work_counter, kWorkCounterWords, 0u, c->stream));
    // one resident set of waves (2048 threads per CU), fewer for small frames
    const uint32_t per_cu = std::max<uint32_t>(1u, 2048u / block);
    const uint32_t grid =
        std::max<uint32_t>(1u, std::min<uint32_t>((fs.n + block - 1) / block, c->num_cus * per_cu));
    if (timed) HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
    HIP_TRY(c, launch_render_refill(o.variant, P, count, grid, block, lds, c->stream));
    if (timed) HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
    return HULK_OK;
}

// Whole frame in one launch (FUSED pipeline): primary hit records to `out`,
// packed pixels to `rgba`.  No allocation, no synchronization.
int render_fused(hulk_ctx* c, const hulk_opts& o, const FrameSet& fs, Work& w, const float* tbound, hulk_hit* out,
                 uint32_t max_bounce, const uint8_t* mirror_dev, uint32_t* rgba, bool count, bool timed) {
    TraceParams P;
    uint32_t block;
    size_t lds;
    int r = trace_setup(c, o, P, block, lds);
    if (r) return r;
    if (fs.frames) {
        P.source = kSourceFrames;
        P.frames = fs.frames;
        P.n_frames = fs.n_frames;
    } else {
        P.source = kSourceCamera;
        P.cam = fs.cam;
        P.W = fs.W;
        P.w_m = udiv_multiplier(fs.W);
        P.tx_m = udiv_multiplier(fs.W >> 3);
        P.ty_m = udiv_multiplier(fs.H >> 3);
        P.fW = static_cast<float>(fs.W);
        P.fH = static_cast<float>(fs.H);
        P.H = fs.H;
    }
    P.n = fs.n;
    P.tbound = tbound;
    P.hits = out;
    P.lights = static_cast<const float4*>(c->d_lights.p);
    P.n_lights = static_cast<uint32_t>(c->lights.size());
    P.mirror = mirror_dev;
    P.max_bounce = mirror_dev ? max_bounce : 0;
    P.rgba = rgba;
    return run_trace(c, o.variant, P, fs.n, block, lds, count, o.persistent != 0, static_cast<uint32_t*>(w.wq.p),
                     timed, true);
}

// The fused frame launch of the options' pipeline (FUSED or REFILL).
int render_frame_launch(hulk_ctx* c, const hulk_opts& o, const FrameSet& fs, Work& w, const float* tbound,
                        hulk_hit* out, uint32_t max_bounce, const uint8_t* mirror_dev, uint32_t* rgba, bool count,
                        bool timed) {
    if (o.pipeline == HULK_PIPELINE_REFILL)
        return render_refill(c, o, fs, w, tbound, out, max_bounce, mirror_dev, rgba, count, timed);
    return render_fused(c, o, fs, w, tbound, out, max_bounce, mirror_dev, rgba, count, timed);
}

// Shading of a primary hit buffer: per generation g, point-light shadow
// rays (compacted queue, nearest hit, t scattered per (ray, light)), Lambert
// accumulation, and mirror rays (compacted queue) traced as generation
// g + 1.  No allocation, no synchronization: safe to capture in a graph.
int shade_pipeline(hulk_ctx* c, const hulk_opts& o, const FrameSet& fs, Work& w, const hulk_hit* hits0,
                   uint32_t max_bounce, const uint8_t* mirror_dev, uint32_t* rgba_dev, bool count) {
    const uint32_t n = fs.n;
    const uint32_t np = c->n_point;
    int r;
    TraceParams P;
    uint32_t block;
    size_t lds;
    if ((r = trace_setup(c, o, P, block, lds))) return r;
    uint32_t* ctr = static_cast<uint32_t*>(w.ctr.p);  // [1] shadow count, [2..3] queue counts
    uint32_t* wq = static_cast<uint32_t*>(w.wq.p);    // persistent-mode work counters
    HIP_TRY(c, launch_fill_f32(static_cast<float*>(w.acc.p), (size_t)n * 3, 0.0f, c->stream));
    const hulk_hit* cur_hits = hits0;
    const hulk_ray* cur_rays = nullptr;
    const uint32_t* cur_pix = nullptr;
    const float* cur_w = nullptr;
    const uint32_t* cur_n = nullptr;
    for (uint32_t g = 0; g <= max_bounce; g++) {
        const bool spawn = g < max_bounce && mirror_dev;
        const int qo = g & 1;  // queue written by this generation (its length is set by the scan)
        GenParams G;
        std::memset(&G, 0, sizeof(G));
        G.hits = cur_hits;
        G.camera_rays = g == 0;
        G.cam = fs.cam;
        G.W = fs.W;
        G.H = fs.H;
        G.frames = fs.frames;
        G.n_frames = fs.n_frames;
        G.gen = g;
        G.rays = cur_rays;
        G.pixel_in = cur_pix;
        G.weight_in = cur_w;
        G.n_dev = g == 0 ? nullptr : cur_n;
        G.n = n;
        G.point_pos = static_cast<const float4*>(c->d_point.p);
        G.n_point = np;
        G.shadow_rays = static_cast<hulk_ray*>(w.shadow_rays.p);
        G.shadow_slot = static_cast<uint32_t*>(w.shadow_slot.p);
        G.shadow_count = ctr + 1;
        G.mirror = mirror_dev;
        if (spawn) {
            G.next_rays = static_cast<hulk_ray*>(w.qrays[qo].p);
            G.pixel_out = static_cast<uint32_t*>(w.qpix[qo].p);
            G.weight_out = static_cast<float*>(w.qw[qo].p);
            G.next_count = ctr + 2 + qo;
        }
        G.block_counts = static_cast<uint32_t*>(w.bcount.p);
        G.block_offsets = static_cast<uint32_t*>(w.boffset.p);
        HIP_TRY(c, launch_gen_secondary(G, n, c->stream));
        if (np) {
            TraceParams S = P;
            S.source = kSourceRays;
            S.rays = static_cast<const hulk_ray*>(w.shadow_rays.p);
            S.n_dev = ctr + 1;
            S.n = n * np;
            S.t_out = static_cast<float*>(w.shadow_t.p);
            S.t_slot = static_cast<const uint32_t*>(w.shadow_slot.p);
            S.stop_at_w = 1;
            if ((r = run_trace(c, o.variant, S, n * np, block, lds, count, o.persistent != 0, wq, false))) return r;
        }
        ShadeParams SP;
        std::memset(&SP, 0, sizeof(SP));
        SP.hits = cur_hits;
        SP.camera_rays = g == 0;
        SP.cam = fs.cam;
        SP.W = fs.W;
        SP.H = fs.H;
        SP.frames = fs.frames;
        SP.n_frames = fs.n_frames;
        SP.gen = g;
        SP.rays = cur_rays;
        SP.pixel_in = cur_pix;
        SP.weight_in = cur_w;
        SP.n_dev = g == 0 ? nullptr : cur_n;
        SP.n = n;
        SP.lights = static_cast<const float4*>(c->d_lights.p);
        SP.n_lights = static_cast<uint32_t>(c->lights.size());
        SP.n_point = np;
        SP.shadow_t = static_cast<const float*>(w.shadow_t.p);
        SP.mirror = mirror_dev;
        SP.spawned = spawn ? 1u : 0u;
        SP.acc = static_cast<float*>(w.acc.p);
        HIP_TRY(c, launch_shade_accumulate(SP, n, c->stream));
        if (!spawn) break;
        TraceParams Q = P;
        Q.source = kSourceRays;
        Q.rays = static_cast<const hulk_ray*>(w.qrays[qo].p);
        Q.n_dev = ctr + 2 + qo;
        Q.n = n;
        hulk_hit* nh = static_cast<hulk_hit*>((g & 1) ? w.hits3.p : w.hits2.p);
        Q.hits = nh;
        if ((r = run_trace(c, o.variant, Q, n, block, lds, count, o.persistent != 0, wq, false))) return r;
        cur_hits = nh;
        cur_rays = static_cast<const hulk_ray*>(w.qrays[qo].p);
        cur_pix = static_cast<const uint32_t*>(w.qpix[qo].p);
        cur_w = static_cast<const float*>(w.qw[qo].p);
        cur_n = ctr + 2 + qo;
    }
    HIP_TRY(c, launch_pack(static_cast<const float*>(w.acc.p), n, rgba_dev, c->stream));
    return HULK_OK;
}

// Resolve an input buffer: device pointers are used as is; host pointers
// are staged through a context buffer.
template <typename T>
int stage_in(hulk_ctx* c, const T* src, size_t count, bool device_ptrs, DevBuf& buf, const T** dst) {
    if (!src) {
        *dst = nullptr;
        return HULK_OK;
    }
    if (device_ptrs) {
        *dst = src;
        return HULK_OK;
    }
    int r = ensure(c, buf, count * sizeof(T));
    if (r) return r;
    HIP_TRY(c, hipMemcpyAsync(buf.p, src, count * sizeof(T), hipMemcpyHostToDevice, c->stream));
    *dst = static_cast<const T*>(buf.p);
    return HULK_OK;
}

}  // namespace


// ------------------------------------------------------------------ ABI --

extern "C" int hulk_create(hulk_ctx** out, const hulk_opts* opts) {
    if (!out) return HULK_E_ILLEGAL_INPUT;
    *out = nullptr;
    hulk_ctx* c = new (std::nothrow) hulk_ctx;
    if (!c) return HULK_E_ALLOC;
    hulk_opts o{};
    o.device = 0;
    o.variant = HULK_VARIANT_FULLSTACK;
    o.block_size = 256;
    o.short_stack_depth = 6;  // c_shortStackDepth, SharedStructs.hpp:19
    o.lds_top_levels = 0;
    o.persistent = 0;
    o.stream = nullptr;
    if (opts) o = *opts;
    int r = check_opts(c, o);
    if (r) {
        delete c;
        return r;
    }
    c->opts = o;
    hipError_t e = hipSetDevice(o.device);
    if (e != hipSuccess) {
        delete c;
        return HULK_E_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, o.device) == hipSuccess) {
        c->num_cus = prop.multiProcessorCount;
        c->max_lds = std::min<size_t>(prop.sharedMemPerBlock, 65536);
    }
    // the device is current here: look up its scratch address (reported by
    // hulk_render_frame_static if the lookup failed)
    c->scratch_err = scratch_address(o.device, &c->scratch);
    if (o.stream) {
        c->stream = static_cast<hipStream_t>(o.stream);
    } else {
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            delete c;
            return HULK_E_HIP;
        }
        c->own_stream = true;
    }
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreate(&c->ev2) != hipSuccess || ensure(c, c->work.ctr, 64) || ensure(c, c->work.wq, kWorkCounterWords * 4) ||
        ensure(c, c->stats, 128) ||
        hipHostMalloc(reinterpret_cast<void**>(&c->host_stats), 128, hipHostMallocDefault) != hipSuccess) {
        hulk_destroy(c);
        return HULK_E_HIP;
    }
    *out = c;
    return HULK_OK;
}

extern "C" void hulk_destroy(hulk_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    DevBuf* bufs[] = {&c->verts, &c->tris, &c->nodes, &c->ref_nodes, &c->lp, &c->recs, &c->rope_nodes, &c->rope_slots, &c->rope_links, &c->selftest,
                      &c->d_lights, &c->d_point, &c->rays, &c->tbound, &c->rgba, &c->mirror, &c->stats, &c->hits_io};
    for (DevBuf* b : bufs) release(*b);
    release_work(c->work);
    kdb_destroy(c->kdb);
    for (DevBuf* b : {&c->s_pool, &c->s_table, &c->s_request, &c->s_list, &c->s_queue[0], &c->s_queue[1],
                      &c->s_state, &c->s_best, &c->s_hits, &c->s_count})
        release(*b);
    if (c->backing) (void)hipHostFree(c->backing);
    if (c->staging) (void)hipHostFree(c->staging);
    if (c->stream_host) (void)hipHostFree(c->stream_host);
    if (c->host_stats) (void)hipHostFree(c->host_stats);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->ev2) (void)hipEventDestroy(c->ev2);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" const char* hulk_last_error(const hulk_ctx* c) { return c ? c->err.c_str() : "null context"; }

extern "C" int hulk_set_options(hulk_ctx* c, const hulk_opts* o) {
    if (!c || !o) return HULK_E_ILLEGAL_INPUT;
    if (o->device != c->opts.device || o->stream != c->opts.stream)
        return fail(c, HULK_E_ILLEGAL_INPUT, "device and stream are fixed at hulk_create");
    int r = check_opts(c, *o);
    if (r) return r;
    c->opts = *o;
    return HULK_OK;
}

extern "C" int hulk_get_options(const hulk_ctx* c, hulk_opts* o) {
    if (!c || !o) return HULK_E_ILLEGAL_INPUT;
    *o = c->opts;
    return HULK_OK;
}

extern "C" void* hulk_stream(hulk_ctx* c) { return c ? static_cast<void*>(c->stream) : nullptr; }

extern "C" int hulk_sync(hulk_ctx* c) {
    if (!c) return HULK_E_ILLEGAL_INPUT;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return HULK_OK;
}

extern "C" int hulk_upload_scene(hulk_ctx* c, const hulk_float4* verts, uint32_t n_verts, const hulk_tri* tris,
                                 uint32_t n_tris, const hulk_box* scene_box) {
    if (!c || !scene_box || (n_verts && !verts) || (n_tris && !tris))
        return fail(c, HULK_E_ILLEGAL_INPUT, "null argument");
    for (uint32_t i = 0; i < n_tris; i++)
        for (int k = 0; k < 3; k++)
            if (tris[i].idx[k] >= n_verts) return fail(c, HULK_E_ILLEGAL_INPUT, "triangle vertex index out of range");
    int r;
    if ((r = ensure(c, c->verts, (size_t)n_verts * 16))) return r;
    if ((r = ensure(c, c->tris, (size_t)n_tris * 16))) return r;
    if (n_verts)
        HIP_TRY(c, hipMemcpyAsync(c->verts.p, verts, (size_t)n_verts * 16, hipMemcpyHostToDevice, c->stream));
    if (n_tris) HIP_TRY(c, hipMemcpyAsync(c->tris.p, tris, (size_t)n_tris * 16, hipMemcpyHostToDevice, c->stream));
    c->n_verts = n_verts;
    c->n_tris = n_tris;
    c->box.mn[0] = scene_box->min.x; c->box.mn[1] = scene_box->min.y; c->box.mn[2] = scene_box->min.z;
    c->box.mx[0] = scene_box->max.x; c->box.mx[1] = scene_box->max.y; c->box.mx[2] = scene_box->max.z;
    c->have_scene = true;
    c->have_tree = false;  // records depend on the vertices
    c->tree_gen++;
    c->ropes_ready = false;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return HULK_OK;
}

extern "C" int hulk_update_vertices(hulk_ctx* c, const hulk_float4* verts, uint32_t n_verts,
                                    const hulk_box* scene_box, uint32_t flags) {
    if (!c || !verts || !scene_box) return fail(c, HULK_E_ILLEGAL_INPUT, "null argument");
    if (!c->have_scene) return fail(c, HULK_E_NO_SCENE, "upload the scene before updating its vertices");
    if (n_verts != c->n_verts) return fail(c, HULK_E_ILLEGAL_INPUT, "vertex count differs from the uploaded scene");
    if (n_verts)
        HIP_TRY(c, hipMemcpyAsync(c->verts.p, verts, (size_t)n_verts * 16,
                                  (flags & HULK_FLAG_DEVICE_PTRS) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                  c->stream));
    c->box.mn[0] = scene_box->min.x; c->box.mn[1] = scene_box->min.y; c->box.mn[2] = scene_box->min.z;
    c->box.mx[0] = scene_box->max.x; c->box.mx[1] = scene_box->max.y; c->box.mx[2] = scene_box->max.z;
    c->have_tree = false;  // records and tree depend on the vertices
    c->ropes_ready = false;
    c->tree_gen++;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return HULK_OK;
}

extern "C" int hulk_upload_kdtree(hulk_ctx* c, const hulk_kd_node* nodes, uint32_t n_nodes,
                                  const uint32_t* leaf_polys, uint32_t n_lp) {
    if (!c) return HULK_E_ILLEGAL_INPUT;
    if (!c->have_scene) return fail(c, HULK_E_NO_SCENE, "upload the scene before the tree");
    if (!nodes || n_nodes == 0 || (n_lp && !leaf_polys)) return fail(c, HULK_E_ILLEGAL_INPUT, "null argument");
    // nodes and rope slots (3 per node) are addressed by 32-bit byte offsets
    if (n_nodes >= (1u << 26)) return fail(c, HULK_E_ILLEGAL_INPUT, "more than 2^26 nodes");
    // triangle records (64 B per leaf reference) are addressed by 32-bit byte offsets
    if (n_lp >= kMaxLeafRefs) return fail(c, HULK_E_ILLEGAL_INPUT, "2^26 or more leaf polygon references");
    // Validate: children in range, every node reached once from the root
    // (no cycles / sharing), leaf ranges inside leaf_polys, references < n_tris.
    std::vector<uint8_t> seen(n_nodes, 0);
    std::vector<uint32_t> depth(n_nodes, 0);
    std::vector<uint32_t> order;
    order.reserve(n_nodes);
    std::deque<uint32_t> q;
    q.push_back(0);
    seen[0] = 1;
    uint32_t max_jdepth = 0;
    std::vector<uint32_t> perm(n_nodes, 0);
    // breadth-first for depth < kTopLevelsMax, depth-first below
    std::vector<uint32_t> dfs;
    uint32_t prefix[kTopLevelsMax + 1] = {0};
    auto visit_children = [&](uint32_t i, auto&& push) -> int {
        const hulk_kd_node& n = nodes[i];
        if ((n.flags & (HULK_NODE_LEAF | HULK_NODE_JUNCTION)) == (HULK_NODE_LEAF | HULK_NODE_JUNCTION))
            return fail(c, HULK_E_ILLEGAL_INPUT, "node is both leaf and junction");
        if (n.flags & HULK_NODE_LEAF) {
            if ((uint64_t)n.u.leaf.start + n.u.leaf.count > n_lp)
                return fail(c, HULK_E_ILLEGAL_INPUT, "leaf range outside leaf_polys");
            return HULK_OK;
        }
        if (!(n.flags & HULK_NODE_JUNCTION)) return fail(c, HULK_E_ILLEGAL_INPUT, "node is neither leaf nor junction");
        if ((n.flags & 3u) == 3u) return fail(c, HULK_E_ILLEGAL_INPUT, "junction axis out of range");
        max_jdepth = std::max(max_jdepth, depth[i] + 1);
        for (uint32_t ch : {n.u.junction.above, n.u.junction.below}) {
            if (ch == 0) continue;
            if (ch >= n_nodes) return fail(c, HULK_E_ILLEGAL_INPUT, "child index out of range");
            if (seen[ch]) return fail(c, HULK_E_ILLEGAL_INPUT, "tree has a cycle or a shared child");
            seen[ch] = 1;
            depth[ch] = depth[i] + 1;
            push(ch);
        }
        return HULK_OK;
    };
    while (!q.empty()) {
        const uint32_t i = q.front();
        q.pop_front();
        if (depth[i] < kTopLevelsMax) {
            perm[i] = static_cast<uint32_t>(order.size());
            order.push_back(i);
            int r = visit_children(i, [&](uint32_t ch) { q.push_back(ch); });
            if (r) return r;
        } else {
            // depth-first pre-order, above child first
            dfs.clear();
            dfs.push_back(i);
            while (!dfs.empty()) {
                const uint32_t k = dfs.back();
                dfs.pop_back();
                perm[k] = static_cast<uint32_t>(order.size());
                order.push_back(k);
                uint32_t kids[2];
                int nk = 0;
                int r = visit_children(k, [&](uint32_t ch) { kids[nk++] = ch; });
                if (r) return r;
                while (nk > 0) dfs.push_back(kids[--nk]);
            }
        }
    }
    if (order.size() != n_nodes) return fail(c, HULK_E_ILLEGAL_INPUT, "tree has nodes unreachable from the root");
    for (uint32_t i = 0; i < n_lp; i++)
        if (leaf_polys[i] >= c->n_tris) return fail(c, HULK_E_ILLEGAL_INPUT, "leaf polygon index out of range");
    for (uint32_t d = 0; d <= kTopLevelsMax; d++) {
        uint32_t k = 0;
        while (k < order.size() && depth[order[k]] < d) k++;
        prefix[d] = k;
    }
    if (max_jdepth > kMaxFullStack) return fail(c, HULK_E_ILLEGAL_INPUT, "tree deeper than 64 junctions");
    // Device word x is normalised: exactly HULK_NODE_LEAF for a leaf and
    // HULK_NODE_JUNCTION | axis for a junction (other flag bits dropped), the
    // form the traversal kernels test with one compare.
    std::vector<uint4> dev(n_nodes);
    for (uint32_t k = 0; k < n_nodes; k++) {
        const hulk_kd_node& n = nodes[order[k]];
        if (n.flags & HULK_NODE_LEAF) {
            dev[k] = make_uint4(HULK_NODE_LEAF, n.u.leaf.start, n.u.leaf.count, n.u.leaf.leaf_id);
        } else {
            uint32_t plane;
            std::memcpy(&plane, &n.u.junction.plane, 4);
            const uint32_t a = n.u.junction.above ? perm[n.u.junction.above] : 0;
            const uint32_t b = n.u.junction.below ? perm[n.u.junction.below] : 0;
            dev[k] = make_uint4(HULK_NODE_JUNCTION | (n.flags & 3u), a, b, plane);
        }
    }
    int r;
    if ((r = ensure(c, c->nodes, (size_t)n_nodes * 16))) return r;
    if ((r = ensure(c, c->ref_nodes, (size_t)n_nodes * 16))) return r;
    if ((r = ensure(c, c->lp, (size_t)std::max<uint32_t>(n_lp, 1) * 4))) return r;
    if ((r = ensure(c, c->recs, (size_t)std::max<uint32_t>(n_lp, 1) * sizeof(TriRec)))) return r;
    HIP_TRY(c, hipMemcpyAsync(c->nodes.p, dev.data(), (size_t)n_nodes * 16, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->ref_nodes.p, nodes, (size_t)n_nodes * 16, hipMemcpyHostToDevice, c->stream));
    if (n_lp) HIP_TRY(c, hipMemcpyAsync(c->lp.p, leaf_polys, (size_t)n_lp * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, launch_build_records(static_cast<const float4*>(c->verts.p), static_cast<const uint4*>(c->tris.p),
                                    static_cast<const uint32_t*>(c->lp.p), n_lp, static_cast<TriRec*>(c->recs.p),
                                    c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->n_nodes = n_nodes;
    c->n_lp = n_lp;
    c->max_jdepth = max_jdepth;
    std::memcpy(c->level_prefix, prefix, sizeof(prefix));
    c->ropes_ready = false;
    c->have_tree = true;
    c->tree_gen++;
    return HULK_OK;
}

extern "C" int hulk_set_lights(hulk_ctx* c, const hulk_light* lights, uint32_t n) {
    if (!c || (n && !lights)) return fail(c, HULK_E_ILLEGAL_INPUT, "null argument");
    c->lights.assign(lights, lights + n);
    c->lights_gen++;
    std::vector<float4> pts;
    for (uint32_t k = 0; k < n; k++)
        if (lights[k].color.w == HULK_LIGHT_POINT)
            pts.push_back(make_float4(lights[k].data.x, lights[k].data.y, lights[k].data.z, 0.0f));
    c->n_point = static_cast<uint32_t>(pts.size());
    int r;
    if ((r = ensure(c, c->d_lights, std::max<size_t>(n, 1) * sizeof(hulk_light)))) return r;
    if ((r = ensure(c, c->d_point, std::max<size_t>(pts.size(), 1) * sizeof(float4)))) return r;
    if (n) HIP_TRY(c, hipMemcpyAsync(c->d_lights.p, lights, n * sizeof(hulk_light), hipMemcpyHostToDevice, c->stream));
    if (!pts.empty())
        HIP_TRY(c, hipMemcpyAsync(c->d_point.p, pts.data(), pts.size() * sizeof(float4), hipMemcpyHostToDevice,
                                  c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return HULK_OK;
}


// ------------------------------------------------------- trace entry points --

extern "C" int hulk_trace_camera(hulk_ctx* c, const hulk_camera* cam, uint32_t W, uint32_t H, const float* tbound,
                                 hulk_hit* hits, uint32_t flags) {
    if (!c || !cam || !hits) return fail(c, HULK_E_ILLEGAL_INPUT, "null argument");
    if ((uint64_t)W * H > 0xFFFFFFC0ull) return fail(c, HULK_E_ILLEGAL_INPUT, "image too large");
    const bool dev = (flags & HULK_FLAG_DEVICE_PTRS) != 0;
    const bool count = (flags & HULK_FLAG_COUNT) != 0;
    const uint32_t n = W * H;
    int r;
    FrameSet fs;
    fs.cam = to_cam(*cam);
    fs.W = W;
    fs.H = H;
    fs.n = n;
    if ((r = ensure(c, c->work.ctr, 64))) return r;
    if ((r = ensure(c, c->work.wq, kWorkCounterWords * 4))) return r;
    const float* tb;
    if ((r = stage_in(c, tbound, n, dev, c->tbound, &tb))) return r;
    hulk_hit* out = hits;
    if (!dev) {
        if ((r = ensure(c, c->hits_io, (size_t)n * sizeof(hulk_hit)))) return r;
        out = static_cast<hulk_hit*>(c->hits_io.p);
    }
    if (count && (r = zero_stats(c))) return r;
    if ((r = trace_primary(c, c->opts, fs, c->work, tb, out, count, true))) return r;
    if (!dev && n) HIP_TRY(c, hipMemcpyAsync(hits, out, (size_t)n * sizeof(hulk_hit), hipMemcpyDeviceToHost, c->stream));
    if (count && (r = read_stats(c))) return r;
    return finish(c, flags, true);
}

extern "C" int hulk_trace_rays(hulk_ctx* c, const hulk_ray* rays, uint32_t n, const float* tbound, hulk_hit* hits,
                               uint32_t flags) {
    if (!c || (n && (!rays || !hits))) return fail(c, HULK_E_ILLEGAL_INPUT, "null argument");
    if (n > 0xFFFFFFC0u) return fail(c, HULK_E_ILLEGAL_INPUT, "too many rays");
    const bool dev = (flags & HULK_FLAG_DEVICE_PTRS) != 0;
    const bool count = (flags & HULK_FLAG_COUNT) != 0;
    TraceParams P;
    uint32_t block;
    size_t lds;
    int r = trace_setup(c, c->opts, P, block, lds);
    if (r) return r;
    if ((r = ensure(c, c->work.ctr, 64))) return r;
    if ((r = ensure(c, c->work.wq, kWorkCounterWords * 4))) return r;
    P.source = kSourceRays;
    if ((r = stage_in(c, rays, n, dev, c->rays, &P.rays))) return r;
    P.n = n;
    if ((r = stage_in(c, tbound, n, dev, c->tbound, &P.tbound))) return r;
    hulk_hit* out = hits;
    if (!dev) {
        if ((r = ensure(c, c->hits_io, (size_t)std::max<uint32_t>(n, 1) * sizeof(hulk_hit)))) return r;
        out = static_cast<hulk_hit*>(c->hits_io.p);
    }
    P.hits = out;
    if (count && (r = zero_stats(c))) return r;
    if ((r = run_trace(c, c->opts.variant, P, n, block, lds, count, c->opts.persistent != 0,
                       static_cast<uint32_t*>(c->work.wq.p), true)))
        return r;
    if (!dev && n) HIP_TRY(c, hipMemcpyAsync(hits, out, (size_t)n * sizeof(hulk_hit), hipMemcpyDeviceToHost, c->stream));
    if (count && (r = read_stats(c))) return r;
    return finish(c, flags, n > 0);
}

extern "C" int hulk_shade(hulk_ctx* c, const hulk_camera* cam, uint32_t W, uint32_t H, const hulk_hit* hits,
                          uint32_t* rgba8, uint32_t mode, uint32_t flags) {
    if (!c || !cam || !hits || !rgba8) return fail(c, HULK_E_ILLEGAL_INPUT, "null argument");
    if (mode != HULK_SHADE_FUSED && mode != HULK_SHADE_PICTURE) return fail(c, HULK_E_ILLEGAL_INPUT, "bad shade mode");
    const bool dev = (flags & HULK_FLAG_DEVICE_PTRS) != 0;
    const bool count = (flags & HULK_FLAG_COUNT) != 0;
    const uint32_t n = W * H;
    int r;
    FrameSet fs;
    fs.cam = to_cam(*cam);
    fs.W = W;
    fs.H = H;
    fs.n = n;
    if ((r = ensure_work(c, c->work, n, c->n_point, false))) return r;
    const hulk_hit* dh;
    if ((r = stage_in(c, hits, n, dev, c->hits_io, &dh))) return r;
    uint32_t* out = rgba8;
    if (!dev) {
        if ((r = ensure(c, c->rgba, (size_t)n * 4))) return r;
        out = static_cast<uint32_t*>(c->rgba.p);
    }
    if (count && (r = zero_stats(c))) return r;
    HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
    if (mode == HULK_SHADE_PICTURE) {
        ShadeParams SP;
        std::memset(&SP, 0, sizeof(SP));
        SP.hits = dh;
        SP.camera_rays = 1;
        SP.cam = fs.cam;
        SP.W = W;
        SP.H = H;
        SP.n = n;
        SP.lights = static_cast<const float4*>(c->d_lights.p);
        SP.n_lights = static_cast<uint32_t>(c->lights.size());
        HIP_TRY(c, launch_shade_picture(SP, out, c->stream));
    } else {
        if ((r = shade_pipeline(c, c->opts, fs, c->work, dh, 0, nullptr, out, count))) return r;
    }
    HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
    if (!dev) HIP_TRY(c, hipMemcpyAsync(rgba8, out, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
    if (count && (r = read_stats(c))) return r;
    if (flags & HULK_FLAG_NO_SYNC) return HULK_OK;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    HIP_TRY(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->cnt.last_shade_ms = ms;
    return HULK_OK;
}

extern "C" int hulk_render_frame(hulk_ctx* c, const hulk_camera* cam, uint32_t W, uint32_t H, uint32_t max_bounce,
                                 const uint8_t* mirror_mask, const float* tbound, uint32_t* rgba8,
                                 hulk_hit* primary_hits, uint32_t flags) {
    if (!c || !cam || !rgba8) return fail(c, HULK_E_ILLEGAL_INPUT, "null argument");
    if (max_bounce > 8) return fail(c, HULK_E_ILLEGAL_INPUT, "max_bounce must be <= 8");
    if ((uint64_t)W * H > 0xFFFFFFC0ull) return fail(c, HULK_E_ILLEGAL_INPUT, "image too large");
    const bool dev = (flags & HULK_FLAG_DEVICE_PTRS) != 0;
    const bool count = (flags & HULK_FLAG_COUNT) != 0;
    const uint32_t n = W * H;
    int r;
    FrameSet fs;
    fs.cam = to_cam(*cam);
    fs.W = W;
    fs.H = H;
    fs.n = n;
    const bool bounces = max_bounce > 0 && mirror_mask;
    if ((r = ensure_work(c, c->work, n, c->n_point, bounces, !use_fused(c->opts, bounces)))) return r;
    const uint8_t* dm = nullptr;
    if (bounces && (r = stage_in(c, mirror_mask, c->n_tris, dev, c->mirror, &dm))) return r;
    const float* dt;
    if ((r = stage_in(c, tbound, n, dev, c->tbound, &dt))) return r;
    hulk_hit* h0 = (dev && primary_hits) ? primary_hits : static_cast<hulk_hit*>(c->work.hits.p);
    uint32_t* out = rgba8;
    if (!dev) {
        if ((r = ensure(c, c->rgba, (size_t)n * 4))) return r;
        out = static_cast<uint32_t*>(c->rgba.p);
    }
    if (count && (r = zero_stats(c))) return r;
    if (use_fused(c->opts, bounces)) {
        if ((r = render_frame_launch(c, c->opts, fs, c->work, dt, h0, bounces ? max_bounce : 0, dm, out, count, true)))
            return r;
    } else {
        if ((r = trace_primary(c, c->opts, fs, c->work, dt, h0, count, true))) return r;
        if ((r = shade_pipeline(c, c->opts, fs, c->work, h0, bounces ? max_bounce : 0, dm, out, count))) return r;
    }
    if (!dev) {
        HIP_TRY(c, hipMemcpyAsync(rgba8, out, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
        if (primary_hits)
            HIP_TRY(c, hipMemcpyAsync(primary_hits, h0, (size_t)n * sizeof(hulk_hit), hipMemcpyDeviceToHost, c->stream));
    }
    if (count && (r = read_stats(c))) return r;
    return finish(c, flags, true);
}

extern "C" int hulk_render_frame_static(hulk_ctx* c, const hulk_camera* cam, uint32_t W, uint32_t H,
                                        uint32_t max_bounce, const uint8_t* mirror_dev, const float* tbound_dev,
                                        uint32_t* rgba_dev, hulk_hit* primary_hits_dev, uint32_t flags) {
    if (!c || !cam || !rgba_dev) return fail(c, HULK_E_ILLEGAL_INPUT, "null argument");
    if (max_bounce > 8) return fail(c, HULK_E_ILLEGAL_INPUT, "max_bounce must be <= 8");
    if ((uint64_t)W * H > 0xFFFFFFC0ull) return fail(c, HULK_E_ILLEGAL_INPUT, "image too large");
    if (!c->scratch) return hip_fail(c, c->scratch_err, "hipGetSymbolAddress(scratch)");
    if (c->opts.pipeline == HULK_PIPELINE_QUEUED)
        return fail(c, HULK_E_UNSUPPORTED,
                    "the queued pipeline keeps per-frame ray queues: use the fused or refill pipeline here");
    const uint32_t n = W * H;
    if (!primary_hits_dev && n > kStaticMaxRays)
        return fail(c, HULK_E_ILLEGAL_INPUT,
                    "without a primary hit buffer a frame may have at most 262144 rays (HULK_STATIC_MAX_RAYS): "
                    "the static scratch holds their hit records");
    if (c->opts.variant == HULK_VARIANT_ROPES && c->have_tree && !c->ropes_ready)
        return fail(c, HULK_E_UNSUPPORTED,
                    "the rope slots of this tree are not built yet (the first rope trace through another entry "
                    "point builds them)");
    const bool count = (flags & HULK_FLAG_COUNT) != 0;
    FrameSet fs;
    fs.cam = to_cam(*cam);
    fs.W = W;
    fs.H = H;
    fs.n = n;
    Work w;  // nothing allocated: only the work counters are used, in the scratch
    w.wq.p = c->scratch;
    w.wq.cap = kScratchCounterBytes;
    hulk_hit* h0 = primary_hits_dev
                       ? primary_hits_dev
                       : reinterpret_cast<hulk_hit*>(static_cast<unsigned char*>(c->scratch) + kScratchCounterBytes);
    const bool bounces = max_bounce > 0 && mirror_dev;
    int r;
    if (count && (r = zero_stats(c))) return r;
    if ((r = render_frame_launch(c, c->opts, fs, w, tbound_dev, h0, bounces ? max_bounce : 0,
                                 bounces ? mirror_dev : nullptr, rgba_dev, count, true)))
        return r;
    if (count && (r = read_stats(c))) return r;
    return finish(c, flags, true);
}

extern "C" int hulk_camera_rays(hulk_ctx* c, const hulk_camera* cam, uint32_t W, uint32_t H, hulk_ray* rays_dev,
                                float* texit_dev, uint32_t flags) {
    if (!c || !cam || !rays_dev) return fail(c, HULK_E_ILLEGAL_INPUT, "null argument");
    if (!c->have_scene) return fail(c, HULK_E_NO_SCENE, "no scene uploaded");
    HIP_TRY(c, launch_camera_rays(to_cam(*cam), W, H, c->box, rays_dev, texit_dev, c->stream));
    return finish(c, flags, false);
}

extern "C" int hulk_selftest_reciprocal(hulk_ctx* c, unsigned long long out[4]) {
    if (!c || !out) return HULK_E_ILLEGAL_INPUT;
    int r;
    if ((r = ensure(c, c->selftest, 8 * sizeof(unsigned long long)))) return r;
    HIP_TRY(c, launch_selftest_rcp(static_cast<unsigned long long*>(c->selftest.p), c->stream));
    HIP_TRY(c, hipMemcpyAsync(out, c->selftest.p, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return HULK_OK;
}

extern "C" int hulk_ctx_scene_box(const hulk_ctx* c, float out[6]) {
    if (!c || !out) return HULK_E_ILLEGAL_INPUT;
    if (!c->have_scene) return HULK_E_NO_SCENE;
    for (int i = 0; i < 3; i++) {
        out[i] = c->box.mn[i];
        out[3 + i] = c->box.mx[i];
    }
    return HULK_OK;
}

// ------------------------------------------------- out-of-core streaming --

namespace {

// Pool allocator of the streaming trace: contiguous runs of records handed
// out in a ring, freed oldest first (the FIFO eviction of
// GLVisitorGPUTrace::clearLeaf, GLVisitorGPUTrace.cpp:491-502, over the
// index / vertex RingBuffers of RingBuffer.h:6-82, here one ring of triangle
// records).  A run never wraps: it starts at the write position or, when
// the tail of the pool is too short, at 0.
struct RecordRing {
    uint32_t cap = 0, head = 0, tail = 0;  // head: oldest run's start; tail: next free
    struct Run {
        uint32_t leaf, off, size, launch;  // launch: the relaunch round that loaded it
    };
    std::deque<Run> runs;
    bool fits(uint32_t size, uint32_t& off) const {
        if (runs.empty()) {
            off = 0;
            return size <= cap;
        }
        if (tail > head) {  // free: [tail, cap) and [0, head)
            if (size <= cap - tail) {
                off = tail;
                return true;
            }
            if (size <= head) {
                off = 0;
                return true;
            }
            return false;
        }
        if (size <= head - tail) {  // wrapped: free [tail, head)
            off = tail;
            return true;
        }
        return false;
    }
    void push(uint32_t leaf, uint32_t off, uint32_t size, uint32_t launch) {
        if (runs.empty()) head = off;
        runs.push_back({leaf, off, size, launch});
        tail = off + size;
    }
    Run pop() {
        Run r = runs.front();
        runs.pop_front();
        head = runs.empty() ? tail : runs.front().off;
        return r;
    }
};

}  // namespace

extern "C" int hulk_trace_camera_streamed(hulk_ctx* c, const hulk_camera* cam, uint32_t W, uint32_t H,
                                          uint32_t pool_records, hulk_hit* hits, uint32_t flags,
                                          hulk_stream_stats* stats) {
    if (!c || !cam || !hits || W == 0 || H == 0) return fail(c, HULK_E_ILLEGAL_INPUT, "null argument or empty image");
    if (!c->have_scene) return fail(c, HULK_E_NO_SCENE, "no scene uploaded");
    if (!c->have_tree) return fail(c, HULK_E_NO_TREE, "no KD tree uploaded");
    if (c->opts.tri_test != HULK_TRI_REFERENCE)
        return fail(c, HULK_E_ILLEGAL_INPUT, "the streamed trace uses the reference triangle test");
    const uint64_t n64 = (uint64_t)W * H;
    if (n64 >= (1ull << 31)) return fail(c, HULK_E_ILLEGAL_INPUT, "image too large");
    const uint32_t n = static_cast<uint32_t>(n64);
    const auto t_start = std::chrono::steady_clock::now();
    int r;
    // per-tree tables: leaf extents from the node array, records -> pinned host memory
    if (c->stream_gen != c->tree_gen) {
        std::vector<uint4> nodes(c->n_nodes);
        HIP_TRY(c, hipMemcpyAsync(nodes.data(), c->nodes.p, (size_t)c->n_nodes * 16, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        uint32_t n_leaves = 0, max_leaf = 0;
        for (const uint4& nd : nodes)
            if (nd.x & HULK_NODE_LEAF) n_leaves = std::max(n_leaves, nd.w + 1);
        if (n_leaves > (1u << 30)) return fail(c, HULK_E_ILLEGAL_INPUT, "leaf ids out of range");
        c->leaf_extent.assign(n_leaves, make_uint2(0, 0));
        for (const uint4& nd : nodes)
            if (nd.x & HULK_NODE_LEAF) {
                c->leaf_extent[nd.w] = make_uint2(nd.y, nd.z);
                max_leaf = std::max(max_leaf, nd.z);
            }
        c->max_leaf = max_leaf;
        const size_t bytes = std::max<size_t>((size_t)c->n_lp * sizeof(TriRec), 64);
        if (bytes > c->backing_cap) {
            if (c->backing) HIP_TRY(c, hipHostFree(c->backing));
            c->backing = nullptr;
            c->backing_cap = 0;
            HIP_TRY(c, hipHostMalloc(reinterpret_cast<void**>(&c->backing), bytes, hipHostMallocDefault));
            c->backing_cap = bytes;
        }
        if (c->n_lp)
            HIP_TRY(c, hipMemcpyAsync(c->backing, c->recs.p, (size_t)c->n_lp * sizeof(TriRec), hipMemcpyDeviceToHost,
                                      c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        c->stream_gen = c->tree_gen;
    }
    if (!c->stream_host)
        HIP_TRY(c, hipHostMalloc(reinterpret_cast<void**>(&c->stream_host), 64, hipHostMallocDefault));
    const uint32_t n_leaves = static_cast<uint32_t>(c->leaf_extent.size());
    const uint32_t pool = pool_records ? pool_records : std::max<uint32_t>(c->n_lp, 1);
    if (pool >= kMaxLeafRefs) return fail(c, HULK_E_ILLEGAL_INPUT, "leaf pool of 2^26 or more records");
    if (pool < c->max_leaf) {
        c->err = "pool of " + std::to_string(pool) + " records is smaller than the largest leaf (" +
                 std::to_string(c->max_leaf) + " records)";
        return HULK_E_ILLEGAL_INPUT;
    }
    const bool dev_out = (flags & HULK_FLAG_DEVICE_PTRS) != 0;
    if ((r = ensure(c, c->s_pool, (size_t)pool * sizeof(TriRec)))) return r;
    if ((r = ensure(c, c->s_table, (size_t)std::max<uint32_t>(n_leaves, 1) * 8))) return r;
    if ((r = ensure(c, c->s_request, (size_t)std::max<uint32_t>(n_leaves, 1) * 4))) return r;
    if ((r = ensure(c, c->s_list, (size_t)std::max<uint32_t>(n_leaves, 1) * 4))) return r;
    for (DevBuf* b : {&c->s_queue[0], &c->s_queue[1]})
        if ((r = ensure(c, *b, (size_t)n * 4))) return r;
    if ((r = ensure(c, c->s_state, (size_t)n * 16))) return r;
    if ((r = ensure(c, c->s_best, (size_t)n * sizeof(Best)))) return r;
    if (!dev_out && (r = ensure(c, c->s_hits, (size_t)n * sizeof(hulk_hit)))) return r;
    if ((r = ensure(c, c->s_count, 64))) return r;
    if ((size_t)pool * sizeof(TriRec) > c->staging_cap) {
        if (c->staging) HIP_TRY(c, hipHostFree(c->staging));
        c->staging = nullptr;
        c->staging_cap = 0;
        HIP_TRY(c, hipHostMalloc(reinterpret_cast<void**>(&c->staging), (size_t)pool * sizeof(TriRec),
                                 hipHostMallocDefault));
        c->staging_cap = (size_t)pool * sizeof(TriRec);
    }
    uint32_t* d_count = static_cast<uint32_t*>(c->s_count.p);
    std::vector<uint2> loaded;  // (pool offset, records) of this round's leaves
    HIP_TRY(c, launch_fill_u32(static_cast<uint32_t*>(c->s_request.p), std::max<uint32_t>(n_leaves, 1), 0u, c->stream));
    std::vector<uint2> table(std::max<uint32_t>(n_leaves, 1), make_uint2(0, 0));
    HIP_TRY(c, hipMemcpyAsync(c->s_table.p, table.data(), table.size() * 8, hipMemcpyHostToDevice, c->stream));

    StreamParams S;
    std::memset(&S, 0, sizeof(S));
    S.nodes = static_cast<const uint4*>(c->nodes.p);
    S.pool = static_cast<const TriRec*>(c->s_pool.p);
    S.pool_bytes = pool * static_cast<uint32_t>(sizeof(TriRec));
    S.leaf_table = static_cast<const uint2*>(c->s_table.p);
    S.request = static_cast<uint32_t*>(c->s_request.p);
    S.box = c->box;
    S.cam = to_cam(*cam);
    S.W = W;
    S.H = H;
    S.queue_in = nullptr;
    S.n_in = n;
    S.state = static_cast<float4*>(c->s_state.p);
    S.best = static_cast<Best*>(c->s_best.p);
    S.hits = dev_out ? hits : static_cast<hulk_hit*>(c->s_hits.p);
    S.guard = static_cast<uint32_t>(std::min<uint64_t>(4ull * c->n_nodes + 4096u, 0x7FFFFFFFull));

    hulk_stream_stats st;
    std::memset(&st, 0, sizeof(st));
    st.pool_records = pool;
    RecordRing ring;
    ring.cap = pool;
    std::vector<uint32_t> req;
    const uint32_t max_launches = 1u << 20;
    uint32_t cur = 0;
    for (;;) {
        if (st.launches >= max_launches) return fail(c, HULK_E_HIP, "streaming trace made no progress");
        S.queue_out = static_cast<uint32_t*>(c->s_queue[cur ^ 1].p);
        S.n_out = d_count;
        HIP_TRY(c, launch_fill_u32(d_count, 2, 0u, c->stream));
        HIP_TRY(c, launch_trace_streamed(S, c->stream));
        HIP_TRY(c, launch_collect_requests(S.request, n_leaves, static_cast<uint32_t*>(c->s_list.p), d_count + 1,
                                           c->stream));
        HIP_TRY(c, hipMemcpyAsync(c->stream_host, d_count, 8, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        st.launches++;
        const uint32_t n_out = c->stream_host[0], n_req = c->stream_host[1];
        if (n_out == 0) break;
        st.max_queue = std::max(st.max_queue, n_out);
        if (n_req == 0 || n_req > n_leaves) return fail(c, HULK_E_HIP, "stopped rays without a leaf request");
        req.resize(n_req);
        HIP_TRY(c, hipMemcpyAsync(req.data(), c->s_list.p, (size_t)n_req * 4, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        std::sort(req.begin(), req.end());  // the device appends in any order
        // admit requested leaves, evicting the oldest; at most half the pool
        // is cleared per launch (GLVisitorGPUTrace.cpp:445-462)
        uint32_t cleared = 0, admitted = 0;
        const uint32_t max_clear = pool / 2;
        // a leaf's records go to the pinned mirror of the pool; the pool is
        // updated per contiguous run of this round's leaves, not per leaf
        loaded.clear();
        auto admit = [&](uint32_t leaf, uint32_t off) -> int {
            const uint2 e = c->leaf_extent[leaf];
            if (e.y) {
                std::memcpy(c->staging + off, c->backing + e.x, (size_t)e.y * sizeof(TriRec));
                loaded.push_back(make_uint2(off, e.y));
            }
            ring.push(leaf, off, e.y, st.launches);
            table[leaf] = make_uint2(off, 1u);
            st.leaves_loaded++;
            st.bytes_loaded += (uint64_t)e.y * sizeof(TriRec);
            admitted++;
            return HULK_OK;
        };
        for (uint32_t leaf : req) {
            if (table[leaf].y) continue;
            const uint32_t size = c->leaf_extent[leaf].y;
            uint32_t off = 0;
            bool ok = ring.fits(size, off);
            // never evict a leaf loaded in this round: its rays have not used it yet
            while (!ok && !ring.runs.empty() && ring.runs.front().launch != st.launches && cleared <= max_clear) {
                const RecordRing::Run old = ring.pop();
                table[old.leaf] = make_uint2(0, 0);
                cleared += old.size;
                st.leaves_evicted++;
                ok = ring.fits(size, off);
            }
            if (!ok || cleared > max_clear) break;
            if ((r = admit(leaf, off))) return r;
        }
        if (admitted == 0) {  // progress: empty the pool for the first request
            while (!ring.runs.empty()) {
                const RecordRing::Run old = ring.pop();
                table[old.leaf] = make_uint2(0, 0);
                st.leaves_evicted++;
            }
            uint32_t off = 0;
            for (uint32_t leaf : req) {
                if (table[leaf].y) continue;
                if (!ring.fits(c->leaf_extent[leaf].y, off)) break;
                if ((r = admit(leaf, off))) return r;
                break;
            }
            if (admitted == 0) return fail(c, HULK_E_HIP, "no requested leaf fits the pool");
        }
        std::sort(loaded.begin(), loaded.end(), [](const uint2& a, const uint2& b) { return a.x < b.x; });
        for (size_t k = 0; k < loaded.size();) {
            uint32_t lo = loaded[k].x, hi = loaded[k].x + loaded[k].y;
            for (k++; k < loaded.size() && loaded[k].x <= hi; k++) hi = std::max(hi, loaded[k].x + loaded[k].y);
            HIP_TRY(c, hipMemcpyAsync(static_cast<TriRec*>(c->s_pool.p) + lo, c->staging + lo,
                                      (size_t)(hi - lo) * sizeof(TriRec), hipMemcpyHostToDevice, c->stream));
        }
        HIP_TRY(c, hipMemcpyAsync(c->s_table.p, table.data(), table.size() * 8, hipMemcpyHostToDevice, c->stream));
        S.queue_in = static_cast<const uint32_t*>(c->s_queue[cur ^ 1].p);
        S.n_in = n_out;
        cur ^= 1;
    }
    if (!dev_out)
        HIP_TRY(c, hipMemcpyAsync(hits, c->s_hits.p, (size_t)n * sizeof(hulk_hit), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    st.ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    if (stats) *stats = st;
    return HULK_OK;
}

// ------------------------------------------------------ device KD build --

extern "C" int hulk_build_kdtree_device(hulk_ctx* c, uint32_t capacity) {
    if (!c) return HULK_E_ILLEGAL_INPUT;
    if (!c->have_scene) return fail(c, HULK_E_NO_SCENE, "upload the scene before building the tree");
    if (!c->kdb && !(c->kdb = kdb_create())) return fail(c, HULK_E_ALLOC, "device KD builder allocation failed");
    KdBuildResult R;
    std::string err;
    int r = kdb_build(c->kdb, static_cast<const float4*>(c->verts.p), c->n_verts, static_cast<const uint4*>(c->tris.p),
                      c->n_tris, c->box, capacity, c->stream, &R, err);
    if (r) {
        c->err = err;
        return r;
    }
    if (R.max_jdepth > kMaxFullStack) return fail(c, HULK_E_ILLEGAL_INPUT, "tree deeper than 64 junctions");
    if (R.n_nodes >= (1u << 26)) return fail(c, HULK_E_ILLEGAL_INPUT, "more than 2^26 nodes");
    if (R.n_lp >= kMaxLeafRefs) return fail(c, HULK_E_ILLEGAL_INPUT, "2^26 or more leaf polygon references");
    // the previous tree stays installed until here
    c->have_tree = false;
    c->ropes_ready = false;
    if ((r = ensure(c, c->nodes, (size_t)R.n_nodes * 16))) return r;
    if ((r = ensure(c, c->ref_nodes, (size_t)R.n_nodes * 16))) return r;
    if ((r = ensure(c, c->lp, (size_t)std::max<uint32_t>(R.n_lp, 1) * 4))) return r;
    if ((r = ensure(c, c->recs, (size_t)std::max<uint32_t>(R.n_lp, 1) * sizeof(TriRec)))) return r;
    HIP_TRY(c, hipMemcpyAsync(c->nodes.p, kdb_dev_nodes(c->kdb), (size_t)R.n_nodes * 16, hipMemcpyDeviceToDevice,
                              c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->ref_nodes.p, kdb_ref_nodes(c->kdb), (size_t)R.n_nodes * 16,
                              hipMemcpyDeviceToDevice, c->stream));
    if (R.n_lp)
        HIP_TRY(c, hipMemcpyAsync(c->lp.p, kdb_leaf_polys(c->kdb), (size_t)R.n_lp * 4, hipMemcpyDeviceToDevice,
                                  c->stream));
    HIP_TRY(c, launch_build_records(static_cast<const float4*>(c->verts.p), static_cast<const uint4*>(c->tris.p),
                                    static_cast<const uint32_t*>(c->lp.p), R.n_lp, static_cast<TriRec*>(c->recs.p),
                                    c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->n_nodes = R.n_nodes;
    c->n_lp = R.n_lp;
    c->max_jdepth = R.max_jdepth;
    static_assert(kKdTopLevels == kTopLevelsMax, "device layout depth");
    std::memcpy(c->level_prefix, R.level_prefix, sizeof(c->level_prefix));
    c->kdb_last = R;
    c->have_tree = true;
    c->tree_gen++;
    return HULK_OK;
}

extern "C" int hulk_download_kdtree(hulk_ctx* c, hulk_kd_node* nodes, uint32_t node_cap, uint32_t* n_nodes,
                                    uint32_t* leaf_polys, uint32_t lp_cap, uint32_t* n_lp) {
    if (!c) return HULK_E_ILLEGAL_INPUT;
    if (!c->have_tree) return fail(c, HULK_E_NO_TREE, "no KD tree uploaded or built");
    if (n_nodes) *n_nodes = c->n_nodes;
    if (n_lp) *n_lp = c->n_lp;
    if ((nodes && node_cap < c->n_nodes) || (leaf_polys && lp_cap < c->n_lp))
        return fail(c, HULK_E_ILLEGAL_INPUT, "download buffer smaller than the tree");
    if (nodes)
        HIP_TRY(c, hipMemcpyAsync(nodes, c->ref_nodes.p, (size_t)c->n_nodes * 16, hipMemcpyDeviceToHost, c->stream));
    if (leaf_polys && c->n_lp)
        HIP_TRY(c, hipMemcpyAsync(leaf_polys, c->lp.p, (size_t)c->n_lp * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return HULK_OK;
}

extern "C" int hulk_device_build_info(const hulk_ctx* c, uint32_t out[4]) {
    if (!c || !out) return HULK_E_ILLEGAL_INPUT;
    out[0] = c->kdb_last.capacity;
    out[1] = c->kdb_last.retries;
    out[2] = c->kdb_last.levels;
    out[3] = c->kdb_last.max_jdepth;
    return HULK_OK;
}

extern "C" int hulk_get_counters(const hulk_ctx* c, hulk_counters* out) {
    if (!c || !out) return HULK_E_ILLEGAL_INPUT;
    *out = c->cnt;
    return HULK_OK;
}

extern "C" int hulk_reset_counters(hulk_ctx* c) {
    if (!c) return HULK_E_ILLEGAL_INPUT;
    c->cnt = hulk_counters{};
    return HULK_OK;
}

// ------------------------------------------------------------------ batch --

struct hulk_batch {
    hulk_ctx* c = nullptr;
    hulk_opts opts{};
    std::vector<FrameDev> frames;
    DevBuf d_frames, d_mirror;
    Work w;
    uint32_t n = 0, max_bounce = 0;
    uint32_t* rgba = nullptr;
    const float* tbound = nullptr;
    // context state the recorded work depends on (tree / vertex pointers,
    // rope slots, light buffers and the queue sizes derived from them)
    uint32_t tree_gen = 0, lights_gen = 0;
    // frames of several scenes (hulk_batch_create_scenes): the context of
    // each scene-table entry, its state at creation, its mirror flags, and
    // the device table; empty for a one-scene batch
    std::vector<hulk_ctx*> scene_ctx;
    std::vector<uint32_t> scene_tree_gen, scene_lights_gen;
    std::vector<DevBuf> scene_mirror;
    DevBuf d_scenes;
    uint32_t max_stack = 1;
    uint32_t max_lds_nodes = 0;  // largest scene's LDS top tree (0: none)
    bool graph_ready = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

// Every frame of a batch over several scenes in one fused launch: the
// options and counters of the first context, the scene of each chunk of 64
// rays from the scene table; an LDS top tree only for static launches whose
// frames hold multiples of the block's rays (each block then lies in one
// frame and stages its scene's top levels).
int render_scenes(hulk_batch* b, bool count) {
    hulk_ctx* c = b->c;
    const hulk_opts& o = b->opts;
    TraceParams P;
    uint32_t block;
    size_t lds;
    int r = trace_setup(c, o, P, block, lds);
    if (r) return r;
    P.lds_node_count = b->max_lds_nodes;
    size_t stack_words = 0;
    if (o.variant == HULK_VARIANT_FULLSTACK) {
        P.stack_depth = b->max_stack;
        stack_words = 2u * P.stack_depth;
    } else if (o.variant == HULK_VARIANT_LDSSTACK) {
        stack_words = 3u * P.stack_depth;
    }
    lds = (size_t)b->max_lds_nodes * 16u + stack_words * 4u * block;
    if (lds > c->max_lds) return fail(c, HULK_E_ILLEGAL_INPUT, "top tree and stack columns exceed the block's LDS");
    P.source = kSourceFrames;
    P.frames = static_cast<const FrameDev*>(b->d_frames.p);
    P.n_frames = static_cast<uint32_t>(b->frames.size());
    P.n = b->n;
    P.tbound = b->tbound;
    P.hits = static_cast<hulk_hit*>(b->w.hits.p);
    P.rgba = b->rgba;
    P.scenes = static_cast<const SceneDev*>(b->d_scenes.p);
    P.persistent = o.persistent ? 1u : 0u;
    P.work_counter = static_cast<uint32_t*>(b->w.wq.p);
    uint32_t grid;
    if (o.persistent) {
        HIP_TRY(c, launch_fill_u32(P.work_counter, kWorkCounterWords, 0u, c->stream));
        const uint32_t per_cu = std::max<uint32_t>(1u, 2048u / block);
        grid = std::max<uint32_t>(1u, std::min<uint32_t>((b->n + block - 1) / block, c->num_cus * per_cu));
        const uint64_t waves = uint64_t(grid) * (block / 64);
        const uint64_t chunks = (uint64_t(b->n) + 63) / 64;
        P.fetch = static_cast<uint32_t>(std::min<uint64_t>(8, std::max<uint64_t>(1, chunks / (waves * 4))));
    } else {
        grid = std::max<uint32_t>(1u, (b->n + block - 1) / block);
    }
    HIP_TRY(c, launch_trace_scenes(o.variant, P, count, P.leaf_batch, grid, block, lds, c->stream));
    return HULK_OK;
}

int batch_record(hulk_batch* b, bool count) {
    if (!b->scene_ctx.empty()) return render_scenes(b, count);
    hulk_ctx* c = b->c;
    FrameSet fs;
    fs.frames = static_cast<const FrameDev*>(b->d_frames.p);
    fs.n_frames = static_cast<uint32_t>(b->frames.size());
    fs.n = b->n;
    fs.frame_w = b->frames[0].W;
    fs.frame_h = b->frames[0].H;
    for (const FrameDev& fd : b->frames)
        if (fd.W != fs.frame_w || fd.H != fs.frame_h) fs.frame_w = fs.frame_h = 0;
    int r;
    hulk_hit* h0 = static_cast<hulk_hit*>(b->w.hits.p);
    const uint8_t* dm = b->max_bounce ? static_cast<const uint8_t*>(b->d_mirror.p) : nullptr;
    if (use_fused(b->opts, dm != nullptr))
        return render_frame_launch(c, b->opts, fs, b->w, b->tbound, h0, dm ? b->max_bounce : 0, dm, b->rgba, count, false);
    if ((r = trace_primary(c, b->opts, fs, b->w, b->tbound, h0, count, false))) return r;
    return shade_pipeline(c, b->opts, fs, b->w, h0, dm ? b->max_bounce : 0, dm, b->rgba, count);
}

}  // namespace

extern "C" int hulk_batch_create(hulk_ctx* c, const hulk_frame_desc* frames, uint32_t n_frames,
                                 const uint8_t* mirror_mask, const float* tbound_dev, uint32_t* rgba_dev,
                                 uint32_t use_graph, hulk_batch** out) {
    if (!c || !frames || !n_frames || !rgba_dev || !out) return fail(c, HULK_E_ILLEGAL_INPUT, "null argument");
    *out = nullptr;
    hulk_batch* b = new (std::nothrow) hulk_batch;
    if (!b) return HULK_E_ALLOC;
    b->c = c;
    b->opts = c->opts;
    uint64_t off = 0;
    for (uint32_t f = 0; f < n_frames; f++) {
        const hulk_frame_desc& d = frames[f];
        if (!d.width || !d.height || d.max_bounce > 8) {
            delete b;
            return fail(c, HULK_E_ILLEGAL_INPUT, "frame needs width, height > 0 and max_bounce <= 8");
        }
        FrameDev fd;
        fd.scene = 0;
        fd.cam = to_cam(d.cam);
        fd.W = d.width;
        fd.H = d.height;
        fd.offset = static_cast<uint32_t>(off);
        fd.max_bounce = d.max_bounce;
        fd.w_m = udiv_multiplier(d.width);
        fd.tx_m = udiv_multiplier(d.width >> 3);
        fd.ty_m = udiv_multiplier(d.height >> 3);
        fd.fW = static_cast<float>(d.width);
        fd.fH = static_cast<float>(d.height);
        b->frames.push_back(fd);
        b->max_bounce = std::max(b->max_bounce, d.max_bounce);
        off += (uint64_t)d.width * d.height;
        if (off > 0xFFFFFFC0ull) {
            delete b;
            return fail(c, HULK_E_ILLEGAL_INPUT, "batch has too many rays");
        }
    }
    b->n = static_cast<uint32_t>(off);
    b->rgba = rgba_dev;
    b->tbound = tbound_dev;
    int r;
    const bool bounces = b->max_bounce > 0 && mirror_mask;
    if (!bounces) b->max_bounce = 0;
    if ((r = ensure(c, b->d_frames, b->frames.size() * sizeof(FrameDev))) ||
        (r = ensure_work(c, b->w, b->n, c->n_point, bounces, !use_fused(b->opts, bounces)))) {
        hulk_batch_destroy(b);
        return r;
    }
    hipError_t e = hipMemcpyAsync(b->d_frames.p, b->frames.data(), b->frames.size() * sizeof(FrameDev),
                                  hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && bounces) {
        if ((r = ensure(c, b->d_mirror, std::max<uint32_t>(c->n_tris, 1)))) {
            hulk_batch_destroy(b);
            return r;
        }
        e = hipMemcpyAsync(b->d_mirror.p, mirror_mask, c->n_tris, hipMemcpyHostToDevice, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        hulk_batch_destroy(b);
        return hip_fail(c, e, "batch upload");
    }
    b->tree_gen = c->tree_gen;
    b->lights_gen = c->lights_gen;
    // one eager run: builds ropes if needed and checks the launch shapes
    if ((r = batch_record(b, false))) {
        hulk_batch_destroy(b);
        return r;
    }
    // a batch is for replays: its rope traversals take pushed ropes from the start
    if (b->opts.variant == HULK_VARIANT_ROPES && (r = ensure_ropes(c))) {
        hulk_batch_destroy(b);
        return r;
    }
    if (use_graph) {
        e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            r = batch_record(b, false);
            hipGraph_t g = nullptr;
            hipError_t e2 = hipStreamEndCapture(c->stream, &g);
            if (r) {
                if (g) (void)hipGraphDestroy(g);
                hulk_batch_destroy(b);
                return r;
            }
            e = e2;
            b->graph = g;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            hulk_batch_destroy(b);
            return hip_fail(c, e, "graph capture");
        }
        b->graph_ready = true;
    }
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        hulk_batch_destroy(b);
        return hip_fail(c, e, "batch warm-up");
    }
    *out = b;
    return HULK_OK;
}

extern "C" int hulk_batch_launch(hulk_batch* b, uint32_t flags) {
    if (!b) return HULK_E_ILLEGAL_INPUT;
    hulk_ctx* c = b->c;
    bool stale = b->tree_gen != c->tree_gen || b->lights_gen != c->lights_gen;
    for (size_t i = 0; i < b->scene_ctx.size(); i++)
        stale = stale || b->scene_tree_gen[i] != b->scene_ctx[i]->tree_gen ||
                b->scene_lights_gen[i] != b->scene_ctx[i]->lights_gen;
    if (stale)
        return fail(c, HULK_E_ILLEGAL_INPUT,
                    "the context's scene, tree or lights changed after the batch was created: create it again");
    const bool count = (flags & HULK_FLAG_COUNT) != 0;
    int r;
    if (count && (r = zero_stats(c))) return r;
    HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
    if (b->graph_ready && !count) {
        HIP_TRY(c, hipGraphLaunch(b->exec, c->stream));
    } else if ((r = batch_record(b, count))) {
        return r;
    }
    HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
    if (count && (r = read_stats(c))) return r;
    return finish(c, flags, true);
}

extern "C" uint32_t hulk_batch_ray_count(const hulk_batch* b) { return b ? b->n : 0; }

extern "C" int hulk_batch_primary_hits(hulk_batch* b, hulk_hit* out, uint32_t flags) {
    if (!b || !out) return HULK_E_ILLEGAL_INPUT;
    hulk_ctx* c = b->c;
    const hipMemcpyKind kind = (flags & HULK_FLAG_DEVICE_PTRS) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    HIP_TRY(c, hipMemcpyAsync(out, b->w.hits.p, (size_t)b->n * sizeof(hulk_hit), kind, c->stream));
    return finish(c, flags, false);
}

extern "C" void hulk_batch_destroy(hulk_batch* b) {
    if (!b) return;
    if (b->c && b->c->stream) (void)hipStreamSynchronize(b->c->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    release(b->d_frames);
    release(b->d_mirror);
    release(b->d_scenes);
    for (DevBuf& m : b->scene_mirror) release(m);
    release_work(b->w);
    delete b;
}

extern "C" int hulk_batch_create_scenes(hulk_ctx* const* ctxs, uint32_t n_ctxs, const hulk_frame_desc* frames,
                                        uint32_t n_frames, const uint8_t* const* mirror_masks,
                                        const float* tbound_dev, uint32_t* rgba_dev, uint32_t use_graph,
                                        hulk_batch** out) {
    if (!ctxs || !n_ctxs || !ctxs[0]) return HULK_E_ILLEGAL_INPUT;
    hulk_ctx* c = ctxs[0];
    if (!frames || !n_frames || !rgba_dev || !out) return fail(c, HULK_E_ILLEGAL_INPUT, "null argument");
    *out = nullptr;
    const hulk_opts& o = c->opts;
    if (o.pipeline != HULK_PIPELINE_FUSED && o.pipeline != HULK_PIPELINE_AUTO)
        return fail(c, HULK_E_UNSUPPORTED, "a batch over several scenes runs the fused frame pipeline");
    if (o.tri_test == HULK_TRI_MOLLER_TRUMBORE)
        return fail(c, HULK_E_UNSUPPORTED, "a batch over several scenes runs the reference triangle test");
    for (uint32_t i = 0; i < n_ctxs; i++) {
        if (!ctxs[i]) return fail(c, HULK_E_ILLEGAL_INPUT, "null context");
        if (ctxs[i]->opts.device != o.device)
            return fail(c, HULK_E_ILLEGAL_INPUT, "the scenes of a batch must live on one device");
    }
    hulk_batch* b = new (std::nothrow) hulk_batch;
    if (!b) return HULK_E_ALLOC;
    b->c = c;
    b->opts = o;
    // an LDS top tree needs every block inside one frame: static blocks and
    // frames holding multiples of the block's rays
    bool lds_ok = !o.persistent;
    for (uint32_t f = 0; f < n_frames; f++)
        lds_ok = lds_ok && ((uint64_t)frames[f].width * frames[f].height) % o.block_size == 0;
    if (!lds_ok) b->opts.lds_top_levels = 0;
    uint64_t off = 0;
    for (uint32_t f = 0; f < n_frames; f++) {
        const hulk_frame_desc& d = frames[f];
        const uint64_t px = (uint64_t)d.width * d.height;
        if (!d.width || !d.height || d.max_bounce > 8 || d.reserved >= n_ctxs || (px % 64) != 0) {
            delete b;
            return fail(c, HULK_E_ILLEGAL_INPUT,
                        "frame needs width, height > 0, a multiple of 64 pixels, max_bounce <= 8 and a scene index "
                        "(reserved) below the number of contexts");
        }
        FrameDev fd;
        fd.cam = to_cam(d.cam);
        fd.W = d.width;
        fd.H = d.height;
        fd.offset = static_cast<uint32_t>(off);
        fd.max_bounce = mirror_masks && mirror_masks[d.reserved] ? d.max_bounce : 0;
        fd.scene = d.reserved;
        fd.w_m = udiv_multiplier(d.width);
        fd.tx_m = udiv_multiplier(d.width >> 3);
        fd.ty_m = udiv_multiplier(d.height >> 3);
        fd.fW = static_cast<float>(d.width);
        fd.fH = static_cast<float>(d.height);
        b->frames.push_back(fd);
        b->max_bounce = std::max(b->max_bounce, fd.max_bounce);
        off += px;
        if (off > 0xFFFFFFC0ull) {
            delete b;
            return fail(c, HULK_E_ILLEGAL_INPUT, "batch has too many rays");
        }
    }
    b->n = static_cast<uint32_t>(off);
    b->rgba = rgba_dev;
    b->tbound = tbound_dev;
    b->tree_gen = c->tree_gen;
    b->lights_gen = c->lights_gen;
    b->scene_ctx.assign(ctxs, ctxs + n_ctxs);
    b->scene_mirror.resize(n_ctxs);
    std::vector<SceneDev> table(n_ctxs);
    int r;
    for (uint32_t i = 0; i < n_ctxs; i++) {
        hulk_ctx* s = ctxs[i];
        TraceParams Pi;
        uint32_t block;
        size_t lds;
        if ((r = trace_setup(s, b->opts, Pi, block, lds))) {  // builds rope slots where needed
            if (s != c) c->err = s->err;
            hulk_batch_destroy(b);
            return r;
        }
        SceneDev& e = table[i];
        std::memset(&e, 0, sizeof(e));
        e.nodes = Pi.nodes;
        e.recs = Pi.recs;
        e.rope_nodes = Pi.rope_nodes;
        e.rope_slots = Pi.rope_slots;
        e.lights = static_cast<const float4*>(s->d_lights.p);
        e.n_lights = static_cast<uint32_t>(s->lights.size());
        e.verts = Pi.verts;
        e.tris = Pi.tris;
        e.n_tris = Pi.n_tris;
        e.box = Pi.box;
        e.node_bytes = Pi.node_bytes;
        e.rec_bytes = Pi.rec_bytes;
        e.rope_node_bytes = Pi.rope_node_bytes;
        e.guard = Pi.guard;
        e.stack_depth = Pi.stack_depth;
        e.lds_node_count = Pi.lds_node_count;
        b->max_stack = std::max(b->max_stack, e.stack_depth);
        b->max_lds_nodes = std::max(b->max_lds_nodes, e.lds_node_count);
        if (mirror_masks && mirror_masks[i] && b->max_bounce) {
            if ((r = ensure(c, b->scene_mirror[i], std::max<uint32_t>(s->n_tris, 1)))) {
                hulk_batch_destroy(b);
                return r;
            }
            hipError_t em = s->n_tris ? hipMemcpyAsync(b->scene_mirror[i].p, mirror_masks[i], s->n_tris,
                                                       hipMemcpyHostToDevice, c->stream)
                                      : hipSuccess;
            if (em != hipSuccess) {
                hulk_batch_destroy(b);
                return hip_fail(c, em, "mirror upload");
            }
            e.mirror = static_cast<const uint8_t*>(b->scene_mirror[i].p);
        }
        b->scene_tree_gen.push_back(s->tree_gen);
        b->scene_lights_gen.push_back(s->lights_gen);
    }
    if ((r = ensure(c, b->d_frames, b->frames.size() * sizeof(FrameDev))) ||
        (r = ensure(c, b->d_scenes, table.size() * sizeof(SceneDev))) ||
        (r = ensure_work(c, b->w, b->n, 0, false, false))) {
        hulk_batch_destroy(b);
        return r;
    }
    hipError_t e = hipMemcpyAsync(b->d_frames.p, b->frames.data(), b->frames.size() * sizeof(FrameDev),
                                  hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->d_scenes.p, table.data(), table.size() * sizeof(SceneDev), hipMemcpyHostToDevice,
                           c->stream);
    for (uint32_t i = 0; i < n_ctxs && e == hipSuccess; i++)
        if (ctxs[i] != c) e = hipStreamSynchronize(ctxs[i]->stream);  // their uploads and rope builds
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        hulk_batch_destroy(b);
        return hip_fail(c, e, "batch upload");
    }
    if ((r = batch_record(b, false))) {
        hulk_batch_destroy(b);
        return r;
    }
    if (use_graph) {
        e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            r = batch_record(b, false);
            hipGraph_t g = nullptr;
            hipError_t e2 = hipStreamEndCapture(c->stream, &g);
            if (r) {
                if (g) (void)hipGraphDestroy(g);
                hulk_batch_destroy(b);
                return r;
            }
            e = e2;
            b->graph = g;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            hulk_batch_destroy(b);
            return hip_fail(c, e, "graph capture");
        }
        b->graph_ready = true;
    }
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        hulk_batch_destroy(b);
        return hip_fail(c, e, "batch warm-up");
    }
    *out = b;
    return HULK_OK;
}

// --------------------------------------------------------------- autotune --

namespace {

struct Cand {
    int32_t variant;
    uint32_t block, depth, lds, persistent, pipeline, leaf_batch, coop;
};

}  // namespace

// Per-scene search over traversal variant, block size, short-stack depth,
// LDS top-tree depth and persistent waves, timing the primary trace of the
// given camera with HIP events: stage 1 times every allowed variant at its
// default shape, stage 2 sweeps the shape of the two fastest variants.
// The fastest candidate becomes the context's options.
extern "C" int hulk_autotune(hulk_ctx* c, const hulk_tune_request* req, hulk_tune_entry* table, uint32_t cap,
                             uint32_t* n_out, hulk_opts* best_out) {
    if (!c || !req) return fail(c, HULK_E_ILLEGAL_INPUT, "null argument");
    const hulk_camera* cam = &req->cam;
    const uint32_t W = req->width, H = req->height;
    uint32_t reps = req->reps, variant_mask = req->variant_mask;
    if (!W || !H) return fail(c, HULK_E_ILLEGAL_INPUT, "width and height must be > 0");
    if (req->objective > HULK_TUNE_FRAME) return fail(c, HULK_E_ILLEGAL_INPUT, "unknown autotune objective");
    if ((uint64_t)W * H > 0xFFFFFFC0ull) return fail(c, HULK_E_ILLEGAL_INPUT, "image too large");
    if (req->max_bounce > 8) return fail(c, HULK_E_ILLEGAL_INPUT, "max_bounce must be <= 8");
    const bool frame = req->objective == HULK_TUNE_FRAME;
    const bool bounces = frame && req->max_bounce > 0 && req->mirror_mask;
    const uint32_t max_bounce = bounces ? req->max_bounce : 0;
    if (!variant_mask) variant_mask = (1u << HULK_VARIANT_FULLSTACK) | (1u << HULK_VARIANT_SHORTSTACK) |
                                      (1u << HULK_VARIANT_LDSSTACK) | (1u << HULK_VARIANT_RESTART) |
                                      (1u << HULK_VARIANT_ROPES);
    reps = std::max<uint32_t>(reps, 1);
    const uint32_t n = W * H;
    FrameSet fs;
    fs.cam = to_cam(*cam);
    fs.W = W;
    fs.H = H;
    fs.n = n;
    int r;
    if ((r = ensure_work(c, c->work, n, c->n_point, bounces, frame))) return r;
    if (frame && (r = ensure(c, c->rgba, (size_t)n * 4))) return r;
    const uint8_t* dm = nullptr;
    if (bounces && (r = stage_in(c, req->mirror_mask, c->n_tris, false, c->mirror, &dm))) return r;
    hulk_hit* out = static_cast<hulk_hit*>(c->work.hits.p);
    uint32_t* px = static_cast<uint32_t*>(c->rgba.p);
    const uint32_t pipe0 = c->opts.pipeline == HULK_PIPELINE_QUEUED ? HULK_PIPELINE_QUEUED : HULK_PIPELINE_FUSED;
    std::vector<std::pair<Cand, float>> done;
    auto measure = [&](const Cand& k, float& ms, uint32_t n_reps, bool record) -> int {
        hulk_opts o = c->opts;
        o.variant = k.variant;
        o.block_size = k.block;
        o.short_stack_depth = k.depth;
        o.lds_top_levels = k.lds;
        o.persistent = k.persistent;
        o.pipeline = k.pipeline;
        o.leaf_batch = k.leaf_batch;
        o.coop_lanes = k.coop;
        TraceParams P;
        uint32_t block;
        size_t lds;
        if (check_opts(c, o) || trace_setup(c, o, P, block, lds)) return 1;  // shape not allowed: skip
        if (o.pipeline == HULK_PIPELINE_REFILL &&
            refill_lds_bytes(lds, block, static_cast<uint32_t>(c->lights.size())) > c->max_lds)
            return 1;
        auto once = [&]() -> int {
            if (!frame) return trace_primary(c, o, fs, c->work, nullptr, out, false, false);
            if (use_fused(o, bounces))
                return render_frame_launch(c, o, fs, c->work, nullptr, out, max_bounce, dm, px, false, false);
            int q = trace_primary(c, o, fs, c->work, nullptr, out, false, false);
            return q ? q : shade_pipeline(c, o, fs, c->work, out, max_bounce, dm, px, false);
        };
        int rr;
        if ((rr = once())) return rr;
        HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
        for (uint32_t i = 0; i < n_reps; i++)
            if ((rr = once())) return rr;
        HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        HIP_TRY(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        ms /= n_reps;
        if (record) done.push_back({k, ms});
        return 0;
    };
    // stage 1
    std::vector<std::pair<float, Cand>> firsts;
    for (int v = HULK_VARIANT_FULLSTACK; v < HULK_VARIANT_COUNT; v++) {
        if (!(variant_mask & (1u << v))) continue;
        Cand k{v, 256, 6, 0, 0, pipe0, 1, 0};
        float ms;
        r = measure(k, ms, reps, true);
        if (r > 1) return r;
        if (r == 0) firsts.push_back({ms, k});
    }
    if (firsts.empty()) return fail(c, HULK_E_ILLEGAL_INPUT, "no variant in the mask can run");
    std::sort(firsts.begin(), firsts.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    // stage 2
    const uint32_t blocks[] = {64, 128, 256};
    const uint32_t depths[] = {4, 6, 8, 12, 16};
    const uint32_t ldss[] = {0, 6, 8, 10, 12};
    for (size_t t = 0; t < std::min<size_t>(2, firsts.size()); t++) {
        const int v = firsts[t].second.variant;
        const bool shortv = v == HULK_VARIANT_SHORTSTACK || v == HULK_VARIANT_LDSSTACK;
        const bool ldsv = v != HULK_VARIANT_ROPES;
        for (uint32_t bs : blocks)
            for (uint32_t dp : depths) {
                if (!shortv && dp != 6) continue;
                for (uint32_t ld : ldss) {
                    if (!ldsv && ld) continue;
                    for (uint32_t ps = 0; ps < 2; ps++)
                        for (uint32_t lb = 1; lb <= 2; lb++) {
                            Cand k{v, bs, dp, ld, ps, pipe0, lb, 0};
                            if (bs == 256 && dp == 6 && ld == 0 && ps == 0 && lb == 1) continue;  // stage 1
                            // persistent + LDS top tree pays off; skip the
                            // non-persistent LDS shapes beyond depth 8
                            if (!ps && ld > 8) continue;
                            // two records in flight: default stack / LDS depths only
                            if (lb == 2 && (dp != 6 || (ld != 0 && ld != 8))) continue;
                            float ms;
                            r = measure(k, ms, reps, true);
                            if (r > 1) return r;
                        }
                }
            }
    }
    size_t bi = 0;
    for (size_t i = 1; i < done.size(); i++)
        if (done[i].second < done[bi].second) bi = i;
    if (frame) {  // stage 3: the other frame pipeline on the best shape, and lane refill (ropes, LDS short stack)
        Cand k = done[bi].first;
        k.pipeline = k.pipeline == HULK_PIPELINE_FUSED ? HULK_PIPELINE_QUEUED : HULK_PIPELINE_FUSED;
        float ms;
        r = measure(k, ms, reps, true);
        if (r > 1) return r;
        const uint32_t coops[] = {0, 16, 64};
        if (variant_mask & (1u << HULK_VARIANT_ROPES))
            for (uint32_t bs : blocks)
                for (uint32_t lb = 1; lb <= 2; lb++)
                    for (uint32_t cl : coops) {
                        Cand kr{HULK_VARIANT_ROPES, bs, 6, 0, 1, HULK_PIPELINE_REFILL, lb, cl};
                        r = measure(kr, ms, reps, true);
                        if (r > 1) return r;
                    }
        if (variant_mask & (1u << HULK_VARIANT_LDSSTACK)) {
            // stack depth and top levels of the best shape when it is the LDS short stack
            const Cand& bk = done[bi].first;
            const bool lds_best = bk.variant == HULK_VARIANT_LDSSTACK;
            const uint32_t depth = lds_best ? bk.depth : 6u, top = lds_best ? bk.lds : 6u;
            const uint32_t coops_s[] = {0, 64};
            for (uint32_t bs : blocks)
                for (uint32_t lb = 1; lb <= 2; lb++)
                    for (uint32_t cl : coops_s) {
                        Cand kr{HULK_VARIANT_LDSSTACK, bs, depth, top, 1, HULK_PIPELINE_REFILL, lb, cl};
                        r = measure(kr, ms, reps, true);
                        if (r > 1) return r;
                    }
        }
    }
    // stage 4: the four fastest again, with four times the repetitions, so
    // that timing noise between near-equal shapes does not decide
    std::vector<size_t> order(done.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = i;
    const size_t top = std::min<size_t>(4, order.size());
    std::partial_sort(order.begin(), order.begin() + top, order.end(),
                      [&](size_t a, size_t b) { return done[a].second < done[b].second; });
    float best_ms = 0.0f;
    for (size_t t = 0; t < top; t++) {
        float ms;
        r = measure(done[order[t]].first, ms, 4 * reps, false);
        if (r > 1) return r;
        if (r == 0 && (t == 0 || ms < best_ms)) {
            best_ms = ms;
            bi = order[t];
        }
    }
    const Cand b = done[bi].first;
    c->opts.variant = b.variant;
    c->opts.block_size = b.block;
    c->opts.short_stack_depth = b.depth;
    c->opts.lds_top_levels = b.lds;
    c->opts.persistent = b.persistent;
    c->opts.leaf_batch = b.leaf_batch;
    c->opts.coop_lanes = b.coop;
    if (frame) c->opts.pipeline = b.pipeline;
    if (best_out) *best_out = c->opts;
    if (table) {
        for (size_t i = 0; i < done.size() && i < cap; i++) {
            table[i].variant = done[i].first.variant;
            table[i].block_size = done[i].first.block;
            table[i].short_stack_depth = done[i].first.depth;
            table[i].lds_top_levels = done[i].first.lds;
            table[i].persistent = done[i].first.persistent;
            table[i].pipeline = done[i].first.pipeline;
            table[i].leaf_batch = done[i].first.leaf_batch;
            table[i].coop_lanes = done[i].first.coop;
            table[i].ms = done[i].second;
        }
    }
    if (n_out) *n_out = static_cast<uint32_t>(done.size());
    return HULK_OK;
}
