/*
 * hulk_abi.h - C ABI of the MI355X (gfx950) traverse-and-shade library
 * (libhulk.so).
 *
 * This header is the drop-in boundary for hulkTracer's GPU trace stage.
 * It replaces the OpenCL seam that `GLVisitorGPUTrace` drives:
 *   - kernel args of `gpu_trace` (HulkEngine/src/GPUTrace_stackless.cl:122-133,
 *     bound at HulkEngine/src/GLVisitorGPUTrace.cpp:342-378),
 *   - `ShadePicture` (HulkEngine/src/shadePicture.cl:7-11, bound at
 *     GLVisitorGPUTrace.cpp:385-398),
 *   - `SetInitialValues` (HulkEngine/src/GPU_setValues.cl:4-5),
 *   - the host frame sequence `traceScene` / `launchTraceIterations`
 *     (GLVisitorGPUTrace.cpp:296-489).
 *
 * Conventions
 *   - Every entry point returns a hulk_status (0 = ok).  The codes extend the
 *     reference's EHulkResult (HulkEngine/public/hulk_assert.h:42-50) with HIP
 *     errors; hulk_last_error() gives a message.  Nothing fails silently.
 *   - The caller owns host buffers; the library owns device buffers.  Output
 *     pointers are host pointers unless HULK_FLAG_DEVICE_PTRS is set, in which
 *     case every pointer argument of that call is a device pointer.
 *   - One context = one HIP stream.  Calls on one context are serialized on
 *     that stream; a context must not be used from two threads at once.
 *     Contexts are independent; the only state shared between them is the
 *     per-GPU scratch array of hulk_render_frame_static (see there).
 *   - Structs below are binary-identical to the reference's shared structs
 *     (HulkEngine/public/SharedStructs.hpp), so the output of the reference's
 *     KDTreeCPU passes through unchanged.
 */
#ifndef HULK_ABI_H
#define HULK_ABI_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- types -- */

/* float4 vertex / vector; vertices carry w = 1 (Scene.hpp:68-70). */
typedef struct hulk_float4 { float x, y, z, w; } hulk_float4;

/* TriangleIndices (SharedStructs.hpp:37-41): three vertex indices + padding. */
typedef struct hulk_tri { uint32_t idx[4]; } hulk_tri;

/* Box (SharedStructs.hpp:65-69). */
typedef struct hulk_box { hulk_float4 min, max; } hulk_box;

/* KDTreeNode (SharedStructs.hpp:113-137): flags bits 0-1 axis
 * (0 = YZ plane / x, 1 = XZ / y, 2 = XY / z), 4 = leaf, 8 = junction.
 * Root = node 0; child index 0 = empty child. */
typedef struct hulk_kd_node {
    uint32_t flags;
    union {
        struct { uint32_t start, count, leaf_id; } leaf;
        struct { uint32_t above, below; float plane; } junction;
    } u;
} hulk_kd_node;

#define HULK_NODE_AXIS_MASK 3u
#define HULK_NODE_LEAF      4u
#define HULK_NODE_JUNCTION  8u

/* CameraPos (SharedStructs.hpp:99-105), field order kept. */
typedef struct hulk_camera {
    hulk_float4 from;
    hulk_float4 vertical_factor;    /* multiplies x / W (GPU_common.cl:24) */
    hulk_float4 horizontal_factor;  /* multiplies y / H (GPU_common.cl:25) */
    hulk_float4 lower_left;
} hulk_camera;

/* Light (SharedStructs.hpp:53-63): type in color.w (0 = directional,
 * 1 = point); data = direction or position. */
typedef struct hulk_light { hulk_float4 color; hulk_float4 data; } hulk_light;

#define HULK_LIGHT_DIRECTIONAL 0.0f
#define HULK_LIGHT_POINT       1.0f

/* Ray (SharedStructs.hpp:71-76).  Directions need not be normalized. */
typedef struct hulk_ray { hulk_float4 start; hulk_float4 dir; } hulk_ray;

/* Hit record.  The reference writes float4(normal.xyz, t)
 * (GPUTrace_stackless.cl:163-165); this record adds the global triangle id
 * (m_leafPolygons[start+i], KDTreeCPU.cpp:267) and the barycentrics
 * u = s1/det, v = s2/det of the triangle test (SharedFunctions.hpp:88-91).
 * Miss: t = FLT_MAX, tri = HULK_NO_HIT. */
typedef struct hulk_hit {
    float t, u, v;
    uint32_t tri;
    float nx, ny, nz;
    uint32_t flags;   /* HULK_HITF_* */
} hulk_hit;

#define HULK_NO_HIT 0xFFFFFFFFu
#define HULK_HITF_BOUNDED_FALLBACK 1u  /* bounded pass missed; re-run unbounded */

/* ------------------------------------------------------------- options -- */

/* Traversal variants.  Names follow the reference --kernelType values
 * (HulkEngine/src/Settings.cpp:50). */
typedef enum hulk_variant {
    HULK_VARIANT_BRUTEFORCE = 0,  /* GPUTrace_bruteforce.cl:27-42          */
    HULK_VARIANT_FULLSTACK  = 1,  /* GPUTrace.cl:6-90                       */
    HULK_VARIANT_SHORTSTACK = 2,  /* GPUTrace_shortStack.cl:6-105 (regs)    */
    HULK_VARIANT_LDSSTACK   = 3,  /* GPUTrace_LDSshortStack.cl:7-105 (LDS)  */
    HULK_VARIANT_RESTART    = 4,  /* GPUTrace_stackless.cl:6-120 kd-restart */
    HULK_VARIANT_ROPES      = 5,  /* stackless rope traversal (new)         */
    HULK_VARIANT_COUNT      = 6
} hulk_variant;

typedef struct hulk_opts {
    int32_t  device;            /* HIP device ordinal                        */
    int32_t  variant;           /* hulk_variant                              */
    uint32_t block_size;        /* threads per workgroup: 64..1024, mult. 64 */
    uint32_t short_stack_depth; /* 1..16, short-stack / LDS-stack variants   */
    uint32_t lds_top_levels;    /* 0..12: top tree levels staged in LDS      */
    uint32_t persistent;        /* 1 = persistent waves with atomic fetch    */
    uint32_t pipeline;          /* hulk_pipeline: frame launch model         */
    uint32_t leaf_batch;        /* triangle records in flight per lane in a
                                   leaf: 0 or 1 = one, 2 = two (for scenes
                                   whose records do not stay in L1)          */
    uint32_t coop_lanes;        /* lane-refill pipeline: when at most this
                                   many lanes of a wave are at a leaf, the
                                   whole wave tests their leaves, the records
                                   of all of them packed 64 per pass;
                                   0 = never, 64 = at every leaf step       */
    uint32_t tri_test;          /* hulk_tri_test: HULK_TRI_REFERENCE (the
                                   reference's checkRayToPoly, bit-exact) or
                                   HULK_TRI_MOLLER_TRUMBORE (optional; one
                                   record in flight, not with the streamed
                                   trace)                                    */
    void*    stream;            /* hipStream_t to use; NULL = own stream     */
} hulk_opts;

/* Triangle test.  REFERENCE restates checkRayToPoly
 * (SharedFunctions.hpp:47-102: plane, then a 2x2 Gram system) bit for bit.
 * MOLLER_TRUMBORE (Moller and Trumbore, JGT 1997, non-culling branch,
 * with the reference's parallel cut |n.d| < FLT_EPSILON in place of the
 * paper's absolute EPSILON) is an option with no counterpart in the
 * reference: its nearest hits agree with the reference test's except at
 * ties and edge grazes, and its t, u, v round differently. */
typedef enum hulk_tri_test {
    HULK_TRI_REFERENCE = 0,
    HULK_TRI_MOLLER_TRUMBORE = 1
} hulk_tri_test;

/* Launch model of a frame (render_frame, batches).  QUEUED: primary trace,
 * then per generation compacted shadow and mirror queues, each traced by
 * its own launch.  FUSED: one launch in which every ray runs its whole
 * path (primary hit, shadow rays, mirror generations) in registers.  Both
 * give identical pixels and hit records.  AUTO: FUSED. */
typedef enum hulk_pipeline {
    HULK_PIPELINE_AUTO = 0,
    HULK_PIPELINE_QUEUED = 1,
    HULK_PIPELINE_FUSED = 2,
    HULK_PIPELINE_REFILL = 3   /* FUSED with lanes refilled from the ray queue
                                  as their paths end (ropes or LDS short-stack
                                  variant, blocks of at most 256) */
} hulk_pipeline;

/* Call flags. */
#define HULK_FLAG_DEVICE_PTRS  1u   /* all pointer args are device pointers */
#define HULK_FLAG_ANYHIT       2u   /* shadow mode: stop at first t < tmax   */
#define HULK_FLAG_COUNT        4u   /* accumulate node / triangle counters   */
#define HULK_FLAG_NO_SYNC      8u   /* do not synchronize the stream on exit */

/* Shading modes for hulk_shade / hulk_render. */
#define HULK_SHADE_FUSED    0u  /* GPUTrace.cl:122-160: shadow rays, miss = 0      */
#define HULK_SHADE_PICTURE  1u  /* shadePicture.cl:14-54: no shadow, miss = 0.1    */

typedef enum hulk_status {
    HULK_OK = 0,
    HULK_E_ALLOC = 1,         /* e_allocFail    */
    HULK_E_FILE_OPEN = 2,     /* e_fileOpenFail */
    HULK_E_FILE_CLOSE = 3,    /* e_fileCloseFail*/
    HULK_E_ILLEGAL_INPUT = 4, /* e_illegalInput */
    HULK_E_NOT_FOUND = 5,     /* e_nodeNotFound */
    HULK_E_HIP = 100,         /* HIP runtime error, message via hulk_last_error */
    HULK_E_NO_SCENE = 101,
    HULK_E_NO_TREE = 102,
    HULK_E_UNSUPPORTED = 103
} hulk_status;

typedef struct hulk_ctx hulk_ctx;

/* Counters of the last counted call (HULK_FLAG_COUNT).  Definitions follow
 * the reference Statistics counters (HulkEngine/src/Statistics.hpp:58-60,
 * 112-120), counted in every variant. */
typedef struct hulk_counters {
    uint64_t rays;          /* nearest-hit rays with a hit record (primary, mirror) */
    uint64_t node_visits;   /* node fetches (all rays)                      */
    uint64_t leaf_visits;
    uint64_t tri_tests;
    uint64_t hits;
    uint64_t fallbacks;     /* bounded rays re-run unbounded                */
    uint64_t shadow_rays;   /* point-light shadow rays (nearest hit, t only) */
    float    last_kernel_ms;/* device time of the last trace kernel (events)*/
    float    last_shade_ms;
} hulk_counters;

/* ------------------------------------------------------- context calls -- */

/* Replaces GLVisitorGPUTrace::initializeVisitor (GLVisitorGPUTrace.cpp:94-161). */
int hulk_create(hulk_ctx** out, const hulk_opts* opts);
void hulk_destroy(hulk_ctx* ctx);
const char* hulk_last_error(const hulk_ctx* ctx);
int hulk_set_options(hulk_ctx* ctx, const hulk_opts* opts);
int hulk_get_options(const hulk_ctx* ctx, hulk_opts* opts);
void* hulk_stream(hulk_ctx* ctx);
int hulk_sync(hulk_ctx* ctx);

/* Scene arrays (Scene::m_vertices / m_vertexIndices / m_sceneBox,
 * Scene.hpp:66-76).  Replaces the vertex/index uploads of
 * GLVisitorGPUTrace::buildTree (GLVisitorGPUTrace.cpp:256-294). */
int hulk_upload_scene(hulk_ctx* ctx, const hulk_float4* verts, uint32_t n_verts,
                      const hulk_tri* tris, uint32_t n_tris, const hulk_box* scene_box);

/* New positions for the uploaded scene's vertices (same count, same
 * triangles), e.g. the next frame of an animation: the vertex write of
 * the per-frame buildTree (GLVisitorGPUTrace.cpp:256-294) without the
 * triangle upload and its index check.  The tree must be uploaded or built
 * again before tracing.  flags: HULK_FLAG_DEVICE_PTRS (verts on the
 * device). */
int hulk_update_vertices(hulk_ctx* ctx, const hulk_float4* verts, uint32_t n_verts, const hulk_box* scene_box,
                         uint32_t flags);

/* KD tree as produced by KDTreeCPU::buildKDTree (m_treeNodes, m_leafPolygons;
 * KDTreeCPU.hpp:128-132).  The library derives its own device layout
 * (per-leaf triangle records, LDS top tree, ropes).  Replaces
 * buildPerLeafData + the node upload (GLVisitorGPUTrace.cpp:163-294).
 * HULK_E_ILLEGAL_INPUT for 2^26 or more nodes or leaf polygon references, a
 * node flagged both leaf and junction, a cycle, a shared or unreachable node,
 * or an index out of range; flag bits above the axis / leaf / junction bits
 * are ignored. */
int hulk_upload_kdtree(hulk_ctx* ctx, const hulk_kd_node* nodes, uint32_t n_nodes,
                       const uint32_t* leaf_polys, uint32_t n_leaf_polys);

/* Lights (Scene::m_lights); replaces the light write of traceScene
 * (GLVisitorGPUTrace.cpp:302-305). */
int hulk_set_lights(hulk_ctx* ctx, const hulk_light* lights, uint32_t n_lights);

/* Nearest-hit primary rays of a pinhole camera, one per pixel, index
 * y*W + x.  Replaces SetInitialValues + gpu_trace (+ relaunch loop)
 * (GLVisitorGPUTrace.cpp:320-382, 412-489).  tbound (nullable, W*H floats):
 * learned bound; rays without a hit below it are re-run unbounded so the
 * result equals the unbounded result. */
int hulk_trace_camera(hulk_ctx* ctx, const hulk_camera* cam, uint32_t width, uint32_t height,
                      const float* tbound, hulk_hit* hits, uint32_t flags);

/* Nearest-hit (or any-hit with HULK_FLAG_ANYHIT) for arbitrary rays. */
int hulk_trace_rays(hulk_ctx* ctx, const hulk_ray* rays, uint32_t n_rays,
                    const float* tbound, hulk_hit* hits, uint32_t flags);

/* Shade a hit buffer produced by hulk_trace_camera with the same camera:
 * Lambert + ambient 0.1, point-light shadow rays from P + 0.001*L
 * (HULK_SHADE_FUSED, GPUTrace.cl:122-160) or the live ShadePicture rule
 * (HULK_SHADE_PICTURE, shadePicture.cl:14-54).  Output packed
 * R | G<<8 | B<<16 per pixel (FrameBuffer.hpp:196-217). */
int hulk_shade(hulk_ctx* ctx, const hulk_camera* cam, uint32_t width, uint32_t height,
               const hulk_hit* hits, uint32_t* rgba8, uint32_t shade_mode, uint32_t flags);

/* Whole frame: primary trace + shadow rays + up to max_bounce mirror
 * generations with per-generation ray compaction, then shading.
 * mirror_mask (nullable, one byte per triangle, 1 = mirror).  Outputs the
 * packed image and (nullable) the primary hit buffer. */
int hulk_render_frame(hulk_ctx* ctx, const hulk_camera* cam, uint32_t width, uint32_t height,
                      uint32_t max_bounce, const uint8_t* mirror_mask,
                      const float* tbound, uint32_t* rgba8, hulk_hit* primary_hits,
                      uint32_t flags);

/* Allocation-free whole frame.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The frame of
 * hulk_render_frame (same pixels, primary hit records and counters) with the
 * context's FUSED or REFILL pipeline; QUEUED options give
 * HULK_E_UNSUPPORTED.  Every pointer is a device pointer; rgba8 is required.
 * The call allocates nothing: the persistent waves' work counters live in a
 * static device array of the library (one per GPU, held until the process
 * ends), which also holds the primary hit records when primary_hits is NULL;
 * such a frame may have at most HULK_STATIC_MAX_RAYS rays (else
 * HULK_E_ILLEGAL_INPUT).  The array is shared by every context on a GPU and
 * keeps the last call's contents: calls that overlap on one GPU (two
 * streams, two threads) race.  With the rope variant the tree's rope slots
 * must already exist (any earlier rope trace on the context builds them),
 * else HULK_E_UNSUPPORTED; the call may launch the build of their pushed
 * ropes (a kernel, no allocation), which a tree's second rope trace runs. */
#define HULK_STATIC_MAX_RAYS 262144u
int hulk_render_frame_static(hulk_ctx* ctx, const hulk_camera* cam, uint32_t width, uint32_t height,
                             uint32_t max_bounce, const uint8_t* mirror_mask_dev, const float* tbound_dev,
                             uint32_t* rgba8_dev, hulk_hit* primary_hits_dev, uint32_t flags);

int hulk_get_counters(const hulk_ctx* ctx, hulk_counters* out);
int hulk_reset_counters(hulk_ctx* ctx);

/* ------------------------------------------------------- batched frames -- */

/* One frame of a batch (camera, resolution, mirror generations). */
typedef struct hulk_frame_desc {
    hulk_camera cam;
    uint32_t width, height;
    uint32_t max_bounce;
    uint32_t reserved;   /* hulk_batch_create_scenes: scene index; else 0 */
} hulk_frame_desc;

typedef struct hulk_batch hulk_batch;

/* Many frames of uneven size and bounce depth on the context's scene, run
 * as one pipeline: one primary launch over every frame's rays, then per
 * generation one compacted shadow-ray launch, one shading pass and one
 * compacted mirror-ray launch for all frames together.  With use_graph the
 * whole pipeline is captured once into a HIP graph and each
 * hulk_batch_launch replays it.  The batch owns its work buffers and keeps
 * the options the context had at creation.  rgba_dev / tbound_dev are
 * device pointers; frame f's pixels start at the sum of the earlier
 * frames' width*height.  Replaces the per-frame traceScene loop
 * (GLVisitorGPUTrace.cpp:296-410) for offline / multi-view rendering.
 * A batch records the context's scene, tree and lights as they are at
 * creation: after hulk_upload_scene, hulk_update_vertices,
 * hulk_upload_kdtree, hulk_build_kdtree_device or hulk_set_lights on its
 * context, hulk_batch_launch returns HULK_E_ILLEGAL_INPUT without running
 * anything, and the batch must be created again. */
int hulk_batch_create(hulk_ctx* ctx, const hulk_frame_desc* frames, uint32_t n_frames,
                      const uint8_t* mirror_mask, const float* tbound_dev, uint32_t* rgba_dev,
                      uint32_t use_graph, hulk_batch** out);
/* Frames from several scenes in one launch: ctxs[k] holds scene k (scene,
 * tree, lights; all on one device), frame f renders scene frames[f].reserved
 * with mirror flags mirror_masks[k] (nullable array, nullable entries).  The
 * batch runs on ctxs[0]'s stream with ctxs[0]'s options, which must use the
 * fused frame pipeline and the reference triangle test (else
 * HULK_E_UNSUPPORTED); every frame needs a multiple of 64 pixels, so that
 * each wave's 64 rays belong to one scene.  The LDS top tree is used only
 * by static launches (persistent = 0) whose frames all hold a multiple of
 * block_size pixels, so that each block lies in one frame; else it is off.
 * Otherwise as hulk_batch_create, including the refusal to launch after any
 * of the contexts' scene, tree or lights changed.  Replaces a traceScene
 * loop over frames of different scenes (GLVisitorGPUTrace.cpp:296-410). */
int hulk_batch_create_scenes(hulk_ctx* const* ctxs, uint32_t n_ctxs, const hulk_frame_desc* frames, uint32_t n_frames,
                             const uint8_t* const* mirror_masks, const float* tbound_dev, uint32_t* rgba_dev,
                             uint32_t use_graph, hulk_batch** out);
int hulk_batch_launch(hulk_batch* batch, uint32_t flags);
uint32_t hulk_batch_ray_count(const hulk_batch* batch);
/* Copy the primary hit records of the last launch (frames back to back). */
int hulk_batch_primary_hits(hulk_batch* batch, hulk_hit* out, uint32_t flags);
void hulk_batch_destroy(hulk_batch* batch);

/* --------------------------------------------------------------- autotune -- */

typedef struct hulk_tune_entry {
    int32_t variant;
    uint32_t block_size, short_stack_depth, lds_top_levels, persistent, pipeline, leaf_batch, coop_lanes;
    float ms;                 /* time per launch of the tuned objective */
} hulk_tune_entry;

/* Autotuner objectives. */
#define HULK_TUNE_TRACE 0u   /* the primary trace of the camera            */
#define HULK_TUNE_FRAME 1u   /* the whole frame (primary, shadow rays,
                                shading) as hulk_render_frame runs it
                                without mirror generations; the fused and
                                queued pipelines are both tried            */

typedef struct hulk_tune_request {
    hulk_camera cam;            /* the frame to time                         */
    uint32_t width, height;
    uint32_t reps;              /* timed launches per candidate (>= 1)        */
    uint32_t variant_mask;      /* bit per hulk_variant; 0 = every tree variant */
    uint32_t objective;         /* HULK_TUNE_TRACE or HULK_TUNE_FRAME         */
    uint32_t max_bounce;        /* frame objective: mirror generations (<= 8) */
    const uint8_t* mirror_mask; /* frame objective: host flag per triangle, or NULL */
} hulk_tune_request;

/* Per-scene autotuner: times the request's objective over traversal variant,
 * block size 64-256, short-stack depth 4-16, LDS top-tree depth 0-12,
 * persistent or not, one or two records in flight (and, for frames, both
 * pipelines), and makes the fastest the context's options.  No reference
 * counterpart (the reference picks the kernel by --kernelType). */
int hulk_autotune(hulk_ctx* ctx, const hulk_tune_request* request, hulk_tune_entry* table, uint32_t table_cap,
                  uint32_t* n_entries, hulk_opts* best);

/* ----------------------------------------------- learned distance bound -- */

/* Small MLP that proposes a per-ray hit-distance bound (no reference
 * counterpart).  Input: ray origin (normalized by the context's scene box)
 * and unit direction; 32 x 6 Fourier matrix -> (sin, cos) 64-dim encoding;
 * three 128-wide ReLU layers; output log t.  BF16 weights and activations,
 * FP32 accumulation on MFMA; FP32 master weights and Adam.  The bound fed
 * to the traversal is exp(y + margin); traversal stays exact through the
 * unbounded fallback of hulk_trace_camera / hulk_render_frame.  Parameter
 * vector order: W1 (128x64), b1, W2 (128x128), b2, W3 (128x128), b3,
 * w4 (128), b4.  All ray / target / bound pointers are device pointers. */
typedef struct hulk_mlp hulk_mlp;
int hulk_mlp_create(hulk_ctx* ctx, uint64_t seed, uint32_t max_batch, hulk_mlp** out);
void hulk_mlp_destroy(hulk_mlp* mlp);
const char* hulk_mlp_last_error(const hulk_mlp* mlp);
uint32_t hulk_mlp_param_count(void);
int hulk_mlp_get_params(hulk_mlp* mlp, float* params, float* fourier /* 32*6 */);
int hulk_mlp_set_params(hulk_mlp* mlp, const float* params, const float* fourier);
int hulk_mlp_forward(hulk_mlp* mlp, const hulk_ray* rays, uint32_t n, float* pred);
/* Loss mean((y^ - y)^2) and its gradient (n multiple of 64, <= max_batch). */
int hulk_mlp_loss_grad(hulk_mlp* mlp, const hulk_ray* rays, const float* target, uint32_t n, float* loss,
                       float* grad_host);
int hulk_mlp_train_step(hulk_mlp* mlp, const hulk_ray* rays, const float* target, uint32_t n, float lr,
                        float* loss);
/* Per-pixel bound of a camera's primary rays: exp(y + margin), FLT_MAX
 * where y + margin >= max_log.  stride s (1..8): the network runs once per
 * s x s pixel block, on the ray of its first pixel, and the block shares
 * the bound. */
int hulk_mlp_bound_camera(hulk_mlp* mlp, const hulk_camera* cam, uint32_t width, uint32_t height, float margin,
                          float max_log, uint32_t stride, float* tbound, uint32_t flags);
/* The bounds of n_cams frames of the same size in one launch; frame f's
 * bounds start at tbound + f * width * height. */
int hulk_mlp_bound_cameras(hulk_mlp* mlp, const hulk_camera* cams, uint32_t n_cams, uint32_t width, uint32_t height,
                           float margin, float max_log, uint32_t stride, float* tbound, uint32_t flags);
/* Primary rays of a camera as ray records (device) and, optionally, each
 * ray's scene-box exit time (0 if it misses the box): training input of
 * the bound (castRay, GPU_common.cl:10-30; box test SharedFunctions.hpp:245-264). */
int hulk_camera_rays(hulk_ctx* ctx, const hulk_camera* cam, uint32_t width, uint32_t height, hulk_ray* rays,
                     float* texit, uint32_t flags);
/* Device self-test of the reciprocal the kernels use in place of 1.0f / x
 * (no reference counterpart): over every finite nonzero float, out[0] =
 * count with 2^-126 <= |x| < 2^126, out[1] = how many of them differ
 * between v_rcp_f32 + one Newton-Raphson step and the IEEE division,
 * out[2] = the same formula's differences outside that range (taken by
 * the IEEE division in the kernels), out[3] = differences of the kernels'
 * reciprocal from 1.0f / x over every x.  Exact parity needs out[1] ==
 * out[3] == 0.  Runs about 4.3e9 evaluations. */
int hulk_selftest_reciprocal(hulk_ctx* ctx, unsigned long long out[4]);

/* Scene box of a context (min xyz, max xyz). */
int hulk_ctx_scene_box(const hulk_ctx* ctx, float out[6]);

/* --------------------------------------------- out-of-core streaming -- */

typedef struct hulk_stream_stats {
    uint32_t launches;        /* trace launches until every ray finished */
    uint32_t leaves_loaded;   /* leaf uploads into the pool */
    uint32_t leaves_evicted;  /* leaves dropped from the pool, oldest first */
    uint32_t max_queue;       /* most rays resumed by one launch */
    uint64_t bytes_loaded;    /* triangle-record bytes copied host -> device */
    float    ms;              /* wall time of the call */
    uint32_t pool_records;    /* pool size used */
} hulk_stream_stats;

/* Primary rays of a camera with the leaves' triangle records paged into a
 * device pool of `pool_records` 64-B records (0 = every record) from a
 * pinned host copy, as rays request them: the residency protocol of
 * GPUTrace_stackless.cl:27-115 (a lane that reaches a non-resident leaf
 * flags it, saves its traversal state and stops) and the relaunch loop of
 * GLVisitorGPUTrace::launchTraceIterations (GLVisitorGPUTrace.cpp:412-489:
 * load the requested leaves, evicting the oldest, at most half the pool per
 * launch; launch again until no ray stops).  Traversal is kd-restart and the
 * hits equal HULK_VARIANT_RESTART's with every leaf resident.  The pool must
 * hold the largest leaf (else HULK_E_ILLEGAL_INPUT).  flags:
 * HULK_FLAG_DEVICE_PTRS (hits is a device pointer). */
int hulk_trace_camera_streamed(hulk_ctx* ctx, const hulk_camera* cam, uint32_t width, uint32_t height,
                               uint32_t pool_records, hulk_hit* hits, uint32_t flags, hulk_stream_stats* stats);

/* ------------------------------------------------ device KD tree build -- */

/* SAH-NlogN KD tree built on the GPU from the scene uploaded to the
 * context, installed as the context's tree (no host round trip).  Same
 * node and leaf-polygon arrays as KDTreeCPU::buildKDTree
 * (KDTreeCPU.cpp:61-94; KDTreeSAH.cpp:247-521) and hulk_kdtree_build, so
 * every trace gives the same hits as with the uploaded CPU tree.  The
 * reference's device builder is an empty stub (KDTreeGPU,
 * KDTreeGPU.hpp:12-23).  `capacity`: polygon references per depth for the
 * first attempt (0 = 6 per triangle); a depth that overflows doubles it and
 * the build starts again.  Returns HULK_E_ILLEGAL_INPUT for a NaN vertex
 * coordinate of a triangle. */
int hulk_build_kdtree_device(hulk_ctx* ctx, uint32_t capacity);

/* The context's current tree (uploaded or built on the device) in the
 * reference layout (m_treeNodes / m_leafPolygons, KDTreeCPU.hpp:128-132).
 * Writes the counts; copies the arrays when the buffers are non-null and
 * large enough (else HULK_E_ILLEGAL_INPUT). */
int hulk_download_kdtree(hulk_ctx* ctx, hulk_kd_node* nodes, uint32_t node_cap, uint32_t* n_nodes,
                         uint32_t* leaf_polys, uint32_t lp_cap, uint32_t* n_lp);

/* Last device build (no reference counterpart): out[0] = capacity it ran
 * with, out[1] = capacity doublings, out[2] = depths holding nodes,
 * out[3] = deepest junction depth + 1. */
int hulk_device_build_info(const hulk_ctx* ctx, uint32_t out[4]);

/* ------------------------------------------ CPU KD tree (input producer) -- */

/* SAH-NlogN KD tree, the reference's default --treeType
 * (KDTreeCPU::buildKDTree, KDTreeCPU.cpp:61-94; KDTreeSAH.cpp:247-521).
 * Host code; produces the node and leaf-reference arrays unchanged. */
typedef struct hulk_kdtree hulk_kdtree;
int hulk_kdtree_build(const hulk_float4* verts, uint32_t n_verts, const hulk_tri* tris,
                      uint32_t n_tris, const hulk_box* scene_box, hulk_kdtree** out);
/* Same tree, built with at most `max_threads` threads (0 = automatic; 1 =
 * the reference's single-threaded order).  Subtrees are built concurrently
 * and concatenated in depth-first order, so the arrays are identical for
 * every thread count. */
int hulk_kdtree_build_threads(const hulk_float4* verts, uint32_t n_verts, const hulk_tri* tris,
                              uint32_t n_tris, const hulk_box* scene_box, uint32_t max_threads,
                              hulk_kdtree** out);
uint32_t hulk_kdtree_node_count(const hulk_kdtree* t);
uint32_t hulk_kdtree_leaf_poly_count(const hulk_kdtree* t);
const hulk_kd_node* hulk_kdtree_nodes(const hulk_kdtree* t);
const uint32_t* hulk_kdtree_leaf_polys(const hulk_kdtree* t);
void hulk_kdtree_free(hulk_kdtree* t);

/* Scene box as BoundingBoxVisitor computes it (BoundingBoxVisitor.cpp:27-79):
 * per-lane min / max over all vertices. */
int hulk_scene_box(const hulk_float4* verts, uint32_t n_verts, hulk_box* out);

/* Camera as GLVisitorCPUTrace::updateCamera builds it
 * (GLVisitorCPUTrace.cpp:105-148) from a position, a rotation matrix
 * (SPMatrix::rotateMatrix(x, y, z, angle), SPMatrix.cpp:54-84, angle in
 * radians as passed to rotateMatrix) and fov (degrees) / aspect.  cos, sin
 * and tan are evaluated in double and rounded to float. */
int hulk_make_camera(const float from[3], const float axis_angle[4], float fov_deg,
                     float aspect, hulk_camera* out);

#ifdef __cplusplus
}
#endif
#endif /* HULK_ABI_H */
