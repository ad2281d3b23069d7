/*
 * hulk_oracle.c - CPU oracle for the traverse-and-shade hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by the
 * tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * never by the product library.  It restates, in plain C, the arithmetic of
 * hulkTracer's GPU path so that the HIP kernels can be compared with it
 * bit for bit.  Every function cites the reference lines it follows.
 *
 * Compiled with -O2 -ffp-contract=off (no FMA contraction); fmaf() is used
 * exactly where the reference writes fma().  Division and sqrt are IEEE.
 *
 * Arithmetic conventions (fixed here because OpenCL leaves them to the
 * implementation):
 *   dot(a, b)      = (a.x*b.x + a.y*b.y) + a.z*b.z   (w lanes are 0 or
 *                    multiply a 0 on every hot-path vector, so they add +-0)
 *   cross(a, b)    = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)
 *   normalize(v)   = v * (1 / sqrt(dot(v, v)))
 *   pow(e01, 2)    = e01 * e01
 *
 * Parity status per function (how each one is pinned) is listed in the
 * header comment of tests/test_oracle.py.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef __SSE__
#include <xmmintrin.h>
#endif
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { float x, y, z, w; } f4;
typedef struct { uint32_t idx[4]; } tri_t;
typedef struct { f4 mn, mx; } box_t;
typedef struct { uint32_t flags, a, b; float c; } node_t; /* c = plane or leaf id bits */
typedef struct { f4 from, vf, hf, ll; } cam_t;
typedef struct { f4 color, data; } light_t;
typedef struct { f4 o, d; } ray_t;
typedef struct { float t, u, v; uint32_t tri; float nx, ny, nz; uint32_t flags; } hit_t;

#define NODE_LEAF 4u
#define NODE_JUNCTION 8u
#define NO_HIT 0xFFFFFFFFu

enum { V_BRUTE = 0, V_FULLSTACK = 1, V_SHORTSTACK = 2, V_LDSSTACK = 3, V_RESTART = 4, V_ROPES = 5 };

/* ------------------------------------------------------------ vectors -- */
static inline f4 v_sub(f4 a, f4 b) { f4 r = {a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; return r; }
static inline f4 v_add(f4 a, f4 b) { f4 r = {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; return r; }
static inline f4 v_scale(f4 a, float s) { f4 r = {a.x * s, a.y * s, a.z * s, a.w * s}; return r; }
static inline float v_dot(f4 a, f4 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline f4 v_cross(f4 a, f4 b) {
    f4 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x, 0.0f};
    return r;
}
static inline f4 v_normalize(f4 v) { return v_scale(v, 1.0f / sqrtf(v_dot(v, v))); }
static inline float v_get(f4 v, uint32_t axis) { return axis == 0 ? v.x : (axis == 1 ? v.y : v.z); }

/* ------------------------------------------------------------ ray gen -- */
/* castRay, HulkEngine/src/GPU_common.cl:10-30.  The normalize() result is
 * discarded in the reference (:27), so the direction is left unnormalized. */
void orc_cast_ray(const cam_t* cam, uint32_t x, uint32_t y, uint32_t W, uint32_t H, ray_t* ray) {
    float rr = (float)y / (float)H;
    float rc = (float)x / (float)W;
    ray->o = cam->from;
    ray->d = v_sub(v_add(v_add(cam->ll, v_scale(cam->vf, rc)), v_scale(cam->hf, rr)), cam->from);
}

/* ------------------------------------------------------------ box test -- */
/* isRayToAbovePlane / isRayToBelowPlane, SharedFunctions.hpp:154-220. */
static inline int above_plane(float p0, float d0, float e0, float* t0, float* t1) {
    if (d0 > 0) {
        float ti = (e0 - p0) / d0;
        if (ti > *t1) return 0;
        if (ti > *t0) *t0 = ti;
        return 1;
    } else if (d0 < 0) {
        float ti = (e0 - p0) / d0;
        if (ti < *t0) return 0;
        if (ti < *t1) *t1 = ti;
        return 1;
    }
    return e0 < p0;
}
static inline int below_plane(float p0, float d0, float e0, float* t0, float* t1) {
    if (d0 < 0) {
        float ti = (e0 - p0) / d0;
        if (ti > *t1) return 0;
        if (ti > *t0) *t0 = ti;
        return 1;
    } else if (d0 > 0) {
        float ti = (e0 - p0) / d0;
        if (ti < *t0) return 0;
        if (ti < *t1) *t1 = ti;
        return 1;
    }
    return e0 > p0;
}
/* checkRayToBoundingBoxGPU, SharedFunctions.hpp:245-264: all six slab tests
 * are evaluated (bitwise &); the interval is written only if all pass. */
int orc_ray_box(const box_t* b, const ray_t* r, float* ts, float* te) {
    float t0 = 0.0f, t1 = FLT_MAX;
    int ok = above_plane(r->o.x, r->d.x, b->mn.x, &t0, &t1);
    ok &= below_plane(r->o.x, r->d.x, b->mx.x, &t0, &t1);
    ok &= above_plane(r->o.y, r->d.y, b->mn.y, &t0, &t1);
    ok &= below_plane(r->o.y, r->d.y, b->mx.y, &t0, &t1);
    ok &= above_plane(r->o.z, r->d.z, b->mn.z, &t0, &t1);
    ok &= below_plane(r->o.z, r->d.z, b->mx.z, &t0, &t1);
    if (ok) { *ts = t0; *te = t1; }
    return ok;
}

/* ------------------------------------------------------- triangle test -- */
typedef struct { float t, s1, s2, det; f4 n; } tri_hit;

/* checkRayToPoly, HulkEngine/src/SharedFunctions.hpp:47-102.
 * Returns 1 on a hit; fills t, s1, s2, det and the normal. */
int orc_tri_test(const ray_t* ray, f4 p1, f4 p2, f4 p3, tri_hit* h) {
    f4 e0 = v_sub(p2, p1);
    f4 e1 = v_sub(p3, p1);
    f4 e2 = v_sub(p3, p2);
    f4 n = v_add(v_add(v_cross(e0, e1), v_cross(e1, e2)), v_cross(e2, e0));
    n = v_normalize(n);
    float den = v_dot(n, ray->d);
    if (fabsf(den) < FLT_EPSILON) return 0;
    den = 1.0f / den;
    float t = v_dot(n, v_sub(p1, ray->o)) * den;
    if (t < 0) return 0;
    f4 P = v_add(ray->o, v_scale(ray->d, t));
    f4 hp = v_sub(P, p1);
    float q0 = v_dot(e0, hp);
    float q1 = v_dot(e1, hp);
    float e00 = v_dot(e0, e0);
    float e01 = v_dot(e0, e1);
    float e11 = v_dot(e1, e1);
    float det = fmaf(e00, e11, -(e01 * e01));
    float s1 = fmaf(e11, q0, -(e01 * q1));
    float s2 = fmaf(e00, q1, -(e01 * q0));
    if (s1 >= 0 && s2 >= 0 && (s1 + s2) <= det) {
        h->t = t; h->s1 = s1; h->s2 = s2; h->det = det; h->n = n;
        return 1;
    }
    return 0;
}

/* Exposed for known-answer tests. */
int orc_tri_test_ex(const ray_t* ray, const f4* p1, const f4* p2, const f4* p3, float out[7]) {
    tri_hit h;
    int r = orc_tri_test(ray, *p1, *p2, *p3, &h);
    if (r) {
        out[0] = h.t; out[1] = h.s1; out[2] = h.s2; out[3] = h.det;
        out[4] = h.n.x; out[5] = h.n.y; out[6] = h.n.z;
    }
    return r;
}

/* Moller-Trumbore test (T. Moller, B. Trumbore, "Fast, minimum storage
 * ray-triangle intersection", JGT 2(1), 1997), non-culling branch.  No
 * counterpart in the reference (its test is checkRayToPoly above); the
 * library's optional HULK_TRI_MOLLER_TRUMBORE mode.  e0 = p2 - p1,
 * e1 = p3 - p1; the hit keeps the reference's unit normal and reports
 * s1 = u, s2 = v, det = 1.  Division is 1.0f / det.  Parallel cut: the
 * reference's |n.d| < FLT_EPSILON, as det^2 < FLT_EPSILON^2 * gram with
 * gram = |e0 x e1|^2 = fma(e00, e11, -(e01 * e01)) (the paper's absolute
 * EPSILON would scale with the triangle's area). */
int orc_tri_test_mt(const ray_t* ray, f4 p1, f4 p2, f4 p3, tri_hit* h) {
    f4 e0 = v_sub(p2, p1);
    f4 e1 = v_sub(p3, p1);
    f4 pvec = v_cross(ray->d, e1);
    float det = v_dot(e0, pvec);
    float gram = fmaf(v_dot(e0, e0), v_dot(e1, e1), -(v_dot(e0, e1) * v_dot(e0, e1)));
    if (det == 0.0f || det * det < (FLT_EPSILON * FLT_EPSILON) * gram) return 0;
    float inv = 1.0f / det;
    f4 tvec = v_sub(ray->o, p1);
    float u = v_dot(tvec, pvec) * inv;
    if (u < 0.0f || u > 1.0f) return 0;
    f4 qvec = v_cross(tvec, e0);
    float v = v_dot(ray->d, qvec) * inv;
    if (v < 0.0f || u + v > 1.0f) return 0;
    float t = v_dot(e1, qvec) * inv;
    if (t < 0.0f) return 0;
    f4 e2 = v_sub(p3, p2);
    f4 n = v_add(v_add(v_cross(e0, e1), v_cross(e1, e2)), v_cross(e2, e0));
    h->t = t; h->s1 = u; h->s2 = v; h->det = 1.0f; h->n = v_normalize(n);
    return 1;
}

int orc_tri_test_mt_ex(const ray_t* ray, const f4* p1, const f4* p2, const f4* p3, float out[7]) {
    tri_hit h;
    int r = orc_tri_test_mt(ray, *p1, *p2, *p3, &h);
    if (r) {
        out[0] = h.t; out[1] = h.s1; out[2] = h.s2; out[3] = h.det;
        out[4] = h.n.x; out[5] = h.n.y; out[6] = h.n.z;
    }
    return r;
}

/* Double-precision twin of the test: used to state the FP32 error of t, u, v. */
int orc_tri_test_f64(const ray_t* ray, const f4* p1, const f4* p2, const f4* p3, double out[3]) {
    double a[3] = {p1->x, p1->y, p1->z}, b[3] = {p2->x, p2->y, p2->z}, c[3] = {p3->x, p3->y, p3->z};
    double o[3] = {ray->o.x, ray->o.y, ray->o.z}, d[3] = {ray->d.x, ray->d.y, ray->d.z};
    double e0[3], e1[3], n[3], w[3];
    for (int i = 0; i < 3; i++) { e0[i] = b[i] - a[i]; e1[i] = c[i] - a[i]; w[i] = a[i] - o[i]; }
    n[0] = e0[1] * e1[2] - e0[2] * e1[1];
    n[1] = e0[2] * e1[0] - e0[0] * e1[2];
    n[2] = e0[0] * e1[1] - e0[1] * e1[0];
    double den = n[0] * d[0] + n[1] * d[1] + n[2] * d[2];
    if (den == 0.0) return 0;
    double t = (n[0] * w[0] + n[1] * w[1] + n[2] * w[2]) / den;
    double h[3];
    for (int i = 0; i < 3; i++) h[i] = o[i] + d[i] * t - a[i];
    double q0 = e0[0] * h[0] + e0[1] * h[1] + e0[2] * h[2];
    double q1 = e1[0] * h[0] + e1[1] * h[1] + e1[2] * h[2];
    double e00 = e0[0] * e0[0] + e0[1] * e0[1] + e0[2] * e0[2];
    double e01 = e0[0] * e1[0] + e0[1] * e1[1] + e0[2] * e1[2];
    double e11 = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
    double det = e00 * e11 - e01 * e01;
    out[0] = t; out[1] = (e11 * q0 - e01 * q1) / det; out[2] = (e00 * q1 - e01 * q0) / det;
    return 1;
}

/* --------------------------------------------------------- scene view -- */
typedef struct {
    const f4* verts;
    const tri_t* tris;
    uint32_t n_tris;
    box_t box;
    const node_t* nodes;
    uint32_t n_nodes;
    const uint32_t* lp;
    /* rope traversal only (orc_build_ropes output) */
    const uint32_t* ropes;
    const box_t* cells;
    const uint32_t* kids;
    uint32_t n_slots;
    int mt; /* 1: the Moller-Trumbore test instead of checkRayToPoly */
} scene_t;

/* Variant argument flag of the public entry points: trace with the
 * Moller-Trumbore test. */
#define ORC_MT 0x100

typedef struct {
    float t, s1, s2, det;
    uint32_t tri;
    f4 n;
} best_t;

typedef struct { uint32_t nodes, leaves, tris; } cnt_t;

static inline void test_tri_into(const scene_t* s, const ray_t* ray, uint32_t tri, best_t* best, cnt_t* c) {
    const tri_t* T = &s->tris[tri];
    tri_hit h;
    c->tris++;
    const f4 p1 = s->verts[T->idx[0]], p2 = s->verts[T->idx[1]], p3 = s->verts[T->idx[2]];
    if (s->mt ? orc_tri_test_mt(ray, p1, p2, p3, &h) : orc_tri_test(ray, p1, p2, p3, &h)) {
        if (h.t < best->t) {
            best->t = h.t; best->s1 = h.s1; best->s2 = h.s2; best->det = h.det;
            best->n = h.n; best->tri = tri;
        }
    }
}

/* testNodePolygons, GPU_common.cl:46-71 / KDTreeCPU.cpp:255-283: leaf order
 * m_leafPolygons[start .. start+count), strict '<' keeps the first tested. */
static inline void test_leaf(const scene_t* s, const ray_t* ray, const node_t* leaf, best_t* best, cnt_t* c) {
    uint32_t start = leaf->a, end = leaf->a + leaf->b;
    c->leaves++;
    for (uint32_t i = start; i < end; i++) test_tri_into(s, ray, s->lp[i], best, c);
}

/* Junction child order, identical in every variant
 * (GPUTrace.cl:112-128, GPUTrace_shortStack.cl:53-69). */
static inline float junction_split(const node_t* n, const ray_t* r, const f4* inv, uint32_t* first, uint32_t* second) {
    uint32_t axis = n->flags & 3u;
    float ro = v_get(r->o, axis), rd = v_get(r->d, axis), id = v_get(*inv, axis);
    float plane = n->c;
    float split = (plane - ro) * id;
    *first = n->a; *second = n->b;
    if (ro < plane || (ro == plane && rd <= 0)) { *first = n->b; *second = n->a; }
    return split;
}

static inline f4 inv_dir(const ray_t* r) {
    f4 i = {1.0f / r->d.x, 1.0f / r->d.y, 1.0f / r->d.z, 0.0f};
    return i;
}

/* Full-stack traversal, trace_ray of GPUTrace.cl:6-90 (twin of
 * KDTreeCPU::searchKDTree, KDTreeCPU.cpp:285-361).  tclip (FLT_MAX = none)
 * clips the box exit. */
static void trav_fullstack(const scene_t* s, const ray_t* ray, float tclip, best_t* best, cnt_t* c) {
    float ts = 0.0f, te = FLT_MAX;
    if (!orc_ray_box(&s->box, ray, &ts, &te)) return;
    if (tclip < te) te = tclip;
    f4 inv = inv_dir(ray);
    struct { float ts, te; uint32_t node; } stack[32];
    uint32_t sp = 0, node = 0;
    do {
        if (best->t < ts) break;
        const node_t* n = &s->nodes[node];
        c->nodes++;
        if (n->flags & NODE_LEAF) {
            test_leaf(s, ray, n, best, c);
            node = 0;
        } else {
            uint32_t first, second;
            float split = junction_split(n, ray, &inv, &first, &second);
            if (split > te || split <= 0) {
                node = first;
            } else if (split < ts) {
                node = second;
            } else {
                if (second != 0) {
                    stack[sp].node = second; stack[sp].ts = split; stack[sp].te = te;
                    sp++;
                }
                node = first;
                te = split;
            }
        }
        if (node == 0 && sp > 0) {
            sp--;
            ts = stack[sp].ts; te = stack[sp].te; node = stack[sp].node;
        }
    } while (node != 0);
}

/* Short stack with restart, trace_ray of GPUTrace_shortStack.cl:6-105
 * (LDS variant GPUTrace_LDSshortStack.cl:7-105 has the same semantics; CPU
 * twin KDTreeCPU::searchKDTreeStackless, KDTreeCPU.cpp:365-447).  depth =
 * c_shortStackDepth (6 on the GPU, SharedStructs.hpp:19; 5 on the CPU). */
static void trav_shortstack(const scene_t* s, const ray_t* ray, float tclip, uint32_t depth, best_t* best, cnt_t* c) {
    float ts = 0.0f, te = FLT_MAX;
    if (!orc_ray_box(&s->box, ray, &ts, &te)) return;
    if (tclip < te) te = tclip;
    f4 inv = inv_dir(ray);
    struct { float ts, te; uint32_t node; } frame[64];
    int size = 0, idx = 0;
    uint32_t node = 0;
    float tmax = ts, tmin = ts;
    while (tmax < te) {
        if (size == 0) {
            tmin = tmax; tmax = te; node = 0;
        } else {
            --size; --idx;
            if (idx == -1) idx += (int)depth;
            tmin = frame[idx].ts; tmax = frame[idx].te; node = frame[idx].node;
        }
        const node_t* n = &s->nodes[node];
        c->nodes++;
        while (n->flags & NODE_JUNCTION) {
            uint32_t first, second;
            float split = junction_split(n, ray, &inv, &first, &second);
            if (split >= tmax || split <= 0) {
                node = first;
            } else if (split <= tmin) {
                node = second;
            } else {
                frame[idx].ts = split; frame[idx].te = tmax; frame[idx].node = second;
                idx = (idx + 1) % (int)depth;
                size++;
                if (size > (int)depth) size = (int)depth;
                node = first;
                tmax = split;
            }
            n = &s->nodes[node];
            c->nodes++;
            if (node == 0) break;
        }
        if (n->flags & NODE_LEAF) {
            test_leaf(s, ray, n, best, c);
            if (best->t < tmax) break;
        }
    }
}

/* kd-restart, trace_ray of GPUTrace_stackless.cl:6-120 with every leaf
 * resident: ThreadContext starts at zero (tmin = tmax = 0; the box entry
 * time is not used), every leaf is followed by a restart from the root, and
 * there is no early exit after a hit. */
static void trav_restart(const scene_t* s, const ray_t* ray, float tclip, best_t* best, cnt_t* c) {
    float ts = 0.0f, te = FLT_MAX;
    if (!orc_ray_box(&s->box, ray, &ts, &te)) return;
    if (tclip < te) te = tclip;
    f4 inv = inv_dir(ray);
    float tmax = 0.0f, tmin = 0.0f;
    if (best->t < tmax) return; /* GPUTrace_stackless.cl:46-51 */
    while (tmax < te) {
        tmin = tmax; tmax = te;
        uint32_t node = 0;
        const node_t* n = &s->nodes[0];
        c->nodes++;
        while (n->flags & NODE_JUNCTION) {
            uint32_t first, second;
            float split = junction_split(n, ray, &inv, &first, &second);
            if (split >= tmax || split <= 0) {
                node = first;
            } else if (split <= tmin) {
                node = second;
            } else {
                node = first;
                tmax = split;
            }
            n = &s->nodes[node];
            c->nodes++;
            if (node == 0) break;
        }
        if (n->flags & NODE_LEAF) test_leaf(s, ray, n, best, c);
    }
}

/* Brute force, GPUTrace_bruteforce.cl:27-42: every triangle in index order. */
static void trav_brute(const scene_t* s, const ray_t* ray, float tclip, best_t* best, cnt_t* c) {
    (void)tclip;
    for (uint32_t i = 0; i < s->n_tris; i++) test_tri_into(s, ray, i, best, c);
}

/* ---------------------------------------------------------------- ropes -- */
/* Stackless rope traversal (Havran 2001; Popov et al. 2007) on the
 * unchanged tree.  Not in the reference (its "stackless" kernel is
 * kd-restart, GPUTrace_stackless.cl:53-117); the rules are fixed here and
 * restated independently in the HIP kernel.
 *
 * Slots: every tree node keeps its index; every empty child (index 0 in the
 * reference tree) gets a virtual empty-leaf slot numbered from n_nodes up,
 * in DFS order (above child before below child).  kids[2s], kids[2s+1] =
 * above / below slot of junction slot s.  cells[s] = the cell box;
 * ropes[6s + f] = neighbour slot across face f (f = 2*axis + 0 for the min
 * face, + 1 for the max face), ROPE_NONE outside the root box.  At a
 * junction on axis a with plane v the above child (cell min[a] = v) gets
 * rope[2a] = below slot and the below child gets rope[2a+1] = above slot;
 * all other ropes are inherited from the junction.  Returns the slot count
 * (at most 2 * n_nodes + 1). */
#define ROPE_NONE 0xFFFFFFFFu

uint32_t orc_build_ropes(const node_t* nodes, uint32_t n_nodes, const box_t* root_box,
                         uint32_t* ropes, box_t* cells, uint32_t* kids) {
    typedef struct { uint32_t slot; box_t cell; uint32_t r[6]; } item;
    item st[256];
    uint32_t sp = 0, slots = n_nodes;
    st[0].slot = 0; st[0].cell = *root_box;
    for (int k = 0; k < 6; k++) st[0].r[k] = ROPE_NONE;
    sp = 1;
    while (sp) {
        item it = st[--sp];
        cells[it.slot] = it.cell;
        for (int k = 0; k < 6; k++) ropes[it.slot * 6 + k] = it.r[k];
        kids[2 * it.slot] = kids[2 * it.slot + 1] = ROPE_NONE;
        if (it.slot >= n_nodes) continue;
        const node_t* n = &nodes[it.slot];
        if (!(n->flags & NODE_JUNCTION)) continue;
        uint32_t a = n->flags & 3u;
        uint32_t above = n->a ? n->a : slots++;
        uint32_t below = n->b ? n->b : slots++;
        kids[2 * it.slot] = above; kids[2 * it.slot + 1] = below;
        item ia = it, ib = it;
        ia.slot = above; ib.slot = below;
        ((float*)&ia.cell.mn)[a] = n->c;
        ((float*)&ib.cell.mx)[a] = n->c;
        ia.r[2 * a] = below;
        ib.r[2 * a + 1] = above;
        st[sp++] = ib;
        st[sp++] = ia;
    }
    return slots;
}

/* Rope walk.  Rules:
 *   entry: [ts, te] from the scene-box test (te clipped by tclip);
 *     t = ts; slot = root.
 *   descent from a slot at time t, at each junction: split and first /
 *     second as in the reference junction step; take first if split <= 0
 *     or split > t, else second (the ray is on the far side after t).
 *   leaf: test its triangles (virtual empty leaves have none).
 *   exit: t_out = min over axes with d[a] != 0 of the slab time of the face
 *     the ray leaves through (d[a] > 0: max face, d[a] < 0: min face),
 *     (face - o[a]) * inv[a]; ties keep the lowest axis.
 *   stop if best.t <= t_out, t_out >= te, no face, or the rope is ROPE_NONE.
 *   next: t = max(t, t_out), slot = rope[face]; at most 4*(slots+8) steps. */
static void trav_ropes(const scene_t* s, const ray_t* ray, float tclip, best_t* best, cnt_t* c) {
    float ts = 0.0f, te = FLT_MAX;
    if (!orc_ray_box(&s->box, ray, &ts, &te)) return;
    if (tclip < te) te = tclip;
    f4 inv = inv_dir(ray);
    float t = ts;
    uint32_t slot = 0, guard = 0;
    const float* o = &ray->o.x;
    const float* d = &ray->d.x;
    const float* iv = &inv.x;
    for (;;) {
        for (;;) {
            c->nodes++;
            if (slot >= s->n_nodes) break;
            const node_t* n = &s->nodes[slot];
            if (!(n->flags & NODE_JUNCTION)) break;
            uint32_t a = n->flags & 3u;
            float ro = o[a], rd = d[a], plane = n->c;
            float split = (plane - ro) * iv[a];
            uint32_t first = s->kids[2 * slot], second = s->kids[2 * slot + 1];
            if (ro < plane || (ro == plane && rd <= 0)) { uint32_t tmp = first; first = second; second = tmp; }
            slot = (split <= 0 || split > t) ? first : second;
        }
        if (slot < s->n_nodes) test_leaf(s, ray, &s->nodes[slot], best, c);
        const float* mn = &s->cells[slot].mn.x;
        const float* mx = &s->cells[slot].mx.x;
        float tout = FLT_MAX;
        int face = -1;
        for (int a = 0; a < 3; a++) {
            float ta;
            int f;
            if (d[a] > 0) { ta = (mx[a] - o[a]) * iv[a]; f = 2 * a + 1; }
            else if (d[a] < 0) { ta = (mn[a] - o[a]) * iv[a]; f = 2 * a; }
            else continue;
            if (ta < tout) { tout = ta; face = f; }
        }
        if (best->t <= tout || tout >= te || face < 0) break;
        uint32_t nxt = s->ropes[slot * 6 + face];
        if (nxt == ROPE_NONE) break;
        if (tout > t) t = tout;
        slot = nxt;
        if (++guard > 4u * (s->n_slots + 8u)) break;
    }
}

/* --------------------------------------------------------- dispatcher -- */
static void trace_one(const scene_t* s, int variant, uint32_t depth, const ray_t* ray, float tclip, hit_t* out, cnt_t* c) {
    best_t best;
    best.t = FLT_MAX; best.tri = NO_HIT; best.s1 = best.s2 = 0; best.det = 1;
    best.n.x = best.n.y = best.n.z = best.n.w = 0;
    switch (variant) {
        case V_BRUTE: trav_brute(s, ray, tclip, &best, c); break;
        case V_FULLSTACK: trav_fullstack(s, ray, tclip, &best, c); break;
        case V_SHORTSTACK:
        case V_LDSSTACK: trav_shortstack(s, ray, tclip, depth, &best, c); break;
        case V_RESTART: trav_restart(s, ray, tclip, &best, c); break;
        case V_ROPES: trav_ropes(s, ray, tclip, &best, c); break;
        default: break;
    }
    out->t = best.t;
    out->tri = best.tri;
    out->flags = 0;
    if (best.tri != NO_HIT) {
        out->u = best.s1 / best.det;
        out->v = best.s2 / best.det;
        out->nx = best.n.x; out->ny = best.n.y; out->nz = best.n.z;
    } else {
        out->u = out->v = 0; out->nx = out->ny = out->nz = 0;
    }
}

/* Bounded trace with exact fallback: a bounded pass with the box exit
 * clipped to tbound * (1 + 2^-10).  If the clip does not shorten the box
 * interval the pass is the unbounded traversal; otherwise a hit with
 * t < tbound is kept and anything else re-runs unbounded (flag bit 0). */
static void trace_bounded(const scene_t* s, int variant, uint32_t depth, const ray_t* ray, float tbound, hit_t* out, cnt_t* c) {
    float tclip = tbound * (1.0f + 1.0f / 1024.0f);
    float bts = 0.0f, bte = 0.0f;
    int clipped = tbound < FLT_MAX && orc_ray_box(&s->box, ray, &bts, &bte) && tclip < bte;
    if (!clipped) { trace_one(s, variant, depth, ray, FLT_MAX, out, c); return; }
    trace_one(s, variant, depth, ray, tclip, out, c);
    if (out->tri != NO_HIT && out->t < tbound) return;
    trace_one(s, variant, depth, ray, FLT_MAX, out, c);
    out->flags |= 1u;
}

/* Rope data handed over from orc_build_ropes (NULL when unused). */
typedef struct { const uint32_t* ropes; const box_t* cells; const uint32_t* kids; uint32_t n_slots; } rope_view_t;

static void make_scene(scene_t* s, const f4* verts, const tri_t* tris, uint32_t n_tris, const box_t* box,
                       const node_t* nodes, uint32_t n_nodes, const uint32_t* lp, const rope_view_t* rv) {
    s->verts = verts; s->tris = tris; s->n_tris = n_tris; s->box = *box;
    s->nodes = nodes; s->n_nodes = n_nodes; s->lp = lp;
    s->ropes = rv ? rv->ropes : NULL; s->cells = rv ? rv->cells : NULL;
    s->kids = rv ? rv->kids : NULL; s->n_slots = rv ? rv->n_slots : 0;
    s->mt = 0;
}

/* Trace arbitrary rays.  counts (nullable): 3 uint32 per ray (nodes, leaves,
 * triangle tests).  tbound (nullable): per-ray learned bound. */
void orc_trace_rays(int variant, uint32_t depth,
                    const f4* verts, const tri_t* tris, uint32_t n_tris, const box_t* box,
                    const node_t* nodes, uint32_t n_nodes, const uint32_t* lp,
                    const rope_view_t* rv,
                    const ray_t* rays, uint32_t n, const float* tbound, hit_t* out, uint32_t* counts,
                    int nthreads) {
    scene_t s;
    make_scene(&s, verts, tris, n_tris, box, nodes, n_nodes, lp, rv);
    s.mt = (variant & ORC_MT) != 0;
    variant &= ~ORC_MT;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256)
#endif
    for (long i = 0; i < (long)n; i++) {
        cnt_t c = {0, 0, 0};
        if (tbound) trace_bounded(&s, variant, depth, &rays[i], tbound[i], &out[i], &c);
        else trace_one(&s, variant, depth, &rays[i], FLT_MAX, &out[i], &c);
        if (counts) { counts[3 * i] = c.nodes; counts[3 * i + 1] = c.leaves; counts[3 * i + 2] = c.tris; }
    }
}

/* Primary rays of a camera (castRay per pixel, index y*W + x). */
void orc_trace_camera(int variant, uint32_t depth,
                      const f4* verts, const tri_t* tris, uint32_t n_tris, const box_t* box,
                      const node_t* nodes, uint32_t n_nodes, const uint32_t* lp,
                      const rope_view_t* rv,
                      const cam_t* cam, uint32_t W, uint32_t H, const float* tbound, hit_t* out,
                      uint32_t* counts, int nthreads) {
    scene_t s;
    make_scene(&s, verts, tris, n_tris, box, nodes, n_nodes, lp, rv);
    s.mt = (variant & ORC_MT) != 0;
    variant &= ~ORC_MT;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (long y = 0; y < (long)H; y++) {
        for (uint32_t x = 0; x < W; x++) {
            size_t i = (size_t)y * W + x;
            ray_t ray;
            orc_cast_ray(cam, x, (uint32_t)y, W, H, &ray);
            cnt_t c = {0, 0, 0};
            if (tbound) trace_bounded(&s, variant, depth, &ray, tbound[i], &out[i], &c);
            else trace_one(&s, variant, depth, &ray, FLT_MAX, &out[i], &c);
            if (counts) { counts[3 * i] = c.nodes; counts[3 * i + 1] = c.leaves; counts[3 * i + 2] = c.tris; }
        }
    }
}

void orc_cast_rays(const cam_t* cam, uint32_t W, uint32_t H, ray_t* out) {
    for (uint32_t y = 0; y < H; y++)
        for (uint32_t x = 0; x < W; x++) orc_cast_ray(cam, x, y, W, H, &out[(size_t)y * W + x]);
}

/* ------------------------------------------------------------- shading -- */
/* Pixel conversion, FrameBuffer::replacePixel (FrameBuffer.hpp:196-217):
 * min(1, c) * 255 rounded to nearest even (cvtps2dq), packed
 * R | G << 8 | B << 16. */
static inline uint32_t to_u8(float c) {
    if (c > 1.0f) c = 1.0f;
    if (!(c > 0.0f)) c = 0.0f;
    float s = c * 255.0f;
    return (uint32_t)nearbyintf(s);
}
uint32_t orc_pack_rgb(float r, float g, float b) { return to_u8(r) | (to_u8(g) << 8) | (to_u8(b) << 16); }

/* Shading of one hit point.  Fused semantics, GPUTrace.cl:122-160 (CPU twin
 * GLVisitorCPUTrace.cpp:165-194): ambient 0.1; directional I = n.dir; point:
 * L = pos - P (w = 0), dist = sqrt(L.L), L = normalize(L), shadow ray
 * (P + 0.001*L, L) traced nearest-hit with the same variant, I = n.L, halved
 * if dist > t_shadow; color += lightColor * max(I, 0).  Miss = 0.
 * Picture semantics, shadePicture.cl:14-54: base 0.1 for every pixel;
 * point I = (n.L)/2 with no shadow ray; directional I = n.dir. */
static void shade_point(const scene_t* s, int variant, uint32_t depth, const light_t* lights, uint32_t n_lights,
                        f4 P, f4 n, int mode, float col[3], cnt_t* c, uint32_t* n_shadow) {
    col[0] = col[1] = col[2] = 0.1f;
    for (uint32_t k = 0; k < n_lights; k++) {
        const light_t* L = &lights[k];
        float intensity = 0.0f;
        if (L->color.w == 0.0f) intensity = v_dot(n, L->data);
        if (L->color.w == 1.0f) {
            f4 tl = v_sub(L->data, P);
            tl.w = 0.0f;
            float dd = v_dot(tl, tl);
            float dist = sqrtf(dd);
            tl = v_scale(tl, 1.0f / sqrtf(dd));
            if (mode == 0) {
                ray_t sr;
                sr.o = v_add(P, v_scale(tl, 0.001f));
                sr.d = tl;
                hit_t sh;
                trace_one(s, variant, depth, &sr, FLT_MAX, &sh, c);
                if (n_shadow) (*n_shadow)++;
                intensity = v_dot(n, tl);
                if (dist > sh.t) intensity *= 0.5f;
            } else {
                intensity = v_dot(n, tl) / 2.0f;
            }
        }
        float m = intensity > 0.0f ? intensity : 0.0f;
        col[0] += L->color.x * m;
        col[1] += L->color.y * m;
        col[2] += L->color.z * m;
    }
}

/* Shade a primary hit buffer: out[y*W+x] packed RGB. */
void orc_shade(int variant, uint32_t depth,
               const f4* verts, const tri_t* tris, uint32_t n_tris, const box_t* box,
               const node_t* nodes, uint32_t n_nodes, const uint32_t* lp,
               const rope_view_t* rv,
               const cam_t* cam, uint32_t W, uint32_t H, const hit_t* hits,
               const light_t* lights, uint32_t n_lights, int mode, uint32_t* out, int nthreads) {
    scene_t s;
    make_scene(&s, verts, tris, n_tris, box, nodes, n_nodes, lp, rv);
    s.mt = (variant & ORC_MT) != 0;
    variant &= ~ORC_MT;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (long y = 0; y < (long)H; y++) {
        for (uint32_t x = 0; x < W; x++) {
            size_t i = (size_t)y * W + x;
            const hit_t* h = &hits[i];
            float col[3] = {0, 0, 0};
            int hit = mode == 0 ? (h->tri != NO_HIT) : (h->t != FLT_MAX && h->t > 0);
            if (hit) {
                ray_t ray;
                orc_cast_ray(cam, x, (uint32_t)y, W, H, &ray);
                f4 P = v_add(ray.o, v_scale(ray.d, h->t));
                f4 n = {h->nx, h->ny, h->nz, 0.0f};
                cnt_t c = {0, 0, 0};
                shade_point(&s, variant, depth, lights, n_lights, P, n, mode, col, &c, NULL);
            } else if (mode == 1) {
                col[0] = col[1] = col[2] = 0.1f;
            }
            out[i] = orc_pack_rgb(col[0], col[1], col[2]);
        }
    }
}

/* ------------------------------------------------- mirror generations -- */
/* Whole frame with mirror bounces (no reference counterpart; rules fixed
 * here and restated in the HIP path):
 *   generation 0 = primary rays; a hit on a mirror triangle with
 *   generation g < max_bounce spawns r = d - 2(d.n)n (d normalized first,
 *   n the unit triangle normal), origin P + 0.001*r, weight w*0.8, and keeps
 *   w*0.2 of its own Lambert shading; other hits keep w*1.0 of it.
 *   Lambert shading = shade_point (fused, with shadow rays).
 *   pixel = sum of weighted contributions in generation order, then packed.
 * Returns the counts of nearest-hit rays and shadow rays traced. */
void orc_render(int variant, uint32_t depth,
                const f4* verts, const tri_t* tris, uint32_t n_tris, const box_t* box,
                const node_t* nodes, uint32_t n_nodes, const uint32_t* lp,
                const rope_view_t* rv,
                const cam_t* cam, uint32_t W, uint32_t H, const light_t* lights, uint32_t n_lights,
                uint32_t max_bounce, const uint8_t* mirror, uint32_t* out, uint64_t ray_counts[2],
                int nthreads) {
    scene_t s;
    make_scene(&s, verts, tris, n_tris, box, nodes, n_nodes, lp, rv);
    s.mt = (variant & ORC_MT) != 0;
    variant &= ~ORC_MT;
    uint64_t n_near = 0, n_shadow = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : n_near, n_shadow)
#endif
    for (long y = 0; y < (long)H; y++) {
        for (uint32_t x = 0; x < W; x++) {
            size_t i = (size_t)y * W + x;
            ray_t ray;
            orc_cast_ray(cam, x, (uint32_t)y, W, H, &ray);
            float acc[3] = {0, 0, 0};
            float w = 1.0f;
            for (uint32_t g = 0; g <= max_bounce; g++) {
                hit_t h;
                cnt_t c = {0, 0, 0};
                trace_one(&s, variant, depth, &ray, FLT_MAX, &h, &c);
                n_near++;
                if (h.tri == NO_HIT) break;
                f4 P = v_add(ray.o, v_scale(ray.d, h.t));
                f4 n = {h.nx, h.ny, h.nz, 0.0f};
                float col[3];
                uint32_t ns = 0;
                shade_point(&s, variant, depth, lights, n_lights, P, n, 0, col, &c, &ns);
                n_shadow += ns;
                int is_mirror = mirror && mirror[h.tri] && g < max_bounce;
                float keep = is_mirror ? w * 0.2f : w;
                acc[0] += col[0] * keep; acc[1] += col[1] * keep; acc[2] += col[2] * keep;
                if (!is_mirror) break;
                f4 dn = v_normalize(ray.d);
                dn.w = 0.0f;
                float k = 2.0f * v_dot(dn, n);
                f4 r = v_sub(dn, v_scale(n, k));
                r.w = 0.0f;
                ray.o = v_add(P, v_scale(r, 0.001f));
                ray.d = r;
                w = w * 0.8f;
            }
            out[i] = orc_pack_rgb(acc[0], acc[1], acc[2]);
        }
    }
    if (ray_counts) { ray_counts[0] = n_near; ray_counts[1] = n_shadow; }
}

/* ---------------------------------------------- reference CPU path port -- */
/* GLVisitorCPUTrace::traceScene (GLVisitorCPUTrace.cpp:149-200) with
 * KDTreeCPU::checkCollision (KDTreeCPU.cpp:230-253): rays normalized with
 * rsqrt + one Newton-Raphson step (SPMatrix.hpp:302-322), SSE dot order
 * (y+w)+(x+z) (SPMatrix.hpp:284-290), short stack of depth 5 with restart
 * (KDTreeCPU.cpp:363-447), cached test checkRayToPolyACC
 * (SharedFunctions.hpp:105-149) over PolyToRayCache (KDTreeCPU.cpp:479-507),
 * one nearest-hit shadow ray per point light.  This is the CPU baseline the
 * GPU is timed against; its t / normal bits are not portable (rsqrt). */
typedef struct { f4 n, e0, e1; float e00, e01, e11, det; } pcache_t;

static inline float sse_dot(f4 a, f4 b) {
    float r0 = a.x * b.x, r1 = a.y * b.y, r2 = a.z * b.z, r3 = a.w * b.w;
    return (r1 + r3) + (r0 + r2);
}
static inline float rsqrt_nr(float a) {
#ifdef __SSE__
    float r0 = _mm_cvtss_f32(_mm_rsqrt_ss(_mm_set_ss(a)));
#else
    float r0 = 1.0f / sqrtf(a);
#endif
    return (0.5f * r0) * (3.0f - (a * r0) * r0);
}
static inline f4 sse_normalize(f4 v) {
    float r0 = v.x * v.x, r1 = v.y * v.y, r2 = v.z * v.z, r3 = v.w * v.w;
    float t = (r1 + r3) + (r0 + r2);
    return v_scale(v, rsqrt_nr(t));
}

void orc_cpu_build_cache(const f4* verts, const tri_t* tris, uint32_t n_tris, pcache_t* cache) {
    for (uint32_t i = 0; i < n_tris; i++) {
        f4 p1 = verts[tris[i].idx[0]], p2 = verts[tris[i].idx[1]], p3 = verts[tris[i].idx[2]];
        f4 a1 = v_sub(p2, p1), a2 = v_sub(p3, p2), a3 = v_sub(p1, p3);
        f4 n = v_add(v_add(v_cross(a1, a2), v_cross(a2, a3)), v_cross(a3, a1));
        n.w = 0.0f;
        cache[i].n = sse_normalize(n);
        cache[i].e0 = v_sub(p2, p1);
        cache[i].e1 = v_sub(p3, p1);
        cache[i].e00 = sse_dot(cache[i].e0, cache[i].e0);
        cache[i].e01 = sse_dot(cache[i].e0, cache[i].e1);
        cache[i].e11 = sse_dot(cache[i].e1, cache[i].e1);
        cache[i].det = (float)((double)(cache[i].e00 * cache[i].e11) - pow((double)cache[i].e01, 2));
    }
}

typedef struct { const scene_t* s; const pcache_t* cache; } cpu_scene_t;

static int cpu_leaf(const cpu_scene_t* cs, const ray_t* ray, const node_t* leaf, float* bt, f4* bp, f4* bn) {
    int ret = 0;
    for (uint32_t i = leaf->a; i < leaf->a + leaf->b; i++) {
        uint32_t pi = cs->s->lp[i];
        const pcache_t* C = &cs->cache[pi];
        f4 p1 = cs->s->verts[cs->s->tris[pi].idx[0]];
        float den = sse_dot(C->n, ray->d);
        if (fabsf(den) < FLT_EPSILON) continue;
        den = 1.0f / den;
        float t = sse_dot(C->n, v_sub(p1, ray->o)) * den;
        if (t < 0) continue;
        f4 P = v_add(ray->o, v_scale(ray->d, t));
        f4 h = v_sub(P, p1);
        float q0 = sse_dot(C->e0, h), q1 = sse_dot(C->e1, h);
        float s1 = (C->e11 * q0) - (C->e01 * q1);
        float s2 = (C->e00 * q1) - (C->e01 * q0);
        if (s1 >= 0 && s2 >= 0 && (s1 + s2) <= C->det) {
            if (t < *bt) { *bt = t; *bp = P; *bn = C->n; ret = 1; }
        }
    }
    return ret;
}

static int cpu_check_collision(const cpu_scene_t* cs, const ray_t* ray, float* bt, f4* bp, f4* bn) {
    const scene_t* s = cs->s;
    float ts = 0, te = INFINITY;
    *bt = INFINITY;
    if (!orc_ray_box(&s->box, ray, &ts, &te)) return 0;
    f4 inv = inv_dir(ray);
    struct { float ts, te; uint32_t node; } frame[5];
    int size = 0, idx = 0, ret = 0;
    float tmin = ts, tmax = ts;
    const node_t* n;
    while (tmax < te) {
        if (size == 0) { n = &s->nodes[0]; tmin = tmax; tmax = te; }
        else {
            --size; --idx; if (idx == -1) idx += 5;
            tmin = frame[idx].ts; tmax = frame[idx].te; n = &s->nodes[frame[idx].node];
        }
        while (n->flags & NODE_JUNCTION) {
            uint32_t first, second;
            float split = junction_split(n, ray, &inv, &first, &second);
            if (split >= tmax || split <= 0) n = &s->nodes[first];
            else if (split <= tmin) n = &s->nodes[second];
            else {
                frame[idx].ts = split; frame[idx].te = tmax; frame[idx].node = second;
                idx = (idx + 1) % 5; size++; if (size > 5) size = 5;
                n = &s->nodes[first]; tmax = split;
            }
            if (n == &s->nodes[0]) break;
        }
        if (n->flags & NODE_LEAF) {
            ret |= cpu_leaf(cs, ray, n, bt, bp, bn);
            if (*bt < tmax) break;
        }
    }
    return ret;
}

/* Returns the number of nearest-hit rays traced (primary + shadow). */
uint64_t orc_cpu_trace_scene(const f4* verts, const tri_t* tris, uint32_t n_tris, const box_t* box,
                             const node_t* nodes, uint32_t n_nodes, const uint32_t* lp, const pcache_t* cache,
                             const cam_t* cam, uint32_t W, uint32_t H, uint32_t row0, uint32_t rows,
                             const light_t* lights, uint32_t n_lights, uint32_t* out, int nthreads) {
    scene_t s;
    make_scene(&s, verts, tris, n_tris, box, nodes, n_nodes, lp, NULL);
    cpu_scene_t cs = {&s, cache};
    uint64_t total = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
#endif
    for (long yi = 0; yi < (long)rows; yi++) {
        uint32_t y = row0 + (uint32_t)yi;
        for (uint32_t x = 0; x < W; x++) {
            /* castRay(j, i) of GLVisitorCPUTrace.cpp:201-216 */
            ray_t ray;
            float rr = (float)x / (float)W, rc = (float)y / (float)H;
            ray.o = cam->from;
            ray.d = v_sub(v_add(v_add(cam->ll, v_scale(cam->vf, rr)), v_scale(cam->hf, rc)), cam->from);
            ray.d = sse_normalize(ray.d);
            float bt; f4 bp, bn;
            total++;
            uint32_t px = 0;
            if (cpu_check_collision(&cs, &ray, &bt, &bp, &bn)) {
                float col[3] = {0.1f, 0.1f, 0.1f};
                for (uint32_t k = 0; k < n_lights; k++) {
                    const light_t* L = &lights[k];
                    if (L->color.w == 0.0f) {
                        float in = sse_dot(bn, L->data);
                        float m = in > 0 ? in : 0;
                        col[0] += L->color.x * m; col[1] += L->color.y * m; col[2] += L->color.z * m;
                    }
                    if (L->color.w == 1.0f) {
                        f4 tl = v_sub(L->data, bp);
                        tl.w = 0;
                        float dist = sqrtf(sse_dot(tl, tl));
                        tl = sse_normalize(tl);
                        ray_t sr;
                        sr.d = tl;
                        sr.o = v_add(bp, v_scale(tl, 0.001f));
                        float st; f4 sp_, sn;
                        int hit = cpu_check_collision(&cs, &sr, &st, &sp_, &sn);
                        total++;
                        float in = sse_dot(bn, tl);
                        if (hit && dist > st) in *= 0.5f;
                        float m = in > 0 ? in : 0;
                        col[0] += L->color.x * m; col[1] += L->color.y * m; col[2] += L->color.z * m;
                    }
                }
                px = orc_pack_rgb(col[0], col[1], col[2]);
            }
            if (out) out[(size_t)y * W + x] = px;
        }
    }
    return total;
}

size_t orc_pcache_size(void) { return sizeof(pcache_t); }
