// hulk_device.hpp - device-side arithmetic and HBM layouts for gfx950.
//
// Every function here restates one piece of the reference GPU arithmetic
// with a fixed operation order, compiled with -ffp-contract=off so that no
// multiply-add is fused unless written as __builtin_fmaf (the reference
// writes fma() in exactly three places, SharedFunctions.hpp:86-89).
// Conventions:
//   dot(a, b)    = (a.x*b.x + a.y*b.y) + a.z*b.z
//   cross(a, b)  = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)
//   normalize(v) = v * (1 / sqrt(dot(v, v)))       (IEEE sqrt and divide)
// Every reciprocal 1 / x goes through rcp_exact, which equals the IEEE
// division 1.0f / x bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hulk {

// Word x of a device node is exactly kNodeLeaf for a leaf and
// kNodeJunction | axis (axis 0-2) for a junction: the host upload and the
// device build write it so.  A junction is then x >= kNodeJunction.
constexpr uint32_t kNodeLeaf = 4u;
constexpr uint32_t kNodeJunction = 8u;
__device__ __forceinline__ bool is_junction(uint32_t x) { return x >= kNodeJunction; }
constexpr uint32_t kNoHit = 0xFFFFFFFFu;
constexpr uint32_t kRopeNone = 0xFFFFFFFFu;
constexpr float kFltMax = 3.402823466e+38f;
constexpr float kFltEpsilon = 1.192092896e-07f;

struct V3 {
    float x, y, z;
};

__device__ __forceinline__ V3 v3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 scale(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// Correctly rounded 1 / x.  For 2^-126 <= |x| < 2^126: v_rcp_f32 (within 1
// ulp) refined by one Newton-Raphson step written with two fma,
// y1 = y0 + y0 (1 - x y0) (Markstein, "IA-64 and Elementary Functions",
// ch. 8); hulk_selftest_reciprocal checks this against the IEEE division
// for every float of that range, with no difference on gfx950.  Zero,
// denormal, huge, infinite and NaN x take the IEEE division itself.
__device__ __forceinline__ float rcp_exact(float x) {
    const float ax = __builtin_fabsf(x);
    if (ax >= 0x1p-126f && ax < 0x1p126f) {
        const float y0 = __builtin_amdgcn_rcpf(x);
        return __builtin_fmaf(__builtin_fmaf(-x, y0, 1.0f), y0, y0);
    }
    return 1.0f / x;
}
// rcp_exact for an x already known not to lie below FLT_EPSILON in
// magnitude (or NaN): only the upper end of the fast range is tested.
__device__ __forceinline__ float rcp_exact_above_eps(float x) {
    if (__builtin_fabsf(x) < 0x1p126f) {
        const float y0 = __builtin_amdgcn_rcpf(x);
        return __builtin_fmaf(__builtin_fmaf(-x, y0, 1.0f), y0, y0);
    }
    return 1.0f / x;
}
__device__ __forceinline__ V3 normalize(V3 v) { return scale(v, rcp_exact(__builtin_sqrtf(dot(v, v)))); }
__device__ __forceinline__ float lane(V3 v, uint32_t axis) { return axis == 0 ? v.x : (axis == 1 ? v.y : v.z); }

// Per-leaf-reference triangle record, 64 B, 16-B aligned, one record per
// entry of m_leafPolygons so a leaf's records are contiguous:
//   a = (n.x, n.y, n.z, triangle id bits)
//   b = (p1.x, p1.y, p1.z, e00)
//   c = (e0.x, e0.y, e0.z, e01)
//   d = (e1.x, e1.y, e1.z, e11)
// n, e0, e1 and the Gram terms are the values checkRayToPoly computes per
// test (SharedFunctions.hpp:53-85), computed once by build_records with the
// same operations, so the per-test arithmetic that depends on the ray is
// unchanged and bit-identical.
struct TriRec {
    float4 a, b, c, d;
};

// Ray state shared by all traversal variants.
struct Ray {
    V3 o, d, inv;
};

struct Best {
    float t, s1, s2, det;
    uint32_t tri;
    float nx, ny, nz;
};

__device__ __forceinline__ void best_init(Best& b) {
    b.t = kFltMax;
    b.s1 = b.s2 = 0.0f;
    b.det = 1.0f;
    b.tri = kNoHit;
    b.nx = b.ny = b.nz = 0.0f;
}

// checkRayToPoly (SharedFunctions.hpp:47-102) on a precomputed record, then
// the strict '<' update of testNodePolygons (GPU_common.cl:62-68).  A test
// whose t is not below best.t cannot be kept, so it stops there.
__device__ __forceinline__ bool test_record(const Ray& r, const TriRec& rec, Best& best) {
    const V3 n = v3(rec.a.x, rec.a.y, rec.a.z);
    float den = dot(n, r.d);
    if (__builtin_fabsf(den) < kFltEpsilon) return false;
    den = rcp_exact_above_eps(den);
    const V3 p1 = v3(rec.b.x, rec.b.y, rec.b.z);
    const float t = dot(n, sub(p1, r.o)) * den;
    if (t < 0.0f) return false;
    if (!(t < best.t)) return false;
    const V3 P = add(r.o, scale(r.d, t));
    const V3 h = sub(P, p1);
    const V3 e0 = v3(rec.c.x, rec.c.y, rec.c.z);
    const V3 e1 = v3(rec.d.x, rec.d.y, rec.d.z);
    const float q0 = dot(e0, h);
    const float q1 = dot(e1, h);
    const float e00 = rec.b.w, e01 = rec.c.w, e11 = rec.d.w;
    const float det = __builtin_fmaf(e00, e11, -(e01 * e01));
    const float s1 = __builtin_fmaf(e11, q0, -(e01 * q1));
    const float s2 = __builtin_fmaf(e00, q1, -(e01 * q0));
    if (s1 >= 0.0f && s2 >= 0.0f && (s1 + s2) <= det) {
        best.t = t;
        best.s1 = s1;
        best.s2 = s2;
        best.det = det;
        best.tri = __float_as_uint(rec.a.w);
        best.nx = n.x;
        best.ny = n.y;
        best.nz = n.z;
        return true;
    }
    return false;
}

// Optional Moller-Trumbore test (T. Moller, B. Trumbore, "Fast, minimum
// storage ray-triangle intersection", JGT 2(1), 1997), the non-culling
// branch, on the record's p1, e0 = p2 - p1 and e1 = p3 - p1.  It is not the
// reference's test (checkRayToPoly solves the plane and a 2x2 Gram system),
// so its t, u, v round differently and its hits agree with the reference's
// up to ties and edge grazes; the hit keeps the reference's unit normal for
// shading and u, v directly (det = 1).  The paper's parallel cut,
// |det| < EPSILON, scales with the triangle's area (on 0.004-wide terrain
// triangles it drops rays up to 3 degrees off the plane), so the cut here is
// the reference's, |n.d| < FLT_EPSILON, written without a normal:
// det = -|e0 x e1| (n.d) and |e0 x e1|^2 = e00 e11 - e01^2 (the Gram
// determinant, from the record's terms), so the ray is cut when
// det^2 < FLT_EPSILON^2 (e00 e11 - e01^2).  The conditions of a hit are a
// conjunction of functions of the inputs, so their order changes no result;
// the cut is tested after u and v, where few triangles are left.  Fixed
// operation order: cross and dot as above, every 1 / x by rcp_exact.
__device__ __forceinline__ bool test_record_mt(const Ray& r, const TriRec& rec, Best& best) {
    const V3 e0 = v3(rec.c.x, rec.c.y, rec.c.z);
    const V3 e1 = v3(rec.d.x, rec.d.y, rec.d.z);
    const V3 pvec = cross(r.d, e1);
    const float det = dot(e0, pvec);
    if (det == 0.0f) return false;
    const float inv = rcp_exact(det);
    const V3 tvec = sub(r.o, v3(rec.b.x, rec.b.y, rec.b.z));
    const float u = dot(tvec, pvec) * inv;
    if (u < 0.0f || u > 1.0f) return false;
    const V3 qvec = cross(tvec, e0);
    const float v = dot(r.d, qvec) * inv;
    if (v < 0.0f || u + v > 1.0f) return false;
    const float gram = __builtin_fmaf(rec.b.w, rec.d.w, -(rec.c.w * rec.c.w));
    if (det * det < (kFltEpsilon * kFltEpsilon) * gram) return false;
    const float t = dot(e1, qvec) * inv;
    if (t < 0.0f) return false;
    if (!(t < best.t)) return false;
    best.t = t;
    best.s1 = u;
    best.s2 = v;
    best.det = 1.0f;
    best.tri = __float_as_uint(rec.a.w);
    best.nx = rec.a.x;
    best.ny = rec.a.y;
    best.nz = rec.a.z;
    return true;
}

// The record test of leaf mode MT (Moller-Trumbore) or the reference test.
template <bool MT>
__device__ __forceinline__ bool test_any(const Ray& r, const TriRec& rec, Best& best) {
    return MT ? test_record_mt(r, rec, best) : test_record(r, rec, best);
}

// The same test straight from three vertices (brute force variant,
// GPUTrace_bruteforce.cl:30-42).
template <bool MT = false>
__device__ __forceinline__ void test_vertices(const Ray& r, V3 p1, V3 p2, V3 p3, uint32_t id, Best& best) {
    const V3 e0 = sub(p2, p1), e1 = sub(p3, p1), e2 = sub(p3, p2);
    const V3 n = normalize(add(add(cross(e0, e1), cross(e1, e2)), cross(e2, e0)));
    TriRec rec;
    rec.a = make_float4(n.x, n.y, n.z, __uint_as_float(id));
    rec.b = make_float4(p1.x, p1.y, p1.z, dot(e0, e0));
    rec.c = make_float4(e0.x, e0.y, e0.z, dot(e0, e1));
    rec.d = make_float4(e1.x, e1.y, e1.z, dot(e1, e1));
    test_any<MT>(r, rec, best);
}

// isRayToAbovePlane / isRayToBelowPlane / checkRayToBoundingBoxGPU
// (SharedFunctions.hpp:154-264).
__device__ __forceinline__ bool above_plane(float p0, float d0, float e0, float& t0, float& t1) {
    if (d0 > 0.0f) {
        float ti = (e0 - p0) / d0;
        if (ti > t1) return false;
        if (ti > t0) t0 = ti;
        return true;
    } else if (d0 < 0.0f) {
        float ti = (e0 - p0) / d0;
        if (ti < t0) return false;
        if (ti < t1) t1 = ti;
        return true;
    }
    return e0 < p0;
}
__device__ __forceinline__ bool below_plane(float p0, float d0, float e0, float& t0, float& t1) {
    if (d0 < 0.0f) {
        float ti = (e0 - p0) / d0;
        if (ti > t1) return false;
        if (ti > t0) t0 = ti;
        return true;
    } else if (d0 > 0.0f) {
        float ti = (e0 - p0) / d0;
        if (ti < t0) return false;
        if (ti < t1) t1 = ti;
        return true;
    }
    return e0 > p0;
}

// The same plane test with one IEEE quotient and no branch: for d > 0 the
// reference tests ti > t1 (miss) and raises t0 to ti; for d < 0 it tests
// ti < t0 (miss) and lowers t1 to ti; for d == 0 the plane only decides by
// the origin's side.  A plane checks only the bound it does not move, so
// the order of check and update within a plane is free, and once a plane
// misses, the box test's result no longer depends on t0 / t1.  Divergent
// rays of a wave then pay one division per plane instead of one for each
// sign of d; the branches of the reference's form hold fewer registers,
// which the frame kernels under a wave cap need.
__device__ __forceinline__ void box_plane(float p0, float d0, float e0, bool above, bool& ok, float& t0,
                                          float& t1) {
    const float ti = (e0 - p0) / d0;
    const bool enter = above ? d0 > 0.0f : d0 < 0.0f;  // ti is where the ray enters the half space
    const bool leave = above ? d0 < 0.0f : d0 > 0.0f;
    const bool side = above ? e0 < p0 : e0 > p0;
    ok &= enter ? !(ti > t1) : (leave ? !(ti < t0) : side);
    t0 = (enter && ti > t0) ? ti : t0;
    t1 = (leave && ti < t1) ? ti : t1;
}

struct SceneBox {
    float mn[3], mx[3];
};

__device__ __forceinline__ bool ray_box(const SceneBox& b, const Ray& r, float& ts, float& te) {
    float t0 = 0.0f, t1 = kFltMax;
    bool ok = above_plane(r.o.x, r.d.x, b.mn[0], t0, t1);
    ok &= below_plane(r.o.x, r.d.x, b.mx[0], t0, t1);
    ok &= above_plane(r.o.y, r.d.y, b.mn[1], t0, t1);
    ok &= below_plane(r.o.y, r.d.y, b.mx[1], t0, t1);
    ok &= above_plane(r.o.z, r.d.z, b.mn[2], t0, t1);
    ok &= below_plane(r.o.z, r.d.z, b.mx[2], t0, t1);
    if (ok) {
        ts = t0;
        te = t1;
    }
    return ok;
}

// ray_box with the branch-free plane test: the same result bit for bit.
__device__ __forceinline__ bool ray_box_nb(const SceneBox& b, const Ray& r, float& ts, float& te) {
    float t0 = 0.0f, t1 = kFltMax;
    bool ok = true;
    box_plane(r.o.x, r.d.x, b.mn[0], true, ok, t0, t1);
    box_plane(r.o.x, r.d.x, b.mx[0], false, ok, t0, t1);
    box_plane(r.o.y, r.d.y, b.mn[1], true, ok, t0, t1);
    box_plane(r.o.y, r.d.y, b.mx[1], false, ok, t0, t1);
    box_plane(r.o.z, r.d.z, b.mn[2], true, ok, t0, t1);
    box_plane(r.o.z, r.d.z, b.mx[2], false, ok, t0, t1);
    if (ok) {
        ts = t0;
        te = t1;
    }
    return ok;
}

// castRay (GPU_common.cl:10-30): d = ((lowerLeft + vF*(x/W)) + hF*(y/H)) - from,
// not normalized (the normalize() result is discarded at :27).
struct Camera {
    float from[3], vf[3], hf[3], ll[3];
};

__device__ __forceinline__ void cast_ray(const Camera& c, uint32_t x, uint32_t y, uint32_t W, uint32_t H, Ray& r) {
    const float rr = (float)y / (float)H;
    const float rc = (float)x / (float)W;
    r.o = v3(c.from[0], c.from[1], c.from[2]);
    const V3 ll = v3(c.ll[0], c.ll[1], c.ll[2]);
    const V3 vf = v3(c.vf[0], c.vf[1], c.vf[2]);
    const V3 hf = v3(c.hf[0], c.hf[1], c.hf[2]);
    r.d = sub(add(add(ll, scale(vf, rc)), scale(hf, rr)), r.o);
}

// castRay with the image size already as floats (fW = (float)W, fH =
// (float)H): the same divisions, with W and H as scalar kernel arguments.
__device__ __forceinline__ void cast_ray_f(const Camera& c, uint32_t x, uint32_t y, float fW, float fH, Ray& r) {
    const float rr = (float)y / fH;
    const float rc = (float)x / fW;
    r.o = v3(c.from[0], c.from[1], c.from[2]);
    const V3 ll = v3(c.ll[0], c.ll[1], c.ll[2]);
    const V3 vf = v3(c.vf[0], c.vf[1], c.vf[2]);
    const V3 hf = v3(c.hf[0], c.hf[1], c.hf[2]);
    r.d = sub(add(add(ll, scale(vf, rc)), scale(hf, rr)), r.o);
}

__device__ __forceinline__ void set_inv(Ray& r) {
    r.inv = v3(rcp_exact(r.d.x), rcp_exact(r.d.y), rcp_exact(r.d.z));
}

}  // namespace hulk
