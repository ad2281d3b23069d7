// scenes.cpp - seeded procedural scenes and camera construction (host).
//
// Camera: GLVisitorCPUTrace::updateCamera (HulkEngine/src/GLVisitorCPUTrace.cpp:
// 105-148) builds CameraPos from a rotation matrix R (SPMatrix::rotateMatrix,
// HulkEngine/src/SPMatrix.cpp:54-84) and fov / aspect.  Under the row-vector
// product of SPMatrix the basis vectors are left = -R.row1, down = -R.row2,
// to = -R.row3 (w cleared), then
//   vf = tan(fov/2), hf = vf * aspect,
//   lowerLeft = ((from + to) + down*vf) + left*hf,
//   verticalFactor = (left*hf) * -2, horizontalFactor = (down*vf) * -2.
// Trigonometry here is evaluated in double with a fixed polynomial so the
// result does not depend on the host libm.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "hulk_abi.h"
#include "hulk_scenes.h"
#include "scene_data.hpp"

namespace {

constexpr double kPi = 3.14159265358979323846;

// sin / cos by argument reduction to [-pi, pi] and a Taylor series to x^27
// (error < 1e-15 on that interval); only + and * so results are host-independent.
double dsin(double x) {
    double k = std::floor(x / (2.0 * kPi) + 0.5);
    x = x - k * (2.0 * kPi);
    double x2 = x * x, term = x, sum = x;
    for (int n = 1; n <= 13; n++) {
        term = -term * x2 / ((2.0 * n) * (2.0 * n + 1.0));
        sum += term;
    }
    return sum;
}
double dcos(double x) { return dsin(x + 0.5 * kPi); }
double dtan(double x) { return dsin(x) / dcos(x); }
// exp / log for the log-uniform radii: Taylor exp after range reduction,
// log by Newton on exp.  Deterministic, not correctly rounded.
double dexp(double x) {
    int k = static_cast<int>(std::floor(x / 0.6931471805599453 + 0.5));
    double r = x - k * 0.6931471805599453;
    double term = 1.0, sum = 1.0;
    for (int n = 1; n <= 25; n++) {
        term = term * r / n;
        sum += term;
    }
    return std::ldexp(sum, k);
}
double dlog(double y) {
    double x = 0.0;
    int e;
    double m = std::frexp(y, &e);
    x = (m - 1.0);
    for (int i = 0; i < 60; i++) x = x - 1.0 + m / dexp(x);
    return x + e * 0.6931471805599453;
}

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {  // splitmix64
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0); }
};

uint64_t hash3(uint64_t seed, int64_t a, int64_t b, int64_t c) {
    uint64_t h = seed ^ 0x243F6A8885A308D3ull;
    for (int64_t v : {a, b, c}) {
        h ^= static_cast<uint64_t>(v) + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
        h = (h ^ (h >> 31)) * 0x7FB5D329728EA185ull;
        h = (h ^ (h >> 27)) * 0x81DADEF4BC2DD44Dull;
        h ^= h >> 33;
    }
    return h;
}

struct V3 { double x, y, z; };
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
V3 normalize(V3 a) { return a * (1.0 / std::sqrt(dot(a, a))); }

hulk_float4 f4(V3 v, float w) {
    return {static_cast<float>(v.x), static_cast<float>(v.y), static_cast<float>(v.z), w};
}

// Value noise: 4 octaves, base lattice 8, smoothstep interpolation, in [0, 1).
double value_noise(uint64_t seed, double x, double z) {
    double sum = 0.0, amp = 1.0, norm = 0.0;
    for (int o = 0; o < 4; o++) {
        double f = 8.0 * static_cast<double>(1 << o);
        double fx = x * f, fz = z * f;
        double ix = std::floor(fx), iz = std::floor(fz);
        double tx = fx - ix, tz = fz - iz;
        tx = tx * tx * (3.0 - 2.0 * tx);
        tz = tz * tz * (3.0 - 2.0 * tz);
        auto lat = [&](double a, double b) {
            return static_cast<double>(hash3(seed + o, static_cast<int64_t>(a), static_cast<int64_t>(b), 0) >> 11) *
                   (1.0 / 9007199254740992.0);
        };
        double v00 = lat(ix, iz), v10 = lat(ix + 1, iz), v01 = lat(ix, iz + 1), v11 = lat(ix + 1, iz + 1);
        double a = v00 + (v10 - v00) * tx, b = v01 + (v11 - v01) * tx;
        sum += amp * (a + (b - a) * tz);
        norm += amp;
        amp *= 0.5;
    }
    return sum / norm;
}

}  // namespace

double hulk_host_sin(double x) { return dsin(x); }
double hulk_host_cos(double x) { return dcos(x); }

namespace {

void add_tri(hulk_scene* s, uint32_t a, uint32_t b, uint32_t c) { s->tris.push_back({{a, b, c, 0}}); }

void gen_random_tris(hulk_scene* s, uint32_t n) {
    Rng rng(s->seed);
    s->verts.reserve(3 * n);
    for (uint32_t i = 0; i < n; i++) {
        V3 c = {rng.uniform(), rng.uniform(), rng.uniform()};
        for (int k = 0; k < 3; k++) {
            V3 off;
            do {
                off = {2.0 * rng.uniform() - 1.0, 2.0 * rng.uniform() - 1.0, 2.0 * rng.uniform() - 1.0};
            } while (dot(off, off) > 1.0);
            s->verts.push_back(f4(c + off * 0.05, 1.0f));
        }
        add_tri(s, 3 * i, 3 * i + 1, 3 * i + 2);
    }
    // mirror flags (used only by renders with mirror generations): p = 0.3
    // per triangle from a separate stream, so the geometry does not depend on it
    Rng mr(s->seed ^ 0x5DEECE66Dull);
    s->mirror.resize(n);
    for (uint32_t i = 0; i < n; i++) s->mirror[i] = mr.uniform() < 0.3 ? 1 : 0;
    s->lights.push_back({{0.9f, 0.9f, 0.9f, HULK_LIGHT_POINT}, {2.0f, 3.0f, 2.0f, 1.0f}});
}

double terrain_height(uint64_t seed, double x, double z) { return 0.2 * value_noise(seed, x, z); }

void add_uv_sphere(hulk_scene* s, V3 c, double r, int stacks, int slices) {
    const uint32_t base = static_cast<uint32_t>(s->verts.size());
    s->verts.push_back(f4(c + V3{0, r, 0}, 1.0f));  // north pole
    for (int i = 1; i < stacks; i++) {
        double phi = kPi * i / stacks;
        double sp = dsin(phi), cp = dcos(phi);
        for (int j = 0; j < slices; j++) {
            double th = 2.0 * kPi * j / slices;
            s->verts.push_back(f4(c + V3{r * sp * dcos(th), r * cp, r * sp * dsin(th)}, 1.0f));
        }
    }
    s->verts.push_back(f4(c + V3{0, -r, 0}, 1.0f));  // south pole
    const uint32_t south = static_cast<uint32_t>(s->verts.size()) - 1;
    auto ring = [&](int i, int j) { return base + 1 + (i - 1) * slices + (j % slices); };
    for (int j = 0; j < slices; j++) add_tri(s, base, ring(1, j + 1), ring(1, j));
    for (int i = 1; i < stacks - 1; i++)
        for (int j = 0; j < slices; j++) {
            add_tri(s, ring(i, j), ring(i, j + 1), ring(i + 1, j + 1));
            add_tri(s, ring(i, j), ring(i + 1, j + 1), ring(i + 1, j));
        }
    for (int j = 0; j < slices; j++) add_tri(s, south, ring(stacks - 1, j), ring(stacks - 1, j + 1));
}

void gen_terrain_spheres(hulk_scene* s, uint32_t n_spheres) {
    const int N = 256;
    s->verts.reserve(N * N + n_spheres * 90);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            double x = static_cast<double>(j) / (N - 1), z = static_cast<double>(i) / (N - 1);
            s->verts.push_back(f4({x, terrain_height(s->seed, x, z), z}, 1.0f));
        }
    for (int i = 0; i + 1 < N; i++)
        for (int j = 0; j + 1 < N; j++) {
            uint32_t v00 = i * N + j, v10 = v00 + 1, v01 = v00 + N, v11 = v01 + 1;
            add_tri(s, v00, v11, v10);
            add_tri(s, v00, v01, v11);
        }
    Rng rng(s->seed ^ 0xA5A5A5A5ull);
    const double lo = dlog(0.01), hi = dlog(0.2);
    for (uint32_t k = 0; k < n_spheres; k++) {
        double r = dexp(lo + rng.uniform() * (hi - lo));
        double x = 0.05 + 0.9 * rng.uniform(), z = 0.05 + 0.9 * rng.uniform();
        double y = terrain_height(s->seed, x, z) + 0.8 * r;
        add_uv_sphere(s, {x, y, z}, r, 8, 12);
    }
    // mirror spheres with p = 0.3 (separate stream); the terrain is diffuse
    Rng mr(s->seed ^ 0x5DEECE66Dull);
    s->mirror.assign(s->tris.size(), 0);
    const size_t terrain = (size_t)2 * (N - 1) * (N - 1), per = (s->tris.size() - terrain) / std::max<uint32_t>(n_spheres, 1);
    for (uint32_t k = 0; k < n_spheres; k++)
        if (mr.uniform() < 0.3) std::memset(&s->mirror[terrain + k * per], 1, per);
    s->lights.push_back({{0.9f, 0.85f, 0.8f, HULK_LIGHT_POINT}, {0.3f, 1.5f, 0.2f, 1.0f}});
    V3 d = normalize({0.3, 0.8, 0.5});
    s->lights.push_back({{0.2f, 0.2f, 0.25f, HULK_LIGHT_DIRECTIONAL}, f4(d, 0.0f)});
}

void gen_animated_tori(hulk_scene* s, uint32_t n_meshes, uint32_t frame) {
    const int U = 100, Vn = 100;
    s->verts.reserve(static_cast<size_t>(n_meshes) * U * Vn + 4);
    s->tris.reserve(static_cast<size_t>(n_meshes) * U * Vn * 2 + 2);
    Rng rng(s->seed);
    std::vector<uint8_t> mesh_mirror(n_meshes);
    for (uint32_t m = 0; m < n_meshes; m++) {
        const double scale = 0.5 + 0.5 * rng.uniform();
        const double R = 0.3 * scale, r = 0.1 * scale;
        const double gx = -4.5 + (m % 10) * 1.0, gz = -2.0 + (m / 10) * 1.0;
        const double phase = 2.0 * kPi * rng.uniform();
        const double tilt = kPi * (rng.uniform() - 0.5);
        mesh_mirror[m] = rng.uniform() < 0.3 ? 1 : 0;
        // rigid motion of frame f: rotation about y by 0.05*f + phase,
        // centre on a circle of radius 0.25 around the grid slot.
        const double ang = 0.05 * frame + phase;
        const double ca = dcos(ang), sa = dsin(ang), ct = dcos(tilt), st = dsin(tilt);
        const V3 centre = {gx + 0.25 * dcos(0.1 * frame + phase), 0.45 + R + r, gz + 0.25 * dsin(0.1 * frame + phase)};
        const uint32_t base = static_cast<uint32_t>(s->verts.size());
        for (int i = 0; i < U; i++) {
            const double a = 2.0 * kPi * i / U;
            for (int j = 0; j < Vn; j++) {
                const double b = 2.0 * kPi * j / Vn;
                V3 p = {(R + r * dcos(b)) * dcos(a), r * dsin(b), (R + r * dcos(b)) * dsin(a)};
                // tilt about x, then spin about y
                V3 q = {p.x, p.y * ct - p.z * st, p.y * st + p.z * ct};
                V3 w = {q.x * ca + q.z * sa, q.y, -q.x * sa + q.z * ca};
                s->verts.push_back(f4(centre + w, 1.0f));
            }
        }
        for (int i = 0; i < U; i++)
            for (int j = 0; j < Vn; j++) {
                uint32_t v00 = base + i * Vn + j;
                uint32_t v10 = base + ((i + 1) % U) * Vn + j;
                uint32_t v01 = base + i * Vn + (j + 1) % Vn;
                uint32_t v11 = base + ((i + 1) % U) * Vn + (j + 1) % Vn;
                add_tri(s, v00, v10, v11);
                add_tri(s, v00, v11, v01);
            }
    }
    const uint32_t g = static_cast<uint32_t>(s->verts.size());
    s->verts.push_back({-6.0f, 0.0f, -4.0f, 1.0f});
    s->verts.push_back({6.0f, 0.0f, -4.0f, 1.0f});
    s->verts.push_back({6.0f, 0.0f, 4.0f, 1.0f});
    s->verts.push_back({-6.0f, 0.0f, 4.0f, 1.0f});
    add_tri(s, g, g + 2, g + 1);
    add_tri(s, g, g + 3, g + 2);
    s->mirror.resize(s->tris.size());
    for (uint32_t m = 0; m < n_meshes; m++)
        std::memset(&s->mirror[static_cast<size_t>(m) * U * Vn * 2], mesh_mirror[m], U * Vn * 2);
    s->mirror[s->tris.size() - 2] = s->mirror[s->tris.size() - 1] = 0;
    s->lights.push_back({{0.85f, 0.85f, 0.8f, HULK_LIGHT_POINT}, {3.0f, 8.0f, 4.0f, 1.0f}});
    V3 d = normalize({-0.4, 0.8, 0.3});
    s->lights.push_back({{0.15f, 0.15f, 0.2f, HULK_LIGHT_DIRECTIONAL}, f4(d, 0.0f)});
}

void camera_from_basis(V3 from, V3 right, V3 upv, V3 fwd, float fov_deg, float aspect, hulk_camera* out) {
    // left = -right, down = -up, to = forward (rows of R: right, up, -forward)
    hulk_float4 F = f4(from, 0.0f);
    hulk_float4 left = f4(right * -1.0, 0.0f), down = f4(upv * -1.0, 0.0f), to = f4(fwd, 0.0f);
    const float fov_rad = fov_deg * 0.017453292519943f;  // DEG_TO_RAD, util.hpp:57
    const float vf = static_cast<float>(dtan(static_cast<double>(fov_rad / 2)));
    const float hf = vf * aspect;
    hulk_float4 dn = {down.x * vf, down.y * vf, down.z * vf, 0.0f};
    hulk_float4 lf = {left.x * hf, left.y * hf, left.z * hf, 0.0f};
    out->from = F;
    out->lower_left = {((F.x + to.x) + dn.x) + lf.x, ((F.y + to.y) + dn.y) + lf.y, ((F.z + to.z) + dn.z) + lf.z, 0.0f};
    out->vertical_factor = {lf.x * -2.0f, lf.y * -2.0f, lf.z * -2.0f, 0.0f};
    out->horizontal_factor = {dn.x * -2.0f, dn.y * -2.0f, dn.z * -2.0f, 0.0f};
}

}  // namespace

extern "C" int hulk_make_camera_lookat(const float from[3], const float target[3], const float up[3],
                                       float fov_deg, float aspect, hulk_camera* out) {
    if (!from || !target || !up || !out) return HULK_E_ILLEGAL_INPUT;
    V3 F = {from[0], from[1], from[2]}, T = {target[0], target[1], target[2]}, U = {up[0], up[1], up[2]};
    V3 fwd = normalize(T - F);
    V3 right = normalize(cross(fwd, U));
    V3 upv = cross(right, fwd);
    camera_from_basis(F, right, upv, fwd, fov_deg, aspect, out);
    return HULK_OK;
}

// rotateMatrix (SPMatrix.cpp:54-84) with cos / sin in double rounded to float.
extern "C" int hulk_make_camera(const float from[3], const float axis_angle[4], float fov_deg, float aspect,
                                hulk_camera* out) {
    if (!from || !axis_angle || !out) return HULK_E_ILLEGAL_INPUT;
    const float x = axis_angle[0], y = axis_angle[1], z = axis_angle[2];
    const float c = static_cast<float>(dcos(axis_angle[3])), s = static_cast<float>(dsin(axis_angle[3]));
    const float oc = 1 - c;
    const float xy = oc * x * y, xz = oc * x * z, yz = oc * y * z;
    const float r11 = c + oc * x * x, r12 = xy - s * z, r13 = xz + s * y;
    const float r21 = xy + s * z, r22 = c + oc * y * y, r23 = yz - s * x;
    const float r31 = xz - s * y, r32 = yz + s * x, r33 = c + oc * z * z;
    hulk_float4 F = {from[0], from[1], from[2], 0.0f};
    hulk_float4 left = {-r11, -r12, -r13, 0.0f}, down = {-r21, -r22, -r23, 0.0f}, to = {-r31, -r32, -r33, 0.0f};
    const float fov_rad = fov_deg * 0.017453292519943f;
    const float vf = static_cast<float>(dtan(static_cast<double>(fov_rad / 2)));
    const float hf = vf * aspect;
    hulk_float4 dn = {down.x * vf, down.y * vf, down.z * vf, 0.0f};
    hulk_float4 lf = {left.x * hf, left.y * hf, left.z * hf, 0.0f};
    out->from = F;
    out->lower_left = {((F.x + to.x) + dn.x) + lf.x, ((F.y + to.y) + dn.y) + lf.y, ((F.z + to.z) + dn.z) + lf.z, 0.0f};
    out->vertical_factor = {lf.x * -2.0f, lf.y * -2.0f, lf.z * -2.0f, 0.0f};
    out->horizontal_factor = {dn.x * -2.0f, dn.y * -2.0f, dn.z * -2.0f, 0.0f};
    return HULK_OK;
}

extern "C" int hulk_scene_create(int kind, uint64_t seed, uint32_t param, uint32_t frame, hulk_scene** out) {
    if (!out) return HULK_E_ILLEGAL_INPUT;
    hulk_scene* s = new (std::nothrow) hulk_scene;
    if (!s) return HULK_E_ALLOC;
    s->kind = kind;
    s->seed = seed;
    try {
        switch (kind) {
            case HULK_SCENE_RANDOM_TRIS: gen_random_tris(s, param ? param : 1000); break;
            case HULK_SCENE_TERRAIN_SPHERES: gen_terrain_spheres(s, param ? param : 200); break;
            case HULK_SCENE_ANIMATED_TORI: gen_animated_tori(s, param ? param : 50, frame); break;
            default: delete s; return HULK_E_ILLEGAL_INPUT;
        }
    } catch (const std::bad_alloc&) {
        delete s;
        return HULK_E_ALLOC;
    }
    *out = s;
    return HULK_OK;
}

extern "C" void hulk_scene_free(hulk_scene* s) { delete s; }
extern "C" uint32_t hulk_scene_vertex_count(const hulk_scene* s) { return s ? (uint32_t)s->verts.size() : 0; }
extern "C" uint32_t hulk_scene_triangle_count(const hulk_scene* s) { return s ? (uint32_t)s->tris.size() : 0; }
extern "C" uint32_t hulk_scene_light_count(const hulk_scene* s) { return s ? (uint32_t)s->lights.size() : 0; }
extern "C" const hulk_float4* hulk_scene_vertices(const hulk_scene* s) { return s ? s->verts.data() : nullptr; }
extern "C" const hulk_tri* hulk_scene_triangles(const hulk_scene* s) { return s ? s->tris.data() : nullptr; }
extern "C" const hulk_light* hulk_scene_lights(const hulk_scene* s) { return s ? s->lights.data() : nullptr; }
extern "C" const uint8_t* hulk_scene_mirror_mask(const hulk_scene* s) { return s ? s->mirror.data() : nullptr; }
extern "C" int hulk_scene_get_box(const hulk_scene* s, hulk_box* out) {
    if (!s) return HULK_E_ILLEGAL_INPUT;
    return hulk_scene_box(s->verts.data(), (uint32_t)s->verts.size(), out);
}

extern "C" int hulk_scene_camera(const hulk_scene* s, uint32_t frame, float aspect, hulk_camera* out) {
    if (!s || !out) return HULK_E_ILLEGAL_INPUT;
    if (s->has_camera) {  // a scene file's camera at its load time; aspect <= 0: the file's
        const float aa[4] = {s->cam_axis[0], s->cam_axis[1], s->cam_axis[2], s->cam_angle};
        return hulk_make_camera(s->cam_from, aa, s->cam_fov, aspect > 0.0f ? aspect : s->cam_aspect, out);
    }
    return hulk_scene_kind_camera(s->kind, frame, aspect, out);
}

extern "C" int hulk_scene_kind_camera(int kind, uint32_t frame, float aspect, hulk_camera* out) {
    if (!out) return HULK_E_ILLEGAL_INPUT;
    float up[3] = {0.0f, 1.0f, 0.0f};
    switch (kind) {
        case HULK_SCENE_RANDOM_TRIS: {
            // from (0.5, 0.5, 2.2) looking down -z, fov 45 (identity rotation)
            float from[3] = {0.5f, 0.5f, 2.2f};
            float aa[4] = {0.0f, 1.0f, 0.0f, 0.0f};
            return hulk_make_camera(from, aa, 45.0f, aspect, out);
        }
        case HULK_SCENE_TERRAIN_SPHERES: {
            // 20 poses on an arc around the terrain, looking at its centre:
            // poses 0-9 at angles k/10 of the arc, poses 10-19 half-way
            // between them (the second pass of the camera path)
            const uint32_t k = frame % 20;
            const double th = -0.5 * kPi + 0.5 * kPi * ((k % 10) + 0.5 * (k / 10)) / 10.0;
            float from[3] = {static_cast<float>(0.5 + 0.75 * dcos(th)), 0.55f,
                             static_cast<float>(0.5 + 0.75 * dsin(th) + 0.0)};
            float target[3] = {0.5f, 0.08f, 0.5f};
            return hulk_make_camera_lookat(from, target, up, 50.0f, aspect, out);
        }
        case HULK_SCENE_ANIMATED_TORI: {
            const double th = 0.02 * frame;
            float from[3] = {static_cast<float>(7.5 * dsin(th)), 4.0f, static_cast<float>(7.5 * dcos(th))};
            float target[3] = {0.0f, 0.4f, 0.0f};
            return hulk_make_camera_lookat(from, target, up, 50.0f, aspect, out);
        }
        default: return HULK_E_ILLEGAL_INPUT;
    }
}
