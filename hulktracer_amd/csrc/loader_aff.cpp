// loader_aff.cpp - BART scenes (AFF files, "A Benchmark for Animated Ray
// Tracing") as the reference's AFFParser reads them (host).
//
// Parsing: AFFParser::viParseFile (BartLoader/src/AFFParser.cpp:1469-1539)
// dispatches on the first character of each entity:
//   m    triangle mesh: "vertices n" xyz..., optional "normals n" xyz...,
//        optional "texturecoords n name" uv..., "triangles n" then per
//        triangle three vertex indices, then three normal / three texture
//        indices when those lists exist (getTriangles, :1275-1344).  Indices
//        are kept as unsigned short, as the reference stores them.
//   xs   static transform: scale xyz, rotation axis xyz and degrees,
//        translation xyz, then '{' (parseXform, :1074-1157);
//   x    key-framed transform: name, then '{'; '}' ends either kind;
//   k    key frames of a named transform: transl / rot / scale (at least 4
//        keys: time, value, tension, continuity, bias) and visibility
//        (parseKeyFrames, :894-1056);
//   l    point light: position, optional colour (parseLight, :178-233);
//   v    viewpoint: from / at / up / angle / hither / resolution; the
//        reference keeps only the angle (fov) and the resolution (:107-156);
//   i    include "detail file", relative to the main file, when detail <=
//        the current detail level; d sets that level (:531-590);
//   #, % comment to the end of the line (at most 999 characters, :56-60);
//   b, f, fm, c, s, p, pp, tt, ttp, tpa, a, am: parsed and dropped - the
//        reference adds no geometry or state for them.
// A malformed entity, which makes the reference exit(1), returns
// HULK_E_ILLEGAL_INPUT here.
//
// Geometry at scene time t: AFFTransformVisitor::preAct
// (BartLoader/src/AFFTransformVisitor.cpp:24-104) walks the transform tree
// with row-vector matrices: a static transform contributes
// S * R(axis, deg) * T * parent, a key-framed one S(t) * R(axis(t),
// -angle(t)) * T(t) * parent from its splines (each factor only where the
// transform has that kind of key); mesh vertices (w = 1) are multiplied by
// the product (VectorMult lane order, SPMatrix.hpp:170-178).  Splines:
// Kochanek-Bartels position splines (kbsplpos.cpp: KB_PosInitialize,
// KB_PosInterpolate0) and the quaternion form for rotations (kbsplrot.cpp:
// KB_RotInitialize, KB_RotInterpolate over the quaternion operations of
// quat.cpp), restated below from those steps.  The camera follows the
// "camera" transform (AFFParser::getTransformation, :1541-1563, and
// GLVisitorCPUTrace::updateCamera, GLVisitorCPUTrace.cpp:105-148).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "hulk_abi.h"
#include "hulk_scenes.h"
#include "scene_data.hpp"

namespace {

// ------------------------------------------------------------- matrices --

struct Mat {
    float m[4][4];
};

Mat identity() {
    Mat a{};
    for (int i = 0; i < 4; i++) a.m[i][i] = 1.0f;
    return a;
}
Mat translate(float x, float y, float z) {
    Mat a = identity();
    a.m[3][0] = x;
    a.m[3][1] = y;
    a.m[3][2] = z;
    return a;
}
Mat scale(float x, float y, float z) {
    Mat a = identity();
    a.m[0][0] = x;
    a.m[1][1] = y;
    a.m[2][2] = z;
    return a;
}
// SPMatrix::rotateMatrix (SPMatrix.cpp:54-84).
Mat rotate(float x, float y, float z, float angle) {
    const float c = static_cast<float>(hulk_host_cos(angle)), s = static_cast<float>(hulk_host_sin(angle));
    const float oc = 1 - c;
    const float xy = oc * x * y, xz = oc * x * z, yz = oc * y * z;
    Mat a{};
    a.m[0][0] = c + oc * x * x; a.m[0][1] = xy - s * z;     a.m[0][2] = xz + s * y;
    a.m[1][0] = xy + s * z;     a.m[1][1] = c + oc * y * y; a.m[1][2] = yz - s * x;
    a.m[2][0] = xz - s * y;     a.m[2][1] = yz + s * x;     a.m[2][2] = c + oc * z * z;
    a.m[3][3] = 1.0f;
    return a;
}
// [A] = [A] * [B], each row as VectorMult: ((a0 B1 + a1 B2) + a2 B3) + a3 B4.
Mat mul(const Mat& A, const Mat& B) {
    Mat r{};
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 4; k++)
            r.m[i][k] = ((A.m[i][0] * B.m[0][k] + A.m[i][1] * B.m[1][k]) + A.m[i][2] * B.m[2][k]) + A.m[i][3] * B.m[3][k];
    return r;
}
hulk_float4 apply(const hulk_float4& v, const Mat& M) {
    float o[4];
    for (int k = 0; k < 4; k++) o[k] = ((v.x * M.m[0][k] + v.y * M.m[1][k]) + v.z * M.m[2][k]) + v.w * M.m[3][k];
    return {o[0], o[1], o[2], o[3]};
}

// --------------------------------------------------- key-frame splines --
//
// Kochanek-Bartels splines (D. Kochanek and R. Bartels, "Interpolating
// splines with local tension, continuity, and bias control", SIGGRAPH 1984).
// A key has a time T, a value V and parameters (tension t, continuity c,
// bias b).  At key i, with the chords a = V_i - V_{i-1} and e = V_{i+1} - V_i,
// the tangent that leaves the key (start of the segment i -> i+1) and the one
// that arrives at it (end of the segment i-1 -> i) are
//   leave  = (1-t)(1+c)(1+b)/2 a + (1-t)(1-c)(1-b)/2 e
//   arrive = (1-t)(1-c)(1+b)/2 a + (1-t)(1+c)(1-b)/2 e
// and, keys being unevenly spaced, each is rescaled to the span of the
// segment it belongs to (section 5 of the paper): leave by
// 2 (T_{i+1} - T_i) / (T_{i+1} - T_{i-1}), arrive by 2 (T_i - T_{i-1}) /
// (T_{i+1} - T_{i-1}).  Only keys with two neighbours get tangents, so n keys
// make n - 3 cubic Hermite segments over [T_1, T_{n-2}]; earlier and later
// times clamp to the ends.
//
// Rotations take the same tangents in the tangent space of unit quaternions
// (chords log(q_{i-1}^-1 q_i)) and are interpolated with Shoemake's
// spherical quadrangle (K. Shoemake, "Animating rotation with quaternion
// curves", SIGGRAPH 1985): between q_i and q_{i+1}, with the chord
// h = log(q_i^-1 q_{i+1}), the inner control points are
// q_i exp((leave_i - h) / 2) and q_{i+1} exp((h - arrive_{i+1}) / 2).
// Everything is evaluated in double precision.

struct Key {
    double t, v[4];  // value: xyz (+ angle in radians for rotations)
    double tension, continuity, bias;
};

// A value in a spline's tangent space: a point or a quaternion log.
struct Vec3 {
    double c[3];
    Vec3 operator+(const Vec3& o) const { return {{c[0] + o.c[0], c[1] + o.c[1], c[2] + o.c[2]}}; }
    Vec3 operator-(const Vec3& o) const { return {{c[0] - o.c[0], c[1] - o.c[1], c[2] - o.c[2]}}; }
    Vec3 operator*(double k) const { return {{c[0] * k, c[1] * k, c[2] * k}}; }
    double norm() const { return std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]); }
};

// Leaving and arriving tangents of key `at` (time-scaled), from the chords
// to and from its neighbours.
struct KeyTangents {
    Vec3 leave, arrive;
};
KeyTangents kb_tangents(const Key& prev, const Key& at, const Key& next, const Vec3& chord_in, const Vec3& chord_out) {
    const double slack = 1.0 - at.tension, cp = 1.0 + at.continuity, cm = 1.0 - at.continuity;
    const double bp = 1.0 + at.bias, bm = 1.0 - at.bias;
    const double span = next.t - prev.t;
    const double to_next = 2.0 * (next.t - at.t) / span, to_prev = 2.0 * (at.t - prev.t) / span;
    KeyTangents k;
    k.leave = (chord_in * (0.5 * slack * cp * bp) + chord_out * (0.5 * slack * cm * bm)) * to_next;
    k.arrive = (chord_in * (0.5 * slack * cm * bp) + chord_out * (0.5 * slack * cp * bm)) * to_prev;
    return k;
}

// Segment index and local parameter in [0, 1] of time `time` over segments
// starting at `starts` (ascending) and the last one ending at `end`.
void find_segment(const std::vector<double>& starts, double end, double time, size_t& seg, double& u) {
    if (!(time > starts.front())) {
        seg = 0;
        u = 0.0;
        return;
    }
    if (!(time < end)) {
        seg = starts.size() - 1;
        u = 1.0;
        return;
    }
    seg = static_cast<size_t>(std::upper_bound(starts.begin(), starts.end(), time) - starts.begin()) - 1;
    const double stop = seg + 1 < starts.size() ? starts[seg + 1] : end;
    u = (time - starts[seg]) / (stop - starts[seg]);
}

// Translation / scale: one cubic p(u) = P + u (S + u (C2 + u C3)) per segment.
struct PosSpline {
    std::vector<double> starts;
    double end = 0.0;
    std::vector<std::array<Vec3, 4>> cubic;  // P, S, C2, C3
    void init(const std::vector<Key>& keys) {
        const size_t n = keys.size();
        auto value = [&](size_t i) { return Vec3{{keys[i].v[0], keys[i].v[1], keys[i].v[2]}}; };
        std::vector<KeyTangents> tan(n);
        for (size_t i = 1; i + 1 < n; i++)
            tan[i] = kb_tangents(keys[i - 1], keys[i], keys[i + 1], value(i) - value(i - 1), value(i + 1) - value(i));
        for (size_t i = 1; i + 2 < n; i++) {
            const Vec3 p0 = value(i), p1 = value(i + 1), step = p1 - p0;
            const Vec3& m0 = tan[i].leave;
            const Vec3& m1 = tan[i + 1].arrive;
            // Hermite to power basis: C2 = 3 (p1 - p0) - 2 m0 - m1, C3 = 2 (p0 - p1) + m0 + m1
            cubic.push_back({p0, m0, step * 3.0 - m0 * 2.0 - m1, step * -2.0 + m0 + m1});
            starts.push_back(keys[i].t);
        }
        end = keys[n - 2].t;
    }
    void eval(double time, double out[3]) const {
        size_t k;
        double u;
        find_segment(starts, end, time, k, u);
        const auto& q = cubic[k];
        for (int a = 0; a < 3; a++) out[a] = q[0].c[a] + u * (q[1].c[a] + u * (q[2].c[a] + u * q[3].c[a]));
    }
};

// Unit quaternion w + (x, y, z).
struct Quaternion {
    double w;
    Vec3 v;
    Quaternion operator*(const Quaternion& o) const {
        const Vec3& a = v;
        const Vec3& b = o.v;
        const Vec3 cross = {{a.c[1] * b.c[2] - a.c[2] * b.c[1], a.c[2] * b.c[0] - a.c[0] * b.c[2],
                             a.c[0] * b.c[1] - a.c[1] * b.c[0]}};
        const double dot = a.c[0] * b.c[0] + a.c[1] * b.c[1] + a.c[2] * b.c[2];
        return {w * o.w - dot, b * w + a * o.w + cross};
    }
    Quaternion conj() const { return {w, v * -1.0}; }
    double dot(const Quaternion& o) const { return w * o.w + v.c[0] * o.v.c[0] + v.c[1] * o.v.c[1] + v.c[2] * o.v.c[2]; }
};

// sin(x) / x, by its series where the quotient loses digits.
double sinc(double x) { return std::fabs(x) < 1e-4 ? 1.0 - x * x / 6.0 : std::sin(x) / x; }

// log of a unit quaternion (cos h, sin h n) = h n, h in [0, pi]; atan2 keeps
// h accurate for small and for nearly opposite rotations alike.
Vec3 qlog(const Quaternion& q) {
    const double s = q.v.norm();
    if (s == 0.0) return Vec3{{0.0, 0.0, 0.0}};
    return q.v * (std::atan2(s, q.w) / s);
}
// exp of a pure quaternion h n = (cos h, sin h n).
Quaternion qexp(const Vec3& x) {
    const double h = x.norm();
    return {std::cos(h), x * sinc(h)};
}
Quaternion from_axis_angle(const Key& k) {
    const double half = 0.5 * k.v[3];
    return {std::cos(half), Vec3{{k.v[0], k.v[1], k.v[2]}} * std::sin(half)};
}
// Great-arc interpolation from a (u = 0) to b (u = 1); nearly equal
// quaternions blend linearly and renormalise.
Quaternion slerp(const Quaternion& a, const Quaternion& b, double u) {
    if (u == 0.0) return a;
    if (u == 1.0) return b;
    const double c = std::max(-1.0, std::min(1.0, a.dot(b)));
    const double ang = std::acos(c);
    double ka, kb;
    if (std::sin(ang) < 1e-9) {
        ka = 1.0 - u;
        kb = u;
    } else {
        ka = std::sin((1.0 - u) * ang) / std::sin(ang);
        kb = std::sin(u * ang) / std::sin(ang);
    }
    Quaternion r{a.w * ka + b.w * kb, a.v * ka + b.v * kb};
    const double len = std::sqrt(r.dot(r));
    return {r.w / len, r.v * (1.0 / len)};
}

struct RotSpline {
    std::vector<double> starts;
    double end = 0.0;
    std::vector<std::array<Quaternion, 4>> quad;  // q_i, inner i, inner i+1, q_{i+1}
    void init(const std::vector<Key>& keys) {
        const size_t n = keys.size();
        std::vector<Quaternion> q(n);
        for (size_t i = 0; i < n; i++) {
            q[i] = from_axis_angle(keys[i]);
            // the same orientation on the near hemisphere: consecutive keys
            // at most 180 degrees apart along the arc
            if (i > 0 && q[i - 1].dot(q[i]) < 0.0) q[i] = {-q[i].w, q[i].v * -1.0};
        }
        std::vector<Vec3> chord(n);  // chord[i] = log(q_{i-1}^-1 q_i)
        for (size_t i = 1; i < n; i++) chord[i] = qlog(q[i - 1].conj() * q[i]);
        std::vector<KeyTangents> tan(n);
        for (size_t i = 1; i + 1 < n; i++) tan[i] = kb_tangents(keys[i - 1], keys[i], keys[i + 1], chord[i], chord[i + 1]);
        for (size_t i = 1; i + 2 < n; i++) {
            const Vec3& h = chord[i + 1];
            quad.push_back({q[i], q[i] * qexp((tan[i].leave - h) * 0.5), q[i + 1] * qexp((h - tan[i + 1].arrive) * 0.5),
                            q[i + 1]});
            starts.push_back(keys[i].t);
        }
        end = keys[n - 2].t;
    }
    // out: axis xyz, angle (radians)
    void eval(double time, double out[4]) const {
        size_t k;
        double u;
        find_segment(starts, end, time, k, u);
        const auto& c = quad[k];
        const Quaternion r = slerp(slerp(c[0], c[3], u), slerp(c[1], c[2], u), 2.0 * u * (1.0 - u));
        const double s = r.v.norm();
        out[3] = 2.0 * std::atan2(s, r.w);
        if (s > 0.0) {
            for (int a = 0; a < 3; a++) out[a] = r.v.c[a] / s;
        } else {
            out[0] = 1.0;
            out[1] = out[2] = 0.0;
        }
    }
};

struct Animation {
    std::string name;
    std::unique_ptr<PosSpline> transl, scale;
    std::unique_ptr<RotSpline> rot;
};

// ------------------------------------------------------------ scene tree --

struct Node {
    enum Kind { kGroup, kStatic, kAnimated, kMesh } kind = kGroup;
    std::vector<std::unique_ptr<Node>> kids;
    float sc[3] = {1, 1, 1}, axis[3] = {0, 0, 0}, deg = 0, tr[3] = {0, 0, 0};  // static transform
    std::string name;                                                        // key-framed transform
    uint32_t first_vertex = 0, n_vertices = 0;                               // mesh: its vertices
};

struct Parser {
    std::string dir;
    int detail = 0;
    std::vector<hulk_float4> local;  // untransformed mesh vertices, w = 1
    std::vector<hulk_tri> tris;
    std::vector<hulk_light> lights;
    std::vector<Animation> anims;
    Node root;
    std::vector<Node*> stack{&root};
    float fov = 45.0f;
    int resx = 800, resy = 600;
    std::string err;
    int depth = 0;

    bool fail(const std::string& m) {
        if (err.empty()) err = m;
        return false;
    }

    static void eat_space(FILE* f) {
        int ch = std::getc(f);
        while (ch == ' ' || ch == '\t' || ch == '\n' || ch == '\f' || ch == '\r') ch = std::getc(f);
        std::ungetc(ch, f);
    }
    static bool next_is(FILE* f, int want) {
        const int ch = std::getc(f);
        if (ch == want) return true;
        std::ungetc(ch, f);
        return false;
    }
    bool floats(FILE* f, float* v, int n) {
        for (int i = 0; i < n; i++)
            if (std::fscanf(f, "%f", &v[i]) != 1) return false;
        return true;
    }

    bool viewpoint(FILE* f) {
        float from[3], at[3], up[3], angle, hither;
        int rx, ry;
        if (std::fscanf(f, " from %f %f %f", &from[0], &from[1], &from[2]) != 3 ||
            std::fscanf(f, " at %f %f %f", &at[0], &at[1], &at[2]) != 3 ||
            std::fscanf(f, " up %f %f %f", &up[0], &up[1], &up[2]) != 3 || std::fscanf(f, " angle %f", &angle) != 1 ||
            std::fscanf(f, " hither %f", &hither) != 1 || std::fscanf(f, " resolution %d %d", &rx, &ry) != 2)
            return fail("viewpoint syntax error");
        fov = angle;
        resx = rx;
        resy = ry;
        return true;
    }
    bool light(FILE* f) {
        if (next_is(f, 'a')) {
            char name[100];
            if (std::fscanf(f, "%99s", name) != 1) return fail("light name syntax error");
        }
        float p[3], c[3] = {1.0f, 1.0f, 1.0f};
        if (std::fscanf(f, "%f %f %f ", &p[0], &p[1], &p[2]) != 3) return fail("light position syntax error");
        const int num = std::fscanf(f, "%f %f %f ", &c[0], &c[1], &c[2]);
        if (num == 0) {
            c[0] = c[1] = c[2] = 1.0f;
        } else if (num != 3) {
            return fail("light colour syntax error");
        }
        hulk_light l;
        l.color = {c[0], c[1], c[2], HULK_LIGHT_POINT};
        l.data = {p[0], p[1], p[2], 1.0f};
        lights.push_back(l);
        return true;
    }
    bool fill(FILE* f) {
        float v[12];
        return floats(f, v, next_is(f, 'm') ? 12 : 8) || fail("fill material syntax error");
    }
    bool poly(FILE* f) {
        const bool patch = next_is(f, 'p');
        int n;
        if (std::fscanf(f, "%d", &n) != 1 || n < 0) return fail("polygon syntax error");
        float v[6];
        for (int q = 0; q < n; q++)
            if (!floats(f, v, patch ? 6 : 3)) return fail("polygon syntax error");
        return true;
    }
    bool include(FILE* f) {
        int level;
        char name[100];
        if (std::fscanf(f, "%d %99s", &level, name) != 2) return fail("include syntax error");
        if (level > detail) return true;  // above the detail level: skipped
        if (++depth > 32) return fail("includes nested deeper than 32");
        const std::string path = dir + name;
        FILE* g = std::fopen(path.c_str(), "r");
        if (!g) return fail("cannot open include file " + path);
        const bool ok = parse(g);
        std::fclose(g);
        depth--;
        return ok;
    }
    bool textured(FILE* f) {
        const int ch = std::getc(f);
        float v[8];
        if (ch == 't') {  // tt / ttp: name, 3 x (xyz [nxyz] uv)
            const bool patch = next_is(f, 'p');
            char name[100];
            if (std::fscanf(f, "%99s", name) != 1) return fail("textured triangle syntax error");
            for (int q = 0; q < 3; q++)
                if (!floats(f, v, patch ? 8 : 5)) return fail("textured triangle syntax error");
            return true;
        }
        if (ch == 'p' && next_is(f, 'a')) {  // tpa: n times, each time + 3 x (xyz nxyz)
            int n;
            if (std::fscanf(f, "%d", &n) != 1 || n < 0) return fail("animated triangle syntax error");
            for (int q = 0; q < n; q++) {
                if (!floats(f, v, 1)) return fail("animated triangle syntax error");
                for (int w = 0; w < 3; w++)
                    if (!floats(f, v, 6)) return fail("animated triangle syntax error");
            }
            return true;
        }
        if (ch == 'p') return true;
        return fail("tt and ttp are the valid t codes");
    }
    bool xform(FILE* f) {
        auto node = std::make_unique<Node>();
        if (next_is(f, 's')) {
            float v[10];
            if (!floats(f, v, 10)) return fail("static transform syntax error");
            node->kind = Node::kStatic;
            for (int a = 0; a < 3; a++) {
                node->sc[a] = v[a];
                node->axis[a] = v[3 + a];
                node->tr[a] = v[7 + a];
            }
            node->deg = v[6];
        } else {
            char name[100];
            if (std::fscanf(f, "%99s", name) != 1) return fail("transform name syntax error");
            node->kind = Node::kAnimated;
            node->name = name;
        }
        eat_space(f);
        if (std::getc(f) != '{') return fail("{ expected after a transform");
        Node* raw = node.get();
        stack.back()->kids.push_back(std::move(node));
        stack.push_back(raw);
        return true;
    }
    bool keyframes(FILE* f) {
        char name[200];
        if (std::fscanf(f, "%199s", name) != 1) return fail("key frame name syntax error");
        eat_space(f);
        if (std::getc(f) != '{') return fail("{ expected after key frame name");
        Animation anim;
        anim.name = name;
        eat_space(f);
        for (int c = std::getc(f); c != '}'; c = std::getc(f)) {
            if (c == EOF) return fail("unterminated key frames");
            std::ungetc(c, f);
            char motion[200];
            int n;
            if (std::fscanf(f, "%199s %d", motion, &n) != 2) return fail("key frame motion syntax error");
            const bool vis = !std::strcmp(motion, "visibility");
            if (n < 4 && !vis) return fail("fewer than 4 key frames");
            if (n < 0) return fail("negative key frame count");
            std::vector<Key> keys;
            for (int i = 0; i < n; i++) {
                float v[8];
                Key k{};
                if (!std::strcmp(motion, "transl") || !std::strcmp(motion, "scale")) {
                    if (!floats(f, v, 7)) return fail("translation / scale key syntax error");
                    k.t = v[0];
                    k.v[0] = v[1];
                    k.v[1] = v[2];
                    k.v[2] = v[3];
                    k.tension = v[4];
                    k.continuity = v[5];
                    k.bias = v[6];
                } else if (!std::strcmp(motion, "rot")) {
                    if (!floats(f, v, 8)) return fail("rotation key syntax error");
                    k.t = v[0];
                    k.v[0] = v[1];
                    k.v[1] = v[2];
                    k.v[2] = v[3];
                    k.v[3] = v[4] * M_PI / 180.0;
                    k.tension = v[5];
                    k.continuity = v[6];
                    k.bias = v[7];
                } else if (vis) {
                    int flag;
                    if (std::fscanf(f, " %f %d", &v[0], &flag) != 2) return fail("visibility key syntax error");
                } else {
                    return fail(std::string("unknown key frame type ") + motion);
                }
                keys.push_back(k);
            }
            if (!std::strcmp(motion, "transl")) {
                anim.transl.reset(new PosSpline);
                anim.transl->init(keys);
            } else if (!std::strcmp(motion, "scale")) {
                anim.scale.reset(new PosSpline);
                anim.scale->init(keys);
            } else if (!std::strcmp(motion, "rot")) {
                anim.rot.reset(new RotSpline);
                anim.rot->init(keys);
            }
            eat_space(f);
        }
        anims.push_back(std::move(anim));
        return true;
    }
    bool mesh(FILE* f) {
        char word[200], tex[200];
        int nv = 0, nn = 0, nt = 0, ntri = 0;
        if (std::fscanf(f, "%199s", word) != 1 || std::strcmp(word, "vertices")) return fail("mesh: 'vertices' expected");
        if (std::fscanf(f, "%d", &nv) != 1 || nv < 0) return fail("mesh: vertex count expected");
        const uint32_t first = static_cast<uint32_t>(local.size());
        for (int q = 0; q < nv; q++) {
            float v[3];
            if (!floats(f, v, 3)) return fail("mesh: vertex syntax error");
            local.push_back({v[0], v[1], v[2], 1.0f});
        }
        if (std::fscanf(f, "%199s", word) != 1) word[0] = '\0';
        if (!std::strcmp(word, "normals")) {
            if (std::fscanf(f, "%d", &nn) != 1 || nn < 0) return fail("mesh: normal count expected");
            float v[3];
            for (int q = 0; q < nn; q++)
                if (!floats(f, v, 3)) return fail("mesh: normal syntax error");
            if (std::fscanf(f, "%199s", word) != 1) word[0] = '\0';
        }
        if (!std::strcmp(word, "texturecoords")) {
            if (std::fscanf(f, "%d", &nt) != 1 || nt < 0) return fail("mesh: texture coordinate count expected");
            if (std::fscanf(f, "%199s", tex) != 1) return fail("mesh: texture name expected");
            float v[2];
            for (int q = 0; q < nt; q++)
                if (!floats(f, v, 2)) return fail("mesh: texture coordinate syntax error");
            if (std::fscanf(f, "%199s", word) != 1) word[0] = '\0';
        }
        if (std::strcmp(word, "triangles")) return fail("mesh: 'triangles' expected");
        if (std::fscanf(f, "%d", &ntri) != 1 || ntri < 0) return fail("mesh: triangle count expected");
        for (int q = 0; q < ntri; q++) {
            int v[3], skip[3];
            if (std::fscanf(f, "%d %d %d", &v[0], &v[1], &v[2]) != 3) return fail("mesh: triangle syntax error");
            if (nn && std::fscanf(f, "%d %d %d", &skip[0], &skip[1], &skip[2]) != 3)
                return fail("mesh: normal index syntax error");
            if (nt && std::fscanf(f, "%d %d %d", &skip[0], &skip[1], &skip[2]) != 3)
                return fail("mesh: texture index syntax error");
            hulk_tri t{{0, 0, 0, 0}};
            for (int c = 0; c < 3; c++) {
                const uint32_t idx = static_cast<unsigned short>(v[c]);  // the reference's index width
                if (idx >= static_cast<uint32_t>(nv)) return fail("mesh: vertex index out of range");
                t.idx[c] = first + idx;
            }
            tris.push_back(t);
        }
        auto node = std::make_unique<Node>();
        node->kind = Node::kMesh;
        node->first_vertex = first;
        node->n_vertices = static_cast<uint32_t>(nv);
        stack.back()->kids.push_back(std::move(node));
        return true;
    }

    bool parse(FILE* f) {
        for (int ch = std::getc(f); ch != EOF; ch = std::getc(f)) {
            bool ok = true;
            float v[8];
            switch (ch) {
                case ' ': case '\t': case '\n': case '\f': case '\r': continue;
                case '#': case '%': {
                    char line[1000];
                    if (!std::fgets(line, 1000, f)) line[0] = '\0';
                    break;
                }
                case 'v': ok = viewpoint(f); break;
                case 'l': ok = light(f); break;
                case 'b': ok = floats(f, v, 3) || fail("background syntax error"); break;
                case 'f': ok = fill(f); break;
                case 'c': ok = floats(f, v, 8) || fail("cone syntax error"); break;
                case 's': ok = floats(f, v, 4) || fail("sphere syntax error"); break;
                case 'p': ok = poly(f); break;
                case 'i': ok = include(f); break;
                case 'd': ok = std::fscanf(f, "%d", &detail) == 1 || fail("detail level syntax error"); break;
                case 't': ok = textured(f); break;
                case 'x': ok = xform(f); break;
                case '}':
                    if (stack.size() <= 1) return fail("} without a transform");
                    stack.pop_back();
                    break;
                case 'a': {
                    int frames;
                    if (next_is(f, 'm'))
                        ok = floats(f, v, 3) || fail("ambient syntax error");
                    else
                        ok = (floats(f, v, 2) && std::fscanf(f, "%d", &frames) == 1) || fail("animation parameters syntax error");
                    break;
                }
                case 'k': ok = keyframes(f); break;
                case 'm': ok = mesh(f); break;
                default: return fail(std::string("unknown AFF entity '") + static_cast<char>(ch) + "'");
            }
            if (!ok) return false;
        }
        return true;
    }

    const Animation* find(const std::string& name) const {
        for (const Animation& a : anims)
            if (a.name == name) return &a;
        return nullptr;
    }

    // AFFTransformVisitor::preAct over the tree at time t.
    void place(const Node& n, const Mat& parent, double t, std::vector<hulk_float4>& out) const {
        Mat m = parent;
        if (n.kind == Node::kStatic) {
            const Mat T = mul(translate(n.tr[0], n.tr[1], n.tr[2]), parent);
            const Mat SR = mul(scale(n.sc[0], n.sc[1], n.sc[2]),
                               rotate(n.axis[0], n.axis[1], n.axis[2], static_cast<float>(n.deg * M_PI / 180.0)));
            m = mul(SR, T);
        } else if (n.kind == Node::kAnimated) {
            if (const Animation* a = find(n.name)) {
                if (a->transl) {
                    double tr[3];
                    a->transl->eval(t, tr);
                    m = mul(translate(float(tr[0]), float(tr[1]), float(tr[2])), m);
                }
                if (a->rot) {
                    double r[4];
                    a->rot->eval(t, r);
                    m = mul(rotate(float(r[0]), float(r[1]), float(r[2]), float(-r[3])), m);
                }
                if (a->scale) {
                    double s[3];
                    a->scale->eval(t, s);
                    m = mul(scale(float(s[0]), float(s[1]), float(s[2])), m);
                }
            }
        } else if (n.kind == Node::kMesh) {
            for (uint32_t i = 0; i < n.n_vertices; i++) out[n.first_vertex + i] = apply(local[n.first_vertex + i], m);
        }
        for (const auto& k : n.kids) place(*k, m, t, out);
    }
};

}  // namespace

extern "C" int hulk_scene_load_aff(const char* path, float time, hulk_scene** out) {
    if (!path || !out) return HULK_E_ILLEGAL_INPUT;
    FILE* f = std::fopen(path, "r");
    if (!f) return HULK_E_FILE_OPEN;
    hulk_scene* s = nullptr;
    try {
        Parser p;
        const char* slash = std::strrchr(path, '/');
        p.dir = slash ? std::string(path, slash + 1) : std::string();
        const bool ok = p.parse(f);
        std::fclose(f);
        if (!ok) return HULK_E_ILLEGAL_INPUT;  // transforms left open at the end stay open, as in the reference
        s = new hulk_scene;
        s->verts.resize(p.local.size());
        p.place(p.root, identity(), time, s->verts);
        s->tris = std::move(p.tris);
        s->lights = std::move(p.lights);
        s->mirror.assign(s->tris.size(), 0);
        // camera: the "camera" transform at time t (zero rotation and
        // translation without one), fov of the viewpoint
        double rot[4] = {0, 0, 0, 0}, tr[3] = {0, 0, 0};
        if (const Animation* a = p.find("camera")) {
            if (a->rot) a->rot->eval(time, rot);
            if (a->transl) a->transl->eval(time, tr);
        }
        s->has_camera = true;
        for (int k = 0; k < 3; k++) s->cam_from[k] = static_cast<float>(tr[k]);
        for (int k = 0; k < 3; k++) s->cam_axis[k] = static_cast<float>(rot[k]);
        s->cam_angle = -static_cast<float>(rot[3]);
        s->cam_fov = p.fov;
        s->cam_aspect = static_cast<float>(p.resx) / static_cast<float>(p.resy);
    } catch (const std::bad_alloc&) {
        delete s;
        return HULK_E_ALLOC;
    }
    *out = s;
    return HULK_OK;
}
