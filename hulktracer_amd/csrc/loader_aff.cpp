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

struct Key {
    double t, v[4];  // value: xyz (+ angle in radians for rotations)
    double tension, continuity, bias;
};

// Tangent multipliers of a Kochanek-Bartels key with the time adjustment
// of unevenly spaced keys (kbsplpos.cpp / kbsplrot.cpp, KB_*Initialize).
void multipliers(const Key& k0, const Key& k1, const Key& k2, const Key& k3, double& out0, double& out1,
                 double& in0, double& in1) {
    const double omt0 = 1 - k1.tension, omc0 = 1 - k1.continuity, opc0 = 1 + k1.continuity;
    const double omb0 = 1 - k1.bias, opb0 = 1 + k1.bias;
    const double adj0 = 2 * (k2.t - k1.t) / (k2.t - k0.t);
    out0 = 0.5 * adj0 * omt0 * opc0 * opb0;
    out1 = 0.5 * adj0 * omt0 * omc0 * omb0;
    const double omt1 = 1 - k2.tension, omc1 = 1 - k2.continuity, opc1 = 1 + k2.continuity;
    const double omb1 = 1 - k2.bias, opb1 = 1 + k2.bias;
    const double adj1 = 2 * (k2.t - k1.t) / (k3.t - k1.t);
    in0 = 0.5 * adj1 * omt1 * omc1 * opb1;
    in1 = 0.5 * adj1 * omt1 * opc1 * omb1;
}

// Segment i covers [keys[i+1].t, keys[i+2].t]; times outside the first /
// last segment clamp to its start / end.
template <class Seg>
void locate(const std::vector<Seg>& seg, double t, size_t& i, double& u) {
    if (seg[0].tmin < t) {
        if (t < seg.back().tmax) {
            for (i = 0; i < seg.size(); i++)
                if (t < seg[i].tmax) break;
            u = (t - seg[i].tmin) / seg[i].trange;
        } else {
            i = seg.size() - 1;
            u = 1.0;
        }
    } else {
        i = 0;
        u = 0.0;
    }
}

struct PosSpline {
    struct Seg {
        double c0[3], c1[3], c2[3], c3[3];
        double tmin, tmax, trange;
    };
    std::vector<Seg> seg;
    void init(const std::vector<Key>& k) {
        for (size_t i0 = 0; i0 + 3 < k.size(); i0++) {
            const Key &k0 = k[i0], &k1 = k[i0 + 1], &k2 = k[i0 + 2], &k3 = k[i0 + 3];
            double out0, out1, in0, in1;
            multipliers(k0, k1, k2, k3, out0, out1, in0, in1);
            Seg s;
            for (int a = 0; a < 3; a++) {
                const double dp = k2.v[a] - k1.v[a];
                const double tout = out1 * dp + out0 * (k1.v[a] - k0.v[a]);
                const double tin = in1 * (k3.v[a] - k2.v[a]) + in0 * dp;
                s.c0[a] = k1.v[a];
                s.c1[a] = tout;
                s.c2[a] = 3 * dp - 2 * tout - tin;
                s.c3[a] = -2 * dp + tout + tin;
            }
            s.tmin = k1.t;
            s.tmax = k2.t;
            s.trange = s.tmax - s.tmin;
            seg.push_back(s);
        }
    }
    void eval(double t, double out[3]) const {
        size_t i;
        double u;
        locate(seg, t, i, u);
        const Seg& s = seg[i];
        for (int a = 0; a < 3; a++) out[a] = s.c0[a] + u * (s.c1[a] + u * (s.c2[a] + u * s.c3[a]));
    }
};

// Quaternions (w, x, y, z) with the operations of quat.cpp.
struct Quat {
    double w, x, y, z;
};
constexpr double kQuatEps = 1e-03;  // sin(angle) cut-off of Log / Exp / Slerp

Quat from_angle_axis(double angle, double ax, double ay, double az) {
    const double h = 0.5 * angle, sn = std::sin(h);
    return {std::cos(h), sn * ax, sn * ay, sn * az};
}
void to_angle_axis(const Quat& q, double& angle, double& ax, double& ay, double& az) {
    const double l2 = q.x * q.x + q.y * q.y + q.z * q.z;
    if (l2 > 0.0) {
        angle = 2.0 * std::acos(q.w);
        const double inv = 1.0 / std::sqrt(l2);
        ax = q.x * inv;
        ay = q.y * inv;
        az = q.z * inv;
    } else {
        angle = 0;
        ax = 1.0;
        ay = 0.0;
        az = 0.0;
    }
}
Quat qmul(const Quat& a, const Quat& b) {
    return {a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
            a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z, a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x};
}
Quat unit_inverse(const Quat& q) { return {q.w, -q.x, -q.y, -q.z}; }
double qdot(const Quat& a, const Quat& b) { return a.w * b.w + a.x * b.x + a.y * b.y + a.z * b.z; }
Quat qscale(const Quat& q, double c) { return {q.w * c, q.x * c, q.y * c, q.z * c}; }
Quat qadd(const Quat& a, const Quat& b) { return {a.w + b.w, a.x + b.x, a.y + b.y, a.z + b.z}; }
Quat qsub(const Quat& a, const Quat& b) { return {a.w - b.w, a.x - b.x, a.y - b.y, a.z - b.z}; }
Quat qlog(const Quat& q) {
    Quat r{0.0, q.x, q.y, q.z};
    if (std::fabs(q.w) < 1.0) {
        const double angle = std::acos(q.w), sn = std::sin(angle);
        if (std::fabs(sn) >= kQuatEps) {
            const double coeff = angle / sn;
            r.x = coeff * q.x;
            r.y = coeff * q.y;
            r.z = coeff * q.z;
        }
    }
    return r;
}
Quat qexp(const Quat& q) {
    const double angle = std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z), sn = std::sin(angle);
    Quat r{std::cos(angle), q.x, q.y, q.z};
    if (std::fabs(sn) >= kQuatEps) {
        const double coeff = sn / angle;
        r.x = coeff * q.x;
        r.y = coeff * q.y;
        r.z = coeff * q.z;
    }
    return r;
}
Quat slerp(double t, const Quat& p, const Quat& q) {
    double c = qdot(p, q);
    c = c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c);
    const double angle = std::acos(c), sn = std::sin(angle);
    if (sn < kQuatEps) return p;
    const double inv = 1.0 / sn;
    const double c0 = std::sin((1.0 - t) * angle) * inv, c1 = std::sin(t * angle) * inv;
    return {c0 * p.w + c1 * q.w, c0 * p.x + c1 * q.x, c0 * p.y + c1 * q.y, c0 * p.z + c1 * q.z};
}
Quat squad(double t, const Quat& p, const Quat& a, const Quat& b, const Quat& q) {
    return slerp(2 * t * (1 - t), slerp(t, p, q), slerp(t, a, b));
}

struct RotSpline {
    struct Seg {
        Quat p, a, b, q;
        double tmin, tmax, trange;
    };
    std::vector<Seg> seg;
    // keys: v = axis xyz, angle (radians); like the reference, a key whose
    // quaternion is flipped to keep consecutive keys within 90 degrees is
    // rewritten in place and the later segments use the rewritten key
    void init(std::vector<Key> k) {
        for (size_t i0 = 0; i0 + 3 < k.size(); i0++) {
            Key &k0 = k[i0], &k1 = k[i0 + 1], &k2 = k[i0 + 2], &k3 = k[i0 + 3];
            const Quat q0 = from_angle_axis(k0.v[3], k0.v[0], k0.v[1], k0.v[2]);
            const Quat q1 = from_angle_axis(k1.v[3], k1.v[0], k1.v[1], k1.v[2]);
            Quat q2 = from_angle_axis(k2.v[3], k2.v[0], k2.v[1], k2.v[2]);
            const Quat q3 = from_angle_axis(k3.v[3], k3.v[0], k3.v[1], k3.v[2]);
            if (qdot(q1, q2) < 0.0) {
                q2 = qscale(q2, -1.0);
                to_angle_axis(q2, k2.v[3], k2.v[0], k2.v[1], k2.v[2]);
            }
            const Quat logdq = qlog(qmul(unit_inverse(q1), q2));
            double out0, out1, in0, in1;
            multipliers(k0, k1, k2, k3, out0, out1, in0, in1);
            const Quat tout = qadd(qscale(logdq, out1), qscale(qlog(qmul(unit_inverse(q0), q1)), out0));
            const Quat tin = qadd(qscale(logdq, in0), qscale(qlog(qmul(unit_inverse(q2), q3)), in1));
            Seg s;
            s.p = q1;
            s.q = q2;
            s.a = qmul(q1, qexp(qscale(qsub(tout, logdq), 0.5)));
            s.b = qmul(q2, qexp(qscale(qsub(logdq, tin), 0.5)));
            s.tmin = k1.t;
            s.tmax = k2.t;
            s.trange = s.tmax - s.tmin;
            seg.push_back(s);
        }
    }
    void eval(double t, double out[4]) const {
        size_t i;
        double u;
        locate(seg, t, i, u);
        const Seg& s = seg[i];
        to_angle_axis(squad(u, s.p, s.a, s.b, s.q), out[3], out[0], out[1], out[2]);
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
