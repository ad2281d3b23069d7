// loaders.cpp - reference scene files (host).
//
// OBJ: OBJLoader::viParseFile (OBJLoader/OBJLoader.cpp:15-174).  The file is
// read as whitespace-separated tokens with the reference's rules:
//   "v x y z"  a vertex (w = 1, SPVector's default, SPMatrix.hpp:78);
//   "f a b c"  exactly three corners; each corner is "v", "v/t", "v//n" or
//              "v/t/n" and only the digits before the first '/' count
//              (atoi, 1-based); a fourth corner of a quad is read as an
//              ordinary token and ignored, as in the reference;
//   "g name"   the group name is skipped;
//   "#"        fgets of at most 99 characters (a longer comment line goes on
//              as tokens, as in the reference);
//   "vn", "vt", "mtllib", "usemtl", "o", "s" and anything else: the keyword is
//              dropped and its arguments are read as further tokens, which
//              match nothing.
// Vertex normals (getVerticesNormals, OBJLoader.cpp:178-206) feed no part of
// the traced path (the triangle test computes face normals) and are not kept.
// At scene time t the vertices are rotated about y by t
// (OBJTransformVisitor::preAct, OBJTransformVisitor.cpp:22-41): row vector
// times SPMatrix::rotateMatrix(0, 1, 0, t) (SPMatrix.cpp:54-84) in the lane
// order of VectorMult (SPMatrix.hpp:170-178).  The scene light is the
// loader's directional light (OBJLoader.cpp:162-169).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "hulk_abi.h"
#include "hulk_scenes.h"
#include "scene_data.hpp"

namespace {

struct Mat4 {
    float r[4][4];
};

// SPMatrix::rotateMatrix (SPMatrix.cpp:54-84); cos / sin of the angle in
// double rounded to float.
Mat4 rotate_matrix(float x, float y, float z, float angle) {
    const float c = static_cast<float>(hulk_host_cos(angle)), s = static_cast<float>(hulk_host_sin(angle));
    const float oc = 1 - c;
    const float xy = oc * x * y, xz = oc * x * z, yz = oc * y * z;
    Mat4 m{};
    m.r[0][0] = c + oc * x * x; m.r[0][1] = xy - s * z;     m.r[0][2] = xz + s * y;     m.r[0][3] = 0.0f;
    m.r[1][0] = xy + s * z;     m.r[1][1] = c + oc * y * y; m.r[1][2] = yz - s * x;     m.r[1][3] = 0.0f;
    m.r[2][0] = xz - s * y;     m.r[2][1] = yz + s * x;     m.r[2][2] = c + oc * z * z; m.r[2][3] = 0.0f;
    m.r[3][0] = 0.0f;           m.r[3][1] = 0.0f;           m.r[3][2] = 0.0f;           m.r[3][3] = 1.0f;
    return m;
}

// VectorMult (SPMatrix.hpp:170-178): ((x*L1 + y*L2) + z*L3) + w*L4 per lane.
hulk_float4 vector_mult(const hulk_float4& v, const Mat4& m) {
    float o[4];
    for (int k = 0; k < 4; k++) o[k] = ((v.x * m.r[0][k] + v.y * m.r[1][k]) + v.z * m.r[2][k]) + v.w * m.r[3][k];
    return {o[0], o[1], o[2], o[3]};
}

// Leading digits of a face corner up to the first '/' (atoi), 1-based.
uint32_t corner_index(const char* tok) {
    char num[32];
    size_t k = 0;
    while (tok[k] != '/' && tok[k] != '\0' && k + 1 < sizeof(num)) {
        num[k] = tok[k];
        k++;
    }
    num[k] = '\0';
    return static_cast<uint32_t>(std::atoi(num)) - 1u;  // 0 and negatives fall out of range
}

}  // namespace

extern "C" int hulk_scene_load_obj(const char* path, float time, hulk_scene** out) {
    if (!path || !out) return HULK_E_ILLEGAL_INPUT;
    FILE* f = std::fopen(path, "r");
    if (!f) return HULK_E_FILE_OPEN;
    std::vector<hulk_float4> verts;
    std::vector<uint32_t> idx;
    char token[256];
    char corner[3][128];
    float x = 0.0f, y = 0.0f, z = 0.0f;  // keep their values when a read fails, as the reference's do
    try {
        while (!std::feof(f)) {
            token[0] = '\0';
            if (std::fscanf(f, "%255s", token) != 1) token[0] = '\0';
            if (!std::strcmp(token, "g")) {
                if (std::fscanf(f, "%255s", token) != 1) token[0] = '\0';
            } else if (!std::strcmp(token, "v")) {
                if (std::fscanf(f, "%f %f %f", &x, &y, &z) < 0) break;
                verts.push_back({x, y, z, 1.0f});
            } else if (!std::strcmp(token, "vn")) {
                if (std::fscanf(f, "%f %f %f", &x, &y, &z) < 0) break;
            } else if (!std::strcmp(token, "f")) {
                for (int i = 0; i < 3; i++)
                    if (std::fscanf(f, "%127s", corner[i]) != 1) corner[i][0] = '\0';
                for (int i = 0; i < 3; i++) idx.push_back(corner_index(corner[i]));
            } else if (!std::strcmp(token, "#")) {
                if (!std::fgets(token, 100, f)) token[0] = '\0';
            }
        }
    } catch (const std::bad_alloc&) {
        std::fclose(f);
        return HULK_E_ALLOC;
    }
    if (std::fclose(f) != 0) return HULK_E_FILE_CLOSE;
    for (uint32_t v : idx)
        if (v >= verts.size()) return HULK_E_ILLEGAL_INPUT;  // the reference would index out of bounds
    hulk_scene* s = new (std::nothrow) hulk_scene;
    if (!s) return HULK_E_ALLOC;
    try {
        const Mat4 rot = rotate_matrix(0.0f, 1.0f, 0.0f, time);
        s->verts.reserve(verts.size());
        for (const hulk_float4& v : verts) s->verts.push_back(vector_mult(v, rot));
        s->tris.reserve(idx.size() / 3);
        for (size_t k = 0; k + 2 < idx.size(); k += 3) s->tris.push_back({{idx[k], idx[k + 1], idx[k + 2], 0u}});
        s->mirror.assign(s->tris.size(), 0);
        // SPVector dir(-1, 1, 1) with w = 1, then SPVector::normalize (SPMatrix.hpp:314-323), which sums the
        // squares of all four lanes: 1 / sqrt(4) = 0.5 per lane.
        hulk_light light;
        light.color = {0.9f, 0.9f, 0.9f, HULK_LIGHT_DIRECTIONAL};
        light.data = {-0.5f, 0.5f, 0.5f, 0.5f};
        s->lights.push_back(light);
    } catch (const std::bad_alloc&) {
        delete s;
        return HULK_E_ALLOC;
    }
    *out = s;
    return HULK_OK;
}
