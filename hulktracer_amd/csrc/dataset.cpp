// dataset.cpp - the on-disk KD dataset ("paris.bin") of the reference (host).
//
// Writer: GLVisitorGPUTrace::saveDataSetToFile (GLVisitorGPUTrace.cpp:
// 638-674) over the per-leaf data of buildPerLeafData (:163-234).  Layout,
// all little-endian:
//   header        nine uint64: headerSize (72), KDTreeSize, KDTreeOffset,
//                 leafExtDataSize, leafExtDataOffset, indexPerLeafSize,
//                 indexPerLeafOffset, vertexPerLeafSize, vertexPerLeafOffset
//                 (KdFileHeader, GLVisitorGPUTrace.hpp:52-63);
//   nodes         the KDTreeNode array (16 B each), at offset 72;
//   leaf table    one LeafExtData {indexOffset, vertexOffset, vertxSize,
//                 indexSize} (uint32) per leaf node, in node-array order;
//   indices       at the next multiple of 512 B: per leaf its triangles as
//                 TriangleIndices into the leaf's own vertex list;
//   vertices      right after the indices: per leaf its vertices (16 B).
// A leaf's vertex list holds each vertex once, numbered in order of first
// use by the leaf's triangles (std::map lookup, :186-199).  Both per-leaf
// lists are padded with empty entries to a multiple of 32
// (c_ObjectPerLeafMult, GLVisitorGPUTrace.hpp:32), and the table's sizes
// include the padding.  The reference leaves the fourth index word and the
// padding entries uninitialized; they are written as zeros here.
//
// Reader: GLVisitorGPUTraceSSD::buildTree (GLVisitorGPUTraceSSD.cpp:43-107)
// reads the header, the nodes and the leaf table and streams the per-leaf
// lists on demand.  With every leaf resident in HBM, hulk_dataset_load turns
// the file into a scene and a tree the context traces directly: the
// vertices are the leaves' vertex lists end to end without their padding (a
// leaf's list is numbered densely from 0, so it holds 1 + its largest local
// index), triangle k is the k-th triangle of the leaves in node-array order,
// and each leaf's references are its own triangles.
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <vector>

#include "hulk_abi.h"
#include "hulk_scenes.h"
#include "kdtree_data.hpp"
#include "scene_data.hpp"

namespace {

constexpr uint32_t kObjectPerLeafMult = 32;
constexpr uint64_t kAlign = 512;

struct FileHeader {
    uint64_t header_size, tree_size, tree_offset, ext_size, ext_offset, index_size, index_offset, vertex_size,
        vertex_offset;
};
static_assert(sizeof(FileHeader) == 72, "KdFileHeader layout");

struct LeafExt {
    uint32_t index_offset, vertex_offset, vertex_size, index_size;
};
static_assert(sizeof(LeafExt) == 16, "LeafExtData layout");

uint32_t round_up(uint32_t n) { return (n + (kObjectPerLeafMult - 1)) & ~(kObjectPerLeafMult - 1); }

bool put(FILE* f, const void* p, size_t bytes) { return bytes == 0 || std::fwrite(p, 1, bytes, f) == bytes; }

}  // namespace

extern "C" int hulk_dataset_write(const char* path, const hulk_float4* verts, uint32_t n_verts, const hulk_tri* tris,
                                  uint32_t n_tris, const hulk_kd_node* nodes, uint32_t n_nodes,
                                  const uint32_t* leaf_polys, uint32_t n_lp) {
    if (!path || !nodes || n_nodes == 0 || (n_lp && !leaf_polys) || (n_tris && !tris) || (n_verts && !verts))
        return HULK_E_ILLEGAL_INPUT;
    std::vector<hulk_tri> idx;
    std::vector<hulk_float4> vtx;
    std::vector<LeafExt> ext;
    try {
        for (uint32_t k = 0; k < n_nodes; k++) {
            const hulk_kd_node& n = nodes[k];
            if (n.flags != HULK_NODE_LEAF) continue;
            if ((uint64_t)n.u.leaf.start + n.u.leaf.count > n_lp) return HULK_E_ILLEGAL_INPUT;
            LeafExt e;
            e.index_offset = static_cast<uint32_t>(idx.size());
            e.vertex_offset = static_cast<uint32_t>(vtx.size());
            std::map<uint32_t, uint32_t> local;
            for (uint32_t i = 0; i < n.u.leaf.count; i++) {
                const uint32_t t = leaf_polys[n.u.leaf.start + i];
                if (t >= n_tris) return HULK_E_ILLEGAL_INPUT;
                hulk_tri out{{0, 0, 0, 0}};
                for (int c = 0; c < 3; c++) {
                    const uint32_t v = tris[t].idx[c];
                    if (v >= n_verts) return HULK_E_ILLEGAL_INPUT;
                    auto it = local.find(v);
                    if (it != local.end()) {
                        out.idx[c] = it->second;
                    } else {
                        out.idx[c] = static_cast<uint32_t>(vtx.size()) - e.vertex_offset;
                        local.emplace(v, out.idx[c]);
                        vtx.push_back(verts[v]);
                    }
                }
                idx.push_back(out);
            }
            idx.resize(e.index_offset + round_up(static_cast<uint32_t>(idx.size()) - e.index_offset), hulk_tri{{0, 0, 0, 0}});
            vtx.resize(e.vertex_offset + round_up(static_cast<uint32_t>(vtx.size()) - e.vertex_offset),
                       hulk_float4{0.0f, 0.0f, 0.0f, 0.0f});
            e.index_size = static_cast<uint32_t>(idx.size()) - e.index_offset;
            e.vertex_size = static_cast<uint32_t>(vtx.size()) - e.vertex_offset;
            ext.push_back(e);
        }
    } catch (const std::bad_alloc&) {
        return HULK_E_ALLOC;
    }
    FileHeader h;
    h.header_size = sizeof(FileHeader);
    h.tree_offset = h.header_size;
    h.tree_size = sizeof(hulk_kd_node) * uint64_t(n_nodes);
    h.ext_offset = h.tree_offset + h.tree_size;
    h.ext_size = sizeof(LeafExt) * uint64_t(ext.size());
    h.index_offset = (h.ext_offset + h.ext_size + (kAlign - 1)) & ~(kAlign - 1);
    h.index_size = sizeof(hulk_tri) * uint64_t(idx.size());
    h.vertex_offset = h.index_offset + h.index_size;
    h.vertex_size = sizeof(hulk_float4) * uint64_t(vtx.size());
    FILE* f = std::fopen(path, "wb");
    if (!f) return HULK_E_FILE_OPEN;
    static const char zeros[kAlign] = {0};
    bool ok = put(f, &h, sizeof(h)) && put(f, nodes, h.tree_size) && put(f, ext.data(), h.ext_size) &&
              put(f, zeros, h.index_offset - (h.ext_offset + h.ext_size)) && put(f, idx.data(), h.index_size) &&
              put(f, vtx.data(), h.vertex_size);
    if (std::fclose(f) != 0) ok = false;
    return ok ? HULK_OK : HULK_E_FILE_CLOSE;
}

extern "C" int hulk_dataset_load(const char* path, hulk_scene** scene_out, hulk_kdtree** tree_out) {
    if (!path || !scene_out || !tree_out) return HULK_E_ILLEGAL_INPUT;
    FILE* f = std::fopen(path, "rb");
    if (!f) return HULK_E_FILE_OPEN;
    std::fseek(f, 0, SEEK_END);
    const long file_size = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    FileHeader h;
    auto in_file = [&](uint64_t off, uint64_t size) { return off <= (uint64_t)file_size && size <= (uint64_t)file_size - off; };
    if (file_size < (long)sizeof(h) || std::fread(&h, sizeof(h), 1, f) != 1 || h.header_size != sizeof(h) ||
        h.tree_size % sizeof(hulk_kd_node) || h.ext_size % sizeof(LeafExt) || h.index_size % sizeof(hulk_tri) ||
        h.vertex_size % sizeof(hulk_float4) || h.tree_size == 0 || !in_file(h.tree_offset, h.tree_size) ||
        !in_file(h.ext_offset, h.ext_size) || !in_file(h.index_offset, h.index_size) ||
        !in_file(h.vertex_offset, h.vertex_size) || h.tree_size / 16 >= (1ull << 30) ||
        h.index_size / 16 >= (1ull << 32) || h.vertex_size / 16 >= (1ull << 32)) {
        std::fclose(f);
        return HULK_E_ILLEGAL_INPUT;
    }
    hulk_scene* s = new (std::nothrow) hulk_scene;
    hulk_kdtree* t = new (std::nothrow) hulk_kdtree;
    std::vector<LeafExt> ext;
    std::vector<hulk_tri> idx;
    int rc = HULK_OK;
    try {
        if (!s || !t) throw std::bad_alloc();
        t->nodes.resize(h.tree_size / sizeof(hulk_kd_node));
        ext.resize(h.ext_size / sizeof(LeafExt));
        idx.resize(h.index_size / sizeof(hulk_tri));
        std::vector<hulk_float4> vtx(h.vertex_size / sizeof(hulk_float4));
        auto get = [&](uint64_t off, void* p, uint64_t bytes) {
            return bytes == 0 || (std::fseek(f, (long)off, SEEK_SET) == 0 && std::fread(p, 1, bytes, f) == bytes);
        };
        if (!get(h.tree_offset, t->nodes.data(), h.tree_size) || !get(h.ext_offset, ext.data(), h.ext_size) ||
            !get(h.index_offset, idx.data(), h.index_size) || !get(h.vertex_offset, vtx.data(), h.vertex_size)) {
            rc = HULK_E_ILLEGAL_INPUT;
        } else {
            // leaves in node-array order take the next table entry
            uint32_t leaf = 0;
            for (hulk_kd_node& n : t->nodes) {
                if (n.flags != HULK_NODE_LEAF) continue;
                if (leaf >= ext.size()) { rc = HULK_E_ILLEGAL_INPUT; break; }
                const LeafExt& e = ext[leaf++];
                const uint32_t count = n.u.leaf.count;
                if (count > e.index_size || (uint64_t)e.index_offset + e.index_size > idx.size() ||
                    (uint64_t)e.vertex_offset + e.vertex_size > vtx.size()) {
                    rc = HULK_E_ILLEGAL_INPUT;
                    break;
                }
                const uint32_t base = static_cast<uint32_t>(s->verts.size());
                uint32_t used = 0;
                n.u.leaf.start = static_cast<uint32_t>(t->leaf_polys.size());
                for (uint32_t i = 0; i < count; i++) {
                    const hulk_tri& loc = idx[e.index_offset + i];
                    hulk_tri g{{0, 0, 0, 0}};
                    for (int c = 0; c < 3; c++) {
                        if (loc.idx[c] >= e.vertex_size) { rc = HULK_E_ILLEGAL_INPUT; break; }
                        g.idx[c] = base + loc.idx[c];
                        if (loc.idx[c] + 1 > used) used = loc.idx[c] + 1;
                    }
                    if (rc) break;
                    t->leaf_polys.push_back(static_cast<uint32_t>(s->tris.size()));
                    s->tris.push_back(g);
                }
                if (rc) break;
                s->verts.insert(s->verts.end(), vtx.begin() + e.vertex_offset, vtx.begin() + e.vertex_offset + used);
            }
            if (!rc && leaf != ext.size()) rc = HULK_E_ILLEGAL_INPUT;
            s->mirror.assign(s->tris.size(), 0);
        }
    } catch (const std::bad_alloc&) {
        rc = HULK_E_ALLOC;
    }
    std::fclose(f);
    if (rc) {
        delete s;
        delete t;
        return rc;
    }
    *scene_out = s;
    *tree_out = t;
    return HULK_OK;
}
