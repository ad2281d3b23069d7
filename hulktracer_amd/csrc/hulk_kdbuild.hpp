// hulk_kdbuild.hpp - SAH-NlogN KD-tree build on the device.
//
// Builds the tree that KDTreeCPU::buildKDTree (KDTreeCPU.cpp:61-94,
// KDTreeSAH.cpp:247-521) builds, with the same arrays, from the scene already
// resident in HBM.  The reference has no device builder (KDTreeGPU,
// KDTreeGPU.hpp:12-23, is an empty subclass of the CPU tree).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "hulk_abi.h"
#include "hulk_device.hpp"

namespace hulk {

constexpr uint32_t kKdLevels = 21;      // depths 0..20 (c_maxKDTreeDepth = 20)
constexpr uint32_t kKdTopLevels = 12;   // breadth-first part of the device layout

struct KdBuildResult {
    uint32_t n_nodes = 0;
    uint32_t n_lp = 0;
    uint32_t max_jdepth = 0;                       // deepest junction depth + 1
    uint32_t level_prefix[kKdTopLevels + 1] = {0};  // device-layout nodes above depth d
    uint32_t levels = 0;                           // levels holding nodes
    uint32_t capacity = 0;                         // element capacity the build ran with
    uint32_t retries = 0;                          // capacity doublings
};

struct KdBuilder;

KdBuilder* kdb_create();
void kdb_destroy(KdBuilder* b);

// Build from device vertices / triangles.  `capacity` = polygon references
// per level the first attempt allocates (0: 6 per triangle); the build
// doubles it and starts again while a level overflows.  On success the
// builder's arrays hold the device-layout nodes (uint4 per node: breadth-first
// above depth 12, depth-first below, the order hulk_upload_kdtree derives),
// the reference-layout nodes (depth-first pre-order) and the leaf polygon
// references, until the next build.  Returns HULK_OK or a hulk_status; `err`
// names the failure.
int kdb_build(KdBuilder* b, const float4* verts, uint32_t n_verts, const uint4* tris, uint32_t n_tris,
              const SceneBox& box, uint32_t capacity, hipStream_t s, KdBuildResult* out, std::string& err);

const uint4* kdb_dev_nodes(const KdBuilder* b);
const hulk_kd_node* kdb_ref_nodes(const KdBuilder* b);
const uint32_t* kdb_leaf_polys(const KdBuilder* b);

}  // namespace hulk
