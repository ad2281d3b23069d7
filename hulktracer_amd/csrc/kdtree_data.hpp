// kdtree_data.hpp - the tree object behind hulk_kdtree_* (host).
#pragma once

#include <stdint.h>

#include <vector>

#include "hulk_abi.h"

// m_treeNodes / m_leafPolygons of KDTreeCPU (KDTreeCPU.hpp:128-132).
struct hulk_kdtree {
    std::vector<hulk_kd_node> nodes;
    std::vector<uint32_t> leaf_polys;
};
