// scene_data.hpp - the scene object behind hulk_scene_* (host).
#pragma once

#include <stdint.h>

#include <vector>

#include "hulk_abi.h"

struct hulk_scene {
    int kind = 0;  // hulk_scene_kind, 0 for a scene read from a file
    uint64_t seed = 0;
    std::vector<hulk_float4> verts;
    std::vector<hulk_tri> tris;
    std::vector<hulk_light> lights;
    std::vector<uint8_t> mirror;
};

// sin / cos of the scene code (double, host-independent polynomial).
double hulk_host_sin(double x);
double hulk_host_cos(double x);
