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
    // camera of a scene read from a file (GLVisitorCPUTrace::updateCamera
    // inputs): position, rotation axis and angle, fov (degrees), aspect
    bool has_camera = false;
    float cam_from[3] = {0, 0, 0}, cam_axis[3] = {0, 0, 0}, cam_angle = 0, cam_fov = 45.0f, cam_aspect = 1.0f;
};

// sin / cos of the scene code (double, host-independent polynomial).
double hulk_host_sin(double x);
double hulk_host_cos(double x);
