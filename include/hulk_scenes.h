/*
 * hulk_scenes.h - seeded procedural scenes and reference scene files for the
 * headless frame driver.
 *
 * The reference fills Scene::m_vertices / m_vertexIndices / m_lights from
 * its BART and OBJ loaders (LoaderInterface, HulkEngine/public/
 * LoaderInterface.hpp:18-27).  These calls have no counterpart in the
 * reference boundary: they synthesize the benchmark scenes (random
 * triangles; heightfield + spheres; animated tori) deterministically from
 * a seed with integer hashing and double arithmetic rounded to float, so
 * the same bytes come out on every x86-64 host.
 */
#ifndef HULK_SCENES_H
#define HULK_SCENES_H

#include "hulk_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum hulk_scene_kind {
    HULK_SCENE_RANDOM_TRIS = 1,     /* n random triangles in the unit cube, edges <= 0.1   */
    HULK_SCENE_TERRAIN_SPHERES = 2, /* 256x256-vertex value-noise heightfield + 200 spheres */
    HULK_SCENE_ANIMATED_TORI = 3    /* 50 tori of 20,000 triangles + ground plane, per frame */
} hulk_scene_kind;

typedef struct hulk_scene hulk_scene;

/* param: triangle count for RANDOM_TRIS (0 = 1000); sphere count for
 * TERRAIN_SPHERES (0 = 200); mesh count for ANIMATED_TORI (0 = 50).
 * frame: animation frame (ANIMATED_TORI) or camera pose (others). */
int hulk_scene_create(int kind, uint64_t seed, uint32_t param, uint32_t frame, hulk_scene** out);
void hulk_scene_free(hulk_scene* s);
uint32_t hulk_scene_vertex_count(const hulk_scene* s);
uint32_t hulk_scene_triangle_count(const hulk_scene* s);
uint32_t hulk_scene_light_count(const hulk_scene* s);
const hulk_float4* hulk_scene_vertices(const hulk_scene* s);
const hulk_tri* hulk_scene_triangles(const hulk_scene* s);
const hulk_light* hulk_scene_lights(const hulk_scene* s);
const uint8_t* hulk_scene_mirror_mask(const hulk_scene* s);  /* one byte per triangle */
int hulk_scene_get_box(const hulk_scene* s, hulk_box* out);
/* Camera of pose / frame `frame` for an image of aspect H / W = aspect. */
int hulk_scene_camera(const hulk_scene* s, uint32_t frame, float aspect, hulk_camera* out);
/* The same camera from the scene kind alone. */
int hulk_scene_kind_camera(int kind, uint32_t frame, float aspect, hulk_camera* out);

/* Scene of an OBJ file as OBJLoader::viParseFile reads it
 * (OBJLoader/OBJLoader.cpp:15-174): triangles "f a b c" over vertices "v",
 * at scene time `time` rotated about y by `time` (OBJTransformVisitor.cpp:
 * 22-41), with the loader's directional light.  HULK_E_FILE_OPEN when the
 * file cannot be read; HULK_E_ILLEGAL_INPUT for a face corner outside the
 * vertex list. */
int hulk_scene_load_obj(const char* path, float time, hulk_scene** out);

/* Scene of a BART AFF file as AFFParser::viParseFile reads it
 * (BartLoader/src/AFFParser.cpp:1469-1539): its triangle meshes at scene
 * time `time` under static and key-framed transforms
 * (AFFTransformVisitor::preAct, AFFTransformVisitor.cpp:24-104), its point
 * lights, and the camera of its "camera" transform and viewpoint fov
 * (hulk_scene_camera with aspect <= 0 uses the file's resolution).  Include
 * files are read relative to the file's directory.  HULK_E_FILE_OPEN when
 * the file cannot be read; HULK_E_ILLEGAL_INPUT for a malformed entity. */
int hulk_scene_load_aff(const char* path, float time, hulk_scene** out);

/* The reference's on-disk KD dataset ("paris.bin"):
 * hulk_dataset_write stores a scene's tree in the layout of
 * GLVisitorGPUTrace::saveDataSetToFile (GLVisitorGPUTrace.cpp:638-674)
 * with the per-leaf vertex / index lists of buildPerLeafData (:163-234);
 * hulk_dataset_load reads one (GLVisitorGPUTraceSSD::buildTree,
 * GLVisitorGPUTraceSSD.cpp:43-107) back into a scene (per-leaf vertices and
 * triangles, triangle k = the k-th triangle of the leaves in node order, no
 * lights) and its tree (free with hulk_scene_free / hulk_kdtree_free). */
int hulk_dataset_write(const char* path, const hulk_float4* verts, uint32_t n_verts, const hulk_tri* tris,
                       uint32_t n_tris, const hulk_kd_node* nodes, uint32_t n_nodes, const uint32_t* leaf_polys,
                       uint32_t n_lp);
int hulk_dataset_load(const char* path, hulk_scene** scene_out, hulk_kdtree** tree_out);

/* Look-at camera: basis in double rounded to float, then the float steps of
 * GLVisitorCPUTrace::updateCamera (GLVisitorCPUTrace.cpp:136-145). */
int hulk_make_camera_lookat(const float from[3], const float target[3], const float up[3],
                            float fov_deg, float aspect, hulk_camera* out);

#ifdef __cplusplus
}
#endif
#endif
