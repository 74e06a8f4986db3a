/* Scene ingestion in front of rs_set_scene: a scene builder and file importers.
 *
 * Replaces, for the modelled pass, Falcor's SceneBuilder
 * (Source/Falcor/Scene/SceneBuilder.cpp:418 addTriangleMesh, addNode / addMeshInstance,
 * addCamera, setEnvMap) and its importer plugins: PBRTImporter
 * (Source/plugins/importers/PBRTImporter/PBRTImporter.cpp:1794-1940 buildScene /
 * importScene) and MitsubaImporter (Source/plugins/importers/MitsubaImporter/
 * MitsubaImporter.cpp:726-835).  Python scene scripts (.pyscene, PythonImporter) are
 * read by the Python host (rs_amd.importers), which drives this builder.
 *
 * The builder holds materials (rs_material), meshes with their 4x4 instance transforms,
 * the camera, the environment map and the film size; rs_builder_get_scene flattens the
 * instances into the world-space triangle soup rs_set_scene takes.  All functions are
 * host-only and do not touch the GPU.  Errors: negative RS_E_* codes with the message in
 * rs_builder_log (which also holds the importers' warnings, one per line). */
#ifndef RS_IMPORT_H
#define RS_IMPORT_H

#include <stdint.h>

#include "rs.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rs_builder rs_builder;

/* A scene or mesh file could not be read or parsed (the message names file and line). */
#define RS_E_IMPORT (-9)

/* Camera of an imported scene (Falcor Camera: position, target, up, focal length on a
 * 24 mm frame, thin lens).  The pass orients its camera with up = +y
 * (RS/CameraManager.slang:105), so `up` is informational. */
typedef struct rs_import_camera {
  float position[3];
  float target[3];
  float up[3];
  float focalLength;    /* mm */
  float frameHeight;    /* mm */
  float apertureRadius; /* scene units */
  float focalDistance;  /* scene units */
  uint32_t present;     /* 0: the scene defines no camera */
} rs_import_camera;

/* Primitive kinds of rs_builder_add_primitive (Falcor TriangleMesh::create*,
 * Source/Falcor/Scene/TriangleMesh.cpp:58-194); params:
 *   QUAD   size.x, size.y                       (XZ plane, normal +y)
 *   DISK   radius, segments                     (XZ plane, normal +y)
 *   CUBE   size.x, size.y, size.z               (centred at the origin)
 *   SPHERE radius, segmentsU, segmentsV */
#define RS_PRIM_QUAD 0u
#define RS_PRIM_DISK 1u
#define RS_PRIM_CUBE 2u
#define RS_PRIM_SPHERE 3u

rs_builder* rs_builder_create(void);
void rs_builder_destroy(rs_builder* b);
/* Warnings and errors so far, newline separated (valid until the next builder call). */
const char* rs_builder_log(const rs_builder* b);

/* Import a scene file into the builder (appends): ".pbrt" (pbrt-v4) or ".xml" (Mitsuba 3),
 * with the subsets and conversions of the reference's importers. */
int32_t rs_builder_import(rs_builder* b, const char* path);

/* Returns the material id (>= 0) or RS_E_*. */
int32_t rs_builder_add_material(rs_builder* b, const rs_material* m);
/* An object-space mesh instanced once with `transform` (row-major 4x4 acting on column
 * vectors; NULL = identity).  normals may be NULL: flat normals are generated.
 * frontFaceCW: the mesh's front faces wind clockwise.  Returns the mesh id or RS_E_*. */
int32_t rs_builder_add_mesh(rs_builder* b, const float* positions, const float* normals, uint32_t numVertices,
                            const uint32_t* indices, uint32_t numTriangles, const float* transform, uint32_t material,
                            uint32_t frontFaceCW);
int32_t rs_builder_add_primitive(rs_builder* b, uint32_t kind, const float* params, const float* transform,
                                 uint32_t material);
/* A PLY or OBJ file (TriangleMesh::createFromFile); smoothNormals selects area-weighted
 * vertex normals where the file has none, else flat normals. */
int32_t rs_builder_add_mesh_file(rs_builder* b, const char* path, const float* transform, uint32_t material,
                                 uint32_t smoothNormals);
/* Another instance of an existing mesh id. */
int32_t rs_builder_add_instance(rs_builder* b, uint32_t mesh, const float* transform, uint32_t material);

int32_t rs_builder_set_camera(rs_builder* b, const rs_import_camera* cam);
/* Lat-long RGBA32F texels (as rs_env_map_desc), copied. */
int32_t rs_builder_set_env_map(rs_builder* b, const rs_env_map_desc* desc);
/* A PFM image: lat-long, or equal-area octahedral (octahedral != 0) converted to
 * lat-long as the pbrt importer does. */
int32_t rs_builder_set_env_map_file(rs_builder* b, const char* path, float intensity, const float tint[3],
                                    const float rotation[3], uint32_t octahedral);
int32_t rs_builder_set_film(rs_builder* b, uint32_t width, uint32_t height);
int32_t rs_builder_set_diffuse_model(rs_builder* b, uint32_t model);

/* The flattened scene; its pointers stay valid until the builder changes. */
int32_t rs_builder_get_scene(rs_builder* b, rs_scene_desc* out);
/* RS_E_INVALID_ARG when the scene has no camera (out->present = 0). */
int32_t rs_builder_get_camera(const rs_builder* b, rs_import_camera* out);
/* RS_E_INVALID_ARG when the scene has no environment map; texels stay owned by the builder. */
int32_t rs_builder_get_env_map(const rs_builder* b, rs_env_map_desc* out);
/* 0 x 0 when the file names no film size. */
int32_t rs_builder_get_film(const rs_builder* b, uint32_t* width, uint32_t* height);

#ifdef __cplusplus
}
#endif

#endif /* RS_IMPORT_H */
