// Scene ingestion: the host-side scene builder and the file importers in front of
// rs_set_scene.  It mirrors Falcor's SceneBuilder (Source/Falcor/Scene/SceneBuilder.cpp)
// and its importer plugins (Source/plugins/importers/{PBRTImporter,MitsubaImporter}): an
// importer parses a file and calls the builder, which keeps materials, mesh instances
// (mesh + 4x4 transform), the camera, the environment map and the film size, and flattens
// the instances into the world-space triangle soup the pass takes.
#pragma once
#include <cmath>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/rs.h"
#include "../../../include/rs_import.h"

namespace ingest {

struct f3 {
  float x = 0.f, y = 0.f, z = 0.f;
};
inline f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline f3 cross(f3 a, f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float length(f3 a) { return std::sqrt(dot(a, a)); }
inline f3 normalize(f3 a) {
  float l = length(a);
  return l > 0.f ? a * (1.f / l) : f3{};
}

// Row-major 4x4 acting on column vectors (p' = M p), Falcor's float4x4 convention.
struct m4 {
  float m[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
  static m4 identity() { return m4{}; }
  static m4 from_rows(const float* r) {
    m4 a;
    for (int i = 0; i < 16; i++) a.m[i / 4][i % 4] = r[i];
    return a;
  }
};
m4 mul(const m4& a, const m4& b);
m4 inverse(const m4& a);
m4 transpose(const m4& a);
f3 xform_point(const m4& a, f3 p);
f3 xform_vector(const m4& a, f3 v);
float det3(const m4& a);
m4 translation(f3 t);
m4 scaling(f3 s);
m4 rotation(float radians, f3 axis);  // Rodrigues, right-handed about `axis`
// View matrix of a left-handed look-at (camera z towards the target), cameraFromWorld.
m4 look_at_lh(f3 eye, f3 target, f3 up);

// One material: the pass's StandardMaterial fields with Falcor's defaults
// (Source/Falcor/Scene/Material/BasicMaterialData.slang:51-60): base colour 1, roughness
// and metallic 0, no emission, IoR 1.5, single-sided.
struct Material {
  std::string name;
  rs_material m{};
  Material() {
    m.baseColor[0] = m.baseColor[1] = m.baseColor[2] = 1.f;
    m.roughness = 0.f;
    m.metallic = 0.f;
    m.ior = 1.5f;
    m.activeLobes = 0xffu;
    m.doubleSided = 0u;
  }
};

// An indexed triangle mesh in object space (Falcor's TriangleMesh).  Normals may be
// absent (empty); frontFaceCW marks clockwise front faces.
struct Mesh {
  std::string name;
  std::vector<f3> positions, normals;
  std::vector<uint32_t> indices;
  bool frontFaceCW = false;
};

// Falcor TriangleMesh generators (Source/Falcor/Scene/TriangleMesh.cpp:58-194).
Mesh mesh_quad(float sx, float sy);
Mesh mesh_disk(float radius, uint32_t segments);
Mesh mesh_cube(f3 size);
Mesh mesh_sphere(float radius, uint32_t segmentsU, uint32_t segmentsV);
// Per-face normals on unshared vertices (each triangle keeps its own three vertices).
void flat_normals(Mesh& mesh);
// Area-weighted vertex normals over shared vertices.
void smooth_normals(Mesh& mesh);

// PLY (ascii, binary little/big endian) and Wavefront OBJ readers; polygons are
// triangulated as fans.  Throw std::runtime_error on malformed files.
Mesh read_ply(const std::string& path);
Mesh read_obj(const std::string& path);
Mesh read_mesh_file(const std::string& path);  // by extension

// PFM (Portable Float Map) reader: rows top to bottom, RGB float.
struct Image {
  uint32_t width = 0, height = 0;
  std::vector<float> rgb;  // 3 floats per texel, row 0 at the top
};
Image read_pfm(const std::string& path);
// Lat-long image of `width` x width/2 texels resampled from an equal-area octahedral map
// (the conversion of Source/plugins/importers/PBRTImporter/EnvMapConverter.cs.slang).
Image octahedral_to_latlong(const Image& oct, uint32_t width);

struct Camera {
  bool present = false;
  f3 position, target{0.f, 0.f, -1.f}, up{0.f, 1.f, 0.f};
  float focalLength = 21.f;  // mm; Falcor's default camera
  float frameHeight = 24.f;
  float apertureRadius = 0.f;
  float focalDistance = 10000.f;
};

class Builder {
 public:
  std::vector<Material> materials;
  struct Instance {
    uint32_t mesh;
    m4 transform;
    uint32_t material;
  };
  std::vector<Mesh> meshes;
  std::vector<Instance> instances;
  Camera camera;
  // Environment map: lat-long RGBA32F texels.
  bool hasEnv = false;
  std::vector<float> envTexels;
  uint32_t envW = 0, envH = 0;
  float envIntensity = 1.f, envTint[3] = {1.f, 1.f, 1.f}, envRotation[3] = {0.f, 0.f, 0.f};
  uint32_t filmW = 0, filmH = 0;
  uint32_t diffuseModel = RS_DIFFUSE_FROSTBITE;
  std::string log;

  uint32_t add_material(const Material& m);
  uint32_t add_mesh(Mesh mesh);
  void add_instance(uint32_t mesh, const m4& transform, uint32_t material);
  void set_env_rgb(const Image& img, float intensity, const float tint[3], const float rotationDeg[3]);
  void warn(const std::string& msg);
  // Flatten the instances into the world-space soup; pointers stay valid until the
  // builder changes.
  void flatten();
  std::vector<float> outPositions, outNormals;
  std::vector<uint32_t> outIndices, outTriMaterial;
  std::vector<rs_material> outMaterials;
  std::vector<rs_instance_desc> outInstances;  // one vertex and triangle range per instance
  bool dirty = true;
};

// Importers (Source/plugins/importers).  Throw std::runtime_error with "file:line: msg".
void import_pbrt(Builder& b, const std::string& path);
void import_mitsuba(Builder& b, const std::string& path);

// Falcor's fovYToFocalLength (Source/Falcor/Scene/Camera/Camera.h): the focal length
// that gives vertical field of view `fovY` (radians) on a frame of height `frameHeight`.
inline float fov_y_to_focal_length(float fovY, float frameHeight) {
  return frameHeight / (2.f * std::tan(0.5f * fovY));
}

// Fresnel reflectance of a conductor with complex IoR eta + ik at cosTheta
// (PBRTImporter.cpp:496-515; the standard dielectric-conductor formula).
f3 fresnel_conductor(f3 eta, f3 k, float cosTheta);

std::string dirname_of(const std::string& path);
std::string join_path(const std::string& dir, const std::string& rel);

}  // namespace ingest
