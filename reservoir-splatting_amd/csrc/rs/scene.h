// Scene data in HBM and the hit -> shading-data path.
//
// The scene is a world-space triangle soup split into instances (contiguous vertex and
// triangle ranges): instance transforms are applied to the vertices when they change
// (instance_point / instance_normal), not at every hit as Falcor's per-instance matrices
// are, so a hit reads world-space vertices directly.  Vertex interpolation follows Source/Falcor/Scene/Scene.slang:
// 440-526 (computeVertexData / getVertexData, face normal :348-355), shading-data setup
// follows Source/Falcor/Scene/Material/MaterialFactory.slang:78-106 and the shading
// frame follows Source/Falcor/Utils/Math/ShadingFrame.slang:76-156 (tangents are absent,
// so the frame is invented with perp_stark).
//
// HitInfo keeps the reference's meaning (Source/Falcor/Scene/HitInfo.slang:92-125): a
// triangle id plus the barycentric weights (b1, b2) of vertices 1 and 2.  Triangle ids
// are global indices into the soup; kInvalidHit marks a miss.
#pragma once
#include "sampling.h"
#include "envmap.h"

namespace rs {

static constexpr uint32_t kInvalidHit = 0xffffffffu;
static constexpr uint32_t kNoInstance = 0xffffffffu;  // a vertex outside every instance

struct HitInfo {
  uint32_t prim;  // kInvalidHit = no hit
  float u, v;     // barycentric weights of vertex 1 and vertex 2
};

RS_HD HitInfo hit_invalid() { return HitInfo{kInvalidHit, 0.f, 0.f}; }
RS_HD bool hit_valid(const HitInfo& h) { return h.prim != kInvalidHit; }

// A caller's PackedHitInfo (Source/Falcor/Scene/HitInfo.slang:92-125, header packing
// :455-487, with the 3 type bits of the uncompressed format, HitInfo.cpp:55-60) as a soup
// triangle: instance `inst` covers triangles instTri[inst].x .. + instTri[inst].y - 1.
// primBits = 0 reads the 128-bit header form (instance in word 0, primitive in word 1),
// primBits = n the 32-bit form (instance above the n primitive bits of word 0).  Hit types
// other than triangles, and indices outside the scene, decode as no hit.
RS_HD HitInfo decode_packed_hit(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3, uint32_t primBits,
                                const uvec2* instTri, uint32_t numInstances) {
  const uint32_t type = w0 >> 29u;
  uint32_t inst, prim;
  if (primBits == 0u) {
    inst = w0 & 0x1fffffffu;
    prim = w1;
  } else {
    inst = (w0 & 0x1fffffffu) >> primBits;
    prim = w0 & ((1u << primBits) - 1u);
  }
  if (type != 1u || inst >= numInstances || prim >= instTri[inst].y) return hit_invalid();
  return HitInfo{instTri[inst].x + prim, asfloat(w2), asfloat(w3)};
}

// Lobe bits, Source/Falcor/Rendering/Materials/LobeType.slang.
enum LobeBits : uint32_t {
  kLobeDiffuseReflection = 0x01,
  kLobeSpecularReflection = 0x02,
  kLobeDeltaReflection = 0x04,
  kLobeDiffuseTransmission = 0x10,
  kLobeSpecularTransmission = 0x20,
  kLobeDeltaTransmission = 0x40,
  kLobeReflection = 0x0f,
  kLobeTransmission = 0xf0,
  kLobeSpecular = 0x22,
  kLobeDelta = 0x44,
  kLobeNonDelta = 0x33,
  kLobeNonDeltaReflection = 0x03,
  kLobeNonDeltaTransmission = 0x30,
  kLobeAll = 0xff,
};

enum DiffuseModel : uint32_t { kDiffuseFrostbite = 0, kDiffuseLambert = 1 };

// One StandardMaterial in the MetalRough shading model (no textures).
struct Material {
  vec3 baseColor;
  float roughness;
  vec3 emissive;  // emissive color times emissive factor
  float metallic;
  float ior;
  uint32_t activeLobes;  // LobeBits subset; kLobeAll for the full material
  uint32_t doubleSided;
  uint32_t _pad;
};

// An emissive triangle of the light collection
// (Source/Falcor/Scene/Lights/LightCollectionShared.slang:41-66).
struct EmissiveTriangle {
  vec3 p0;
  float area;
  vec3 p1;
  uint32_t prim;  // triangle id in the soup
  vec3 p2;
  float flux;
  vec3 normal;
  uint32_t materialID;
  vec3 Le;
  float _pad;
};

struct SceneView {
  const float* positions;      // 3 floats per vertex
  const float* normals;        // 3 floats per vertex
  const uint32_t* indices;     // 3 per triangle
  const uint32_t* triMaterial; // material id per triangle
  const int32_t* triToLight;   // emissive-triangle index or -1
  const Material* materials;
  const EmissiveTriangle* lights;
  const uvec2* aliasTable;     // packed 16-bit threshold + 2 x 24-bit indices
  uint32_t numTriangles;
  uint32_t numVertices;
  uint32_t numMaterials;
  uint32_t numLights;
  float invWeightsSum;  // 1 / sum of light fluxes
  uint32_t diffuseModel;
  uint32_t _pad0, _pad1;
  EnvMapView env;       // env.present == 0 without an environment map
};

// Three consecutive floats of element i.  The byte offset is formed in 32 bits (the vertex
// arrays stay below 4 GiB; rs_set_scene checks it), so that the GPU build addresses them from
// the array's scalar base with one lane offset instead of a 64-bit address per float.
RS_HD vec3 load3(const float* p, uint32_t i) {
  const float* q = reinterpret_cast<const float*>(reinterpret_cast<const char*>(p) + i * 12u);
  return vec3{q[0], q[1], q[2]};
}

// The three vertex indices of triangle t, addressed the same way.
RS_HD void load_indices(const uint32_t* indices, uint32_t t, uint32_t& i0, uint32_t& i1, uint32_t& i2) {
  const uint32_t* q = reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(indices) + t * 12u);
  i0 = q[0];
  i1 = q[1];
  i2 = q[2];
}

// Rigid (affine) instance transform, m a row-major 3x4 matrix: the point m * (p, 1).
RS_HD vec3 instance_point(const float* m, vec3 p) {
  return vec3{m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
              m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]};
}

// Its normal transform: the inverse transpose of the 3x3 part, as the cofactor matrix times
// the sign of the determinant (the inverse transpose without the division), renormalized;
// a normal that maps to zero stays zero.
RS_HD vec3 instance_normal(const float* m, vec3 n) {
  const float c00 = m[5] * m[10] - m[6] * m[9], c01 = m[6] * m[8] - m[4] * m[10], c02 = m[4] * m[9] - m[5] * m[8];
  const float c10 = m[2] * m[9] - m[1] * m[10], c11 = m[0] * m[10] - m[2] * m[8], c12 = m[1] * m[8] - m[0] * m[9];
  const float c20 = m[1] * m[6] - m[2] * m[5], c21 = m[2] * m[4] - m[0] * m[6], c22 = m[0] * m[5] - m[1] * m[4];
  const float det = m[0] * c00 + m[1] * c01 + m[2] * c02;
  vec3 r{c00 * n.x + c01 * n.y + c02 * n.z, c10 * n.x + c11 * n.y + c12 * n.z, c20 * n.x + c21 * n.y + c22 * n.z};
  if (det < 0.f) r = -r;
  const float len = length(r);
  return len > 0.f ? r / len : r;
}

struct VertexData {
  vec3 posW;
  vec3 normalW;
  vec3 faceNormalW;
};

RS_HD VertexData get_vertex_data(const SceneView& s, uint32_t prim, float u, float v) {
  uint32_t i0, i1, i2;
  load_indices(s.indices, prim, i0, i1, i2);
  vec3 p0 = load3(s.positions, i0), p1 = load3(s.positions, i1), p2 = load3(s.positions, i2);
  vec3 n0 = load3(s.normals, i0), n1 = load3(s.normals, i1), n2 = load3(s.normals, i2);
  float b0 = 1.f - u - v;
  VertexData vd;
  vd.posW = p0 * b0;
  vd.posW += p1 * u;
  vd.posW += p2 * v;
  vd.normalW = n0 * b0;
  vd.normalW += n1 * u;
  vd.normalW += n2 * v;
  vd.faceNormalW = normalize(cross(p1 - p0, p2 - p0));
  vd.normalW = normalize(vd.normalW);
  return vd;
}

struct ShadingData {
  vec3 posW;
  vec3 V;  // view direction (towards the previous vertex); may be zero (reference quirk)
  vec3 T, B, N;  // shading frame
  vec3 faceN;
  bool frontFacing;
  uint32_t materialID;
  uint32_t lobeMask;     // ReSTIR lobe mask of the material header (0 = all)
  uint32_t activeLobes;  // material's active lobes
  bool doubleSided;
  float IoR;

  RS_HD vec3 toLocal(vec3 v) const { return vec3{dot(v, T), dot(v, B), dot(v, N)}; }
  RS_HD vec3 fromLocal(vec3 v) const { return T * v.x + B * v.y + N * v.z; }
  RS_HD vec3 orientedFaceNormal() const { return frontFacing ? faceN : -faceN; }
  RS_HD vec3 rayOrigin(bool viewside = true) const {
    return compute_ray_origin(posW, (frontFacing == viewside) ? faceN : -faceN);
  }
};

// loadShadingData(hit, rayOrigin, rayDir): RS/LoadShadingData.slang:134-142.  An invalid
// hit loads zeroed vertex data and material 0, as loadVertexData does for a hit of type
// None (the shifts can be handed one after a failed gather shift); the frame built from
// the zero normal is NaN, exactly like normalize(0) in the reference.
RS_HD ShadingData load_shading_data(const SceneView& s, const HitInfo& hit, vec3 rayDir) {
  const bool valid = hit_valid(hit) && hit.prim < s.numTriangles;
  VertexData vd = valid ? get_vertex_data(s, hit.prim, hit.u, hit.v) : VertexData{v3s(0.f), v3s(0.f), v3s(0.f)};
  ShadingData sd;
  sd.posW = vd.posW;
  sd.V = -rayDir;
  sd.N = normalize(vd.normalW);
  sd.T = perp_stark(sd.N);
  sd.B = cross(sd.N, sd.T);
  sd.faceN = vd.faceNormalW;
  sd.frontFacing = dot(sd.V, sd.faceN) >= 0.f;
  uint32_t mid = valid ? s.triMaterial[hit.prim] : 0u;
  sd.materialID = mid;
  const Material& m = s.materials[mid];
  sd.lobeMask = 0;
  sd.activeLobes = m.activeLobes;
  sd.doubleSided = m.doubleSided != 0;
  sd.IoR = 1.f;
  return sd;
}

}  // namespace rs
