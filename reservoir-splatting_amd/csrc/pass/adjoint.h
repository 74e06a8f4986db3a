// Frozen-reservoir replay and its adjoint: gradients of an image loss with respect to the
// materials' base colour and roughness.
//
// The reference has no adjoint; this follows the differentiable-rendering use of ReSTIR
// reservoirs: after a frame, every pixel's final reservoir holds one selected path (its
// reconnection record) and the image value colour = integrand * UCW (RS/ResolveReSTIR.cs.slang
// :45-63).  The replay keeps that selection, its resampling weight, every sampling pdf and
// MIS weight, and the radiance arriving at the reconnection vertex from the rest of the
// path (all stored in the record, RS/ReconnectionData.slang:89-166) fixed, and re-evaluates
// the BSDF factors of the path prefix at new material parameters theta:
//   R_c(theta) = colour_c * prod_k f_k,c(theta) / f_k,c(theta0),
// over the path's prefix vertices k: the primary hit x1 (for paths of length >= 2, also
// when the second vertex is the environment map) and the reconnection vertex x2 (length
// >= 3).  f_k is the BSDF value the integrand used
// there: the full mixture for a next-event estimate at that vertex, or the sampled lobe
// alone (RS/PathTracer.slang:331-382, lobe mask 1 << component type) for a BSDF-sampled
// continuation.  At theta0 the replay reproduces the frame's image exactly.  The adjoint
// kernel sums dL/dR_c * dR_c/dtheta over pixels into per-material gradients.
#pragma once
#include "pathtracer.h"

namespace rs {

// Per-pixel frozen data: the final image value and the prefix BSDF values at theta0.
struct FrozenPixel {
  float color[3];
  uint32_t flags;  // bit 0: x1 factor, bit 1: x2 factor
  float f1[3];
  uint32_t mat1;
  float f2[3];
  uint32_t mat2;
};
static_assert(sizeof(FrozenPixel) == 48, "FrozenPixel layout");

// Material parameters of the replay: base colour RGB and roughness per material.
struct ReplayParams {
  const float* params;  // 4 per material
  uint32_t numMaterials;
};

// The prefix vertices of a selected path, with the shading data and lobe mask the forward
// pass used there.
struct ReplayVertex {
  bool active;
  ShadingData sd;
  vec3 wo;
  bool adjust;
};

RS_HD void replay_vertices(const FrameParams& p, const SceneView& s, const ReconnectionData& rd, ReplayVertex& v1,
                           ReplayVertex& v2) {
  v1.active = false;
  v2.active = false;
  if (rd.pathLength < 2 || !hit_valid(rd.firstHit) || rd.firstHit.prim >= s.numTriangles) return;
  if (rd.pathLength == 2 && rd.lightIsDistant) {
    // the second vertex is the environment: x1 scatters towards the stored direction
    v1.sd = load_shading_data(s, rd.firstHit, -rd.firstWi);
    v1.wo = rd.secondWo;
    v1.sd.lobeMask = rd.lightIsNEE ? 0u : (1u << rd.firstBSDFComponentType);
    v1.adjust = p.adjustPrimary != 0;
    v1.active = true;
    return;
  }
  if (!hit_valid(rd.secondHit) || rd.secondHit.prim >= s.numTriangles) return;
  v1.sd = load_shading_data(s, rd.firstHit, -rd.firstWi);
  vec3 x2 = get_vertex_data(s, rd.secondHit.prim, rd.secondHit.u, rd.secondHit.v).posW;
  vec3 d12 = normalize(x2 - v1.sd.posW);
  v1.wo = d12;
  v1.sd.lobeMask = (rd.pathLength == 2 && rd.lightIsNEE) ? 0u : (1u << rd.firstBSDFComponentType);
  v1.adjust = p.adjustPrimary != 0;
  v1.active = true;
  if (rd.pathLength >= 3) {
    v2.sd = load_shading_data(s, rd.secondHit, d12);
    v2.wo = rd.secondWo;
    v2.sd.lobeMask = (rd.pathLength == 3 && rd.lightIsNEE) ? 0u : (1u << rd.secondBSDFComponentType);
    v2.adjust = p.adjustSecondary != 0;
    v2.active = true;
  }
}

RS_HD BSDFGrad replay_eval(const SceneView& s, const ReplayVertex& v, const float* params) {
  const float* q = params + 4 * v.sd.materialID;
  MaterialInstance mi = make_material_instance_with(s, v.sd, v.adjust, vec3{q[0], q[1], q[2]}, q[3]);
  return mi_eval_grad(v.sd, mi, v.wo, s.materials[v.sd.materialID].metallic);
}

RS_HD float replay_ratio(float f, float f0) { return f0 != 0.f ? f / f0 : 0.f; }

}  // namespace rs
