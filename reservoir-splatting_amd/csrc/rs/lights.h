// Emissive-triangle light sampling proportional to power.
//
// Follows Source/Falcor/Rendering/Lights/EmissivePowerSampler.slang:61-124 (uniform table
// slot, fp16 alias threshold, two 24-bit indices; selection pdf = flux / sum),
// Source/Falcor/Rendering/Lights/EmissiveLightSamplerHelpers.slang:54-129 (uniform-area
// point on the triangle, solid-angle pdf, geometry term).  The table itself is built on
// the host by a deterministic Vose pass (no shuffle) and packed exactly like
// Source/Falcor/Rendering/Lights/EmissivePowerSampler.cpp:157-164.
#pragma once
#include "scene.h"
#include "rng.h"

namespace rs {

struct TriangleLightSample {
  uint32_t triangleIndex;
  vec3 posW;
  vec3 normalW;
  vec3 dir;
  float distance;
  vec3 Le;
  float pdf;
  vec2 uv;  // barycentric weights of vertices 1 and 2
  float geometryTerm;
};

RS_HD bool sample_triangle_light(const SceneView& s, vec3 posW, uint32_t li, vec2 u, TriangleLightSample& ls) {
  ls.triangleIndex = li;
  const EmissiveTriangle& tri = s.lights[li];
  vec3 b = sample_triangle(u);
  ls.posW = tri.p0 * b.x + tri.p1 * b.y + tri.p2 * b.z;
  vec3 toLight = ls.posW - posW;
  float distSqr = fmax_(kFltMin, dot(toLight, toLight));
  ls.distance = fsqrt(distSqr);
  ls.dir = toLight / ls.distance;
  ls.normalW = tri.normal;
  float cosTheta = dot(ls.normalW, -ls.dir);
  if (cosTheta <= 0.f) return false;
  ls.Le = tri.Le;
  float denom = fmax_(kFltMin, cosTheta * tri.area);
  ls.pdf = distSqr / denom;
  float su = fsqrt(u.x);
  ls.uv = vec2{1.f - su, u.y * su};
  ls.geometryTerm = cosTheta / distSqr;
  return true;
}

RS_HD bool sample_emissive(const SceneView& s, vec3 posW, Rng& sg, TriangleLightSample& ls) {
  if (s.numLights == 0) return false;
  float uLight = next1(sg);
  uint32_t n = s.numLights;
  uint32_t idx = (uint32_t)(uLight * (float)n);
  idx = idx < n - 1 ? idx : n - 1;
  uvec2 packed = s.aliasTable[idx];
  float threshold = f16tof32(packed.x >> 16u);
  uint32_t selectAbove = ((packed.x & 0xFFFFu) << 8u) | ((packed.y >> 24u) & 0xFFu);
  uint32_t selectBelow = packed.y & 0xFFFFFFu;
  idx = (next1(sg) >= threshold) ? selectAbove : selectBelow;
  float selPdf = s.lights[idx].flux * s.invWeightsSum;
  vec2 u = next2(sg);
  if (!sample_triangle_light(s, posW, idx, u, ls)) return false;
  ls.pdf *= selPdf;
  return true;
}

// Solid-angle pdf at posW of a light hit at hitPos on light `li` with oriented normal n.
RS_HD float emissive_eval_pdf(const SceneView& s, vec3 posW, uint32_t li, vec3 hitPos, vec3 hitNormal) {
  float selPdf = s.lights[li].flux * s.invWeightsSum;
  vec3 toLight = hitPos - posW;
  float distSqr = dot(toLight, toLight);
  if (distSqr <= kFltMin) return selPdf * 0.f;
  vec3 L = toLight / fsqrt(distSqr);
  float cosTheta = dot(hitNormal, -L);
  if (cosTheta <= 0.f) return selPdf * 0.f;
  float denom = fmax_(kFltMin, cosTheta * s.lights[li].area);
  return selPdf * (distSqr / denom);
}

}  // namespace rs
