// GRIS shift mappings with the reconnection shift.
//
// Follows Source/RenderPasses/ReservoirSplatting/ShiftMapping.slang: circle-of-confusion
// helpers :43-80, gatherLensVertexCopyShift :90-180, gatherPrimaryHitReconnectionShift
// :183-296, scatterReprojectionShift :301-421 and pathReconnectionShift :430-513.
// The sample generator is passed by value, as in the reference: a shift derives its
// child stream from a copy, so the caller's stream does not advance.
#pragma once
#include "pathtracer.h"

namespace rs {

RS_HD float coc_primary_hit(const FrameParams& p, const CameraData& cam, vec3 x1) {
  float lensRadius = cam.apertureRadius;
  vec3 camDir = normalize(cam.cameraW);
  float camZ = dot(x1 - cam.posW, camDir);
  float filmZ = camZ - length(cam.cameraW);
  float filmRadius = fabs_(lensRadius / camZ * filmZ);
  float normalizedFilm = length(cam.cameraU / (float)p.W + cam.cameraV / (float)p.H);
  return filmRadius / normalizedFilm;
}

RS_HD vec2 dof_gather_probabilities(float r) {
  float gamma = 0.2f + 6.2f / (r - 5.6f);
  gamma = (r <= 0.2f) ? 1.f : fclamp(gamma, 0.2f, 1.f);
  return vec2{gamma, 1.f - gamma};
}

#if defined(__HIPCC__)

// pathReconnectionShift: RS/ShiftMapping.slang:430-513.
__device__ RS_OUTLINE vec3 path_reconnection_shift(const DevCtx& c, const ReconnectionData& rd, PathState& ps,
                                               float& dstJ) {
  const FrameParams& p = *c.p;
  dstJ = 1.f;
  handle_reconnection_hit(c, ps, rd, dstJ);
  if (!ps.active()) return v3s(0.f);
  ShadingData sd2 = load_shading_data(*c.s, rd.secondHit, ps.dir);
  MaterialInstance mi = make_material_instance(*c.s, sd2, p.adjustSecondary != 0);
  sd2.lobeMask = 1u << rd.secondBSDFComponentType;
  vec3 lobeBsdf = mi_eval(sd2, mi, rd.secondWo);
  float lobePdf = mi_eval_pdf(sd2, mi, rd.secondWo);
  sd2.lobeMask = 0xF;
  vec3 totalBsdf = mi_eval(sd2, mi, rd.secondWo);
  float totalPdf = mi_eval_pdf(sd2, mi, rd.secondWo);
  if (rd.pathLength > 2 && !any_pos(totalBsdf)) return v3s(0.f);
  float misWeight = 1.f;
  if (rd.pathLength == 3) {
    bool valid = rd.lightIsNEE ? true : (lobePdf > 0.f);
    if (!valid) return v3s(0.f);
    misWeight = rd.lightIsNEE ? eval_mis(p, 1.f, rd.lightPdf, 1.f, totalPdf) : eval_mis(p, 1.f, totalPdf, 1.f, rd.lightPdf);
    vec3 lightThp = rd.lightIsNEE ? totalBsdf / rd.lightPdf : lobeBsdf / lobePdf;
    ps.thp *= lightThp;
    dstJ *= rd.lightIsNEE ? 1.f : lobePdf;
  } else if (rd.pathLength > 3) {
    if (!(lobePdf > 0.f)) return v3s(0.f);
    ps.thp *= lobeBsdf / lobePdf;
    dstJ *= lobePdf;
  }
  return misWeight * ps.thp * rd.irradiance;
}

// gatherLensVertexCopyShift: RS/ShiftMapping.slang:90-180.
__device__ RS_OUTLINE ShiftedPathData gather_lens_vertex_copy_shift(const DevCtx& c, Rng sg, const ReconnectionData& rd,
                                                                float time, vec2 fp, vec2 lensSample) {
  const FrameParams& p = *c.p;
  ShiftedPathData sh = shifted_default();
  if (!fpixel_in(p, fp)) return sh;
  const bool hitEnvMap = (rd.pathLength == 1 && rd.lightIsDistant);
  PathState ps;
  generate_path_with_params(c, ps, asuint(next1(sg)), time, fp, lensSample);
  next_hit(c, ps);
  CameraData cam = camera_at(p, time);
  ps.time = time;
  vec2 lensLocal = cam.apertureRadius * lensSample;
  vec3 lensWorld = lensLocal.x * normalize(cam.cameraU) + lensLocal.y * normalize(cam.cameraV);
  ps.origin = cam.posW + lensWorld;
  sh.primaryHit = ps.hit;
  sh.firstRayDir = ps.dir;
  sh.fractionalPixel = fp;
  sh.lensSample = lensSample;
  if (ps.has(kFlagHit)) {
    if (hitEnvMap) return shifted_default();
    ShadingData sd = load_shading_data(*c.s, ps.hit, ps.dir);
    vec3 x1 = sd.posW;
    float dist = length(x1 - ps.origin);
    primary_jacobians(x1, sd.faceN, ps.dir, dist, cam, sh.subPixelJacobian, sh.lensVertexJacobian);
  } else {
    // A miss can only match an environment sample.
    if (!hitEnvMap) return shifted_default();
    float cosSensor = dot(normalize(cam.cameraW), ps.dir);
    sh.subPixelJacobian = 1.f / fabs_(cosSensor * cosSensor * cosSensor);
    float d0 = length(cam.cameraW) / cosSensor;
    sh.lensVertexJacobian = (d0 * d0) / fabs_(cosSensor);
    sh.radiance = reconnection_env(c, ps);
    return sh;
  }
  if (rd.pathLength == 1) {
    sh.radiance = reconnection_primary_light(c, ps);
    return sh;
  }
  sh.radiance = path_reconnection_shift(c, rd, ps, sh.secondaryPathJacobian);
  return sh;
}

// gatherPrimaryHitReconnectionShift: RS/ShiftMapping.slang:183-296.
__device__ RS_OUTLINE ShiftedPathData gather_primary_hit_reconnection_shift(const DevCtx& c, Rng sg,
                                                                        const ReconnectionData& rd, float time,
                                                                        vec2 fp, HitInfo primaryHit) {
  const FrameParams& p = *c.p;
  ShiftedPathData sh = shifted_default();
  const bool hitEnvMap = (rd.pathLength == 1 && rd.lightIsDistant);
  PathState ps;
  path_clear(ps);
  ps.set(kFlagActive);
  ps.thp = v3s(1.f);
  CameraData cam = camera_at(p, time);
  ps.time = time;
  vec2 q{fp.x / (float)p.W, fp.y / (float)p.H};
  vec2 ndc{2.f * q.x + -1.f, -2.f * q.y + 1.f};
  vec3 filmWorld = cam.posW + (ndc.x * cam.cameraU + ndc.y * cam.cameraV + cam.cameraW);
  vec3 x1, n;
  if (hitEnvMap) {
    ps.dir = -rd.firstWi;
    x1 = filmWorld + ps.dir;
    n = v3s(0.f);
  } else {
    ShadingData sd = load_shading_data(*c.s, primaryHit, v3s(0.f));
    x1 = sd.posW;
    ps.dir = normalize(x1 - filmWorld);
    n = sd.faceN;
  }
  vec3 camDir = normalize(cam.cameraW);
  float camZ = dot(x1 - cam.posW, camDir);
  float rayT = camZ / dot(ps.dir, camDir);
  ps.origin = x1 - rayT * ps.dir;
  ps.dir = hitEnvMap ? ps.dir : normalize(x1 - ps.origin);
  vec3 lensOffsetWorld = ps.origin - cam.posW;
  vec2 lensLocalN = vec2{dot(lensOffsetWorld, normalize(cam.cameraU)), dot(lensOffsetWorld, normalize(cam.cameraV))} /
                    cam.apertureRadius;
  if (length(lensLocalN) > 1.f) return shifted_default();
  float dist = length(x1 - ps.origin);
  float rayMin = hitEnvMap ? 0.001f : 0.001f * dist;
  float rayMax = hitEnvMap ? kRayTMax : 0.999f * dist;
  Ray vr{ps.origin, rayMin, ps.dir, rayMax};
  if (!trace_visible(*c.tr, vr)) return shifted_default();
  sh.firstRayDir = ps.dir;
  sh.fractionalPixel = fp;
  sh.lensSample = lensLocalN;
  sh.primaryHit = primaryHit;
  ps.sg = rng_make((uint32_t)fp.x, (uint32_t)fp.y, asuint(next1(sg)));
  ps.flags += 1;
  ps.hit = primaryHit;
  ps.set(kFlagHit);
  if (hitEnvMap) {
    float cosSensor = dot(camDir, ps.dir);
    sh.subPixelJacobian = 1.f / fabs_(cosSensor * cosSensor * cosSensor);
    float d0 = length(cam.cameraW) / dot(camDir, ps.dir);
    sh.lensVertexJacobian = (d0 * d0) / fabs_(dot(camDir, ps.dir));
    sh.radiance = reconnection_env(c, ps);
    return sh;
  } else {
    float dd = length(x1 - ps.origin);
    float distSquared = dd * dd;
    float cosNormal = dot(ps.dir, n);
    float cosSensor = dot(camDir, ps.dir);
    sh.subPixelJacobian = fabs_(cosNormal / distSquared) / fabs_(cosSensor * cosSensor * cosSensor);
    float d0 = (length(cam.cameraW) / fabs_(camZ) * rayT);
    float d1 = rayT - d0;
    sh.lensVertexJacobian = (d0 * d0) / (d1 * d1) * (fabs_(cosNormal)) / (fabs_(cosSensor));
  }
  if (rd.pathLength == 1) {
    sh.radiance = reconnection_primary_light(c, ps);
    return sh;
  }
  sh.radiance = path_reconnection_shift(c, rd, ps, sh.secondaryPathJacobian);
  return sh;
}

// scatterReprojectionShift: RS/ShiftMapping.slang:301-421.
__device__ RS_OUTLINE ShiftedPathData scatter_reprojection_shift(const DevCtx& c, Rng sg, const ReconnectionData& rd,
                                                             float time, HitInfo primaryHit, vec2 lensSample,
                                                             bool skipVisibility) {
  const FrameParams& p = *c.p;
  ShiftedPathData sh = shifted_default();
  const bool hitEnvMap = (rd.pathLength == 1) && rd.lightIsDistant;
  PathState ps;
  path_clear(ps);
  ps.set(kFlagActive);
  ps.thp = v3s(1.f);
  CameraData cam = camera_at(p, time);
  ps.time = time;
  vec2 lensLocal = cam.apertureRadius * lensSample;
  vec3 camU = normalize(cam.cameraU), camV = normalize(cam.cameraV);
  vec3 lensWorld = lensLocal.x * camU + lensLocal.y * camV;
  ps.origin = cam.posW + lensWorld;
  vec3 x1 = v3s(0.f), n = v3s(0.f);
  if (hitEnvMap) {
    ps.dir = -rd.firstWi;
  } else {
    ShadingData sd = load_shading_data(*c.s, primaryHit, v3s(0.f));
    x1 = sd.posW;
    n = sd.faceN;
    ps.dir = normalize(x1 - ps.origin);
  }
  vec3 camDir = normalize(cam.cameraW);
  vec3 camRay{dot(camU, ps.dir), dot(camV, ps.dir), dot(camDir, ps.dir)};
  if (camRay.z <= 0.001f) return shifted_default();
  float lw = length(cam.cameraW);
  vec2 film = lensLocal + lw * (vec2{camRay.x, camRay.y} / camRay.z);
  vec2 ndc = film / vec2{length(cam.cameraU), length(cam.cameraV)};
  vec2 nfp{(0.5f * ndc.x + 0.5f) * (float)p.W, (-0.5f * ndc.y + 0.5f) * (float)p.H};
  if (!fpixel_in(p, nfp)) return shifted_default();
  float dist = length(x1 - ps.origin);
  float rayMax = hitEnvMap ? kRayTMax : 0.999f * dist;
  Ray vr{ps.origin, 0.001f, ps.dir, rayMax};
  if (!skipVisibility && !trace_visible(*c.tr, vr)) return shifted_default();
  sh.lensSample = lensSample;
  sh.firstRayDir = ps.dir;
  sh.fractionalPixel = nfp;
  sh.primaryHit = primaryHit;
  ps.sg = rng_make((uint32_t)nfp.x, (uint32_t)nfp.y, asuint(next1(sg)));
  ps.flags += 1;
  ps.hit = primaryHit;
  ps.set(kFlagHit);
  if (hitEnvMap) {
    float cosSensor = dot(camDir, ps.dir);
    sh.subPixelJacobian = 1.f / fabs_(cosSensor * cosSensor * cosSensor);
    float d0 = lw / dot(camDir, ps.dir);
    sh.lensVertexJacobian = (d0 * d0) / fabs_(cosSensor);
    sh.radiance = reconnection_env(c, ps);
    return sh;
  }
  primary_jacobians(x1, n, ps.dir, dist, cam, sh.subPixelJacobian, sh.lensVertexJacobian);
  if (rd.pathLength == 1) {
    sh.radiance = reconnection_primary_light(c, ps);
    return sh;
  }
  sh.radiance = path_reconnection_shift(c, rd, ps, sh.secondaryPathJacobian);
  return sh;
}

#endif

}  // namespace rs
