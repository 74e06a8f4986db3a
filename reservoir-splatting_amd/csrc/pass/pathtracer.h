// Device path tracer with reconnection-data recording (initial candidates).
//
// Follows Source/RenderPasses/ReservoirSplatting/PathTracer.slang: generatePath :194-221,
// generatePathWithParameters :230-255, nextHit :278-299, generateScatterRay :331-382 and
// :391-471, evalMIS :480-508, emissive light sampling :543-565 and :634-701, Russian
// roulette :743-754, handleHit :783-1168, handleReconnectionHit :1177-1397,
// handleReconnectionPrimaryLight :1399-1410, handleMiss :1433-1524 and
// handlePrimaryMiss :1529-1569; path state flags RS/PathStates.slang:38-219.
// The environment map is optional (SceneView::env); light sampling picks between it and the
// emissive triangles with equal probability when both exist (:604-701).  Materials are
// opaque, so the dielectric interior list is always empty.
#pragma once
#include "types.h"
#include "trace.h"
#include "../rs/lights.h"

namespace rs {

// Runtime parameters of one frame (RS/Params.slang:81-139 plus static params).
struct FrameParams {
  uint32_t W, H;
  uint32_t frameCount, seed;
  uint32_t useFixedSeed, fixedSeed;
  uint32_t spp;
  uint32_t maxSurfaceBounces, maxDiffuseBounces, maxSpecularBounces, maxTransmissionBounces;
  uint32_t useRussianRoulette;
  uint32_t misHeuristic;  // 0 balance, 1 power-2, 2 power-exp
  float misPowerExponent;
  float specularRoughnessThreshold;
  uint32_t adjustPrimary;    // GBUFFER_ADJUST_SHADING_NORMALS
  uint32_t adjustSecondary;  // kAdjustShadingNormals
  uint32_t tileShift;        // pixel tiles of the per-pixel kernels are (1 << tileShift) wide
  uint32_t firstSample;      // the sample index that starts a pixel's reservoir (see candidate_finish)
  CameraData camera;  // lens / film parameters; orientation comes from the keyframes
  CameraManager cm;
  // orient_camera results for two sample times, computed once per frame with the same
  // code (bit-identical); every time of a zero-shutter frame is one of these two.
  float cachedTime[2];
  uint32_t numCached;
  uint32_t quirkMask;  // reference behaviour reproduced (kQuirk* bits)
  CameraData cached[2];
};

// Bits of FrameParams::quirkMask (the values of RS_QUIRK_* in include/rs.h).
static constexpr uint32_t kQuirkRobustGrey = 0x1u;      // Robust collect stores a grey integrand
static constexpr uint32_t kQuirkPathTimeZero = 0x2u;    // primary-hit Jacobians at t = 0
static constexpr uint32_t kQuirkMultiProjection = 0x4u; // MultiReproject's lens-less projection
static constexpr uint32_t kQuirkMisDropsLightHits = 0x8u; // BSDF hits on lights dropped after NEE
static constexpr uint32_t kQuirkSppLastSample = 0x10u;    // spp > 1 keeps only the last sample's candidate

RS_HD uint32_t pass_seed(const FrameParams& p, uint32_t id) {
  return p.useFixedSeed ? p.fixedSeed + id : 8u * (p.frameCount + p.seed) + id;
}
RS_HD bool pixel_in(const FrameParams& p, int x, int y) { return x >= 0 && y >= 0 && x < (int)p.W && y < (int)p.H; }
RS_HD bool fpixel_in(const FrameParams& p, vec2 f) {
  return f.x >= 0.f && f.y >= 0.f && f.x < (float)p.W && f.y < (float)p.H;
}

RS_HD float eval_mis(const FrameParams& p, float n0, float p0, float n1, float p1) {
  if (p.misHeuristic == 0) {
    float q0 = n0 * p0, q1 = n1 * p1;
    return q0 / (q0 + q1);
  } else if (p.misHeuristic == 1) {
    float q0 = (n0 * p0) * (n0 * p0), q1 = (n1 * p1) * (n1 * p1);
    return q0 / (q0 + q1);
  } else if (p.misHeuristic == 2) {
    float q0 = fpow(n0 * p0, p.misPowerExponent), q1 = fpow(n1 * p1, p.misPowerExponent);
    return q0 / (q0 + q1);
  }
  return 0.f;
}

RS_HD CameraData camera_at(const FrameParams& p, float time) {
  for (uint32_t i = 0; i < p.numCached; i++)
    if (time == p.cachedTime[i]) return p.cached[i];
  CameraData c = p.camera;
  orient_camera(p.cm, c, p.frameCount, time);
  return c;
}

// Fill the per-frame orientation cache for times 0 and one frame time.
RS_HD void fill_camera_cache(FrameParams& p) {
  p.numCached = 0;
  const float times[2] = {0.f, p.cm.artificialFrameTime};
  for (int i = 0; i < 2; i++) {
    CameraData c = p.camera;
    orient_camera(p.cm, c, p.frameCount, times[i]);
    p.cached[i] = c;
    p.cachedTime[i] = times[i];
  }
  p.numCached = 2;
}

RS_HD vec3 lens_offset(const CameraData& c, vec2 ls) {
  return (ls.x * normalize(c.cameraU) + ls.y * normalize(c.cameraV)) * c.apertureRadius;
}

// |cos x1| / d^2 / |cos sensor|^3 and the lens-vertex Jacobian, shared by the primary
// hit of a path (RS/PathTracer.slang:821-835) and by the shifts (RS/ShiftMapping.slang).
RS_HD void primary_jacobians(vec3 x1, vec3 n, vec3 dir, float dist, const CameraData& c, float& subPixelJ,
                             float& lensJ) {
  float distSquared = dist * dist;
  float cosNormal = dot(-dir, n);
  vec3 camDir = normalize(c.cameraW);
  float cosSensor = dot(camDir, dir);
  subPixelJ = fabs_(cosNormal / distSquared) / fabs_(cosSensor * cosSensor * cosSensor);
  float camZ = dot(x1 - c.posW, camDir);
  float d0 = (length(c.cameraW) / fabs_(camZ) * dist);
  float d1 = dist - d0;
  lensJ = (d0 * d0) / (d1 * d1) * fabs_(cosNormal) / fabs_(cosSensor);
}

// ---------------------------------------------------------------- path state

enum PathFlagBits : uint32_t {
  kFlagActive = 0x0001,
  kFlagHit = 0x0002,
  kFlagTransmission = 0x0004,
  kFlagSpecular = 0x0008,
  kFlagDelta = 0x0010,
  kFlagLightSampledUpper = 0x0040,
  kFlagLightSampledLower = 0x0080,
};
static constexpr uint32_t kVertexBits = 10;
static constexpr uint32_t kVertexMask = (1u << kVertexBits) - 1u;

struct PathState {
  uint32_t id;
  uint32_t flags;  // flags << 10 | vertex index
  uint32_t bounces;
  float time;
  vec3 origin, dir;
  float pdf, totalPdf;
  vec3 normal;
  HitInfo hit;
  vec3 prevHitPos;
  vec3 thp;
  Rng sg;
  ReconnectionData reconnection;
  CandidateReservoir reservoir;
  ReconnectionData selected;

  RS_HD bool has(uint32_t f) const { return (flags & (f << kVertexBits)) != 0; }
  RS_HD void set(uint32_t f, bool v = true) {
    if (v) flags |= (f << kVertexBits);
    else flags &= ~(f << kVertexBits);
  }
  RS_HD bool active() const { return has(kFlagActive); }
  RS_HD void terminate() { set(kFlagActive, false); }
  RS_HD uint32_t vertex() const { return flags & kVertexMask; }
  RS_HD bool lightSampled() const { return has(kFlagLightSampledUpper | kFlagLightSampledLower); }
  RS_HD bool lightSamplable() const { return !has(kFlagDelta); }
  RS_HD uint32_t getBounces(uint32_t type) const { return (bounces >> (type << 3)) & 0xffu; }
  RS_HD void incBounces(uint32_t type) { bounces += (1u << (type << 3)); }
  RS_HD uint32_t pixelX() const { return id & 0xfffu; }
  RS_HD uint32_t pixelY() const { return (id >> 12) & 0xfffu; }
  RS_HD uint32_t sampleIdx() const { return id >> 24; }
};

RS_HD void path_clear(PathState& ps) {
  ps.id = 0;
  ps.flags = 0;
  ps.bounces = 0;
  ps.time = 0.f;
  ps.origin = v3s(0.f);
  ps.dir = v3s(0.f);
  ps.pdf = 0.f;
  ps.totalPdf = 0.f;
  ps.normal = v3s(0.f);
  ps.hit = hit_invalid();
  ps.prevHitPos = v3s(0.f);
  ps.thp = v3s(0.f);
  ps.sg = Rng{0};
  ps.reconnection = recon_default();
  ps.reservoir = CandidateReservoir{0.f, v3s(0.f)};
  ps.selected = recon_default();
}

RS_HD bool finished_bounces(const FrameParams& p, const PathState& ps) {
  uint32_t d = ps.getBounces(0), s = ps.getBounces(1), t = ps.getBounces(2);
  return (d + s + t > p.maxSurfaceBounces) || d > p.maxDiffuseBounces || s > p.maxSpecularBounces ||
         t > p.maxTransmissionBounces;
}

#if defined(__HIPCC__)

struct DevCtx {
  const FrameParams* p;
  const SceneView* s;
  Tracer* tr;
  // s->env.present, or a compile-time false in kernels instantiated for scenes without an
  // environment map, so that their code carries no environment-map branches
  bool env;
};

__device__ inline void generate_path(const DevCtx& c, PathState& ps, uint32_t seed, uint32_t pathID) {
  const FrameParams& p = *c.p;
  path_clear(ps);
  ps.set(kFlagActive);
  ps.id = pathID;
  ps.thp = v3s(1.f);
  ps.sg = rng_make(ps.pixelX(), ps.pixelY(), seed);
  ps.reconnection.time = next1(ps.sg) * p.camera.shutterSpeed;
  ps.reconnection.subPixel = next2(ps.sg);
  ps.reconnection.lensSample = (p.camera.apertureRadius > 0.f) ? sample_disk(next2(ps.sg)) : vec2{0.f, 0.f};
  if (!(p.quirkMask & kQuirkPathTimeZero)) ps.time = ps.reconnection.time;
  CameraData cam = camera_at(p, ps.reconnection.time);
  vec2 fp{(float)ps.pixelX() + ps.reconnection.subPixel.x, (float)ps.pixelY() + ps.reconnection.subPixel.y};
  Ray r = camera_ray_subpixel(cam, p.W, p.H, fp, ps.reconnection.lensSample);
  ps.origin = r.origin;
  ps.dir = r.dir;
  ps.prevHitPos = cam.posW + lens_offset(cam, ps.reconnection.lensSample);
}

__device__ inline void generate_path_with_params(const DevCtx& c, PathState& ps, uint32_t seed, float time, vec2 fp,
                                                 vec2 lensSample) {
  const FrameParams& p = *c.p;
  path_clear(ps);
  ps.set(kFlagActive);
  ps.thp = v3s(1.f);
  ps.sg = rng_make((uint32_t)fp.x, (uint32_t)fp.y, seed);
  CameraData cam = camera_at(p, time);
  ps.time = time;
  Ray r = camera_ray_subpixel(cam, p.W, p.H, fp, lensSample);
  ps.origin = r.origin;
  ps.dir = r.dir;
  ps.prevHitPos = cam.posW + lens_offset(cam, ps.reconnection.lensSample);
}

__device__ inline void next_hit(const DevCtx& c, PathState& ps) {
  ps.flags += 1;
  Ray r{ps.origin, 0.f, ps.dir, kRayTMax};
  float t;
  HitInfo h = trace_closest(*c.tr, r, t);
  if (hit_valid(h)) {
    ps.hit = h;
    ps.set(kFlagHit);
  } else {
    ps.set(kFlagHit, false);
  }
}

// generateScatterRay(BSDFSample, ...): RS/PathTracer.slang:391-471.
__device__ inline bool scatter_from_sample(const DevCtx& c, const BSDFSample& bs, const ShadingData& sd,
                                           const MaterialInstance& mi, PathState& ps) {
  ps.dir = bs.wo;
  ps.thp *= bs.weight;
  ps.set(kFlagTransmission | kFlagSpecular | kFlagDelta, false);
  if (bs.lobeType & kLobeReflection) {
    float roughness = mi_roughness(mi);
    bool isDiffuse = (bs.lobeType & kLobeDiffuseReflection) || roughness > c.p->specularRoughnessThreshold;
    if (isDiffuse) {
      ps.incBounces(0);
    } else {
      ps.incBounces(1);
      ps.set(kFlagSpecular);
    }
  }
  if (bs.lobeType & kLobeDelta) ps.set(kFlagDelta);
  ps.normal = sd.orientedFaceNormal();
  return any_pos(ps.thp);
}

// generateScatterRay(ShadingData, ...): RS/PathTracer.slang:331-382.
__device__ inline bool generate_scatter_ray(const DevCtx& c, ShadingData& sd, const MaterialInstance& mi,
                                            PathState& ps) {
  BSDFSample r;
  sd.lobeMask = 1;
  bool valid = mi_sample(sd, mi, ps.sg, r);
  ps.pdf = r.pdf;
  uint32_t comp = 0;
  if (r.lobeType & kLobeDiffuseReflection) {
    if ((r.lobeType & kLobeSpecular) || (r.lobeType & kLobeDelta)) {
      comp = 3;
      r.lobeType &= ~(uint32_t)kLobeDiffuseReflection;
    }
  } else {
    if (r.lobeType & kLobeDiffuseTransmission) comp = 1;
    else if ((r.lobeType & kLobeSpecular) || (r.lobeType & kLobeDelta)) comp = 2;
  }
  uint32_t v = ps.vertex();
  if (v == 1) {
    ps.reconnection.firstBSDFComponentType = comp;
  } else if (v == 2) {
    ps.reconnection.secondBSDFComponentType = comp;
    ps.reconnection.secondWo = r.wo;
  }
  sd.lobeMask = 0xF;
  ps.totalPdf = mi_eval_pdf(sd, mi, r.wo);
  if (valid) scatter_from_sample(c, r, sd, mi, ps);
  if (v == 1 || v == 2) {
    ps.reconnection.earlyThroughput *= r.weight;
    ps.thp = v3s(1.f);
  }
  return valid;
}

struct LightSample {
  vec3 Li;
  float pdf;
  vec3 origin;
  float distance;
  vec3 dir;
  uint32_t triangleId;
  vec2 barycentrics;
  float geometryTerm;
  bool distant;  // sampled on the environment map (LightType::EnvMap)
};

// getLightTypeSelectionProbabilities (RS/PathTracer.slang:604-622): equal weights over the
// light types present (environment map, emissive triangles; no analytic lights), normalised.
RS_HD void light_type_probs(const SceneView& s, bool env, float& pEnv, float& pEmissive) {
  pEnv = env ? 1.f : 0.f;
  pEmissive = s.numLights > 0 ? 1.f : 0.f;
  float sum = pEnv + pEmissive + 0.f;
  if (sum == 0.f) return;
  float invSum = 1.f / sum;
  pEnv *= invSum;
  pEmissive *= invSum;
}

RS_HD vec3 vertex_ray_origin(vec3 pos, vec3 faceN, vec3 rayDir) {
  return compute_ray_origin(pos, dot(faceN, rayDir) >= 0.f ? faceN : -faceN);
}

// generateLightSample: RS/PathTracer.slang:663-701, with generateEnvMapSample :516-534 and
// generateEmissiveSample :543-565.
__device__ inline bool generate_light_sample(const DevCtx& c, vec3 pos, vec3 faceN, bool frontFacing, bool upper,
                                             bool lower, Rng& sg, LightSample& ls) {
  ls.Li = v3s(0.f);
  ls.pdf = 0.f;
  ls.distant = false;
  // selectLightType (:630-655): the uniform number is drawn even when no type can be chosen.
  float u = next1(sg);
  float pEnv, pEmissive;
  light_type_probs(*c.s, c.env, pEnv, pEmissive);
  float selectionPdf;
  bool env;
  if (c.env && u < pEnv) {  // pEnv is 0 without a map; c.env lets that fold away
    env = true;
    selectionPdf = pEnv;
  } else {
    u -= pEnv;
    if (!(u < pEmissive)) return false;
    env = false;
    selectionPdf = pEmissive;
  }
  if (env) {
    vec3 dir, Le;
    float pdf;
    if (!env_sample(c.s->env, next2(sg), dir, pdf, Le)) return false;
    ls.Li = pdf > 0.f ? Le / pdf : v3s(0.f);
    ls.pdf = pdf;
    ls.origin = vertex_ray_origin(pos, faceN, dir);
    ls.distance = kRayTMax;
    ls.dir = dir;
    ls.distant = true;
  } else {
    TriangleLightSample tls;
    if (!sample_emissive(*c.s, pos, sg, tls)) return false;
    ls.Li = tls.pdf > 0.f ? tls.Le / tls.pdf : v3s(0.f);
    ls.pdf = tls.pdf;
    vec3 lightPos = compute_ray_origin(tls.posW, tls.normalW);
    ls.origin = vertex_ray_origin(pos, faceN, lightPos - pos);
    vec3 toLight = lightPos - ls.origin;
    ls.distance = length(toLight);
    ls.dir = normalize(toLight);
    ls.barycentrics = tls.uv;
    ls.triangleId = tls.triangleIndex;
    ls.geometryTerm = tls.geometryTerm;
  }
  if (!any_pos(ls.Li)) return false;
  float NdotL = dot(frontFacing ? faceN : -faceN, ls.dir);
  if ((!upper && NdotL >= -kMinCosTheta) || (!lower && NdotL <= kMinCosTheta)) return false;
  ls.pdf *= selectionPdf;
  ls.Li /= selectionPdf;
  return true;
}

// handleHit: RS/PathTracer.slang:783-1168.
__device__ RS_OUTLINE void handle_hit(const DevCtx& c, PathState& ps) {
  const FrameParams& p = *c.p;
  const SceneView& s = *c.s;
  const bool isPrimaryHit = ps.vertex() == 1;
  ShadingData sd = load_shading_data(s, ps.hit, ps.dir);
  float distToPrev = length(sd.posW - ps.prevHitPos);
  float geometryFactor = fabs_(dot(sd.faceN, sd.V)) / (distToPrev * distToPrev);
  float prevScatterPdf = ps.pdf;
  ps.prevHitPos = sd.posW;
  uint32_t v = ps.vertex();
  if (v == 1) {
    ps.reconnection.firstHit = ps.hit;
    ps.reconnection.firstWi = -ps.dir;
    CameraData cam = camera_at(p, ps.time);
    vec3 x0 = cam.posW + lens_offset(cam, ps.reconnection.lensSample);
    float dist = length(sd.posW - x0);
    primary_jacobians(sd.posW, sd.faceN, ps.dir, dist, cam, ps.reconnection.subPixelJacobian,
                      ps.reconnection.lensVertexJacobian);
  } else if (v == 2) {
    ps.reconnection.secondHit = ps.hit;
    ps.reconnection.secondaryPathJacobian = prevScatterPdf * geometryFactor;
  } else if (v == 3) {
    ps.reconnection.secondaryPathJacobian *= prevScatterPdf;
  }

  MaterialInstance mi = make_material_instance(s, sd, isPrimaryHit ? p.adjustPrimary != 0 : p.adjustSecondary != 0);
  const bool useEmissive = s.numLights > 0;
  const bool isLightSamplable = ps.lightSamplable();
  const bool dropsLightHits = (p.quirkMask & kQuirkMisDropsLightHits) != 0;
  bool computeEmissive = isPrimaryHit || (useEmissive && (!dropsLightHits || !ps.lightSampled() || !isLightSamplable));
  if (computeEmissive && any_pos(mi.emission)) {
    float misWeight = 1.f;
    float lightPdf = 0.f;
    if (useEmissive && !isPrimaryHit && ps.lightSampled() && isLightSamplable) {
      float pEnv, pEmissive;
      light_type_probs(s, c.env, pEnv, pEmissive);
      int32_t li = s.triToLight[ps.hit.prim];
      if (li >= 0) lightPdf = pEmissive * emissive_eval_pdf(s, ps.origin, (uint32_t)li, sd.posW, sd.orientedFaceNormal());
      misWeight = eval_mis(p, 1.f, ps.totalPdf, 1.f, lightPdf);
    }
    bool sel = candidate_add_vertex(ps.reservoir, ps.sg, 1.f,
                                    ps.reconnection.earlyThroughput * ps.thp * mi.emission * misWeight);
    if (sel) {
      ps.selected = ps.reconnection;
      ps.selected.pathLength = v;
      ps.selected.lightIsNEE = false;
      ps.selected.lightIsDistant = false;
      ps.selected.lightPdf = lightPdf;
      if (v == 2) ps.selected.lightPdf *= geometryFactor;
      if (v == 2 || v == 3) ps.selected.irradiance = mi.emission;
      else ps.selected.irradiance = ps.thp * mi.emission * misWeight;
    }
  }

  if (finished_bounces(p, ps)) {
    ps.terminate();
    return;
  }
  ps.origin = sd.rayOrigin();
  const uint32_t lt = bsdf_lobe_types(mi.data);
  const bool hasNonDelta = (lt & kLobeNonDelta) != 0;
  ps.set(kFlagLightSampledUpper | kFlagLightSampledLower, false);
  if (hasNonDelta) {
    bool upper = (lt & kLobeNonDeltaReflection) != 0;
    bool lower = (lt & kLobeNonDeltaTransmission) != 0;
    LightSample ls;
    bool valid = generate_light_sample(c, sd.posW, sd.faceN, sd.frontFacing, upper, lower, ps.sg, ls);
    ps.set(kFlagLightSampledUpper, upper);
    ps.set(kFlagLightSampledLower, lower);
    if (valid) {
      vec3 Le = ls.Li * ls.pdf;
      float scatterPdf = mi_eval_pdf(sd, mi, ls.dir);
      float misWeight = eval_mis(p, 1.f, ls.pdf, 1.f, scatterPdf);
      vec3 weight = mi_eval(sd, mi, ls.dir);
      vec3 Lr = weight * misWeight * ls.Li;
      if (any_pos(Lr)) {
        Ray r{ls.origin, 0.f, ls.dir, ls.distance};
        if (trace_visible(*c.tr, r)) {
          bool sel = candidate_add_vertex(ps.reservoir, ps.sg, 1.f,
                                          ps.reconnection.earlyThroughput * ps.thp * weight * ls.Li * misWeight);
          if (sel) {
            ps.selected = ps.reconnection;
            ps.selected.pathLength = v + 1;
            ps.selected.lightIsNEE = true;
            ps.selected.irradiance = Le;
            if (v == 1) ps.selected.transmissionEvent = dot(sd.faceN, sd.V) * dot(sd.faceN, ls.dir) < 0.f;
            ps.selected.lightIsDistant = ls.distant;
            if (v == 1 && ls.distant) {
              // A distant second vertex is a direction: the reference packs it into the
              // second hit's payload; this record keeps it in secondWo (unused at length 2).
              ps.selected.secondHit = hit_invalid();
              ps.selected.secondWo = ls.dir;
              ps.selected.secondaryPathJacobian = ls.pdf;
              ps.selected.lightPdf = ls.pdf;
            } else if (v == 1) {
              ps.selected.secondHit = HitInfo{s.lights[ls.triangleId].prim, ls.barycentrics.x, ls.barycentrics.y};
              ps.selected.secondaryPathJacobian = ls.pdf * ls.geometryTerm;
              ps.selected.lightPdf = ls.pdf * ls.geometryTerm;
            } else if (v == 2) {
              ps.selected.secondWo = ls.dir;
              ps.selected.lightPdf = ls.pdf;
            } else {
              ps.selected.irradiance = ps.thp * Lr;
            }
          }
        }
      }
    }
  }

  if (p.useRussianRoulette) {
    float u = next1(ps.sg);
    float rr = luminance(ps.thp);
    float prob = fmax_(0.f, 1.f - rr);
    if (u < prob) {
      ps.terminate();
      return;
    }
    ps.thp /= 1.f - prob;
  }
  bool valid = generate_scatter_ray(c, sd, mi, ps);
  bool isLast = finished_bounces(p, ps);
  if (isLast && ps.lightSamplable()) valid = false;
  if (!valid) ps.terminate();
}

// handleMiss: RS/PathTracer.slang:1433-1524.  The environment map's radiance joins the
// candidate reservoir, MIS-weighted against environment-map light sampling.
__device__ inline void handle_miss(const DevCtx& c, PathState& ps) {
  const SceneView& s = *c.s;
  float prevScatterPdf = ps.pdf;
  uint32_t v = ps.vertex();
  if (v == 2) ps.reconnection.secondaryPathJacobian = prevScatterPdf * 1.f;
  else if (v == 3) ps.reconnection.secondaryPathJacobian *= prevScatterPdf;
  const bool isLightSamplable = ps.lightSamplable();
  const bool dropsLightHits = (c.p->quirkMask & kQuirkMisDropsLightHits) != 0;
  const bool computeEnv = c.env && (!dropsLightHits || !ps.lightSampled() || !isLightSamplable);
  if (computeEnv) {
    float misWeight = 1.f;
    float lightPdf = 0.f;
    if (ps.lightSampled() && isLightSamplable) {
      float pEnv, pEmissive;
      light_type_probs(s, c.env, pEnv, pEmissive);
      lightPdf = pEnv * env_eval_pdf(s.env, ps.dir);
      misWeight = eval_mis(*c.p, 1.f, ps.totalPdf, 1.f, lightPdf);
    }
    vec3 Le = env_eval(s.env, ps.dir);
    vec3 emitterRadiance = misWeight * Le;
    bool sel = candidate_add_vertex(ps.reservoir, ps.sg, 1.f, ps.reconnection.earlyThroughput * ps.thp * Le * misWeight);
    if (sel) {
      ps.selected = ps.reconnection;
      ps.selected.pathLength = v;
      ps.selected.lightIsNEE = false;
      ps.selected.lightIsDistant = true;
      ps.selected.lightPdf = lightPdf;
      if (v == 2) {
        ps.selected.secondHit = hit_invalid();  // the direction, as for a distant NEE sample
        ps.selected.secondWo = ps.dir;
      }
      if (v == 2 || v == 3) ps.selected.irradiance = Le;
      else ps.selected.irradiance = ps.thp * emitterRadiance;
    }
  }
  ps.terminate();
}

// handlePrimaryMiss: RS/PathTracer.slang:1529-1569.  The environment seen directly is a
// candidate that no shift can reconnect.
__device__ inline void handle_primary_miss(const DevCtx& c, PathState& ps) {
  const SceneView& s = *c.s;
  if (c.env) {
    vec3 Le = env_eval(s.env, ps.dir);
    bool sel = candidate_add_vertex(ps.reservoir, ps.sg, 1.f, Le);
    if (sel) {
      ps.selected = ps.reconnection;
      ps.selected.lightIsDistant = true;
      CameraData cam = camera_at(*c.p, ps.time);
      vec3 camDir = normalize(cam.cameraW);
      float cosSensor = dot(camDir, ps.dir);
      ps.selected.subPixelJacobian = 1.f / fabs_(cosSensor * cosSensor * cosSensor);
      float focalPlaneZ = length(cam.cameraW);
      float d0 = focalPlaneZ / dot(camDir, ps.dir);
      ps.selected.lensVertexJacobian = (d0 * d0) / fabs_(cosSensor);
      ps.selected.firstWi = -ps.dir;
      ps.selected.pathLength = 1;
      ps.selected.irradiance = Le;
    }
  }
  ps.terminate();
}

// handleReconnectionHit: RS/PathTracer.slang:1177-1397.
__device__ RS_OUTLINE void handle_reconnection_hit(const DevCtx& c, PathState& ps, const ReconnectionData& rd,
                                               float& dstJ) {
  const FrameParams& p = *c.p;
  const SceneView& s = *c.s;
  const bool isPrimaryHit = ps.vertex() == 1;
  ShadingData sd = load_shading_data(s, ps.hit, ps.dir);
  MaterialInstance mi = make_material_instance(s, sd, isPrimaryHit ? p.adjustPrimary != 0 : p.adjustSecondary != 0);
  ps.origin = sd.rayOrigin();
  if (rd.transmissionEvent) ps.origin = sd.rayOrigin(false);
  const bool secondHitIsDistant = (rd.pathLength == 2) && rd.lightIsDistant;
  dstJ = 1.f;
  vec3 nextDir = v3s(0.f);
  float rayLength = kRayTMax;
  if (secondHitIsDistant) {
    nextDir = rd.secondWo;  // the stored direction of the distant vertex
  } else {
    ShadingData sd2 = load_shading_data(s, rd.secondHit, v3s(0.f));
    vec3 off = sd2.posW - sd.posW;
    float d2 = dot(off, off);
    float cosv = fabs_(dot(sd2.faceN, -normalize(off)));
    dstJ = cosv / d2;
    vec3 adj = sd2.posW - ps.origin;
    nextDir = normalize(adj);
    rayLength = length(adj) * 0.999f;
  }
  sd.lobeMask = 1u << rd.firstBSDFComponentType;
  vec3 lobeBsdf = mi_eval(sd, mi, nextDir);
  float lobePdf = mi_eval_pdf(sd, mi, nextDir);
  sd.lobeMask = 0;
  vec3 totalBsdf = mi_eval(sd, mi, nextDir);
  float totalPdf = mi_eval_pdf(sd, mi, nextDir);
  ps.dir = nextDir;
  ps.pdf = lobePdf;
  if (rd.pathLength == 2) {
    float lightPdf = rd.lightPdf / dstJ;
    float misWeight = 1.f;
    if (rd.lightIsNEE) {
      if (!any_pos(totalBsdf) || totalPdf == 0.f) {
        ps.terminate();
        return;
      }
      dstJ *= lightPdf;
      misWeight = eval_mis(p, 1.f, lightPdf, 1.f, totalPdf);
      ps.thp *= totalBsdf / lightPdf;
    } else {
      if (!any_pos(lobeBsdf) || lobePdf == 0.f) {
        ps.terminate();
        return;
      }
      dstJ *= lobePdf;
      misWeight = eval_mis(p, 1.f, totalPdf, 1.f, lightPdf);
      ps.thp *= lobeBsdf / lobePdf;
    }
    ps.thp *= misWeight;
  } else {
    if (!any_pos(lobeBsdf) || lobePdf == 0.f) {
      ps.terminate();
      return;
    }
    dstJ *= lobePdf;
    ps.thp *= lobeBsdf / lobePdf;
  }
  bool valid = any_pos(ps.thp);
  if (!valid) {
    ps.terminate();
  } else {
    Ray r{ps.origin, 0.f, ps.dir, rayLength};
    if (!trace_visible(*c.tr, r)) ps.terminate();
  }
}

// handleReconnectionEnvMap: RS/PathTracer.slang:1412-1416.
__device__ inline vec3 reconnection_env(const DevCtx& c, const PathState& ps) {
  return c.env ? env_eval(c.s->env, ps.dir) : v3s(0.f);
}

__device__ inline vec3 reconnection_primary_light(const DevCtx& c, const PathState& ps) {
  ShadingData sd = load_shading_data(*c.s, ps.hit, ps.dir);
  MaterialInstance mi = make_material_instance(*c.s, sd, c.p->adjustPrimary != 0);
  return mi.emission;
}

#endif

}  // namespace rs
