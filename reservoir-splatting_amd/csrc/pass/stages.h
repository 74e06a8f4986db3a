// Per-pixel bodies of the pass kernels (one lane per pixel).
//
// Each function follows one reference program:
//   candidate_*         RS/InitialCandidates.cs.slang:52-141 (run by k_initial)
//   primary_vbuffer     Source/RenderPasses/GBuffer/VBuffer/VBufferRT.slang:58-150 and
//                       calcMotionVector, Source/Falcor/Utils/Math/MathHelpers.slang:429-440
//   collect_temporal    RS/CollectTemporalSamples.cs.slang:70-257 (Fast, Clamped, Robust)
//   robust_shift_task   RS/RobustReuseOptimization.rt.slang:63-89 (one sample, one ring pixel)
//   reproject           RS/ReprojectTemporalSamples.rt.slang:64-175
//   multi_reproject     RS/MultiReprojectTemporalSamples.rt.slang:64-168
//   resolve             RS/ResolveReSTIR.cs.slang:45-63
// The ScatterOnly, ScatterBackup and MultiScatter resampling programs and spatial reuse
// are in tasks.h.
// Reservoirs and reconnection records live in HBM in the packed layouts of types.h.
#pragma once
#include "shift.h"

namespace rs {

enum GatherMechanism : uint32_t { kGatherFast = 0, kGatherClamped = 1, kGatherRobust = 2 };
enum ScatterBackupMIS : uint32_t { kBackupBalance = 0, kBackupPairwise = 1 };
static constexpr uint32_t kSpatialNeighborSamples = 8192;
static constexpr uint32_t kNoSplat = 0xffffffffu;
// Confidence of every pixel's initial-candidate reservoir (addCandidateReservoir sets 1,
// RS/InitialCandidates.cs.slang:52-76).
static constexpr float kInitialConfidence = 1.f;
static constexpr uint32_t kMaxTimePartitions = 64;

struct PassBuffers {
  PackedReservoir* currR;
  PackedRecon* currC;
  PackedReservoir* prevR;
  PackedRecon* prevC;
  PackedReservoir* backupR;  // intermediate (collect output)
  PackedRecon* backupC;
  PackedReservoir* outR;  // spatial output
  PackedRecon* outC;
  float* mvec;             // 2 floats per pixel
  HitInfo* vbuffer;        // primary hit per pixel
  float* floatingCoords;   // 2 floats per pixel
  // Splat lists, K partitions of P entries each (K = 1 except for MultiScatter); a cell
  // index is partition * P + pixel.
  uint32_t* splatDst;      // per (partition, source pixel): destination cell or kNoSplat
  uint32_t* cellCount;     // per cell
  uint32_t* cellOffset;    // per cell
  uint32_t* splatRank;     // per (partition, source pixel): the entry's rank in its cell, from the
                           // reprojection's atomic count; afterwards the long-cell sort's temporary
  uint32_t* sorted;        // source pixels grouped by cell, ascending within a cell
  PackedShifted* shiftedPaths;    // 8 per pixel: the previous sample shifted to its 3x3 ring (Robust only)
  const int8_t* neighborOffsets;  // 2 x 8192
  float* color;            // RGBA32F output
  // The caller's VBuffer (PackedHitInfo, 4 words per pixel) or null, its header form and the
  // instance triangle ranges that resolve its instance ids (decode_packed_hit).
  const uint32_t* vbufferIn;
  uint32_t vbufferPrimBits;
  uint32_t numInstances;
  const uvec2* instTri;
  uint32_t* rayCountOut;   // per-pixel ray counts of the initial candidates, or null
};

struct TemporalOptions {
  uint32_t gatherMechanism;
  uint32_t backupMIS;
  uint32_t useConfidence;
  uint32_t spatialNeighbors;
  uint32_t useConfidenceSpatial;
  uint32_t spatialIteration;
  float spatialRadius;
  uint32_t numPartitions;  // time partitions K of MultiScatter (1 otherwise)
};

#if defined(__HIPCC__)

__device__ __forceinline__ PathReservoir ldR(const PackedReservoir* b, uint32_t i) {
  const float4* q = reinterpret_cast<const float4*>(b + i);
  float4 a = q[0], c = q[1];
  PathReservoir r;
  r.confidence = a.x;
  r.totalWeight = a.y;
  r.integrand = vec3{a.z, a.w, c.x};
  r.subPixel = vec2{c.y, c.z};
  return r;
}
__device__ __forceinline__ void stR(PackedReservoir* b, uint32_t i, const PathReservoir& r) {
  float4* q = reinterpret_cast<float4*>(b + i);
  q[0] = make_float4(r.confidence, r.totalWeight, r.integrand.x, r.integrand.y);
  q[1] = make_float4(r.integrand.z, r.subPixel.x, r.subPixel.y, 0.f);
}
__device__ __forceinline__ ReconnectionData ldC(const PackedRecon* b, uint32_t i) {
  const uint4* q = reinterpret_cast<const uint4*>(b + i);
  PackedRecon p;
#pragma unroll
  for (int k = 0; k < 7; k++) {
    uint4 v = q[k];
    p.w[4 * k + 0] = v.x;
    p.w[4 * k + 1] = v.y;
    p.w[4 * k + 2] = v.z;
    p.w[4 * k + 3] = v.w;
  }
  return unpack_recon(p);
}
__device__ __forceinline__ void stC(PackedRecon* b, uint32_t i, const ReconnectionData& r) {
  PackedRecon p = pack_recon(r);
  uint4* q = reinterpret_cast<uint4*>(b + i);
#pragma unroll
  for (int k = 0; k < 7; k++) q[k] = make_uint4(p.w[4 * k], p.w[4 * k + 1], p.w[4 * k + 2], p.w[4 * k + 3]);
}

__device__ __forceinline__ ShiftedPathData ldS(const PackedShifted* b, uint32_t i) {
  const uint4* q = reinterpret_cast<const uint4*>(b + i);
  PackedShifted p;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    uint4 v = q[k];
    p.w[4 * k + 0] = v.x;
    p.w[4 * k + 1] = v.y;
    p.w[4 * k + 2] = v.z;
    p.w[4 * k + 3] = v.w;
  }
  return unpack_shifted(p);
}
__device__ __forceinline__ void stS(PackedShifted* b, uint32_t i, const ShiftedPathData& s) {
  PackedShifted p = pack_shifted(s);
  uint4* q = reinterpret_cast<uint4*>(b + i);
#pragma unroll
  for (int k = 0; k < 4; k++) q[k] = make_uint4(p.w[4 * k], p.w[4 * k + 1], p.w[4 * k + 2], p.w[4 * k + 3]);
}

__device__ __forceinline__ float conf_or_one(bool use, float c) { return use ? c : 1.f; }

__device__ __forceinline__ vec2 prev_pixel_from_mvec(const FrameParams& p, const PassBuffers& b, uint32_t x, uint32_t y) {
  uint32_t idx = y * p.W + x;
  vec2 mv{b.mvec[2 * idx], b.mvec[2 * idx + 1]};
  mv = (length(mv) < 1e-06f) ? vec2{0.f, 0.f} : mv;
  return vec2{(float)x, (float)y} + mv * vec2{(float)p.W, (float)p.H};
}

// tracePath (RS/InitialCandidates.cs.slang:81-116) as three steps, so that a lane can run
// paths of different pixels one after another: candidate_start generates the path,
// candidate_step traces one segment and handles its hit or miss (a primary miss ends the
// path after handlePrimaryMiss), and candidate_finish merges the
// path's candidate into the pixel's reservoir (addCandidateReservoir, :52-76).

// pathID of sample s: the pixel in bits 0-23, the sample index above for spp > 1 (:123-140).
__device__ __forceinline__ uint32_t candidate_path_id(const FrameParams& p, uint32_t x, uint32_t y, uint32_t s) {
  return p.spp == 1 ? (x | (y << 12)) : (x | (y << 12) | (s << 24));
}

__device__ __forceinline__ void candidate_start(const DevCtx& c, PathState& ps, uint32_t pathID) {
  generate_path(c, ps, pass_seed(*c.p, 0), pathID);
}

__device__ __forceinline__ void candidate_step(const DevCtx& c, PathState& ps, bool primary) {
  next_hit(c, ps);
  if (ps.has(kFlagHit)) handle_hit(c, ps);
  else if (primary) handle_primary_miss(c, ps);
  else handle_miss(c, ps);
}

__device__ inline void candidate_finish(const DevCtx& c, const PassBuffers& b, PathState& ps) {
  const FrameParams& p = *c.p;
  // The reference starts the pixel's reservoir at sample spp - 1, its last (quirk); corrected,
  // at sample 0, so that every sample's candidate is merged.  The host picks the index
  // (FrameParams::firstSample), which keeps the choice out of the path tracer's registers.
  const bool isFirstSample = ps.sampleIdx() == p.firstSample;
  const uint32_t idx = ps.pixelY() * p.W + ps.pixelX();
  PathReservoir ex = isFirstSample ? reservoir_default() : ldR(b.currR, idx);
  bool sel = reservoir_add_candidate(ex, ps.sg, 1.f / (float)p.spp, ps.reservoir);
  ex.subPixel = sel ? ps.reconnection.subPixel : ex.subPixel;
  ex.confidence = kInitialConfidence;
  stR(b.currR, idx, ex);
  if (sel) stC(b.currC, idx, ps.selected);
  else if (isFirstSample) stC(b.currC, idx, recon_default());
}

// Pixel-center primary hit and screen-space motion vector (previous minus current
// normalized position) against the previous frame's camera keyframe.
__device__ inline void primary_vbuffer(const DevCtx& c, const PassBuffers& b, uint32_t x, uint32_t y) {
  const FrameParams& p = *c.p;
  uint32_t idx = y * p.W + x;
  CameraData cur = camera_at(p, 0.f);
  vec2 q{((float)x + 0.5f) / (float)p.W, ((float)y + 0.5f) / (float)p.H};
  HitInfo h;
  if (b.vbufferIn) {
    const uint4 w = reinterpret_cast<const uint4*>(b.vbufferIn)[idx];
    h = decode_packed_hit(w.x, w.y, w.z, w.w, b.vbufferPrimBits, b.instTri, b.numInstances);
  } else {
    vec2 ndc{2.f * q.x + -1.f, -2.f * q.y + 1.f};
    Ray r;
    r.origin = cur.posW;
    r.dir = normalize(ndc.x * cur.cameraU + ndc.y * cur.cameraV + cur.cameraW);
    r.tMin = 0.f;
    r.tMax = kRayTMax;
    float t;
    h = trace_closest(*c.tr, r, t);
    c.tr->rays_closest--;  // the primary-visibility buffer is not part of the pass's ray count
  }
  b.vbuffer[idx] = h;
  vec2 mv{0.f, 0.f};
  if (hit_valid(h)) {
    VertexData vd = get_vertex_data(*c.s, h.prim, h.u, h.v);
    CameraData prv = camera_at(p, p.cm.artificialFrameTime);
    vec3 d = vd.posW - prv.posW;
    vec3 camRay{dot(normalize(prv.cameraU), d), dot(normalize(prv.cameraV), d), dot(normalize(prv.cameraW), d)};
    float lw = length(prv.cameraW);
    vec2 film = lw * (vec2{camRay.x, camRay.y} / camRay.z);
    vec2 pndc = film / vec2{length(prv.cameraU), length(prv.cameraV)};
    vec2 prevCrd{0.5f * pndc.x + 0.5f, -0.5f * pndc.y + 0.5f};
    mv = prevCrd - q;
  }
  b.mvec[2 * idx] = mv.x;
  b.mvec[2 * idx + 1] = mv.y;
}

// Shift k of source pixel ni as the collect reads it: the stored one when the source's
// sample is live, otherwise the default that the per-pixel program stores for it.
__device__ __forceinline__ ShiftedPathData robust_slot(const PassBuffers& b, uint32_t ni, uint32_t k, bool live) {
  return live ? ldS(b.shiftedPaths, 8 * ni + k) : unpack_shifted(pack_shifted(shifted_default()));
}

// CollectTemporalSamples::execute (Fast and Clamped).
__device__ inline void collect_temporal(const DevCtx& c, const PassBuffers& b, const TemporalOptions& o, uint32_t x,
                                        uint32_t y) {
  const FrameParams& p = *c.p;
  Rng sg = rng_make(x, y, pass_seed(p, 2));
  uint32_t idx = y * p.W + x;
  vec2 prevPixel = prev_pixel_from_mvec(p, b, x, y);
  PathReservoir dst = reservoir_default();
  ReconnectionData dstC = recon_default();
  vec2 dstFC{-1.f, -1.f};
  if (!fpixel_in(p, prevPixel)) {
    stR(b.backupR, idx, dst);
    stC(b.backupC, idx, dstC);
    b.floatingCoords[2 * idx] = dstFC.x;
    b.floatingCoords[2 * idx + 1] = dstFC.y;
    return;
  }
  int tlx = (int)ffloor(prevPixel.x), tly = (int)ffloor(prevPixel.y);
  vec2 frac{fsaturate(prevPixel.x - (float)tlx), fsaturate(prevPixel.y - (float)tly)};
  if (o.gatherMechanism == kGatherFast) {
    dstFC = prevPixel;
    float totalConfidence = 0.f;
    for (int xx = 0; xx < 2; xx++) {
      for (int yy = 0; yy < 2; yy++) {
        int nx = tlx + xx, ny = tly + yy;
        if (!pixel_in(p, nx, ny)) continue;
        uint32_t ni = (uint32_t)ny * p.W + (uint32_t)nx;
        PathReservoir nr = ldR(b.prevR, ni);
        vec2 rel = (vec2{(float)xx, (float)yy} + nr.subPixel) - frac;
        float bw = flerp((float)(1 - xx), (float)xx, frac.x) * flerp((float)(1 - yy), (float)yy, frac.y);
        totalConfidence += bw * conf_or_one(o.useConfidence, nr.confidence);
        if (rel.x >= 0.f && rel.y >= 0.f && rel.x < 1.f && rel.y < 1.f) {
          nr.subPixel = rel;
          bool sel = reservoir_add_from(dst, sg, 1.f, nr.integrand, 1.f, nr);
          if (sel) dstC = ldC(b.prevC, ni);
        }
      }
    }
    dst.confidence = totalConfidence;
  } else if (o.gatherMechanism == kGatherClamped) {
    vec2 rp{fround(prevPixel.x), fround(prevPixel.y)};
    uint32_t rx = (uint32_t)rp.x, ry = (uint32_t)rp.y;
    dstFC = vec2{(float)rx, (float)ry};
    if (pixel_in(p, (int)rx, (int)ry)) {
      uint32_t pi = ry * p.W + rx;
      dst = ldR(b.prevR, pi);
      dstC = ldC(b.prevC, pi);
    }
  } else {
    // Robust: every 2x2 neighbour fits the floating pixel, through its precomputed shift
    // to the ring pixel that contains it; generalized balance MIS over the bilinear
    // footprint.  The resampled integrand is the shifted radiance's luminance on all
    // three channels, as in the reference (its scalar argument widens to float3).
    dstFC = prevPixel;
    float totalConfidence = 0.f;
    for (int xx = 0; xx < 2; xx++) {
      for (int yy = 0; yy < 2; yy++) {
        int nx = tlx + xx, ny = tly + yy;
        if (!pixel_in(p, nx, ny)) continue;
        uint32_t ni = (uint32_t)ny * p.W + (uint32_t)nx;
        PathReservoir nr = ldR(b.prevR, ni);
        ReconnectionData nc = ldC(b.prevC, ni);
        const bool live = any_pos(nr.integrand);  // whether its ring shifts were stored
        vec2 rel = (vec2{(float)xx, (float)yy} + nr.subPixel) - frac;
        float bw = flerp((float)(1 - xx), (float)xx, frac.x) * flerp((float)(1 - yy), (float)yy, frac.y);
        totalConfidence += bw * conf_or_one(o.useConfidence, nr.confidence);
        int ox = 0, oy = 0;
        if (rel.x < 0.f) ox += 1;
        if (rel.x >= 1.f) ox -= 1;
        if (rel.y < 0.f) oy += 1;
        if (rel.y >= 1.f) oy -= 1;
        rel = rel + vec2{(float)ox, (float)oy};
        uint32_t k = (uint32_t)(ox + 1) + 3u * (uint32_t)(oy + 1);
        const bool noShift = k == 4;
        k = k > 4 ? k - 1 : k;
        ShiftedPathData sp;
        if (noShift) {
          sp.primaryHit = nc.firstHit;
          sp.fractionalPixel = vec2{(float)nx, (float)ny} + nr.subPixel;
          sp.lensSample = nc.lensSample;
          sp.firstRayDir = -nc.firstWi;
          sp.subPixelJacobian = nc.subPixelJacobian;
          sp.lensVertexJacobian = nc.lensVertexJacobian;
          sp.secondaryPathJacobian = nc.secondaryPathJacobian;
          sp.radiance = nr.integrand;
        } else {
          sp = robust_slot(b, ni, k, live);
        }
        float J = noShift ? 1.f : sp.secondaryPathJacobian / nc.secondaryPathJacobian;
        float sourceWeight = bw * luminance(nr.integrand) * conf_or_one(o.useConfidence, nr.confidence);
        float totalPHat = sourceWeight;
        for (int tx = 0; tx < 2; ++tx) {
          for (int ty = 0; ty < 2; ++ty) {
            int px = tlx + tx, py = tly + ty;
            if (!pixel_in(p, px, py)) continue;
            int dx = tx - xx, dy = ty - yy;
            if (dx == 0 && dy == 0) continue;
            float tbw = flerp((float)(1 - tx), (float)tx, frac.x) * flerp((float)(1 - ty), (float)ty, frac.y);
            uint32_t ti = (uint32_t)(dx + 1) + 3u * (uint32_t)(dy + 1);
            ti = ti > 4 ? ti - 1 : ti;
            ShiftedPathData tp = robust_slot(b, ni, ti, live);
            float tPHat = luminance(tp.radiance);
            float tJ = tp.secondaryPathJacobian / nc.secondaryPathJacobian;
            float tConf = conf_or_one(o.useConfidence, ldR(b.prevR, (uint32_t)py * p.W + (uint32_t)px).confidence);
            totalPHat += tbw * tPHat * tJ * tConf;
          }
        }
        float mis = (totalPHat > 0.f) ? sourceWeight / totalPHat : 0.f;
        nr.subPixel = rel;
        recon_update(nc, sp);
        nc.subPixel = rel;
        float l = luminance(sp.radiance);
        vec3 integrand = (p.quirkMask & kQuirkRobustGrey) ? v3s(l) : sp.radiance;
        if (reservoir_add_from(dst, sg, mis, integrand, J, nr)) dstC = nc;
      }
    }
    dst.confidence = totalConfidence;
  }
  if (fisnan(dst.totalWeight)) {
    dst.integrand = v3s(0.f);
    dst.totalWeight = 0.f;
    dstC.subPixelJacobian = 1.f;
    dstC.lensVertexJacobian = 1.f;
    dstC.secondaryPathJacobian = 1.f;
    dstC.pathLength = 0;
  }
  stR(b.backupR, idx, dst);
  stC(b.backupC, idx, dstC);
  b.floatingCoords[2 * idx] = dstFC.x;
  b.floatingCoords[2 * idx + 1] = dstFC.y;
}

// RobustReuseOptimization::run (RS/RobustReuseOptimization.rt.slang:63-89) shifts each
// pixel's previous sample (lens-vertex copy, into the current frame's time) to the 8
// pixels of its 3x3 ring, k = 0..7 in row order without the centre.  Only a sample whose
// integrand has a positive component is shifted; the others store the default shift.
// Here one task is one (live sample, ring pixel) pair, run by k_robust_tasks over the
// compacted list of live samples; the slots of samples that are not live are never
// written, and the collect reads them as the default shift (robust_slot), so the values it
// sees are the per-pixel program's.  Every shift of a pixel starts from the same sample
// generator state, so the tasks need no order.
__device__ __forceinline__ void robust_shift_task(const DevCtx& c, const PassBuffers& b, uint32_t idx, uint32_t k) {
  const FrameParams& p = *c.p;
  const uint32_t x = idx % p.W, y = idx / p.W;
  const uint32_t m = k < 4u ? k : k + 1u;  // position in the 3x3 ring, skipping the centre
  const vec2 np{(float)((int)x + (int)(m % 3u) - 1), (float)((int)y + (int)(m / 3u) - 1)};
  ShiftedPathData sh = shifted_default();
  if (fpixel_in(p, np)) {
    Rng sg = rng_make(x, y, pass_seed(p, 1));
    PathReservoir pr = ldR(b.prevR, idx);
    ReconnectionData pc = ldC(b.prevC, idx);
    sh = gather_lens_vertex_copy_shift(c, sg, pc, pc.time + p.cm.artificialFrameTime, np + pr.subPixel, pc.lensSample);
  }
  stS(b.shiftedPaths, 8 * idx + k, sh);
}


// ReprojectTemporalSamples::run: forward splat of the previous sample's primary hit.
// Writes the destination cell of source pixel (x, y), or kNoSplat.
// Splat target of a previous sample at `time`, through its lens sample onto the film
// (RS/ReprojectTemporalSamples.rt.slang:85-150): kNoSplat when it faces away, is occluded
// or leaves the frame.
__device__ inline uint32_t reproject_through_lens(const DevCtx& c, const ReconnectionData& pc, float time) {
  const FrameParams& p = *c.p;
  CameraData cur = camera_at(p, time);
  vec2 lensLocal = cur.apertureRadius * pc.lensSample;
  vec3 camU = normalize(cur.cameraU), camV = normalize(cur.cameraV);
  vec3 lensWorld = lensLocal.x * camU + lensLocal.y * camV;
  Ray r;
  r.origin = cur.posW + lensWorld;
  if (hit_valid(pc.firstHit)) {
    ShadingData sd = load_shading_data(*c.s, pc.firstHit, v3s(0.f));
    vec3 x1 = sd.posW;
    r.dir = normalize(x1 - r.origin);
    r.tMin = 0.001f * length(x1 - r.origin);
    r.tMax = 0.999f * length(x1 - r.origin);
  } else {
    if (pc.lightIsDistant) {
      r.dir = -pc.firstWi;
      r.tMin = 0.001f;
      r.tMax = kRayTMax;
    } else {
      return kNoSplat;
    }
  }
  vec3 camDir{dot(camU, r.dir), dot(camV, r.dir), dot(normalize(cur.cameraW), r.dir)};
  if (camDir.z <= 0.001f) return kNoSplat;
  if (!trace_visible(*c.tr, r)) return kNoSplat;
  vec2 film = lensLocal + length(cur.cameraW) * (vec2{camDir.x, camDir.y} / camDir.z);
  vec2 ndc = film / vec2{length(cur.cameraU), length(cur.cameraV)};
  vec2 nfp{(0.5f * ndc.x + 0.5f) * (float)p.W, (-0.5f * ndc.y + 0.5f) * (float)p.H};
  if (!fpixel_in(p, nfp)) return kNoSplat;
  uint32_t nx = (uint32_t)ffloor(nfp.x), ny = (uint32_t)ffloor(nfp.y);
  return ny * p.W + nx;
}

__device__ inline uint32_t reproject(const DevCtx& c, const PassBuffers& b, uint32_t x, uint32_t y) {
  const FrameParams& p = *c.p;
  uint32_t idx = y * p.W + x;
  PathReservoir pr = ldR(b.prevR, idx);
  if (all_zero(pr.integrand)) return kNoSplat;
  ReconnectionData pc = ldC(b.prevC, idx);
  return reproject_through_lens(c, pc, pc.time);
}

// Bilinear confidence of the previous frame at the motion-vector position.
__device__ inline float bilinear_confidence(const FrameParams& p, const PassBuffers& b, vec2 prevPixel, bool useConf,
                                            float newConfidence) {
  int tlx = (int)ffloor(prevPixel.x), tly = (int)ffloor(prevPixel.y);
  vec2 frac{fsaturate(prevPixel.x - (float)tlx), fsaturate(prevPixel.y - (float)tly)};
  for (int xx = 0; xx < 2; xx++) {
    for (int yy = 0; yy < 2; yy++) {
      int nx = tlx + xx, ny = tly + yy;
      if (!pixel_in(p, nx, ny)) continue;
      PathReservoir nr = ldR(b.prevR, (uint32_t)ny * p.W + (uint32_t)nx);
      float bw = flerp((float)(1 - xx), (float)xx, frac.x) * flerp((float)(1 - yy), (float)yy, frac.y);
      newConfidence += bw * conf_or_one(useConf, nr.confidence);
    }
  }
  return newConfidence;
}

__device__ __forceinline__ float scatter_jacobian(const ShiftedPathData& s, const ReconnectionData& r) {
  return (s.subPixelJacobian * s.secondaryPathJacobian) / (r.subPixelJacobian * r.secondaryPathJacobian);
}

// MultiReprojectTemporalSamples::run: the previous sample's primary hit is splatted once
// per time partition, at times (t / shutter + i) * shutter / K.  The projection is the
// reference's tanFovY form from the camera position (no lens offset), and the
// visibility ray runs along the unnormalized direction to 0.999 of the hit.

// Partition i's splat of the previous sample pc of pixel idx (primary hit x1 when haveHit).
__device__ __forceinline__ void multi_reproject_partition(const DevCtx& c, const PassBuffers& b, uint32_t K,
                                                          uint32_t idx, uint32_t i, const ReconnectionData& pc,
                                                          bool haveHit, vec3 x1) {
  const FrameParams& p = *c.p;
  const uint32_t P = p.W * p.H;
  const float shutter = p.camera.shutterSpeed;
  const float fractionalTime = pc.time / shutter;
  const float timePartitions = shutter / (float)K;
  const float newTime = (fractionalTime + (float)i) * timePartitions;
  if (!(p.quirkMask & kQuirkMultiProjection)) {
    // corrected: the single-splat reprojection through the sample's lens point
    uint32_t cell = reproject_through_lens(c, pc, newTime);
    if (cell != kNoSplat) {
      b.splatDst[i * P + idx] = i * P + cell;
      b.splatRank[i * P + idx] = atomicAdd(&b.cellCount[i * P + cell], 1u);
    }
    return;
  }
  CameraData cur = camera_at(p, newTime);
  vec3 rayDir = haveHit ? x1 - cur.posW : -pc.firstWi;
  float rayLength = haveHit ? 0.999f : kRayTMax;
  vec3 nd = normalize(rayDir);
  float d = dot(nd, normalize(cur.cameraW));
  if (d <= 0.001f) return;
  Ray vr{cur.posW, 0.001f, rayDir, rayLength};
  if (!trace_visible(*c.tr, vr)) return;
  float tanFovY = cur.focalLength == 0.f ? 0.f : cur.frameHeight / cur.focalLength;
  d *= tanFovY;
  vec2 off{dot(nd, normalize(cur.cameraU)), -dot(nd, normalize(cur.cameraV))};
  vec2 ndc = vec2{0.5f, 0.5f} + off / vec2{d * cur.aspectRatio, d};
  vec2 nfp = ndc * vec2{(float)p.W, (float)p.H};
  if (!fpixel_in(p, nfp)) return;
  uint32_t cell = i * P + (uint32_t)ffloor(nfp.y) * p.W + (uint32_t)ffloor(nfp.x);
  b.splatDst[i * P + idx] = cell;
  b.splatRank[i * P + idx] = atomicAdd(&b.cellCount[cell], 1u);
}

// The previous sample of pixel idx and its primary hit; false when it splats nowhere.
__device__ __forceinline__ bool multi_reproject_sample(const DevCtx& c, const PassBuffers& b, uint32_t idx,
                                                       ReconnectionData& pc, bool& haveHit, vec3& x1) {
  PathReservoir pr = ldR(b.prevR, idx);
  if (all_zero(pr.integrand)) return false;
  pc = ldC(b.prevC, idx);
  // The previous frame's camera only provides loadShadingData's view point, which the
  // hit position does not depend on.
  haveHit = hit_valid(pc.firstHit);
  if (!haveHit && !pc.lightIsDistant) return false;
  x1 = haveHit ? load_shading_data(*c.s, pc.firstHit, v3s(0.f)).posW : v3s(0.f);
  return true;
}

// Every partition of one pixel (one lane per pixel).
__device__ __forceinline__ void multi_reproject(const DevCtx& c, const PassBuffers& b, uint32_t K, uint32_t x, uint32_t y) {
  const FrameParams& p = *c.p;
  const uint32_t P = p.W * p.H;
  const uint32_t idx = y * p.W + x;
  for (uint32_t i = 0; i < K; i++) b.splatDst[i * P + idx] = kNoSplat;
  ReconnectionData pc;
  bool haveHit;
  vec3 x1;
  if (!multi_reproject_sample(c, b, idx, pc, haveHit, x1)) return;
  for (uint32_t i = 0; i < K; i++) multi_reproject_partition(c, b, K, idx, i, pc, haveHit, x1);
}

// One partition of one listed pixel (the compacted form: the pixel's destinations were
// cleared when it was listed).
__device__ __forceinline__ void multi_reproject_task(const DevCtx& c, const PassBuffers& b, uint32_t K, uint32_t idx,
                                                     uint32_t i) {
  ReconnectionData pc;
  bool haveHit;
  vec3 x1;
  if (multi_reproject_sample(c, b, idx, pc, haveHit, x1)) multi_reproject_partition(c, b, K, idx, i, pc, haveHit, x1);
}

// Lens-vertex-copy / primary-hit-reconnection split from the circle of confusion at the
// pixel's primary hit (RS/GatherTemporalResampling.rt.slang:75-96,
// RS/SpatialResampling.rt.slang:76-96); (1, 0) without depth of field.
__device__ inline vec2 dof_probabilities(const DevCtx& c, const PassBuffers& b, uint32_t idx) {
  const FrameParams& p = *c.p;
  if (!(p.camera.apertureRadius > 0.f)) return vec2{1.f, 0.f};
  HitInfo h = b.vbuffer[idx];
  CameraData cam = camera_at(p, 0.f);
  float coc;
  if (hit_valid(h)) {
    ShadingData sd = load_shading_data(*c.s, h, v3s(0.f));
    coc = coc_primary_hit(p, cam, sd.posW);
  } else {
    coc = cam.apertureRadius / length(cam.cameraU / (float)p.W + cam.cameraV / (float)p.H);
  }
  return dof_gather_probabilities(coc);
}

// GatherTemporalResampling::run (RS/GatherTemporalResampling.rt.slang:68-235): two-source
// GRIS between the current sample and the collected previous sample (backup buffers),
// with a second GRIS pass through the primary-hit reconnection shift under depth of field.
__device__ __forceinline__ void gather_temporal(const DevCtx& c, const PassBuffers& b, const TemporalOptions& o, uint32_t x,
                                       uint32_t y) {
  const FrameParams& p = *c.p;
  const bool uc = o.useConfidence != 0;
  const float dt = p.cm.artificialFrameTime;
  Rng sg = rng_make(x, y, pass_seed(p, 3));
  uint32_t idx = y * p.W + x;
  const bool hasDoF = p.camera.apertureRadius > 0.f;
  vec2 probs = dof_probabilities(c, b, idx);
  const bool doLensCopy = probs.x > 0.f;
  const bool doPrimaryRecon = probs.y > 0.f;
  PathReservoir dst = reservoir_default();
  ReconnectionData dstC = recon_default();
  vec2 floatingCoord{b.floatingCoords[2 * idx], b.floatingCoords[2 * idx + 1]};
  vec2 pix{(float)x, (float)y};
  if (doLensCopy) {
    PathReservoir pr = ldR(b.backupR, idx), cr = ldR(b.currR, idx);
    ReconnectionData pc = ldC(b.backupC, idx), cc = ldC(b.currC, idx);
    float currMIS = 0.f;
    vec3 currPHat = v3s(0.f);
    if (any_pos(cr.integrand)) {
      float m1 = luminance(cr.integrand) * conf_or_one(uc, cr.confidence);
      currPHat = cr.integrand;
      ShiftedPathData sh = gather_lens_vertex_copy_shift(c, sg, cc, cc.time + dt, floatingCoord + cr.subPixel, cc.lensSample);
      float J = sh.secondaryPathJacobian / cc.secondaryPathJacobian;
      float m2 = luminance(sh.radiance) * J * conf_or_one(uc, pr.confidence) * probs.x;
      m2 = fisnan(m2) ? 0.f : m2;
      currMIS = (m1 + m2 == 0.f) ? 0.f : m1 / (m1 + m2);
    }
    if (reservoir_add_from(dst, sg, currMIS, currPHat, 1.f, cr)) dstC = cc;
    float prevMIS = 0.f;
    vec3 prevPHat = v3s(0.f);
    float J = 1.f;
    if (any_pos(pr.integrand)) {
      ShiftedPathData sh = gather_lens_vertex_copy_shift(c, sg, pc, pc.time, pix + pr.subPixel, pc.lensSample);
      J = sh.secondaryPathJacobian / pc.secondaryPathJacobian;
      float m1 = luminance(sh.radiance) * J * conf_or_one(uc, cr.confidence);
      prevPHat = fisnan(m1) ? v3s(0.f) : sh.radiance;
      J = fisnan(m1) ? 1.f : J;
      m1 = fisnan(m1) ? 0.f : m1;
      recon_update(pc, sh);
      float m2 = luminance(pr.integrand) * conf_or_one(uc, pr.confidence) * probs.x;
      prevMIS = ((m1 + m2) == 0.f) ? 0.f : m2 / (m1 + m2);
    }
    if (reservoir_add_from(dst, sg, prevMIS, prevPHat, J, pr)) dstC = pc;
  }
  if (hasDoF && doPrimaryRecon) {
    PathReservoir pr = ldR(b.backupR, idx);
    PathReservoir cr = doLensCopy ? dst : ldR(b.currR, idx);
    ReconnectionData pc = ldC(b.backupC, idx);
    ReconnectionData cc = doLensCopy ? dstC : ldC(b.currC, idx);
    dst = reservoir_default();
    dstC = recon_default();
    float currMIS = 0.f;
    vec3 currPHat = v3s(0.f);
    if (any_pos(cr.integrand)) {
      float m1 = luminance(cr.integrand) * conf_or_one(uc, cr.confidence);
      currPHat = cr.integrand;
      ShiftedPathData sh =
          gather_primary_hit_reconnection_shift(c, sg, cc, cc.time + dt, floatingCoord + cr.subPixel, cc.firstHit);
      float J = (sh.lensVertexJacobian * sh.secondaryPathJacobian) / (cc.lensVertexJacobian * cc.secondaryPathJacobian);
      float m2 = luminance(sh.radiance) * J * conf_or_one(uc, pr.confidence) * probs.y;
      m2 = fisnan(m2) ? 0.f : m2;
      currMIS = (m1 + m2 == 0.f) ? 0.f : m1 / (m1 + m2);
    }
    if (reservoir_add_from(dst, sg, currMIS, currPHat, 1.f, cr)) dstC = cc;
    float prevMIS = 0.f;
    vec3 prevPHat = v3s(0.f);
    float J = 1.f;
    if (any_pos(pr.integrand)) {
      ShiftedPathData sh = gather_primary_hit_reconnection_shift(c, sg, pc, pc.time, pix + pr.subPixel, pc.firstHit);
      J = (sh.lensVertexJacobian * sh.secondaryPathJacobian) / (pc.lensVertexJacobian * pc.secondaryPathJacobian);
      float m1 = luminance(sh.radiance) * J * conf_or_one(uc, cr.confidence);
      prevPHat = fisnan(m1) ? v3s(0.f) : sh.radiance;
      J = fisnan(m1) ? 1.f : J;
      m1 = fisnan(m1) ? 0.f : m1;
      recon_update(pc, sh);
      float m2 = luminance(pr.integrand) * conf_or_one(uc, pr.confidence) * probs.y;
      prevMIS = ((m1 + m2) == 0.f) ? 0.f : m2 / (m1 + m2);
    }
    if (reservoir_add_from(dst, sg, prevMIS, prevPHat, J, pr)) dstC = pc;
  }
  stR(b.currR, idx, dst);
  stC(b.currC, idx, dstC);
}

__device__ inline void resolve(const FrameParams& p, const PassBuffers& b, const PackedReservoir* R, uint32_t x,
                               uint32_t y) {
  uint32_t idx = y * p.W + x;
  PathReservoir r = ldR(R, idx);
  float ucw = reservoir_ucw(r);
  vec3 col = r.integrand * ucw;
  float4* o = reinterpret_cast<float4*>(b.color) + idx;
  *o = make_float4(col.x, col.y, col.z, 1.f);
}

#endif

}  // namespace rs
