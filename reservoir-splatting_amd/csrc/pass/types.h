// Reservoir, reconnection and shifted-path records of the pass, with their HBM layouts.
//
// Field meaning follows Source/RenderPasses/ReservoirSplatting/Reservoir.slang:35-132
// (CandidateReservoir, PathReservoir; confidence cap 20) and ReconnectionData.slang:
// 35-166 (ShiftedPathData, ReconnectionData and its update()).
//
// HBM layouts (little-endian 32-bit words):
//   PathReservoir, 32 B = 2 x 16 B:
//     [0] confidence  [1] totalWeight  [2..4] integrand.xyz  [5..6] subPixel.xy  [7] 0
//   ReconnectionData, 112 B = 7 x 16 B:
//     [0..1] subPixel  [2..3] lensSample  [4] time
//     [5] bits: pathLength (0-7) | firstComp (8-11) | secondComp (12-15) |
//               transmissionEvent (16) | lightIsNEE (17) | lightIsDistant (18)
//     [6] lightPdf  [7] subPixelJacobian
//     [8] firstHit.prim  [9] firstHit.u  [10] firstHit.v  [11] lensVertexJacobian
//     [12..14] firstWi  [15] secondaryPathJacobian
//     [16] secondHit.prim  [17] secondHit.u  [18] secondHit.v  [19] irradiance.x
//     [20..22] secondWo  [23] irradiance.y
//     [24] irradiance.z  [25..27] earlyThroughput
//   ShiftedPathData, 64 B = 4 x 16 B (the Robust gather's precomputed shifts):
//     [0] primaryHit.prim  [1] primaryHit.u  [2] primaryHit.v  [3] subPixelJacobian
//     [4..5] fractionalPixel  [6..7] lensSample  [8..10] firstRayDir  [11] lensVertexJacobian
//     [12..14] radiance  [15] secondaryPathJacobian
#pragma once
#include "../rs/bsdf.h"

namespace rs {

static constexpr float kConfidenceCap = 20.f;

struct ReconnectionData {
  vec2 subPixel;
  vec2 lensSample;
  float time;
  uint32_t pathLength;
  HitInfo firstHit;
  uint32_t firstBSDFComponentType;
  vec3 firstWi;
  HitInfo secondHit;
  uint32_t secondBSDFComponentType;
  vec3 secondWo;
  bool transmissionEvent;
  bool lightIsNEE;
  bool lightIsDistant;
  float lightPdf;
  float subPixelJacobian;
  float lensVertexJacobian;
  float secondaryPathJacobian;
  vec3 irradiance;
  vec3 earlyThroughput;
};

RS_HD ReconnectionData recon_default() {
  ReconnectionData r;
  r.subPixel = vec2{0.5f, 0.5f};
  r.lensSample = vec2{0.f, 0.f};
  r.time = 0.f;
  r.pathLength = 0;
  r.firstHit = hit_invalid();
  r.firstBSDFComponentType = 0;
  r.firstWi = v3s(0.f);
  r.secondHit = hit_invalid();
  r.secondBSDFComponentType = 0;
  r.secondWo = v3s(0.f);
  r.transmissionEvent = false;
  r.lightIsNEE = false;
  r.lightIsDistant = false;
  r.lightPdf = 0.f;
  r.subPixelJacobian = 1.f;
  r.lensVertexJacobian = 1.f;
  r.secondaryPathJacobian = 1.f;
  r.irradiance = v3s(0.f);
  r.earlyThroughput = v3s(1.f);
  return r;
}

struct ShiftedPathData {
  HitInfo primaryHit;
  vec2 fractionalPixel;
  vec2 lensSample;
  vec3 firstRayDir;
  float subPixelJacobian;
  float lensVertexJacobian;
  float secondaryPathJacobian;
  vec3 radiance;
};

RS_HD ShiftedPathData shifted_default() {
  ShiftedPathData s;
  s.primaryHit = hit_invalid();
  s.fractionalPixel = vec2{-1.f, -1.f};
  s.lensSample = vec2{0.f, 0.f};
  s.firstRayDir = v3s(0.f);
  s.subPixelJacobian = 1.f;
  s.lensVertexJacobian = 1.f;
  s.secondaryPathJacobian = 1.f;
  s.radiance = v3s(0.f);
  return s;
}

// ReconnectionData::update(ShiftedPathData).
RS_HD void recon_update(ReconnectionData& r, const ShiftedPathData& s) {
  r.firstHit = s.primaryHit;
  r.subPixel = s.fractionalPixel - vfloor(s.fractionalPixel);
  r.lensSample = s.lensSample;
  r.firstWi = -s.firstRayDir;
  r.subPixelJacobian = s.subPixelJacobian;
  r.lensVertexJacobian = s.lensVertexJacobian;
  r.secondaryPathJacobian = s.secondaryPathJacobian;
}

struct PathReservoir {
  float confidence;
  float totalWeight;
  vec3 integrand;
  vec2 subPixel;
};

RS_HD PathReservoir reservoir_default() {
  PathReservoir r;
  r.confidence = 0.f;
  r.totalWeight = 0.f;
  r.integrand = v3s(0.f);
  r.subPixel = vec2{0.5f, 0.5f};
  return r;
}

RS_HD float reservoir_ucw(const PathReservoir& r) {
  float pHat = luminance(r.integrand);
  return pHat == 0.f ? 0.f : r.totalWeight / pHat;
}

// PathReservoir::addSampleFromReservoir.
RS_HD bool reservoir_add_from(PathReservoir& dst, Rng& sg, float mis, vec3 integrand, float jacobian,
                              const PathReservoir& other) {
  float pHat = luminance(integrand);
  float w = mis * pHat * reservoir_ucw(other) * jacobian;
  dst.totalWeight = dst.totalWeight + w;
  dst.confidence = fmin_(dst.confidence + other.confidence, kConfidenceCap);
  float u = next1(sg);
  bool selected = (u * dst.totalWeight < w);
  if (selected) {
    dst.integrand = integrand;
    dst.subPixel = other.subPixel;
  }
  return selected;
}

struct CandidateReservoir {
  float totalWeight;
  vec3 integrand;
};

// CandidateReservoir::addVertex.
RS_HD bool candidate_add_vertex(CandidateReservoir& c, Rng& sg, float mis, vec3 f) {
  float pHat = luminance(f);
  float w = mis * pHat;
  w = fisnan(w) ? 0.f : w;
  c.totalWeight += w;
  float u = next1(sg);
  bool selected = (u * c.totalWeight < w);
  if (selected) c.integrand = f;
  return selected;
}

// PathReservoir::add(CandidateReservoir).
RS_HD bool reservoir_add_candidate(PathReservoir& dst, Rng& sg, float mis, const CandidateReservoir& c) {
  float w = mis * c.totalWeight;
  dst.totalWeight = dst.totalWeight + w;
  dst.confidence = fmin_(dst.confidence + 1.f, kConfidenceCap);
  float u = next1(sg);
  bool selected = (u * dst.totalWeight < w);
  if (selected) dst.integrand = c.integrand;
  return selected;
}

// ------------------------------------------------------------------ packing

struct PackedReservoir {
  float w[8];
};
struct PackedRecon {
  uint32_t w[28];
};

struct PackedShifted {
  uint32_t w[16];
};

RS_HD PackedShifted pack_shifted(const ShiftedPathData& s) {
  PackedShifted p;
  p.w[0] = s.primaryHit.prim;
  p.w[1] = asuint(s.primaryHit.u);
  p.w[2] = asuint(s.primaryHit.v);
  p.w[3] = asuint(s.subPixelJacobian);
  p.w[4] = asuint(s.fractionalPixel.x);
  p.w[5] = asuint(s.fractionalPixel.y);
  p.w[6] = asuint(s.lensSample.x);
  p.w[7] = asuint(s.lensSample.y);
  p.w[8] = asuint(s.firstRayDir.x);
  p.w[9] = asuint(s.firstRayDir.y);
  p.w[10] = asuint(s.firstRayDir.z);
  p.w[11] = asuint(s.lensVertexJacobian);
  p.w[12] = asuint(s.radiance.x);
  p.w[13] = asuint(s.radiance.y);
  p.w[14] = asuint(s.radiance.z);
  p.w[15] = asuint(s.secondaryPathJacobian);
  return p;
}

RS_HD ShiftedPathData unpack_shifted(const PackedShifted& p) {
  ShiftedPathData s;
  s.primaryHit = HitInfo{p.w[0], asfloat(p.w[1]), asfloat(p.w[2])};
  s.subPixelJacobian = asfloat(p.w[3]);
  s.fractionalPixel = vec2{asfloat(p.w[4]), asfloat(p.w[5])};
  s.lensSample = vec2{asfloat(p.w[6]), asfloat(p.w[7])};
  s.firstRayDir = vec3{asfloat(p.w[8]), asfloat(p.w[9]), asfloat(p.w[10])};
  s.lensVertexJacobian = asfloat(p.w[11]);
  s.radiance = vec3{asfloat(p.w[12]), asfloat(p.w[13]), asfloat(p.w[14])};
  s.secondaryPathJacobian = asfloat(p.w[15]);
  return s;
}

RS_HD PackedReservoir pack_reservoir(const PathReservoir& r) {
  PackedReservoir p;
  p.w[0] = r.confidence;
  p.w[1] = r.totalWeight;
  p.w[2] = r.integrand.x;
  p.w[3] = r.integrand.y;
  p.w[4] = r.integrand.z;
  p.w[5] = r.subPixel.x;
  p.w[6] = r.subPixel.y;
  p.w[7] = 0.f;
  return p;
}
RS_HD PathReservoir unpack_reservoir(const PackedReservoir& p) {
  PathReservoir r;
  r.confidence = p.w[0];
  r.totalWeight = p.w[1];
  r.integrand = vec3{p.w[2], p.w[3], p.w[4]};
  r.subPixel = vec2{p.w[5], p.w[6]};
  return r;
}

RS_HD PackedRecon pack_recon(const ReconnectionData& r) {
  PackedRecon p;
  p.w[0] = asuint(r.subPixel.x);
  p.w[1] = asuint(r.subPixel.y);
  p.w[2] = asuint(r.lensSample.x);
  p.w[3] = asuint(r.lensSample.y);
  p.w[4] = asuint(r.time);
  p.w[5] = (r.pathLength & 0xffu) | ((r.firstBSDFComponentType & 0xfu) << 8) | ((r.secondBSDFComponentType & 0xfu) << 12) |
           ((r.transmissionEvent ? 1u : 0u) << 16) | ((r.lightIsNEE ? 1u : 0u) << 17) | ((r.lightIsDistant ? 1u : 0u) << 18);
  p.w[6] = asuint(r.lightPdf);
  p.w[7] = asuint(r.subPixelJacobian);
  p.w[8] = r.firstHit.prim;
  p.w[9] = asuint(r.firstHit.u);
  p.w[10] = asuint(r.firstHit.v);
  p.w[11] = asuint(r.lensVertexJacobian);
  p.w[12] = asuint(r.firstWi.x);
  p.w[13] = asuint(r.firstWi.y);
  p.w[14] = asuint(r.firstWi.z);
  p.w[15] = asuint(r.secondaryPathJacobian);
  p.w[16] = r.secondHit.prim;
  p.w[17] = asuint(r.secondHit.u);
  p.w[18] = asuint(r.secondHit.v);
  p.w[19] = asuint(r.irradiance.x);
  p.w[20] = asuint(r.secondWo.x);
  p.w[21] = asuint(r.secondWo.y);
  p.w[22] = asuint(r.secondWo.z);
  p.w[23] = asuint(r.irradiance.y);
  p.w[24] = asuint(r.irradiance.z);
  p.w[25] = asuint(r.earlyThroughput.x);
  p.w[26] = asuint(r.earlyThroughput.y);
  p.w[27] = asuint(r.earlyThroughput.z);
  return p;
}

RS_HD ReconnectionData unpack_recon(const PackedRecon& p) {
  ReconnectionData r;
  r.subPixel = vec2{asfloat(p.w[0]), asfloat(p.w[1])};
  r.lensSample = vec2{asfloat(p.w[2]), asfloat(p.w[3])};
  r.time = asfloat(p.w[4]);
  uint32_t b = p.w[5];
  r.pathLength = b & 0xffu;
  r.firstBSDFComponentType = (b >> 8) & 0xfu;
  r.secondBSDFComponentType = (b >> 12) & 0xfu;
  r.transmissionEvent = ((b >> 16) & 1u) != 0;
  r.lightIsNEE = ((b >> 17) & 1u) != 0;
  r.lightIsDistant = ((b >> 18) & 1u) != 0;
  r.lightPdf = asfloat(p.w[6]);
  r.subPixelJacobian = asfloat(p.w[7]);
  r.firstHit = HitInfo{p.w[8], asfloat(p.w[9]), asfloat(p.w[10])};
  r.lensVertexJacobian = asfloat(p.w[11]);
  r.firstWi = vec3{asfloat(p.w[12]), asfloat(p.w[13]), asfloat(p.w[14])};
  r.secondaryPathJacobian = asfloat(p.w[15]);
  r.secondHit = HitInfo{p.w[16], asfloat(p.w[17]), asfloat(p.w[18])};
  r.irradiance = vec3{asfloat(p.w[19]), asfloat(p.w[23]), asfloat(p.w[24])};
  r.secondWo = vec3{asfloat(p.w[20]), asfloat(p.w[21]), asfloat(p.w[22])};
  r.earlyThroughput = vec3{asfloat(p.w[25]), asfloat(p.w[26]), asfloat(p.w[27])};
  return r;
}

}  // namespace rs
