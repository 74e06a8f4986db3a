// Environment-map light: lookup, importance map and importance sampling, shared by the GPU
// kernels and the CPU oracle.
//
// Follows Falcor's EnvMap and EnvMapSampler:
//   * lookup: Source/Falcor/Scene/Lights/EnvMap.slang:60-118 (eval(dir) = texture(
//     worldToUv(dir)) * intensity * tint; the sampler is bilinear, wrapping in u and
//     clamping in v, Source/Falcor/Scene/Lights/EnvMap.cpp:128-136);
//   * rotation: EnvMap.cpp:62-72 with matrixFromRotationXYZ,
//     Source/Falcor/Utils/Math/MatrixMath.h:649-680;
//   * importance map: EnvMapSamplerSetup.cs.slang:43-75 (each texel of a square
//     equal-area octahedral map averages the luminance of 8x8 subsamples of the lat-long
//     texture; 512x512 texels, EnvMapSampler.cpp:37-41, 80-114), then a box-filtered mip
//     chain down to 1x1;
//   * sampling: EnvMapSampler.slang:80-140 (hierarchical warping through the mips, pdf =
//     texel / mean / 4 pi) and evalPdf :147-163 (point lookup with clamping);
//   * mappings: Source/Falcor/Utils/Math/MathHelpers.slang:111-137 (lat-long) and :213-255
//     (equal-area octahedral, Clarberg 2008).
// Two details are this build's own, so that host and device agree bit for bit:
//   * the lat-long mapping evaluates atan2 and acos in fp32 with the transcendentals of
//     fmath.h (the reference converts doubles to fp32, a difference of an ulp or so);
//   * bilinear weights are exact fp32 (texture units round them to 8 fraction bits), and a
//     mip texel is ((a + b) + (c + d)) * 0.25 of its four children.
#pragma once
#include "fmath.h"

namespace rs {

static constexpr uint32_t kEnvImportanceDim = 512;   // EnvMapSampler.cpp:40
static constexpr uint32_t kEnvImportanceSpp = 64;    // 8 x 8 subsamples, EnvMapSampler.cpp:41
static constexpr uint32_t kEnvImportanceLog2 = 9;
static constexpr float kInv2Pi = 0.159154943091895335769f;
static constexpr float kInv4Pi = 0.0795774715459476678844f;      // M_1_4PI
static constexpr float k4Pi = 12.5663706143591729539f;           // M_4PI
static constexpr float kPiOver4 = 0.785398163397448309616f;
static constexpr float k2OverPi = 0.636619772367581343076f;       // M_2_PI

struct EnvMapView {
  const float* texels;      // RGBA32F lat-long texture, W x H, row 0 at v = 0
  uint32_t W, H;
  vec3 scale;               // intensity * tint
  float toWorld[9];         // row-major 3x3 (EnvMapData::transform)
  float toLocal[9];         // its transpose (the inverse of a rotation)
  const float* importance;  // mip chain: level k has (dim >> k)^2 texels, levels concatenated
  uint32_t present;         // 0: the scene has no environment map
};

RS_HD vec3 mat3_mul(const float* m, vec3 v) {
  return vec3{m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
              m[6] * v.x + m[7] * v.y + m[8] * v.z};
}

// Rotation from Euler angles in degrees (EnvMap::setRotation), and its transpose.
RS_HD void env_rotation(const float degreesXYZ[3], float toWorld[9], float toLocal[9]) {
  const float kDeg = 0.0174532925199432957692f;
  float ax = -(degreesXYZ[0] * kDeg), ay = -(degreesXYZ[1] * kDeg), az = -(degreesXYZ[2] * kDeg);
  float c1 = fcos(ax), c2 = fcos(ay), c3 = fcos(az);
  float s1 = fsin(ax), s2 = fsin(ay), s3 = fsin(az);
  toWorld[0] = c2 * c3;
  toWorld[1] = c2 * s3;
  toWorld[2] = -s2;
  toWorld[3] = -c1 * s3 + s1 * s2 * c3;
  toWorld[4] = c1 * c3 + s1 * s2 * s3;
  toWorld[5] = s1 * c2;
  toWorld[6] = s1 * s3 + c1 * s2 * c3;
  toWorld[7] = -s1 * c3 + c1 * s2 * s3;
  toWorld[8] = c1 * c2;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) toLocal[3 * i + j] = toWorld[3 * j + i];
}

RS_HD float fsign(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }

// atan2(y, x) in (-pi, pi].
RS_HD float fatan2(float y, float x) {
  if (x == 0.f) return y > 0.f ? 0.5f * kPi : (y < 0.f ? -0.5f * kPi : 0.f);
  float a = fatan(y / x);
  if (x > 0.f) return a;
  return y >= 0.f ? a + kPi : a - kPi;
}

// world_to_latlong_map: u from atan2(x, -z), v from acos(y).
RS_HD vec2 world_to_latlong(vec3 dir) {
  vec3 p = normalize(dir);
  return vec2{fatan2(p.x, -p.z) * kInv2Pi + 0.5f, facos(p.y) * kInvPi};
}

RS_HD vec3 latlong_to_world(vec2 uv) {
  float phi = kPi * (2.f * fsaturate(uv.x) - 1.f);
  float theta = kPi * fsaturate(uv.y);
  float sinTheta = fsin(theta), cosTheta = fcos(theta), sinPhi = fsin(phi), cosPhi = fcos(phi);
  return vec3{sinTheta * sinPhi, cosTheta, -sinTheta * cosPhi};
}

// ndir_to_oct_equal_area_unorm.
RS_HD vec2 ndir_to_oct_ea(vec3 n) {
  float r = fsqrt(1.f - fabs_(n.z));
  float phi = fatan2(fabs_(n.y), fabs_(n.x));
  vec2 p;
  p.y = r * phi * k2OverPi;
  p.x = r - p.y;
  if (n.z < 0.f) p = vec2{1.f - p.y, 1.f - p.x};
  p = vec2{p.x * fsign(n.x), p.y * fsign(n.y)};
  return vec2{p.x * 0.5f + 0.5f, p.y * 0.5f + 0.5f};
}

// oct_to_ndir_equal_area_unorm.
RS_HD vec3 oct_ea_to_ndir(vec2 p) {
  p = vec2{p.x * 2.f - 1.f, p.y * 2.f - 1.f};
  float d = 1.f - (fabs_(p.x) + fabs_(p.y));
  float r = 1.f - fabs_(d);
  float phi = (r > 0.f) ? ((fabs_(p.y) - fabs_(p.x)) / r + 1.f) * kPiOver4 : 0.f;
  float f = r * fsqrt(2.f - r * r);
  float x = f * fsign(p.x) * fcos(phi);
  float y = f * fsign(p.y) * fsin(phi);
  float z = fsign(d) * (1.f - r * r);
  return vec3{x, y, z};
}

RS_HD vec3 env_texel(const EnvMapView& e, int x, int y) {
  int W = (int)e.W, H = (int)e.H;
  x %= W;
  if (x < 0) x += W;
  y = y < 0 ? 0 : (y >= H ? H - 1 : y);
  const float* t = e.texels + 4 * ((size_t)y * e.W + (size_t)x);
  return vec3{t[0], t[1], t[2]};
}

// Texture2D.SampleLevel(bilinear, wrap u, clamp v, lod 0).rgb.
RS_HD vec3 env_lookup(const EnvMapView& e, vec2 uv) {
  float fx = uv.x * (float)e.W - 0.5f, fy = uv.y * (float)e.H - 0.5f;
  float x0f = ffloor(fx), y0f = ffloor(fy);
  float tx = fx - x0f, ty = fy - y0f;
  int x0 = (int)x0f, y0 = (int)y0f;
  vec3 a = env_texel(e, x0, y0), b = env_texel(e, x0 + 1, y0);
  vec3 c = env_texel(e, x0, y0 + 1), d = env_texel(e, x0 + 1, y0 + 1);
  vec3 top = a + (b - a) * tx, bot = c + (d - c) * tx;
  return top + (bot - top) * ty;
}

// EnvMap::eval(dir): radiance arriving from world direction dir.
RS_HD vec3 env_eval(const EnvMapView& e, vec3 dir) {
  vec2 uv = world_to_latlong(mat3_mul(e.toLocal, dir));
  return env_lookup(e, uv) * e.scale;
}

// First texel of importance mip level k.
RS_HD size_t env_mip_offset(uint32_t k) {
  size_t off = 0;
  for (uint32_t i = 0; i < k; i++) off += (size_t)(kEnvImportanceDim >> i) * (kEnvImportanceDim >> i);
  return off;
}

RS_HD float env_imp(const EnvMapView& e, uint32_t mip, uint32_t x, uint32_t y) {
  return e.importance[env_mip_offset(mip) + (size_t)y * (kEnvImportanceDim >> mip) + x];
}

// One texel of the base importance level: mean luminance of 8 x 8 subsamples.
RS_HD float env_importance_texel(const EnvMapView& e, uint32_t px, uint32_t py) {
  const uint32_t ns = 8, dimS = kEnvImportanceDim * ns;
  float L = 0.f;
  for (uint32_t y = 0; y < ns; y++) {
    for (uint32_t x = 0; x < ns; x++) {
      vec2 p{((float)(px * ns + x) + 0.5f) / (float)dimS, ((float)(py * ns + y) + 0.5f) / (float)dimS};
      vec3 dir = oct_ea_to_ndir(p);
      L += luminance(env_lookup(e, world_to_latlong(dir)));
    }
  }
  return L * (1.f / (float)(ns * ns));
}

// Texel (x, y) of mip level k >= 1 from level k - 1.
RS_HD float env_importance_mip_texel(const float* prev, uint32_t prevDim, uint32_t x, uint32_t y) {
  float a = prev[(size_t)(2 * y) * prevDim + 2 * x], b = prev[(size_t)(2 * y) * prevDim + 2 * x + 1];
  float c = prev[(size_t)(2 * y + 1) * prevDim + 2 * x], d = prev[(size_t)(2 * y + 1) * prevDim + 2 * x + 1];
  return ((a + b) + (c + d)) * 0.25f;
}

// EnvMapSampler::sample: direction towards the light, its solid-angle pdf and radiance.
RS_HD bool env_sample(const EnvMapView& e, vec2 rnd, vec3& dir, float& pdf, vec3& Le) {
  vec2 p = rnd;
  uint32_t px = 0, py = 0;
  for (int mip = (int)kEnvImportanceLog2 - 1; mip >= 0; mip--) {
    px *= 2;
    py *= 2;
    float w0 = env_imp(e, (uint32_t)mip, px, py), w1 = env_imp(e, (uint32_t)mip, px + 1, py);
    float w2 = env_imp(e, (uint32_t)mip, px, py + 1), w3 = env_imp(e, (uint32_t)mip, px + 1, py + 1);
    float q0 = w0 + w2, q1 = w1 + w3;
    uint32_t ox, oy;
    float d = q0 / (q0 + q1);
    if (p.x < d) {
      ox = 0;
      p.x = p.x / d;
    } else {
      ox = 1;
      p.x = (p.x - d) / (1.f - d);
    }
    float ev = ox == 0 ? (w0 / q0) : (w1 / q1);
    if (p.y < ev) {
      oy = 0;
      p.y = p.y / ev;
    } else {
      oy = 1;
      p.y = (p.y - ev) / (1.f - ev);
    }
    px += ox;
    py += oy;
  }
  const float invDim = 1.f / (float)kEnvImportanceDim;
  vec2 uv{((float)px + p.x) * invDim, ((float)py + p.y) * invDim};
  vec3 local = oct_ea_to_ndir(uv);
  float avg = env_imp(e, kEnvImportanceLog2, 0, 0);
  float texPdf = env_imp(e, 0, px, py) / avg;
  dir = mat3_mul(e.toWorld, local);
  pdf = texPdf * kInv4Pi;
  Le = env_eval(e, dir);
  return true;
}

// EnvMapSampler::evalPdf.
RS_HD float env_eval_pdf(const EnvMapView& e, vec3 dir) {
  vec2 uv = ndir_to_oct_ea(mat3_mul(e.toLocal, dir));
  int x = (int)ffloor(uv.x * (float)kEnvImportanceDim), y = (int)ffloor(uv.y * (float)kEnvImportanceDim);
  const int last = (int)kEnvImportanceDim - 1;
  x = x < 0 ? 0 : (x > last ? last : x);
  y = y < 0 ? 0 : (y > last ? last : y);
  float avg = env_imp(e, kEnvImportanceLog2, 0, 0);
  float pdf = env_imp(e, 0, (uint32_t)x, (uint32_t)y) / avg;
  return pdf * (1.f / k4Pi);
}

}  // namespace rs
