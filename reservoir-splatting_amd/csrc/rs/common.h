// Shared host/device foundation: vector types and exactly-rounded scalar helpers.
//
// Every header under csrc/rs/ is compiled twice: by hipcc for gfx950 kernels and by
// g++ for the CPU oracle.  Both builds use -ffp-contract=off and IEEE division and
// square root, so every function here returns bit-identical results on host and device.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RS_HD __host__ __device__ __forceinline__
#define RS_D __device__ __forceinline__
// Large device routines (traversal, shifts, shading) are inlined at every call site by
// default.  RS_OUTLINE_CALLS=1 emits each once and calls it instead: kernels shrink from
// tens of thousands of instructions to a few thousand, but the calls spill live state to
// scratch; at 1080p the inlined build is ~12% faster overall (resampling kernels ~25%).
#if defined(RS_OUTLINE_CALLS) && RS_OUTLINE_CALLS
#define RS_OUTLINE __noinline__
#else
#define RS_OUTLINE inline
#endif
#else
#include <math.h>
#include <string.h>
#define RS_HD inline
#define RS_D inline
#endif

#if defined(__clang__)
#pragma clang fp contract(off)
#endif

namespace rs {

// ---------------------------------------------------------------- scalar helpers

RS_HD float fsqrt(float x) { return __builtin_sqrtf(x); }
RS_HD float fabs_(float x) { return __builtin_fabsf(x); }
RS_HD float ffloor(float x) { return __builtin_floorf(x); }
// Round to nearest, ties to even (HLSL/Slang `round`).
RS_HD float fround(float x) { return __builtin_rintf(x); }
RS_HD float fmin_(float a, float b) { return a < b ? a : b; }
RS_HD float fmax_(float a, float b) { return a > b ? a : b; }
RS_HD float fclamp(float x, float lo, float hi) { return fmin_(fmax_(x, lo), hi); }
RS_HD float fsaturate(float x) { return fclamp(x, 0.f, 1.f); }
RS_HD bool fisnan(float x) { return x != x; }
RS_HD float flerp(float a, float b, float t) { return a + (b - a) * t; }

RS_HD uint32_t asuint(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  return u;
}
RS_HD float asfloat(uint32_t u) {
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}
RS_HD int32_t asint(float f) { return (int32_t)asuint(f); }
RS_HD float asfloat_i(int32_t i) { return asfloat((uint32_t)i); }

static constexpr float kPi = 3.14159265358979323846f;
static constexpr float k2Pi = 6.28318530717958647693f;
static constexpr float kInvPi = 0.318309886183790671538f;
static constexpr float kFltMin = 1.175494351e-38f;

// ---------------------------------------------------------------- vectors

struct vec2 {
  float x, y;
};
struct vec3 {
  float x, y, z;
};
struct uvec2 {
  uint32_t x, y;
};

RS_HD vec2 v2(float x, float y) { return vec2{x, y}; }
RS_HD vec3 v3(float x, float y, float z) { return vec3{x, y, z}; }
RS_HD vec3 v3s(float s) { return vec3{s, s, s}; }

RS_HD vec2 operator+(vec2 a, vec2 b) { return {a.x + b.x, a.y + b.y}; }
RS_HD vec2 operator-(vec2 a, vec2 b) { return {a.x - b.x, a.y - b.y}; }
RS_HD vec2 operator*(vec2 a, vec2 b) { return {a.x * b.x, a.y * b.y}; }
RS_HD vec2 operator/(vec2 a, vec2 b) { return {a.x / b.x, a.y / b.y}; }
RS_HD vec2 operator*(float s, vec2 a) { return {s * a.x, s * a.y}; }
RS_HD vec2 operator*(vec2 a, float s) { return {a.x * s, a.y * s}; }
// Division of a vector by a scalar is one IEEE division and a multiply per component
// (both builds), instead of a division per component: the correctly rounded divide is a
// ten-instruction sequence on the GPU.
RS_HD vec2 operator/(vec2 a, float s) {
  float r = 1.f / s;
  return {a.x * r, a.y * r};
}

RS_HD vec3 operator+(vec3 a, vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
RS_HD vec3 operator-(vec3 a, vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
RS_HD vec3 operator-(vec3 a) { return {-a.x, -a.y, -a.z}; }
RS_HD vec3 operator*(vec3 a, vec3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
RS_HD vec3 operator/(vec3 a, vec3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
RS_HD vec3 operator*(float s, vec3 a) { return {s * a.x, s * a.y, s * a.z}; }
RS_HD vec3 operator*(vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
RS_HD vec3 operator/(vec3 a, float s) {
  float r = 1.f / s;
  return {a.x * r, a.y * r, a.z * r};
}
RS_HD vec3& operator+=(vec3& a, vec3 b) { a = a + b; return a; }
RS_HD vec3& operator*=(vec3& a, vec3 b) { a = a * b; return a; }
RS_HD vec3& operator*=(vec3& a, float s) { a = a * s; return a; }
RS_HD vec3& operator/=(vec3& a, float s) { a = a / s; return a; }

// Dot products evaluate left to right: (x*x' + y*y') + z*z'.
RS_HD float dot(vec2 a, vec2 b) { return a.x * b.x + a.y * b.y; }
RS_HD float dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RS_HD vec3 cross(vec3 a, vec3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
RS_HD float length(vec3 a) { return fsqrt(dot(a, a)); }
RS_HD float length(vec2 a) { return fsqrt(dot(a, a)); }
RS_HD vec3 normalize(vec3 a) {
  float l = length(a);
  return a / l;
}
RS_HD vec3 lerp(vec3 a, vec3 b, float t) { return a + (b - a) * t; }
RS_HD bool any_pos(vec3 a) { return a.x > 0.f || a.y > 0.f || a.z > 0.f; }
RS_HD bool all_zero(vec3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }
RS_HD vec3 vabs(vec3 a) { return {fabs_(a.x), fabs_(a.y), fabs_(a.z)}; }
RS_HD vec2 vfloor(vec2 a) { return {ffloor(a.x), ffloor(a.y)}; }
RS_HD vec3 vmin(vec3 a, vec3 b) { return {fmin_(a.x, b.x), fmin_(a.y, b.y), fmin_(a.z, b.z)}; }
RS_HD vec3 vmax(vec3 a, vec3 b) { return {fmax_(a.x, b.x), fmax_(a.y, b.y), fmax_(a.z, b.z)}; }

// Rec.709 luminance; defines the target function p-hat of every resampling step
// (Source/Falcor/Utils/Color/ColorHelpers.slang:45-49).
RS_HD float luminance(vec3 c) { return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f; }

// ---------------------------------------------------------------- fp16 <-> fp32

// IEEE binary16 from binary32 with round-to-nearest-even (used for the alias table
// threshold, Source/Falcor/Rendering/Lights/EmissivePowerSampler.cpp:157-164).
RS_HD uint32_t f32tof16(float f) {
  uint32_t x = asuint(f);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t absx = x & 0x7fffffffu;
  if (absx >= 0x7f800000u) return sign | (absx > 0x7f800000u ? 0x7e00u : 0x7c00u);
  if (absx >= 0x477ff000u) return sign | 0x7c00u;  // rounds to >= 65520 -> inf
  if (absx < 0x38800000u) {                         // result is subnormal or zero
    if (absx < 0x33000000u) return sign;            // < 2^-25 rounds to 0
    uint32_t e = absx >> 23;
    uint32_t m = (absx & 0x7fffffu) | 0x800000u;
    uint32_t shift = 126u - e;                      // 14..24 (target exponent -14 at 10 frac bits)
    uint32_t h = m >> (shift);
    uint32_t rem = m & ((1u << shift) - 1u);
    uint32_t half = 1u << (shift - 1u);
    if (rem > half || (rem == half && (h & 1u))) h += 1u;
    return sign | h;
  }
  uint32_t e = (absx >> 23) - 112u;  // rebias 127 -> 15
  uint32_t m = absx & 0x7fffffu;
  uint32_t h = (e << 10) | (m >> 13);
  uint32_t rem = m & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h += 1u;
  return sign | h;
}

RS_HD float f16tof32(uint32_t h) {
  uint32_t sign = (h & 0x8000u) << 16;
  uint32_t e = (h >> 10) & 0x1fu;
  uint32_t m = h & 0x3ffu;
  if (e == 0) {
    // subnormal: m * 2^-24, exact in fp32
    float v = (float)m * 5.9604644775390625e-8f;
    return asfloat(asuint(v) | sign);
  }
  if (e == 31) return asfloat(sign | 0x7f800000u | (m << 13));
  return asfloat(sign | ((e + 112u) << 23) | (m << 13));
}

}  // namespace rs
