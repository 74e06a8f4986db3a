// Sampling and geometry helpers used along the pass.
//
// Follows Source/Falcor/Utils/Math/MathHelpers.slang:269-277 (sample_disk),
// :323-354 (concentric disk / cosine hemisphere), :391-396 (sample_triangle),
// :542-554 (perp_stark), Source/Falcor/Utils/Math/Quaternion.slang:39-59 and
// Source/Falcor/Utils/Geometry/GeometryHelpers.slang:54-68 (computeRayOrigin, the
// integer-ulp offset of Waechter & Binder, Ray Tracing Gems ch. 6).
#pragma once
#include "fmath.h"

namespace rs {

RS_HD vec2 sample_disk(vec2 u) {
  float r = fsqrt(u.x);
  float phi = k2Pi * u.y;
  float s, c;
  fsincos(phi, s, c);
  return vec2{r * c, r * s};
}

RS_HD vec2 sample_disk_concentric(vec2 u) {
  u = vec2{2.f * u.x - 1.f, 2.f * u.y - 1.f};
  if (u.x == 0.f && u.y == 0.f) return u;
  float phi, r;
  if (fabs_(u.x) > fabs_(u.y)) {
    r = u.x;
    phi = (u.y / u.x) * 0.785398163397448309616f;
  } else {
    r = u.y;
    phi = 1.57079632679489661923f - (u.x / u.y) * 0.785398163397448309616f;
  }
  float s, c;
  fsincos(phi, s, c);
  return vec2{r * c, r * s};
}

RS_HD vec3 sample_cosine_hemisphere_concentric(vec2 u, float& pdf) {
  vec2 d = sample_disk_concentric(u);
  float z = fsqrt(fmax_(0.f, 1.f - dot(d, d)));
  pdf = z * kInvPi;
  return vec3{d.x, d.y, z};
}

// Barycentrics (1-u-v, u, v) of a uniform point on a triangle.
RS_HD vec3 sample_triangle(vec2 u) {
  float su = fsqrt(u.x);
  float bx = 1.f - su, by = u.y * su;
  return vec3{1.f - bx - by, bx, by};
}

RS_HD vec3 perp_stark(vec3 u) {
  vec3 a = vabs(u);
  uint32_t uyx = (a.x - a.y) < 0.f ? 1u : 0u;
  uint32_t uzx = (a.x - a.z) < 0.f ? 1u : 0u;
  uint32_t uzy = (a.y - a.z) < 0.f ? 1u : 0u;
  uint32_t xm = uyx & uzx;
  uint32_t ym = (1u ^ xm) & uzy;
  uint32_t zm = 1u ^ (xm | ym);
  return normalize(cross(u, vec3{(float)xm, (float)ym, (float)zm}));
}

// Quaternion (xyz, w) helpers.
struct quat {
  float x, y, z, w;
};
RS_HD quat qmul(quat a, quat b) {
  vec3 av{a.x, a.y, a.z}, bv{b.x, b.y, b.z};
  vec3 xyz = bv * a.w + av * b.w + cross(av, bv);
  return quat{xyz.x, xyz.y, xyz.z, a.w * b.w - dot(av, bv)};
}
RS_HD vec3 rotate_vector(vec3 v, quat r) {
  quat rc{-r.x, -r.y, -r.z, r.w};
  quat t = qmul(r, qmul(quat{v.x, v.y, v.z, 0.f}, rc));
  return vec3{t.x, t.y, t.z};
}
RS_HD quat rotate_angle_axis(float angle, vec3 axis) {
  float sn, cs;
  fsincos(angle * 0.5f, sn, cs);
  return quat{axis.x * sn, axis.y * sn, axis.z * sn, cs};
}

// Offset a hit position along the face normal by a few ulps (integer offset far
// from the origin, a fixed float offset close to it).
RS_HD vec3 compute_ray_origin(vec3 pos, vec3 n) {
  const float origin = 1.f / 16.f;
  const float fScale = 3.f / 65536.f;
  const float iScale = 3.f * 256.f;
  float p[3] = {pos.x, pos.y, pos.z};
  float nn[3] = {n.x, n.y, n.z};
  float out[3];
  for (int i = 0; i < 3; i++) {
    int32_t iOff = (int32_t)(nn[i] * iScale);
    float iPos = asfloat_i(asint(p[i]) + (p[i] < 0.f ? -iOff : iOff));
    float fOff = nn[i] * fScale;
    out[i] = fabs_(p[i]) < origin ? p[i] + fOff : iPos;
  }
  return vec3{out[0], out[1], out[2]};
}

}  // namespace rs
