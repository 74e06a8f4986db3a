// Deterministic fp32 transcendentals for host and device.
//
// The pass evaluates sin/cos (disk and hemisphere sampling, GGX VNDF, quaternion
// slerp), acos/atan/tan (CameraManager, Source/RenderPasses/ReservoirSplatting/
// CameraManager.slang:50-82) and pow (PowerExp MIS, RS/PathTracer.slang:501-502).
// Vendor libm results differ between glibc and the GPU's device library in the last
// ulp, which would break CPU/GPU bit-exactness, so these are computed here from
// first principles with a fixed operation order:
//   * sin/cos: Cody-Waite reduction by pi/2 with a three-part split constant, then
//     truncated Taylor series on [-pi/4, pi/4] (degree 11 for sin, 12 for cos).
//   * atan: reduction to |t| <= tan(pi/8) via atan(x) = pi/2 - atan(1/x) and
//     atan(x) = pi/4 + atan((x-1)/(x+1)), then the Taylor series of atan to t^21.
//   * acos(x) = 2*atan(sqrt((1-x)/(1+x))).
//   * tan = sin/cos.
//   * log2 via frexp-style exponent split and the atanh series
//     ln(m) = 2*atanh((m-1)/(m+1)); exp2 via integer split and the Taylor series of e^(x ln2).
// All series coefficients are exact rational numbers 1/n! or 1/(2k+1) rounded to
// fp32; accuracy is a few ulp, which the tests check against libm in double.
#pragma once
#include "common.h"

namespace rs {

// pi/2 split into three parts whose products with small integers are exact.
static constexpr float kPio2Hi = 1.5703125f;               // 8 significant bits
static constexpr float kPio2Mid = 4.83751296997070312e-4f;  // next bits
static constexpr float kPio2Lo = 7.54978995489188216e-8f;   // remainder

RS_HD float sin_poly(float r) {
  // r in [-pi/4, pi/4]; Horner in r^2.
  float r2 = r * r;
  float p = -2.5052108385441718775e-08f;        // -1/11!
  p = p * r2 + 2.7557319223985890653e-06f;      // 1/9!
  p = p * r2 + -1.9841269841269841270e-04f;     // -1/7!
  p = p * r2 + 8.3333333333333333333e-03f;      // 1/5!
  p = p * r2 + -1.6666666666666666667e-01f;     // -1/3!
  return r + r * (r2 * p);
}

RS_HD float cos_poly(float r) {
  float r2 = r * r;
  float p = 2.0876756987868098979e-09f;         // 1/12!
  p = p * r2 + -2.7557319223985890653e-07f;     // -1/10!
  p = p * r2 + 2.4801587301587301587e-05f;      // 1/8!
  p = p * r2 + -1.3888888888888888889e-03f;     // -1/6!
  p = p * r2 + 4.1666666666666666667e-02f;      // 1/4!
  p = p * r2 + -0.5f;                           // -1/2!
  return 1.f + r2 * p;
}

// Reduce x to r in [-pi/4, pi/4] with x = q*pi/2 + r.  Valid for |x| < ~1e5.
RS_HD float reduce_pio2(float x, int& q) {
  float qf = fround(x * 0.636619772367581343076f);
  q = (int)qf;
  float r = x - qf * kPio2Hi;
  r = r - qf * kPio2Mid;
  r = r - qf * kPio2Lo;
  return r;
}

RS_HD float fsin(float x) {
  int q;
  float r = reduce_pio2(x, q);
  float s = sin_poly(r), c = cos_poly(r);
  switch (q & 3) {
    case 0: return s;
    case 1: return c;
    case 2: return -s;
    default: return -c;
  }
}

RS_HD float fcos(float x) {
  int q;
  float r = reduce_pio2(x, q);
  float s = sin_poly(r), c = cos_poly(r);
  switch (q & 3) {
    case 0: return c;
    case 1: return -s;
    case 2: return -c;
    default: return s;
  }
}

RS_HD void fsincos(float x, float& s_out, float& c_out) {
  int q;
  float r = reduce_pio2(x, q);
  float s = sin_poly(r), c = cos_poly(r);
  switch (q & 3) {
    case 0: s_out = s; c_out = c; break;
    case 1: s_out = c; c_out = -s; break;
    case 2: s_out = -s; c_out = -c; break;
    default: s_out = -c; c_out = s; break;
  }
}

RS_HD float ftan(float x) {
  float s, c;
  fsincos(x, s, c);
  return s / c;
}

// atan on |t| <= tan(pi/8) ~ 0.4142 by the alternating Taylor series to t^21.
RS_HD float atan_poly(float t) {
  float t2 = t * t;
  float p = 1.f / 21.f;
  p = p * t2 - 1.f / 19.f;
  p = p * t2 + 1.f / 17.f;
  p = p * t2 - 1.f / 15.f;
  p = p * t2 + 1.f / 13.f;
  p = p * t2 - 1.f / 11.f;
  p = p * t2 + 1.f / 9.f;
  p = p * t2 - 1.f / 7.f;
  p = p * t2 + 1.f / 5.f;
  p = p * t2 - 1.f / 3.f;
  return t + t * (t2 * p);
}

RS_HD float fatan(float x) {
  float ax = fabs_(x);
  float base = 0.f;
  bool inv = ax > 1.f;
  float t = inv ? 1.f / ax : ax;
  float res;
  if (t > 0.41421356237309504880f) {
    res = 0.785398163397448309616f + atan_poly((t - 1.f) / (t + 1.f));
  } else {
    res = atan_poly(t);
  }
  if (inv) res = 1.57079632679489661923f - res;
  res = res + base;
  return x < 0.f ? -res : res;
}

RS_HD float facos(float x) {
  if (x >= 1.f) return 0.f;
  if (x <= -1.f) return kPi;
  return 2.f * fatan(fsqrt((1.f - x) / (1.f + x)));
}

// log2 for x > 0 finite.
RS_HD float flog2(float x) {
  if (!(x > 0.f)) return x == 0.f ? -__builtin_inff() : __builtin_nanf("");
  if (x == __builtin_inff()) return x;
  uint32_t u = asuint(x);
  int e;
  if ((u >> 23) == 0) {  // subnormal: scale by 2^23
    x = x * 8388608.f;
    u = asuint(x);
    e = (int)(u >> 23) - 127 - 23;
  } else {
    e = (int)(u >> 23) - 127;
  }
  float m = asfloat((u & 0x7fffffu) | 0x3f800000u);  // [1,2)
  if (m > 1.41421356237309504880f) {
    m = m * 0.5f;
    e += 1;
  }
  // ln(m) = 2 atanh(s), s = (m-1)/(m+1), |s| <= 0.1716
  float s = (m - 1.f) / (m + 1.f);
  float s2 = s * s;
  float p = 1.f / 13.f;
  p = p * s2 + 1.f / 11.f;
  p = p * s2 + 1.f / 9.f;
  p = p * s2 + 1.f / 7.f;
  p = p * s2 + 1.f / 5.f;
  p = p * s2 + 1.f / 3.f;
  float ln = 2.f * (s + s * (s2 * p));
  return (float)e + ln * 1.44269504088896340736f;
}

RS_HD float fexp2(float x) {
  if (x != x) return x;
  if (x >= 128.f) return __builtin_inff();
  if (x < -150.f) return 0.f;
  float n = ffloor(x + 0.5f);
  float f = x - n;  // [-0.5, 0.5]
  float y = f * 0.693147180559945309417f;
  // e^y, |y| <= 0.347, Taylor to y^8
  float p = 1.f / 40320.f;
  p = p * y + 1.f / 5040.f;
  p = p * y + 1.f / 720.f;
  p = p * y + 1.f / 120.f;
  p = p * y + 1.f / 24.f;
  p = p * y + 1.f / 6.f;
  p = p * y + 0.5f;
  p = p * y + 1.f;
  p = p * y + 1.f;
  int ni = (int)n;
  // scale by 2^ni in two steps to cover the subnormal range
  int n1 = ni / 2, n2 = ni - n1;
  float s1 = asfloat((uint32_t)(n1 + 127) << 23);
  float s2 = asfloat((uint32_t)(n2 + 127) << 23);
  return p * s1 * s2;
}

RS_HD float fpow(float x, float y) {
  if (y == 0.f) return 1.f;
  if (x == 0.f) return y > 0.f ? 0.f : __builtin_inff();
  if (x < 0.f) return __builtin_nanf("");
  return fexp2(y * flog2(x));
}

}  // namespace rs
