// Temporal accumulation kernels: Source/RenderPasses/AccumulatePass/Accumulate.cs.slang
// accumulateSingle (:82-125), accumulateSingleCompensated (:130-170) and accumulateDouble
// (:175-235), one lane per pixel over RGBA32F images.
//
// The running sums live in HBM next to the images: float4 per pixel (Single), float4 sum +
// float4 correction (SingleCompensated) or double4 (Double).  The reference clears them
// when the count is 0 before the dispatch; here a count of 0 reads them as zero instead,
// which skips the clear pass and gives the same values.  The kernels are streaming: 64 B
// per pixel (Single) or 96 B (SingleCompensated, Double) of HBM traffic.
#pragma once
#include "../rs/common.h"

namespace rs {

struct AccumArgs {
  const float4* src;
  float4* dst;
  float4* sum;    // Single, SingleCompensated
  float4* corr;   // SingleCompensated
  double2* sumD;  // Double: two double2 per pixel
  uint32_t pixels;
  uint32_t count;          // frames accumulated before this one (gAccumCount)
  uint32_t accumulate;     // gAccumulate
  uint32_t view;           // gViewAccumulation
  uint32_t movingAverage;  // gMovingAverageMode (maxFrameCount > 0)
};

#if defined(__HIPCC__)

__device__ __forceinline__ float4 f4_zero() { return make_float4(0.f, 0.f, 0.f, 0.f); }
__device__ __forceinline__ float4 f4_add(float4 a, float4 b) { return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
__device__ __forceinline__ float4 f4_sub(float4 a, float4 b) { return make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
__device__ __forceinline__ float4 f4_mul(float4 a, float s) { return make_float4(a.x * s, a.y * s, a.z * s, a.w * s); }
// Component-wise IEEE division (not a reciprocal multiply).
__device__ __forceinline__ float4 f4_div(float4 a, float s) { return make_float4(a.x / s, a.y / s, a.z / s, a.w / s); }
// lerp(x, y, s) = x + s * (y - x), without fused multiply-add.
__device__ __forceinline__ float4 f4_lerp(float4 x, float4 y, float s) { return f4_add(x, f4_mul(f4_sub(y, x), s)); }

__global__ __launch_bounds__(256) void k_accum_single(AccumArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.pixels) return;
  const float4 cur = a.src[i];
  float4 out;
  if (a.accumulate) {
    const float w = 1.f / (float)(a.count + 1u);
    const float4 last = a.count == 0 ? f4_zero() : a.sum[i];
    if (a.movingAverage) {
      out = f4_lerp(last, cur, w);
      a.sum[i] = out;
    } else {
      const float4 s = f4_add(last, cur);
      out = f4_mul(s, w);
      a.sum[i] = s;
    }
  } else if (a.view) {
    out = f4_div(a.count == 0 ? f4_zero() : a.sum[i], (float)a.count);
  } else {
    out = cur;
  }
  a.dst[i] = out;
}

__global__ __launch_bounds__(256) void k_accum_compensated(AccumArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.pixels) return;
  const float4 cur = a.src[i];
  float4 out;
  if (a.accumulate) {
    const float4 s = a.count == 0 ? f4_zero() : a.sum[i];
    const float4 c = a.count == 0 ? f4_zero() : a.corr[i];
    const float4 y = f4_sub(cur, c);
    const float4 sn = f4_add(s, y);
    out = f4_div(sn, (float)(a.count + 1u));
    a.sum[i] = sn;
    a.corr[i] = f4_sub(f4_sub(sn, s), y);
  } else if (a.view) {
    out = f4_div(a.count == 0 ? f4_zero() : a.sum[i], (float)a.count);
  } else {
    out = cur;
  }
  a.dst[i] = out;
}

__global__ __launch_bounds__(256) void k_accum_double(AccumArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.pixels) return;
  const float4 cur = a.src[i];
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  if (a.count != 0 && (a.accumulate || a.view)) {
    const double2 lo = a.sumD[2 * i], hi = a.sumD[2 * i + 1];
    s[0] = lo.x;
    s[1] = lo.y;
    s[2] = hi.x;
    s[3] = hi.y;
  }
  const double c[4] = {(double)cur.x, (double)cur.y, (double)cur.z, (double)cur.w};
  float o[4];
  if (a.accumulate) {
    const double w = 1.0 / (double)(a.count + 1u);
    for (int k = 0; k < 4; k++) {
      if (a.movingAverage) {
        s[k] = s[k] * (1.0 - w) + c[k] * w;
        o[k] = (float)s[k];
      } else {
        s[k] += c[k];
        o[k] = (float)(s[k] * w);
      }
    }
    a.sumD[2 * i] = make_double2(s[0], s[1]);
    a.sumD[2 * i + 1] = make_double2(s[2], s[3]);
  } else if (a.view) {
    const double w = 1.0 / (double)a.count;
    for (int k = 0; k < 4; k++) o[k] = (float)(s[k] * w);
  } else {
    o[0] = cur.x;
    o[1] = cur.y;
    o[2] = cur.z;
    o[3] = cur.w;
  }
  a.dst[i] = make_float4(o[0], o[1], o[2], o[3]);
}

#endif

}  // namespace rs
