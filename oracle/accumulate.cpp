// CPU ORACLE of temporal accumulation (AccumulatePass) -- test infrastructure only.
//
// Only tests/ load it, as the checker of rs_accum_*.  A scalar restatement, written
// apart from the HIP kernels, of:
//   Source/RenderPasses/AccumulatePass/AccumulatePass.cpp:130-198  execute(): overflow at
//       the frame limit (Stop returns, Reset restarts, EMA continues), reset on a change
//       of frame size, the disabled blit
//   AccumulatePass.cpp:228-290  accumulate(): buffers cleared at count 0, the constants,
//       the frame-count update (not past the limit except in SingleCompensated)
//   AccumulatePass.cpp:390-420  prepareAccumulation(): per-mode buffers
//   Accumulate.cs.slang:82-125  accumulateSingle (running sum times 1/(n+1), or the
//       moving average lerp(last, cur, 1/(n+1)))
//   Accumulate.cs.slang:130-170 accumulateSingleCompensated (Kahan summation, divided by
//       n+1)
//   Accumulate.cs.slang:175-235 accumulateDouble (double sums)
// The lerp is x + s * (y - x) without contraction.  Host images are RGBA32F.
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/rs.h"

namespace {

struct CpuAccum {
  rs_accum_options opt;
  uint32_t frameCount = 0;
  uint32_t W = 0, H = 0;
  std::vector<float> sum, corr;  // 4 per pixel; cleared (empty) when unused
  std::vector<double> sumD;      // 4 per pixel
};

}  // namespace

extern "C" {

int32_t rs_cpu_accum_create(const rs_accum_options* opts, void** out) {
  if (!out) return RS_E_INVALID_ARG;
  CpuAccum* a = new CpuAccum();
  a->opt = rs_accum_options{1, RS_ACCUM_SINGLE, 0, RS_ACCUM_OVERFLOW_STOP};
  if (opts) a->opt = *opts;
  *out = a;
  return RS_OK;
}

void rs_cpu_accum_destroy(void* p) { delete static_cast<CpuAccum*>(p); }

int32_t rs_cpu_accum_set_options(void* p, const rs_accum_options* opts) {
  CpuAccum* a = static_cast<CpuAccum*>(p);
  a->opt = *opts;
  a->frameCount = 0;
  return RS_OK;
}

int32_t rs_cpu_accum_reset(void* p) {
  static_cast<CpuAccum*>(p)->frameCount = 0;
  return RS_OK;
}

uint32_t rs_cpu_accum_frame_count(void* p) { return static_cast<CpuAccum*>(p)->frameCount; }

int32_t rs_cpu_accum_execute(void* p, const float* src, float* dst, uint32_t W, uint32_t H, int shouldAccumulate,
                             int viewAccumulation) {
  CpuAccum* a = static_cast<CpuAccum*>(p);
  const rs_accum_options& o = a->opt;
  if (o.maxFrameCount > 0 && a->frameCount == o.maxFrameCount) {
    if (o.overflowMode == RS_ACCUM_OVERFLOW_STOP) return RS_OK;
    if (o.overflowMode == RS_ACCUM_OVERFLOW_RESET) a->frameCount = 0;
  }
  if (W != a->W || H != a->H) {
    a->W = W;
    a->H = H;
    a->frameCount = 0;
  }
  const size_t n = (size_t)W * H * 4;
  if (!o.enabled) {
    std::memcpy(dst, src, n * sizeof(float));
    return RS_OK;
  }
  // prepareAccumulation: (re)create the mode's buffers, reset on creation, clear at 0.
  auto prep = [&](auto& v, bool used) {
    if (!used) {
      v.clear();
      return;
    }
    if (v.size() != n) {
      v.assign(n, 0);
      a->frameCount = 0;
    }
  };
  prep(a->sum, o.precision != RS_ACCUM_DOUBLE);
  prep(a->corr, o.precision == RS_ACCUM_SINGLE_COMPENSATED);
  prep(a->sumD, o.precision == RS_ACCUM_DOUBLE);
  if (a->frameCount == 0) {
    for (auto& x : a->sum) x = 0.f;
    for (auto& x : a->corr) x = 0.f;
    for (auto& x : a->sumD) x = 0.0;
  }
  const uint32_t count = a->frameCount;
  const bool movingAverage = o.maxFrameCount > 0;
  if (o.maxFrameCount == 0 || o.precision == RS_ACCUM_SINGLE_COMPENSATED || a->frameCount < o.maxFrameCount)
    a->frameCount += shouldAccumulate ? 1u : 0u;
  for (size_t i = 0; i < n; i++) {
    const float cur = src[i];
    float out;
    if (o.precision == RS_ACCUM_SINGLE) {
      if (shouldAccumulate) {
        const float w = 1.0f / (float)(count + 1);
        if (movingAverage) {
          const float last = a->sum[i];
          out = last + (cur - last) * w;
          a->sum[i] = out;
        } else {
          const float s = a->sum[i] + cur;
          out = s * w;
          a->sum[i] = s;
        }
      } else if (viewAccumulation) {
        out = a->sum[i] / (float)count;
      } else {
        out = cur;
      }
    } else if (o.precision == RS_ACCUM_SINGLE_COMPENSATED) {
      if (shouldAccumulate) {
        const float s = a->sum[i], c = a->corr[i];
        const float y = cur - c;
        const float sn = s + y;
        out = sn / (float)(count + 1);
        a->sum[i] = sn;
        a->corr[i] = (sn - s) - y;
      } else if (viewAccumulation) {
        out = a->sum[i] / (float)count;
      } else {
        out = cur;
      }
    } else {
      if (shouldAccumulate) {
        const double w = 1.0 / (double)(count + 1);
        double s = a->sumD[i];
        if (movingAverage) {
          s = s * (1.0 - w) + (double)cur * w;
          out = (float)s;
        } else {
          s += (double)cur;
          out = (float)(s * w);
        }
        a->sumD[i] = s;
      } else if (viewAccumulation) {
        out = (float)(a->sumD[i] * (1.0 / (double)count));
      } else {
        out = cur;
      }
    }
    dst[i] = out;
  }
  return RS_OK;
}

}  // extern "C"
