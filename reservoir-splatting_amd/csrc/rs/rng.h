// Per-pixel pseudorandom streams: TEA-seeded 32-bit LCG (TinyUniformSampleGenerator).
//
// Follows Source/Falcor/Utils/Sampling/TinyUniformSampleGenerator.slang:51-69,
// Source/Falcor/Utils/Math/HashUtils.slang:62-74 (TEA, 16 rounds, fixed key),
// Source/Falcor/Utils/Math/BitTricks.slang:47-62 (32-bit Morton interleave),
// Source/Falcor/Utils/Sampling/Pseudorandom/LCG.slang:50-67 (a = 1664525, c = 1013904223)
// and Source/Falcor/Utils/Sampling/SampleGeneratorInterface.slang:50-64 (upper 24 bits).
#pragma once
#include "common.h"

namespace rs {

RS_HD uint32_t interleave_32bit(uint32_t vx, uint32_t vy) {
  uint32_t x = vx & 0xffffu;
  x = (x | (x << 8)) & 0x00FF00FFu;
  x = (x | (x << 4)) & 0x0F0F0F0Fu;
  x = (x | (x << 2)) & 0x33333333u;
  x = (x | (x << 1)) & 0x55555555u;
  uint32_t y = vy & 0xffffu;
  y = (y | (y << 8)) & 0x00FF00FFu;
  y = (y | (y << 4)) & 0x0F0F0F0Fu;
  y = (y | (y << 2)) & 0x33333333u;
  y = (y | (y << 1)) & 0x55555555u;
  return x | (y << 1);
}

// TEA block cipher (Wheeler & Needham 1994) with an explicit 128-bit key.
RS_HD uvec2 tea_key(uint32_t v0, uint32_t v1, const uint32_t* k, uint32_t iterations) {
  uint32_t sum = 0;
  const uint32_t delta = 0x9e3779b9u;
  for (uint32_t i = 0; i < iterations; i++) {
    sum += delta;
    v0 += ((v1 << 4) + k[0]) ^ (v1 + sum) ^ ((v1 >> 5) + k[1]);
    v1 += ((v0 << 4) + k[2]) ^ (v0 + sum) ^ ((v0 >> 5) + k[3]);
  }
  return uvec2{v0, v1};
}

// blockCipherTEA with Falcor's fixed key and 16 rounds.
RS_HD uvec2 tea(uint32_t v0, uint32_t v1, uint32_t iterations = 16) {
  const uint32_t k[4] = {0xa341316cu, 0xc8013ea4u, 0xad90777du, 0x7e95761eu};
  return tea_key(v0, v1, k, iterations);
}

struct Rng {
  uint32_t state;
};

RS_HD Rng rng_from_seed(uint32_t s) { return Rng{s}; }
// SampleGenerator(uint2 pixel, uint sampleNumber).
RS_HD Rng rng_make(uint32_t px, uint32_t py, uint32_t n) {
  return Rng{tea(interleave_32bit(px, py), n).x};
}
RS_HD uint32_t rng_next(Rng& r) {
  r.state = 1664525u * r.state + 1013904223u;
  return r.state;
}
// The generator after n more draws.
RS_HD Rng rng_skip(Rng r, uint32_t n) {
  for (uint32_t i = 0; i < n; i++) rng_next(r);
  return r;
}
RS_HD float next1(Rng& r) { return (float)(rng_next(r) >> 8) * 5.9604644775390625e-8f; }
RS_HD vec2 next2(Rng& r) {
  float a = next1(r);
  float b = next1(r);
  return vec2{a, b};
}

}  // namespace rs
