// Portable random streams for synthetic scenes.
//
// PCG32 (O'Neill, "PCG: A Family of Simple Fast Space-Efficient Statistically
// Good Algorithms for Random Number Generation", 2014; pcg32 = XSH-RR 64/32)
// and MT19937 (Matsumoto & Nishimura 1998).  Both produce the same integer
// stream on every platform; floats are derived with exact operations only.
#pragma once

#include <stdint.h>
#include <math.h>

#include "../re_math.hpp"

namespace re {

struct Pcg32 {
  uint64_t state = 0, inc = 1;
  Pcg32(uint64_t initstate, uint64_t initseq) {
    state = 0u;
    inc = (initseq << 1u) | 1u;
    next();
    state += initstate;
    next();
  }
  uint32_t next() {
    uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
  }
  // Uniform in [0, 1) with 24 random bits (exact in FP32).
  float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
  float uniform(float lo, float hi) { return lo + (hi - lo) * uniform(); }
  // Uniform unit vector by rejection from the cube (no transcendentals).
  f3 unit_vector() {
    for (;;) {
      f3 v{2.0f * uniform() - 1.0f, 2.0f * uniform() - 1.0f, 2.0f * uniform() - 1.0f};
      float l2 = dot(v, v);
      if (l2 > 1e-6f && l2 <= 1.0f) return normalize(v);
    }
  }
  // Standard normal, Marsaglia polar method (double precision internally).
  double normal() {
    for (;;) {
      double u = 2.0 * ((double)next() / 4294967296.0) - 1.0;
      double v = 2.0 * ((double)next() / 4294967296.0) - 1.0;
      double s = u * u + v * v;
      if (s > 0.0 && s < 1.0) return u * sqrt(-2.0 * log(s) / s);
    }
  }
};

struct Mt19937 {
  uint32_t mt[624];
  int idx = 624;
  explicit Mt19937(uint32_t seed) {
    mt[0] = seed;
    for (int i = 1; i < 624; ++i) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
  }
  uint32_t next() {
    if (idx >= 624) {
      for (int i = 0; i < 624; ++i) {
        uint32_t y = (mt[i] & 0x80000000u) | (mt[(i + 1) % 624] & 0x7fffffffu);
        uint32_t v = mt[(i + 397) % 624] ^ (y >> 1);
        if (y & 1u) v ^= 0x9908b0dfu;
        mt[i] = v;
      }
      idx = 0;
    }
    uint32_t y = mt[idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
  }
};

}  // namespace re
