// Perlin permutation table (host).
//
// RPerlinNoise (RayEngine/src/World/PerlinNoise.h:16-52) seeds std::mt19937
// and binds it into two distributions with std::bind, which stores a COPY of
// the engine each time.  The swap loop therefore draws outputs 1..256 of a
// freshly seeded engine; uniform_int_distribution<unsigned> over the full
// 32-bit range returns the raw engine output, so swap i exchanges P[i] with
// P[draw_i & 255].  The gradients the constructor also draws are not used by
// the noise primitive and are not produced here.
#include "errors.hpp"
#include "rng.hpp"

namespace re {

void perlin_table(uint32_t seed, uint32_t* perm512) {
  Mt19937 g(seed);
  uint32_t P[256];
  for (uint32_t i = 0; i < 256; ++i) P[i] = i;
  for (uint32_t i = 0; i < 256; ++i) {
    uint32_t j = g.next() & 255u;
    uint32_t tmp = P[i];
    P[i] = P[j];
    P[j] = tmp;
  }
  for (uint32_t i = 0; i < 256; ++i) {
    perm512[i] = P[i];
    perm512[256 + i] = P[i];
  }
}

}  // namespace re

extern "C" RE_API re_status re_perlin_table(uint32_t seed, uint32_t* perm) {
  if (!perm) return re::fail(RE_ERR_INVALID, "re_perlin_table: perm is null");
  re::perlin_table(seed, perm);
  return RE_OK;
}
