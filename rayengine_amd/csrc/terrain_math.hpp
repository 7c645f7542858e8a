// Gradient noise, fBm terrain field, bricked volume addressing and trilinear
// sampling, shared by the noise-fill and sphere-tracing kernels.
//
// The noise primitive follows eval / gradientDotV / quintic
// (RayEngine/src/cuda-src/sdf_functions.cuh:25-135) with its exact operation
// order.  The reference hashes through a 2D texture whose contents are
// undefined (a 2 KB table copied as 256 rows of 1 KB); this build defines the
// hash as P[P[P[x] + y] + z] on the doubled 512-entry permutation table that
// RPerlinNoise prepares (RayEngine/src/World/PerlinNoise.h:48-51).
#pragma once

#include "re_math.hpp"

namespace re {

// Brick of 8x8x8 voxels, Morton-ordered inside: one 128-byte cache line holds
// a 4x4x2 block, so the 8 corners of a trilinear cell usually share 1-2 lines
// and a ray marching through a brick stays in 2 KB.
constexpr int kBrick = 8;
constexpr int kBrickVox = kBrick * kBrick * kBrick;

RE_HD uint32_t spread3(uint32_t v) { return (v & 1u) | ((v & 2u) << 2) | ((v & 4u) << 4); }
RE_HD uint32_t compact3(uint32_t m) { return (m & 1u) | ((m >> 2) & 2u) | ((m >> 4) & 4u); }

struct VolView {
  const float* data;
  int32_t dim[3];   // voxels per axis
  int32_t bdim[3];  // bricks per axis
  float spacing[3];
};

RE_HD int64_t vox_index(const int32_t* bdim, int x, int y, int z) {
  int64_t b = ((int64_t)(z >> 3) * bdim[1] + (y >> 3)) * bdim[0] + (x >> 3);
  uint32_t m = spread3((uint32_t)x & 7u) | (spread3((uint32_t)y & 7u) << 1) | (spread3((uint32_t)z & 7u) << 2);
  return b * kBrickVox + m;
}

// 16-case gradient dot product (sdf_functions.cuh:25-48) as s1*A + s2*B with
// A, B in {x, y, z} and s1, s2 = +-1, which rounds identically to each case
// (x - y == x + (-y) and multiplying by +-1 is exact).
RE_HD float grad_dot(uint32_t h, float x, float y, float z) {
  switch (h & 15u) {
    case 0: return x + y;
    case 1: return -x + y;
    case 2: return x - y;
    case 3: return -x - y;
    case 4: return x + z;
    case 5: return -x + z;
    case 6: return x - z;
    case 7: return -x - z;
    case 8: return y + z;
    case 9: return -y + z;
    case 10: return y - z;
    case 11: return -y - z;
    case 12: return y + x;
    case 13: return -x + y;
    case 14: return -y + z;
    default: return -y - z;
  }
}

RE_HD float quintic(float t) { return t * t * t * (t * (t * 6 - 15) + 10); }

// One octave of noise_eval split in two: the cell-relative coordinates and
// fades of p in its lattice cell (fx, fy, fz = floor of p), and the blend of
// the eight corner gradients given the corners' hashes h[0..7] (corner order
// 000, 100, 010, 110, 001, 101, 011, 111 in x, y, z).
struct NoiseCell {
  float x0, x1, y0, y1, z0, z1, u, v, w;
};

RE_HD NoiseCell noise_cell(f3 p, int fx, int fy, int fz) {
  NoiseCell c;
  const float tx = p.x - (float)fx, ty = p.y - (float)fy, tz = p.z - (float)fz;
  c.u = quintic(tx);
  c.v = quintic(ty);
  c.w = quintic(tz);
  c.x0 = tx;
  c.x1 = tx - 1;
  c.y0 = ty;
  c.y1 = ty - 1;
  c.z0 = tz;
  c.z1 = tz - 1;
  return c;
}

RE_HD float noise_blend(const uint32_t* h, const NoiseCell& c) {
  float a = grad_dot(h[0], c.x0, c.y0, c.z0);
  float b = grad_dot(h[1], c.x1, c.y0, c.z0);
  float cc = grad_dot(h[2], c.x0, c.y1, c.z0);
  float d = grad_dot(h[3], c.x1, c.y1, c.z0);
  float e = grad_dot(h[4], c.x0, c.y0, c.z1);
  float f = grad_dot(h[5], c.x1, c.y0, c.z1);
  float g = grad_dot(h[6], c.x0, c.y1, c.z1);
  float hh = grad_dot(h[7], c.x1, c.y1, c.z1);
  float k0 = a;
  float k1 = (b - a);
  float k2 = (cc - a);
  float k3 = (e - a);
  float k4 = (a + d - b - cc);
  float k5 = (a + f - b - e);
  float k6 = (a + g - cc - e);
  float k7 = (b + cc + e + hh - a - d - f - g);
  const float u = c.u, v = c.v, w = c.w;
  return k0 + k1 * u + k2 * v + k3 * w + k4 * u * v + k5 * u * w + k6 * v * w + k7 * u * v * w;
}

// The corners' hashes of lattice cell (fx, fy, fz).
template <typename Table>
RE_HD void noise_hashes(const Table& P, int fx, int fy, int fz, uint32_t* h) {
  const int xi0 = fx & 255, yi0 = fy & 255, zi0 = fz & 255;
  const int xi1 = (xi0 + 1) & 255, yi1 = (yi0 + 1) & 255, zi1 = (zi0 + 1) & 255;
  const uint32_t px0 = P[xi0], px1 = P[xi1];
  const uint32_t p00 = P[px0 + yi0], p10 = P[px1 + yi0], p01 = P[px0 + yi1], p11 = P[px1 + yi1];
  h[0] = P[p00 + zi0];
  h[1] = P[p10 + zi0];
  h[2] = P[p01 + zi0];
  h[3] = P[p11 + zi0];
  h[4] = P[p00 + zi1];
  h[5] = P[p10 + zi1];
  h[6] = P[p01 + zi1];
  h[7] = P[p11 + zi1];
}

// P is the doubled table (512 entries, values 0..255).
template <typename Table>
RE_HD float noise_eval(const Table& P, f3 p) {
  const int fx = (int)floorf(p.x), fy = (int)floorf(p.y), fz = (int)floorf(p.z);
  uint32_t h[8];
  noise_hashes(P, fx, fy, fz, h);
  return noise_blend(h, noise_cell(p, fx, fy, fz));
}

struct FbmParams {
  int32_t octaves;
  float lacunarity, gain, base_freq, y_offset;
};

// Terrain field value: (p.y - y_offset) - sum_o gain^o * noise(p * base_freq *
// lacunarity^o), i.e. minus the density fBm(p) - (p.y - y_offset).  Positive
// above the surface, like the reference's terrain distance
// (RayEngine/src/cuda-src/sdf_functions.cuh:476-503).
// noise(q) is one octave of the noise at q (noise_eval, or an equivalent
// evaluation of it).
template <typename Noise>
RE_HD float fbm_field(const FbmParams& fp, f3 p, const Noise& noise) {
  float sum = 0.0f, amp = 1.0f, freq = fp.base_freq;
  for (int o = 0; o < fp.octaves; ++o) {
    sum = sum + amp * noise(f3{p.x * freq, p.y * freq, p.z * freq});
    freq = freq * fp.lacunarity;
    amp = amp * fp.gain;
  }
  return (p.y - fp.y_offset) - sum;
}

template <typename Table>
RE_HD float terrain_field(const Table& P, const FbmParams& fp, f3 p) {
  return fbm_field(fp, p, [&](f3 q) { return noise_eval(P, q); });
}

// Trilinear sample at world point p: grid coordinate c = p / spacing, cell
// index clamped to [0, dim-2], lerps in x then y then z
// (interpolate / bilinear, RayEngine/src/cuda-src/sdf_functions.cuh:209-265).
RE_HD float trilinear(const VolView& V, f3 p) {
  float gx = p.x / V.spacing[0], gy = p.y / V.spacing[1], gz = p.z / V.spacing[2];
  int ix = (int)gx, iy = (int)gy, iz = (int)gz;
  ix = ix < 0 ? 0 : (ix > V.dim[0] - 2 ? V.dim[0] - 2 : ix);
  iy = iy < 0 ? 0 : (iy > V.dim[1] - 2 ? V.dim[1] - 2 : iy);
  iz = iz < 0 ? 0 : (iz > V.dim[2] - 2 ? V.dim[2] - 2 : iz);
  float tx = gx - (float)ix, ty = gy - (float)iy, tz = gz - (float)iz;
  const float* D = V.data;
  float c000 = D[vox_index(V.bdim, ix, iy, iz)];
  float c100 = D[vox_index(V.bdim, ix + 1, iy, iz)];
  float c010 = D[vox_index(V.bdim, ix, iy + 1, iz)];
  float c110 = D[vox_index(V.bdim, ix + 1, iy + 1, iz)];
  float c001 = D[vox_index(V.bdim, ix, iy, iz + 1)];
  float c101 = D[vox_index(V.bdim, ix + 1, iy, iz + 1)];
  float c011 = D[vox_index(V.bdim, ix, iy + 1, iz + 1)];
  float c111 = D[vox_index(V.bdim, ix + 1, iy + 1, iz + 1)];
  float a0 = c000 * (1.f - tx) + c100 * tx;
  float b0 = c010 * (1.f - tx) + c110 * tx;
  float e = a0 * (1.f - ty) + b0 * ty;
  float a1 = c001 * (1.f - tx) + c101 * tx;
  float b1 = c011 * (1.f - tx) + c111 * tx;
  float f = a1 * (1.f - ty) + b1 * ty;
  return e * (1 - tz) + f * tz;
}

}  // namespace re
