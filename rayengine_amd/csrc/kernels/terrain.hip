// Noise-volume generation and sphere tracing of the terrain field on gfx950.
//
// noise_fill_kernel: one work-group per 8x8x8 brick (256 lanes, 2 voxels
// each), the 512-entry permutation table staged in LDS, every lane writing
// its voxel's field value at its Morton slot so a wave's stores are one
// contiguous 256-byte run.  Replaces generate_noise
// (RayEngine/src/cuda-src/renderer.cu:105-126), which used 8-thread blocks
// and a linear float2 grid.
//
// sdf_trace_kernel: one ray per lane.  Clip to the volume box, then the
// reference march (single_ray_sphere_trace,
// RayEngine/src/cuda-src/sphere_tracing.cuh:182-216): sample, step
// t += d * scale, then test d <= eps * t on the stepped t, at most max_steps
// samples.  Normals by forward differences of the sampled field with
// delta = min(0.001 t, 0.002) (compute_sdf_normal,
// RayEngine/src/cuda-src/sdf_functions.cuh:506-516).
#include "../device_paths.hpp"
#include "../terrain_math.hpp"

namespace re {
namespace {

struct LdsTable {
  const uint16_t* t;
  __device__ uint32_t operator[](uint32_t i) const { return t[i]; }
};

__device__ __forceinline__ uint32_t wave_uniform(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

// One octave of noise_eval for the active lanes of a wave.  When they all lie
// in one lattice cell, the cell's corner hashes are looked up once and kept in
// scalar registers, so the eight gradient selections are scalar branches
// taken by the whole wave; otherwise every lane looks up its own.  Either way
// each lane's arithmetic is noise_eval's.  A wave covers a 4x4x4 voxel block;
// on the 256^3 unit-cube terrain the finest octave's cells are four voxels
// wide and aligned with those blocks, so every wave of it takes the first
// path.
__device__ __forceinline__ float noise_eval_wave(const LdsTable& P, f3 p) {
  const int fx = (int)floorf(p.x), fy = (int)floorf(p.y), fz = (int)floorf(p.z);
  const int sx = __builtin_amdgcn_readfirstlane(fx), sy = __builtin_amdgcn_readfirstlane(fy),
            sz = __builtin_amdgcn_readfirstlane(fz);
  const NoiseCell c = noise_cell(p, fx, fy, fz);
  uint32_t h[8];
  if (__builtin_amdgcn_ballot_w64(fx != sx || fy != sy || fz != sz) == 0ull) {
    noise_hashes(P, sx, sy, sz, h);
    for (int k = 0; k < 8; ++k) h[k] = wave_uniform(h[k]);
    return noise_blend(h, c);
  }
  noise_hashes(P, fx, fy, fz, h);
  return noise_blend(h, c);
}

__global__ void __launch_bounds__(256) noise_fill_kernel(float* __restrict__ out, const uint32_t* __restrict__ perm,
                                                         int32_t dx, int32_t dy, int32_t dz, int32_t bx, int32_t by,
                                                         float sx, float sy, float sz, float ox, float oy,
                                                         float oz, FbmParams fp) {
  __shared__ uint16_t P[512];
  for (int i = threadIdx.x; i < 512; i += blockDim.x) P[i] = (uint16_t)perm[i];
  __syncthreads();
  const int64_t brick = blockIdx.x;
  const int gx = (int)(brick % bx), gy = (int)((brick / bx) % by), gz = (int)(brick / ((int64_t)bx * by));
  LdsTable T{P};
  for (int m = threadIdx.x; m < kBrickVox; m += blockDim.x) {
    const int x = gx * kBrick + (int)compact3((uint32_t)m);
    const int y = gy * kBrick + (int)compact3((uint32_t)m >> 1);
    const int z = gz * kBrick + (int)compact3((uint32_t)m >> 2);
    float v = 0.0f;
    if (x < dx && y < dy && z < dz) {
      const f3 p{ox + (float)x * sx, oy + (float)y * sy, oz + (float)z * sz};
      v = fbm_field(fp, p, [&](f3 q) { return noise_eval_wave(T, q); });
    }
    out[brick * kBrickVox + m] = v;
  }
}

template <bool NORMALS>
__global__ void __launch_bounds__(256) sdf_trace_kernel(VolView V, float bx0, float by0, float bz0, float bx1,
                                                        float by1, float bz1, const float* __restrict__ ray_o,
                                                        const float* __restrict__ ray_d, int64_t n, int32_t max_steps,
                                                        float eps, float step_scale, int32_t step_rule,
                                                        float* __restrict__ t_out, uint8_t* __restrict__ hit_out,
                                                        int32_t* __restrict__ steps_out, float* __restrict__ nrm_out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const f3 o{ray_o[3 * i], ray_o[3 * i + 1], ray_o[3 * i + 2]};
  const f3 d{ray_d[3 * i], ray_d[3 * i + 1], ray_d[3 * i + 2]};
  const float bmin[3] = {bx0, by0, bz0}, bmax[3] = {bx1, by1, bz1};
  int32_t steps = 0;
  float t = kInf;
  const bool hit = plain_march(V, bmin, bmax, o, d, max_steps, eps, step_scale, step_rule, t, steps);
  t_out[i] = hit ? t : kInf;
  hit_out[i] = hit ? 1 : 0;
  if (steps_out) steps_out[i] = steps;
  if (NORMALS) {
    f3 nrm{0.0f, 0.0f, 0.0f};
    if (hit) {
      const float delta = fminf((float)(0.001 * (double)t), 0.002f);
      const f3 p = add(o, smul(t, d));
      const float c = trilinear(V, p);
      nrm = normalize(f3{trilinear(V, add(p, f3{delta, 0.0f, 0.0f})) - c,
                         trilinear(V, add(p, f3{0.0f, delta, 0.0f})) - c,
                         trilinear(V, add(p, f3{0.0f, 0.0f, delta})) - c});
    }
    nrm_out[3 * i] = nrm.x;
    nrm_out[3 * i + 1] = nrm.y;
    nrm_out[3 * i + 2] = nrm.z;
  }
}

// Brush distances (sphere_distance / aabb_distance,
// RayEngine/src/cuda-src/sdf_functions.cuh:422-437): |p| - r, and the box
// distance |max(|p| - s, 0)| + min(max(d.x, d.y, d.z), 0).
__device__ __forceinline__ float vlen(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }

__device__ __forceinline__ float box_distance(float px, float py, float pz, const float* s) {
  const float dx = fabsf(px) - s[0], dy = fabsf(py) - s[1], dz = fabsf(pz) - s[2];
  return vlen(fmaxf(dx, 0.0f), fmaxf(dy, 0.0f), fmaxf(dz, 0.0f)) + fminf(fmaxf(dx, fmaxf(dy, dz)), 0.0f);
}

// One brush applied to every voxel (insert_sphere_to_texture,
// RayEngine/src/cuda-src/renderer.cu:30-92): offset p = voxel point - brush
// point, optionally snapped in x and z; then, on (distance, material) pairs,
// SPHERE_ADD: sdf_smin (h = max(k - |a - b|, 0) / k, the pair of the smaller
// distance with h^2 k / 4 taken off its distance), SPHERE_SUBTRACT and
// CUBE_SUBTRACT: the pair of the larger distance against the negated brush
// distance (sdf_fmaxf), CUBE_ADD: the componentwise minimum (fminf on
// float2) -- sdf_functions.cuh:270-289.  One work-group per brick, as in the
// noise fill; the kernel streams the volume once (8 bytes in and out per
// voxel).
__global__ void __launch_bounds__(256) volume_edit_kernel(EditLaunch L) {
  const int64_t brick = blockIdx.x;
  const int gx = (int)(brick % L.bdim[0]), gy = (int)((brick / L.bdim[0]) % L.bdim[1]);
  const int gz = (int)(brick / ((int64_t)L.bdim[0] * L.bdim[1]));
  for (int m = threadIdx.x; m < kBrickVox; m += blockDim.x) {
    const int x = gx * kBrick + (int)compact3((uint32_t)m);
    const int y = gy * kBrick + (int)compact3((uint32_t)m >> 1);
    const int z = gz * kBrick + (int)compact3((uint32_t)m >> 2);
    if (x >= L.dim[0] || y >= L.dim[1] || z >= L.dim[2]) continue;
    const int64_t i = brick * kBrickVox + m;
    float px = (float)x * L.spacing[0] - L.center[0];
    const float py = (float)y * L.spacing[1] - L.center[1];
    float pz = (float)z * L.spacing[2] - L.center[2];
    if (L.snap) {
      const float snap = L.extent[0];
      px = floorf(px - fabsf(fmodf(px, snap)));
      pz = floorf(pz - fabsf(fmodf(pz, snap)));
    }
    const float ad = L.data[i], am = L.mat[i];
    float rd = ad, rm = am;
    if (L.type == 0) {
      const float bd = vlen(px, py, pz) - L.extent[0];
      const float h = fmaxf(L.k - fabsf(ad - bd), 0.0f) / L.k;
      const bool a_min = ad < bd ? true : false;
      rd = (a_min ? ad : bd) - h * h * L.k * 0.25f;
      rm = a_min ? am : L.material;
    } else if (L.type == 1) {
      const float bd = -(vlen(px, py, pz) - L.extent[0]);
      if (!(ad > bd)) {
        rd = bd;
        rm = L.material;
      }
    } else if (L.type == 2) {
      const float bd = box_distance(px, py, pz, L.extent);
      rd = fminf(ad, bd);
      rm = fminf(am, L.material);
    } else {
      const float bd = -box_distance(px, py, pz, L.extent);
      if (!(ad > bd)) {
        rd = bd;
        rm = L.material;
      }
    }
    L.data[i] = rd;
    L.mat[i] = rm;
  }
}

}  // namespace

hipError_t launch_volume_edit(const EditLaunch& L, hipStream_t s) {
  const int64_t bricks = (int64_t)L.bdim[0] * L.bdim[1] * L.bdim[2];
  if (bricks <= 0) return hipSuccess;
  hipLaunchKernelGGL(volume_edit_kernel, dim3((unsigned)bricks), dim3(256), 0, s, L);
  return hipGetLastError();
}

hipError_t launch_noise_fill(const NoiseLaunch& L, hipStream_t s) {
  const int64_t bricks = (int64_t)L.bdim[0] * L.bdim[1] * L.bdim[2];
  hipLaunchKernelGGL(noise_fill_kernel, dim3((unsigned)bricks), dim3(256), 0, s, L.out, L.perm, L.dim[0], L.dim[1],
                     L.dim[2], L.bdim[0], L.bdim[1], L.spacing[0], L.spacing[1], L.spacing[2], L.origin[0], L.origin[1],
                     L.origin[2], L.fbm);
  return hipGetLastError();
}

hipError_t launch_sdf_trace(const SdfLaunch& L, hipStream_t s) {
  if (L.n <= 0) return hipSuccess;
  const int block = L.block > 0 ? L.block : 256;
  const int64_t grid = (L.n + block - 1) / block;
#define RE_LAUNCH(NRM)                                                                                        \
  hipLaunchKernelGGL((sdf_trace_kernel<NRM>), dim3((unsigned)grid), dim3(block), 0, s, L.vol, L.box_min[0],  \
                     L.box_min[1], L.box_min[2], L.box_max[0], L.box_max[1], L.box_max[2], L.ray_o, L.ray_d, \
                     L.n, L.max_steps, L.eps, L.step_scale, L.step_rule, L.t, L.hit, L.steps, L.nrm)
  if (L.nrm) RE_LAUNCH(true); else RE_LAUNCH(false);
#undef RE_LAUNCH
  return hipGetLastError();
}

}  // namespace re
