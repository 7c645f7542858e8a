// Frame output: the per-pixel tail of render_sphere_trace
// (RayEngine/src/cuda-src/renderer.cu:257-283).  One thread per pixel: sky
// for a miss, the normal-visualisation material for a hit, then gamma,
// vignetting, clipping and the RGBA8 pack.  The sky is the reference's
// single-scattering model (compute_incident_light,
// material_functions.cuh:295-344), after scratchapixel's "simulating the
// colors of the sky", with its constants and its final expression as
// written there (sumR betaR phaseR + sumM + 100 betaM phaseM, times 20).
// Like the reference, built with fast math, the sky's inner loops use the
// hardware exponential, reciprocal and square root (a few ulp each) instead
// of the correctly rounded library sequences; hit pixels use exact
// arithmetic.
#include "../internal.hpp"

namespace re {
namespace {

struct V3 {
  float x, y, z;
};
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float saturate(float x) { return x > 0.0f ? (x < 1.0f ? x : 1.0f) : 0.0f; }

// Hardware forms: e^x as 2^(x log2 e), 1/x, sqrt.
__device__ __forceinline__ float hexp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }
__device__ __forceinline__ float hrcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float hsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float hlen(V3 a) { return hsqrt(dot(a, a)); }

// solve_quadratic / gpu_ray_sphere_intersect (ray_functions.cuh:95-126).
__device__ bool ray_sphere(V3 o, V3 d, float radius, float& t0, float& t1) {
  const float a = dot(d, d);
  const float b = 2 * dot(d, o);
  const float c = dot(o, o) - radius * radius;
  const float discr = b * b - 4.f * a * c;
  if (discr < 0.0f) return false;
  if (discr == 0.0f) {
    t0 = t1 = -0.5f * b * hrcp(a);
  } else {
    const float q = (b > 0.0f) ? -0.5f * (b + hsqrt(discr)) : -0.5f * (b - hsqrt(discr));
    t0 = q * hrcp(a);
    t1 = c * hrcp(q);
  }
  if (t0 > t1) {
    const float t = t0;
    t0 = t1;
    t1 = t;
  }
  return true;
}

__device__ V3 incident_light(const FrameLaunch& L, V3 orig, V3 dir) {
  const V3 sun{L.sun[0], L.sun[1], L.sun[2]};
  const V3 betaR{3.8e-6f, 13.5e-6f, 33.1e-6f};
  const float betaM = 21e-6f;
  float tmin = 0.0f, tmax = kInf;
  float t0, t1;
  if (!ray_sphere(orig, dir, L.ar, t0, t1) || t1 < 0) return {1.0f, 1.0f, 1.0f};
  if (t0 > tmin && t0 > 0) tmin = t0;
  if (t1 < tmax) tmax = t1;
  const int num_samples = 32, num_samples_light = 8;
  const float inv_hr = 1.0f / L.hr, inv_hm = 1.0f / L.hm;
  const float seg = (tmax - tmin) / (float)num_samples;
  float t_cur = tmin;
  V3 sum_r{0, 0, 0}, sum_m{0, 0, 0};
  float od_r = 0, od_m = 0;
  const float mu = dot(dir, sun);
  // the phase constants are formed in double as the reference's M_PI makes them
  const float phase_r = (float)(3.f / (16.f * 3.14159265358979323846) * (double)(1 + mu * mu));
  const float g = 0.76f;
  const float phase_m = (float)(3.f / (8.f * 3.14159265358979323846) * (double)((1.f - g * g) * (1.f + mu * mu)) /
                                (double)((2.f + g * g) * powf(1.f + g * g - 2.f * g * mu, 1.5f)));
  for (int i = 0; i < num_samples; ++i) {
    const V3 sp = orig + (t_cur + seg * 0.5f) * dir;
    const float height = hlen(sp) - L.er;
    const float hr = hexp(-height * inv_hr) * seg;
    const float hm = hexp(-height * inv_hm) * seg;
    od_r += hr;
    od_m += hm;
    float t0l = 0.0f, t1l = 0.0f;
    ray_sphere(sp, sun, L.ar, t0l, t1l);
    const float seg_l = t1l / (float)num_samples_light;
    float t_cur_l = 0;
    float od_lr = 0, od_lm = 0;
    int j;
    for (j = 0; j < num_samples_light; ++j) {
      const V3 spl = sp + (t_cur_l + seg_l * 0.5f) * sun;
      const float hl = hlen(spl) - L.er;
      if (hl < 0) break;
      od_lr += hexp(-hl * inv_hr) * seg_l;
      od_lm += hexp(-hl * inv_hm) * seg_l;
      t_cur_l += seg_l;
    }
    if (j == num_samples_light) {
      const float m = betaM * 1.1f * (od_m + od_lm);
      const V3 tau{betaR.x * (od_r + od_lr) + m, betaR.y * (od_r + od_lr) + m, betaR.z * (od_r + od_lr) + m};
      const V3 att{hexp(-tau.x), hexp(-tau.y), hexp(-tau.z)};
      sum_r = sum_r + V3{att.x * hr, att.y * hr, att.z * hr};
      sum_m = sum_m + V3{att.x * hm, att.y * hm, att.z * hm};
    }
    t_cur += seg;
  }
  const float km = 100 * betaM * phase_m;
  return {(sum_r.x * betaR.x * phase_r + sum_m.x + km) * 20, (sum_r.y * betaR.y * phase_r + sum_m.y + km) * 20,
          (sum_r.z * betaR.z * phase_r + sum_m.z + km) * 20};
}

__device__ __forceinline__ float sky_tone(float c) {
  return c < 1.413f ? powf(c * 0.38317f, 1.0f / 2.2f) : 1.0f - expf(-c);
}

__global__ void __launch_bounds__(256) frame_kernel(FrameLaunch L) {
  const int64_t n = (int64_t)L.width * L.height;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int x = (int)(i % L.width), y = (int)(i / L.width);
  const float u = x / (float)L.width, v = y / (float)L.height;
  const V3 d{L.ray_d[3 * i], L.ray_d[3 * i + 1], L.ray_d[3 * i + 2]};
  V3 c;
  if (L.hit[i]) {
    // simple_shade: mix(colour, facing ratio / 2, 0.5), mix = saturate(a + t b)
    const V3 nrm{L.normal[3 * i], L.normal[3 * i + 1], L.normal[3 * i + 2]};
    const float f = fmaxf(0.f, dot(nrm, V3{-d.x, -d.y, -d.z}) / 2);
    c = {saturate(L.base[0] + 0.5f * f), saturate(L.base[1] + 0.5f * f), saturate(L.base[2] + 0.5f * f)};
  } else {
    const V3 sky = incident_light(L, V3{L.origin[0], L.origin[1], L.origin[2]}, d);
    c = {sky_tone(sky.x), sky_tone(sky.y), sky_tone(sky.z)};
  }
  if (L.gamma) c = {powf(c.x, 1.0f / 2.2f), powf(c.y, 1.0f / 2.2f), powf(c.z, 1.0f / 2.2f)};
  if (L.vignetting) {
    const float w = 0.5f + 0.5f * powf(16.0f * u * v * (1.0f - u) * (1.0f - v), L.vignetting_k);
    c = {c.x * w, c.y * w, c.z * w};
  }
  c = {saturate(c.x), saturate(c.y), saturate(c.z)};
  L.rgba[i] = (255u << 24) | ((uint32_t)(c.z * 255) << 16) | ((uint32_t)(c.y * 255) << 8) | (uint32_t)(c.x * 255);
}

}  // namespace

hipError_t launch_frame(const FrameLaunch& L, hipStream_t s) {
  const int64_t n = (int64_t)L.width * L.height;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(frame_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, L);
  return hipGetLastError();
}

}  // namespace re
