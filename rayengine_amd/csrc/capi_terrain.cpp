// C-ABI entry points of the terrain path: noise-volume generation, volume
// upload/readback, .rsdf files, brush edits, and sphere tracing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "capi_util.hpp"
#include "host/rsdf.hpp"

using namespace re;

namespace {

bool valid_dims(const int32_t* dim, const float* spacing) { return valid_volume_dims(dim, spacing); }

re_status alloc_volume(re_ctx* ctx, const int32_t* dim, const float* spacing, DevVolume& v) {
  for (int a = 0; a < 3; ++a) {
    v.dim[a] = dim[a];
    v.bdim[a] = (dim[a] + kBrick - 1) / kBrick;
    v.spacing[a] = spacing[a];
    v.box_min[a] = 0.0f;
    // The reference volume box spans dim * spacing (renderer.cu:115).
    v.box_max[a] = (float)dim[a] * spacing[a];
  }
  v.bytes = (int64_t)v.bdim[0] * v.bdim[1] * v.bdim[2] * kBrickVox * (int64_t)sizeof(float);
  HIP_TRY(ctx, hipMalloc((void**)&v.data, (size_t)v.bytes), "volume hipMalloc");
  return RE_OK;
}

// Linear x-fastest <-> bricked (Morton inside 8^3 bricks) copies.
void to_bricked(const DevVolume& v, const float* linear, std::vector<float>& bricked) {
  bricked.assign((size_t)(v.bytes / 4), 0.0f);
  for (int z = 0; z < v.dim[2]; ++z)
    for (int y = 0; y < v.dim[1]; ++y)
      for (int x = 0; x < v.dim[0]; ++x)
        bricked[vox_index(v.bdim, x, y, z)] = linear[((int64_t)z * v.dim[1] + y) * v.dim[0] + x];
}

void to_linear(const DevVolume& v, const std::vector<float>& bricked, float* linear) {
  for (int z = 0; z < v.dim[2]; ++z)
    for (int y = 0; y < v.dim[1]; ++y)
      for (int x = 0; x < v.dim[0]; ++x)
        linear[((int64_t)z * v.dim[1] + y) * v.dim[0] + x] = bricked[vox_index(v.bdim, x, y, z)];
}

re_status read_channel(re_ctx* ctx, const DevVolume& v, const float* src, float* linear) {
  const int64_t n = (int64_t)v.dim[0] * v.dim[1] * v.dim[2];
  if (!src) {
    std::fill(linear, linear + n, 0.0f);
    return RE_OK;
  }
  std::vector<float> bricked((size_t)(v.bytes / 4));
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  HIP_TRY(ctx, hipMemcpyAsync(bricked.data(), src, (size_t)v.bytes, hipMemcpyDeviceToHost, ctx->stream), "volume read");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "volume read sync");
  to_linear(v, bricked, linear);
  return RE_OK;
}

// Gives a volume a zero material channel if it has none.
re_status ensure_material(re_ctx* ctx, DevVolume& v) {
  if (v.mat) return RE_OK;
  HIP_TRY(ctx, hipMalloc((void**)&v.mat, (size_t)v.bytes), "material hipMalloc");
  HIP_TRY(ctx, hipMemsetAsync(v.mat, 0, (size_t)v.bytes, ctx->stream), "material clear");
  return RE_OK;
}

}  // namespace

extern "C" {

RE_API void re_noise_default_params(re_noise_params* p) {
  if (!p) return;
  for (int a = 0; a < 3; ++a) {
    p->dim[a] = 256;
    p->spacing[a] = 1.0f / 256.0f;
  }
  p->seed = 2016;
  p->octaves = 5;
  p->lacunarity = 2.0f;
  p->gain = 0.5f;
  p->base_freq = 4.0f;
  p->y_offset = 0.4f;
  for (float& o : p->origin) o = 0.0f;
}

RE_API re_status re_noise_generate(re_ctx* ctx, const re_noise_params* params, int32_t* vol_id) {
  if (!ctx || !vol_id) return fail(RE_ERR_INVALID, "re_noise_generate: null argument");
  re_noise_params p;
  re_noise_default_params(&p);
  if (params) p = *params;
  if (!valid_dims(p.dim, p.spacing) || p.octaves < 0 || p.octaves > 16 || !std::isfinite(p.origin[0]) ||
      !std::isfinite(p.origin[1]) || !std::isfinite(p.origin[2]))
    return set_err(ctx, RE_ERR_INVALID,
                   "re_noise_generate: dims in [2,4096], spacing > 0, octaves in [0,16], finite origin");
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  DevVolume v;
  re_status st = alloc_volume(ctx, p.dim, p.spacing, v);
  if (st) return st;
  uint32_t perm[512];
  perlin_table(p.seed, perm);
  st = ensure_stage(ctx, sizeof(perm) + 256);
  if (st) {
    (void)hipFree(v.data);
    return st;
  }
  hipError_t e = hipMemcpyAsync(ctx->stage, perm, sizeof(perm), hipMemcpyHostToDevice, ctx->stream);
  NoiseLaunch L;
  L.out = v.data;
  L.perm = (const uint32_t*)ctx->stage;
  for (int a = 0; a < 3; ++a) {
    L.dim[a] = v.dim[a];
    L.bdim[a] = v.bdim[a];
    L.spacing[a] = v.spacing[a];
    L.origin[a] = p.origin[a];
  }
  L.fbm = FbmParams{p.octaves, p.lacunarity, p.gain, p.base_freq, p.y_offset};
  if (e == hipSuccess) e = ctx->launch_begin();
  if (e == hipSuccess) e = launch_noise_fill(L, ctx->stream);
  if (e == hipSuccess) e = ctx->launch_end();
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    (void)hipFree(v.data);
    return hip_err(ctx, e, "re_noise_generate");
  }
  int32_t id = ctx->next_id++;
  ctx->volumes[id] = v;
  *vol_id = id;
  return RE_OK;
}

RE_API re_status re_volume_upload(re_ctx* ctx, const int32_t* dim, const float* spacing, const float* linear,
                                  int32_t* vol_id) {
  if (!ctx || !dim || !spacing || !linear || !vol_id) return fail(RE_ERR_INVALID, "re_volume_upload: null argument");
  if (!valid_dims(dim, spacing)) return set_err(ctx, RE_ERR_INVALID, "re_volume_upload: dims in [2,4096], spacing > 0");
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  DevVolume v;
  re_status st = alloc_volume(ctx, dim, spacing, v);
  if (st) return st;
  std::vector<float> bricked((size_t)(v.bytes / 4), 0.0f);
  for (int z = 0; z < dim[2]; ++z)
    for (int y = 0; y < dim[1]; ++y)
      for (int x = 0; x < dim[0]; ++x)
        bricked[vox_index(v.bdim, x, y, z)] = linear[((int64_t)z * dim[1] + y) * dim[0] + x];
  hipError_t e = hipMemcpyAsync(v.data, bricked.data(), (size_t)v.bytes, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    (void)hipFree(v.data);
    return hip_err(ctx, e, "re_volume_upload");
  }
  int32_t id = ctx->next_id++;
  ctx->volumes[id] = v;
  *vol_id = id;
  return RE_OK;
}

RE_API re_status re_volume_info(re_ctx* ctx, int32_t vol_id, int32_t* dim, float* spacing, int64_t* bytes) {
  if (!ctx) return fail(RE_ERR_INVALID, "re_volume_info: null ctx");
  auto it = ctx->volumes.find(vol_id);
  if (it == ctx->volumes.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "re_volume_info: unknown volume");
  for (int a = 0; a < 3; ++a) {
    if (dim) dim[a] = it->second.dim[a];
    if (spacing) spacing[a] = it->second.spacing[a];
  }
  if (bytes) *bytes = it->second.bytes;
  return RE_OK;
}

RE_API re_status re_volume_read(re_ctx* ctx, int32_t vol_id, float* linear) {
  if (!ctx || !linear) return fail(RE_ERR_INVALID, "re_volume_read: null argument");
  auto it = ctx->volumes.find(vol_id);
  if (it == ctx->volumes.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "re_volume_read: unknown volume");
  const DevVolume& v = it->second;
  std::vector<float> bricked((size_t)(v.bytes / 4));
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  HIP_TRY(ctx, hipMemcpyAsync(bricked.data(), v.data, (size_t)v.bytes, hipMemcpyDeviceToHost, ctx->stream), "volume read");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "volume read sync");
  for (int z = 0; z < v.dim[2]; ++z)
    for (int y = 0; y < v.dim[1]; ++y)
      for (int x = 0; x < v.dim[0]; ++x)
        linear[((int64_t)z * v.dim[1] + y) * v.dim[0] + x] = bricked[vox_index(v.bdim, x, y, z)];
  return RE_OK;
}

RE_API re_status re_volume_release(re_ctx* ctx, int32_t vol_id) {
  if (!ctx) return fail(RE_ERR_INVALID, "re_volume_release: null ctx");
  auto it = ctx->volumes.find(vol_id);
  if (it == ctx->volumes.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "re_volume_release: unknown volume");
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(it->second.data);
  if (it->second.mat) (void)hipFree(it->second.mat);
  ctx->volumes.erase(it);
  return RE_OK;
}

RE_API re_status re_volume_save_rsdf(re_ctx* ctx, int32_t vol_id, const char* path) {
  if (!ctx || !path) return fail(RE_ERR_INVALID, "re_volume_save_rsdf: null argument");
  auto it = ctx->volumes.find(vol_id);
  if (it == ctx->volumes.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "re_volume_save_rsdf: unknown volume");
  const DevVolume& v = it->second;
  const int64_t n = (int64_t)v.dim[0] * v.dim[1] * v.dim[2];
  std::vector<float> dist((size_t)n), mat((size_t)n);
  re_status st = read_channel(ctx, v, v.data, dist.data());
  if (!st) st = read_channel(ctx, v, v.mat, mat.data());
  if (st) return st;
  std::vector<float> pairs((size_t)(2 * n));
  for (int64_t i = 0; i < n; ++i) {
    pairs[(size_t)(2 * i)] = dist[(size_t)i];
    pairs[(size_t)(2 * i + 1)] = mat[(size_t)i];
  }
  FILE* f = std::fopen(path, "wb");
  if (!f) return set_err(ctx, RE_ERR_IO, std::string("re_volume_save_rsdf: cannot open ") + path);
  bool ok = std::fwrite(v.dim, sizeof(int32_t), 3, f) == 3 && std::fwrite(v.spacing, sizeof(float), 3, f) == 3 &&
            std::fwrite(pairs.data(), sizeof(float), pairs.size(), f) == pairs.size();
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return set_err(ctx, RE_ERR_IO, std::string("re_volume_save_rsdf: write failed: ") + path);
  return RE_OK;
}

}  // extern "C"

namespace {

re_status open_rsdf(re_ctx* ctx, const char* path, const char* who, FILE** out, int32_t dim[3], float spacing[3]) {
  const char* why = nullptr;
  const re_status st = rsdf_open(path, out, dim, spacing, &why);
  if (st) return set_err(ctx, st, std::string(who) + ": " + why + ": " + path);
  return RE_OK;
}

// A new volume from linear x-fastest (distance, material) pairs; box_max_last
// bounds the box by the point of the last voxel (load_map, Grid.cpp:61)
// instead of dim * spacing.
re_status upload_pairs(re_ctx* ctx, const int32_t* dim, const float* spacing, const float* pairs, bool box_max_last,
                       int32_t* vol_id) {
  const int64_t n = (int64_t)dim[0] * dim[1] * dim[2];
  std::vector<float> dist((size_t)n), mat((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    dist[(size_t)i] = pairs[(size_t)(2 * i)];
    mat[(size_t)i] = pairs[(size_t)(2 * i + 1)];
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  DevVolume v;
  re_status st = alloc_volume(ctx, dim, spacing, v);
  if (st) return st;
  if (box_max_last)
    for (int a = 0; a < 3; ++a) v.box_max[a] = (float)(dim[a] - 1) * spacing[a];
  std::vector<float> bd, bm;
  to_bricked(v, dist.data(), bd);
  to_bricked(v, mat.data(), bm);
  hipError_t e = hipMalloc((void**)&v.mat, (size_t)v.bytes);
  if (e == hipSuccess) e = hipMemcpyAsync(v.data, bd.data(), (size_t)v.bytes, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(v.mat, bm.data(), (size_t)v.bytes, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    (void)hipFree(v.data);
    if (v.mat) (void)hipFree(v.mat);
    v.mat = nullptr;
    return hip_err(ctx, e, "volume upload");
  }
  int32_t id = ctx->next_id++;
  ctx->volumes[id] = v;
  *vol_id = id;
  return RE_OK;
}

}  // namespace

extern "C" {

RE_API re_status re_volume_upload_pairs(re_ctx* ctx, const int32_t* dim, const float* spacing, const float* pairs,
                                        int32_t* vol_id) {
  if (!ctx || !dim || !spacing || !pairs || !vol_id) return fail(RE_ERR_INVALID, "re_volume_upload_pairs: null argument");
  if (!valid_dims(dim, spacing))
    return set_err(ctx, RE_ERR_INVALID, "re_volume_upload_pairs: dims in [2,4096], spacing > 0");
  return guarded(ctx, "re_volume_upload_pairs",
                 [&]() -> re_status { return upload_pairs(ctx, dim, spacing, pairs, false, vol_id); });
}

RE_API re_status re_volume_load_rsdf(re_ctx* ctx, const char* path, int32_t* vol_id) {
  if (!ctx || !path || !vol_id) return fail(RE_ERR_INVALID, "re_volume_load_rsdf: null argument");
  return guarded(ctx, "re_volume_load_rsdf", [&]() -> re_status {
    FILE* f = nullptr;
    int32_t dim[3];
    float spacing[3];
    re_status st0 = open_rsdf(ctx, path, "re_volume_load_rsdf", &f, dim, spacing);
    if (st0) return st0;
    const int64_t n = (int64_t)dim[0] * dim[1] * dim[2];
    std::vector<float> pairs((size_t)(2 * n));
    const bool ok = std::fread(pairs.data(), sizeof(float), pairs.size(), f) == pairs.size();
    std::fclose(f);
    if (!ok) return set_err(ctx, RE_ERR_IO, "re_volume_load_rsdf: file shorter than its header says");
    return upload_pairs(ctx, dim, spacing, pairs.data(), true, vol_id);
  });
}

RE_API re_status re_volume_read_material(re_ctx* ctx, int32_t vol_id, float* linear) {
  if (!ctx || !linear) return fail(RE_ERR_INVALID, "re_volume_read_material: null argument");
  auto it = ctx->volumes.find(vol_id);
  if (it == ctx->volumes.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "re_volume_read_material: unknown volume");
  return read_channel(ctx, it->second, it->second.mat, linear);
}

RE_API void re_brush_default(re_brush* b) {
  if (!b) return;
  std::memset(b, 0, sizeof(*b));
  b->type = RE_BRUSH_SPHERE_ADD;
  b->extent[0] = b->extent[1] = b->extent[2] = 1.0f;
  b->smooth_k = 0.1f;
}

RE_API re_status re_volume_edit(re_ctx* ctx, int32_t vol_id, const re_brush* brush) {
  if (!ctx || !brush) return fail(RE_ERR_INVALID, "re_volume_edit: null argument");
  auto it = ctx->volumes.find(vol_id);
  if (it == ctx->volumes.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "re_volume_edit: unknown volume");
  const re_brush& b = *brush;
  bool ok = b.type >= RE_BRUSH_SPHERE_ADD && b.type <= RE_BRUSH_CUBE_SUBTRACT && std::isfinite(b.material);
  for (int a = 0; a < 3; ++a) ok = ok && std::isfinite(b.center[a]) && std::isfinite(b.extent[a]);
  if (b.type == RE_BRUSH_SPHERE_ADD) ok = ok && b.smooth_k > 0.0f && std::isfinite(b.smooth_k);
  if (b.snap_to_grid) ok = ok && b.extent[0] != 0.0f;
  if (!ok)
    return set_err(ctx, RE_ERR_INVALID,
                   "re_volume_edit: brush type in [0,3], finite center/extent/material, smooth_k > 0 for "
                   "SPHERE_ADD, extent[0] != 0 when snapping");
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  DevVolume& v = it->second;
  re_status st = ensure_material(ctx, v);
  if (st) return st;
  EditLaunch L;
  L.data = v.data;
  L.mat = v.mat;
  for (int a = 0; a < 3; ++a) {
    L.dim[a] = v.dim[a];
    L.bdim[a] = v.bdim[a];
    L.spacing[a] = v.spacing[a];
    L.center[a] = b.center[a];
    L.extent[a] = b.extent[a];
  }
  L.type = b.type;
  L.material = b.material;
  L.snap = b.snap_to_grid ? 1 : 0;
  L.k = b.smooth_k;
  HIP_TRY(ctx, ctx->launch_begin(), "event");
  HIP_TRY(ctx, launch_volume_edit(L, ctx->stream), "volume edit launch");
  HIP_TRY(ctx, ctx->launch_end(), "event");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "volume edit sync");
  return RE_OK;
}

RE_API void re_sdf_default_params(re_sdf_params* p) {
  if (!p) return;
  p->max_steps = 256;
  p->eps = 1e-3f;
  p->step_scale = 0.8f;
  p->step_rule = 0;
  p->block = 0;
}

RE_API re_status re_sdf_trace(re_ctx* ctx, int32_t vol_id, const re_sdf_params* params, const float* ray_o,
                              const float* ray_d, int64_t n, float* t, uint8_t* hit, int32_t* steps, float* normal,
                              uint32_t flags) {
  if (!ctx) return fail(RE_ERR_INVALID, "re_sdf_trace: null ctx");
  if (n < 0 || (n > 0 && (!ray_o || !ray_d || !t || !hit))) return set_err(ctx, RE_ERR_INVALID, "re_sdf_trace: bad arguments");
  auto it = ctx->volumes.find(vol_id);
  if (it == ctx->volumes.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "re_sdf_trace: unknown volume");
  re_sdf_params p;
  re_sdf_default_params(&p);
  if (params) p = *params;
  if (p.max_steps < 1 || p.max_steps > (1 << 20) || !(p.block == 0 || p.block == 64 || p.block == 128 || p.block == 256) ||
      !(p.step_rule == 0 || p.step_rule == 1))
    return set_err(ctx, RE_ERR_INVALID, "re_sdf_trace: max_steps in [1,2^20], block in {0,64,128,256}, step_rule 0/1");
  if (n == 0) return RE_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  const DevVolume& v = it->second;
  SdfLaunch L;
  L.vol = v.view();
  for (int a = 0; a < 3; ++a) {
    L.box_min[a] = v.box_min[a];
    L.box_max[a] = v.box_max[a];
  }
  L.n = n;
  L.max_steps = p.max_steps;
  L.eps = p.eps;
  L.step_scale = p.step_scale;
  L.step_rule = p.step_rule;
  L.block = p.block;
  const bool dev = (flags & RE_DEVICE_PTRS) != 0;
  if (!dev) {
    size_t rb = (size_t)n * 12;
    re_status st = ensure_stage(ctx, stage_size({rb, rb, (size_t)n * 4, (size_t)n, (size_t)n * 4, (size_t)n * 12}));
    if (st) return st;
    StageCursor c{(char*)ctx->stage};
    float* d_o = c.take<float>(3 * n);
    float* d_d = c.take<float>(3 * n);
    L.t = c.take<float>(n);
    L.hit = c.take<uint8_t>(n);
    int32_t* d_s = c.take<int32_t>(n);
    float* d_n = c.take<float>(3 * n);
    L.steps = steps ? d_s : nullptr;
    L.nrm = normal ? d_n : nullptr;
    HIP_TRY(ctx, hipMemcpyAsync(d_o, ray_o, rb, hipMemcpyHostToDevice, ctx->stream), "copy ray_o");
    HIP_TRY(ctx, hipMemcpyAsync(d_d, ray_d, rb, hipMemcpyHostToDevice, ctx->stream), "copy ray_d");
    L.ray_o = d_o;
    L.ray_d = d_d;
  } else {
    L.ray_o = ray_o;
    L.ray_d = ray_d;
    L.t = t;
    L.hit = hit;
    L.steps = steps;
    L.nrm = normal;
  }
  HIP_TRY(ctx, ctx->launch_begin(), "event");
  HIP_TRY(ctx, launch_sdf_trace(L, ctx->stream), "sdf_trace launch");
  HIP_TRY(ctx, ctx->launch_end(), "event");
  if (!dev) {
    HIP_TRY(ctx, hipMemcpyAsync(t, L.t, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream), "copy t");
    HIP_TRY(ctx, hipMemcpyAsync(hit, L.hit, (size_t)n, hipMemcpyDeviceToHost, ctx->stream), "copy hit");
    if (steps) HIP_TRY(ctx, hipMemcpyAsync(steps, L.steps, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream), "copy steps");
    if (normal) HIP_TRY(ctx, hipMemcpyAsync(normal, L.nrm, (size_t)n * 12, hipMemcpyDeviceToHost, ctx->stream), "copy nrm");
  }
  if (!(flags & RE_ASYNC) || !dev) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "sdf sync");
  return RE_OK;
}

}  // extern "C"
