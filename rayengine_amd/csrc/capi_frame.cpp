// C-ABI entry points of frame output: colouring a traced frame and writing
// it as a PPM file.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <vector>

#include "capi_util.hpp"

using namespace re;

extern "C" {

RE_API void re_frame_default_params(re_frame_params* p) {
  if (!p) return;
  p->sun_dir[0] = 0.0f;
  p->sun_dir[1] = 1.0f;
  p->sun_dir[2] = 0.0f;
  p->earth_radius = 6360e3f;
  p->atmosphere_radius = 6420e3f;
  p->hr = 79945.0f;
  p->hm = 1200.0f;
  p->sky_origin[0] = 0.0f;
  p->sky_origin[1] = 6360e3f + 10000.0f;
  p->sky_origin[2] = 300000.0f;
  p->base_color[0] = p->base_color[1] = p->base_color[2] = 0.0f;
  p->gamma = 0;
  p->vignetting = 0;
  p->vignetting_k = 0.25f;
}

RE_API re_status re_frame_shade(re_ctx* ctx, const float* ray_d, const uint8_t* hit, const float* normal,
                                int32_t width, int32_t height, const re_frame_params* params, uint32_t* rgba,
                                uint32_t flags) {
  if (!ctx) return fail(RE_ERR_INVALID, "re_frame_shade: null ctx");
  if (width < 0 || height < 0) return set_err(ctx, RE_ERR_INVALID, "re_frame_shade: negative size");
  const int64_t n = (int64_t)width * height;
  if (n > 0 && (!ray_d || !hit || !normal || !rgba)) return set_err(ctx, RE_ERR_INVALID, "re_frame_shade: null array");
  re_frame_params p;
  re_frame_default_params(&p);
  if (params) p = *params;
  bool ok = p.earth_radius > 0.0f && p.atmosphere_radius > p.earth_radius && p.hr > 0.0f && p.hm > 0.0f &&
            std::isfinite(p.atmosphere_radius) && std::isfinite(p.vignetting_k);
  for (int a = 0; a < 3; ++a) ok = ok && std::isfinite(p.sun_dir[a]) && std::isfinite(p.sky_origin[a]);
  if (!ok)
    return set_err(ctx, RE_ERR_INVALID, "re_frame_shade: radii must satisfy 0 < earth < atmosphere, scale heights > 0");
  if (n == 0) return RE_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  FrameLaunch L{};
  L.width = width;
  L.height = height;
  for (int a = 0; a < 3; ++a) {
    L.sun[a] = p.sun_dir[a];
    L.origin[a] = p.sky_origin[a];
    L.base[a] = p.base_color[a];
  }
  L.er = p.earth_radius;
  L.ar = p.atmosphere_radius;
  L.hr = p.hr;
  L.hm = p.hm;
  L.gamma = p.gamma ? 1 : 0;
  L.vignetting = p.vignetting ? 1 : 0;
  L.vignetting_k = p.vignetting_k;
  const bool dev = (flags & RE_DEVICE_PTRS) != 0;
  if (!dev) {
    const size_t b3 = (size_t)n * 12;
    re_status st = ensure_stage(ctx, stage_size({b3, b3, (size_t)n, (size_t)n * 4}));
    if (st) return st;
    StageCursor c{(char*)ctx->stage};
    float* dd = c.take<float>(3 * n);
    float* dn = c.take<float>(3 * n);
    uint8_t* dh = c.take<uint8_t>(n);
    L.rgba = c.take<uint32_t>(n);
    HIP_TRY(ctx, hipMemcpyAsync(dd, ray_d, b3, hipMemcpyHostToDevice, ctx->stream), "copy ray_d");
    HIP_TRY(ctx, hipMemcpyAsync(dn, normal, b3, hipMemcpyHostToDevice, ctx->stream), "copy normal");
    HIP_TRY(ctx, hipMemcpyAsync(dh, hit, (size_t)n, hipMemcpyHostToDevice, ctx->stream), "copy hit");
    L.ray_d = dd;
    L.normal = dn;
    L.hit = dh;
  } else {
    L.ray_d = ray_d;
    L.normal = normal;
    L.hit = hit;
    L.rgba = rgba;
  }
  HIP_TRY(ctx, ctx->launch_begin(), "event");
  HIP_TRY(ctx, launch_frame(L, ctx->stream), "frame launch");
  HIP_TRY(ctx, ctx->launch_end(), "event");
  if (!dev) HIP_TRY(ctx, hipMemcpyAsync(rgba, L.rgba, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream), "copy rgba");
  if (!(flags & RE_ASYNC) || !dev) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "frame sync");
  return RE_OK;
}

RE_API re_status re_write_ppm(const char* path, const uint32_t* rgba, int32_t width, int32_t height) {
  if (!path || width <= 0 || height <= 0 || !rgba) return fail(RE_ERR_INVALID, "re_write_ppm: bad arguments");
  const int64_t n = (int64_t)width * height;
  std::vector<unsigned char> rgb((size_t)(3 * n));
  for (int64_t i = 0; i < n; ++i) {
    rgb[(size_t)(3 * i)] = (unsigned char)(rgba[i] & 0xffu);
    rgb[(size_t)(3 * i + 1)] = (unsigned char)((rgba[i] >> 8) & 0xffu);
    rgb[(size_t)(3 * i + 2)] = (unsigned char)((rgba[i] >> 16) & 0xffu);
  }
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(RE_ERR_IO, "re_write_ppm: cannot open %s", path);
  bool ok = std::fprintf(f, "P6\n%d %d\n255\n", width, height) > 0 &&
            std::fwrite(rgb.data(), 1, rgb.size(), f) == rgb.size();
  ok = (std::fclose(f) == 0) && ok;
  return ok ? RE_OK : fail(RE_ERR_IO, "re_write_ppm: write failed: %s", path);
}

}  // extern "C"
