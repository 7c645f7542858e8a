// The allocation-free closest-hit entry point.  Its counters, ray counter and
// host-pointer staging live in one file-scope device array, `scratch`, of
// fixed size, resolved to its device address once per device when a context
// is created there (re_create); the call itself allocates nothing.  Every
// call on a device shares that device's array and leaves the last call's
// contents in it, so calls that overlap on one device (two contexts or two
// host threads) race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "capi_util.hpp"

using namespace re;

namespace {

// Rays a host-pointer call may stage: per ray 12 + 12 B in, 4 + 4 + 4 B out,
// plus the staging layout's alignment slack.
constexpr int64_t kNoAllocMaxRays = (int64_t)1 << 18;
constexpr size_t kScratchBytes = kScratchHeader + (size_t)kNoAllocMaxRays * 36 + 1024;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

}  // namespace

namespace re {

re_status noalloc_setup(re_ctx* ctx) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return set_err(ctx, RE_ERR_UNSUPPORTED, "re_create: device index beyond the fixed-area table");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[ctx->device]) {
    void* p = nullptr;
    // `scratch` names a host-side stand-in; the kernels get this address.
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_err(ctx, e, "re_create: address of the fixed area");
    g_scratch[ctx->device] = p;
  }
  ctx->noalloc_area = g_scratch[ctx->device];
  return RE_OK;
}

}  // namespace re

extern "C" {

RE_API int64_t re_trace_closest_noalloc_max_rays(void) { return kNoAllocMaxRays; }

RE_API re_status re_trace_closest_noalloc(re_ctx* ctx, int32_t scene_id, const float* ray_o, const float* ray_d,
                                          int64_t n, const re_trace_params* params, int32_t* prim_id, float* t,
                                          int32_t* obj_id, uint64_t counts[4], uint32_t flags) {
  if (!ctx) return fail(RE_ERR_INVALID, "re_trace_closest_noalloc: null ctx");
  if (!ctx->noalloc_area) return set_err(ctx, RE_ERR_INVALID, "re_trace_closest_noalloc: fixed area not set up");
  if (!(flags & RE_DEVICE_PTRS) && n > kNoAllocMaxRays)
    return set_err(ctx, RE_ERR_INVALID,
                   "re_trace_closest_noalloc: at most 262144 rays through host pointers (device pointers: any n)");
  const ScratchArea sa{ctx->noalloc_area, kScratchBytes};
  TraceCounts c{0, 0, 0, 0};
  re_status st =
      trace_common(ctx, scene_id, ray_o, ray_d, n, params, prim_id, t, obj_id, flags, counts ? &c : nullptr, &sa);
  if (counts) {
    counts[0] = c.entries;
    counts[1] = c.leaves;
    counts[2] = c.tris;
    counts[3] = c.rays_hit;
  }
  return st;
}

}  // extern "C"
