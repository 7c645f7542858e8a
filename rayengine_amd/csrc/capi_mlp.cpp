// C-ABI entry points of the SDF surrogate: network state, forward pass,
// batch gradient, Adam training on samples drawn from a terrain volume, the
// signed distance transform and the sampler.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "capi_util.hpp"
#include "host/rng.hpp"
#include "mlp_layout.hpp"

using namespace re;
using namespace re::mlp;

namespace {

void free_mlp(DevMlp& m) {
  for (void* p : {(void*)m.params, (void*)m.m1, (void*)m.m2, m.pack, (void*)m.small, (void*)m.partial,
                  (void*)m.grad, (void*)m.loss})
    if (p) (void)hipFree(p);
  m = DevMlp{};
}

re_status get_mlp(re_ctx* ctx, int32_t id, DevMlp** out) {
  auto it = ctx->mlps.find(id);
  if (it == ctx->mlps.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "unknown surrogate network id");
  *out = &it->second;
  return RE_OK;
}

// Rewrites the bf16 operand pack and the small FP32 parameters from params.
re_status repack(re_ctx* ctx, DevMlp& m) {
  HIP_TRY(ctx, launch_mlp_update(m.params, m.m1, m.m2, nullptr, 0, 0, 0.f, 0.f, 0.f, 0.f, 0, nullptr, m.pack,
                                 m.small, nullptr, ctx->stream), "mlp pack");
  return RE_OK;
}

int32_t train_blocks(re_ctx* ctx) {
  int cus = 0;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
  return cus > 0 ? cus : 256;
}

}  // namespace

extern "C" {

RE_API int32_t re_mlp_param_count(void) { return kParams; }

RE_API re_status re_mlp_create(re_ctx* ctx, uint64_t seed, int32_t* mlp_id) {
  if (!ctx || !mlp_id) return fail(RE_ERR_INVALID, "re_mlp_create: null argument");
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  hipDeviceProp_t prop;
  HIP_TRY(ctx, hipGetDeviceProperties(&prop, ctx->device), "device properties");
  if (prop.sharedMemPerBlock < mlp_train_lds_bytes())
    return set_err(ctx, RE_ERR_UNSUPPORTED, "re_mlp_create: the training kernel needs more LDS than the device offers");
  DevMlp m;
  m.partial_blocks = train_blocks(ctx);
  const size_t pb = (size_t)kParams * 4;
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = hipMalloc((void**)&m.params, pb);
  if (e == hipSuccess) e = hipMalloc((void**)&m.m1, pb);
  if (e == hipSuccess) e = hipMalloc((void**)&m.m2, pb);
  if (e == hipSuccess) e = hipMalloc(&m.pack, (size_t)kPackAlloc * 2);
  if (e == hipSuccess) e = hipMalloc((void**)&m.small, (size_t)kSmallAlloc * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&m.partial, (size_t)m.partial_blocks * (kParams + 1) * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&m.grad, pb);
  if (e == hipSuccess) e = hipMalloc((void**)&m.loss, 4 * 1024);
  if (e != hipSuccess) {
    free_mlp(m);
    return hip_err(ctx, e, "re_mlp_create");
  }
  m.loss_cap = 1024;
  // Uniform initialisation with variance 1/fan_in for the weights, zero biases.
  std::vector<float> P(kParams, 0.0f);
  Pcg32 rng(seed, 0x4d4c5030ull);
  auto fill = [&](int off, int fan_out, int fan_in, float gain) {
    const float bound = gain * std::sqrt(3.0f / (float)fan_in);
    for (int i = 0; i < fan_out * fan_in; ++i) P[off + i] = (2.0f * rng.uniform() - 1.0f) * bound;
  };
  fill(oW0, kH, kIn, 1.0f);
  fill(oW1, kH, kH, 1.0f);
  fill(oW2, kH, kH, 1.0f);
  fill(oW3, kH, kH, 1.0f);
  fill(oW4, 1, kH, 0.1f);
  std::vector<float> zeros(kParams, 0.0f);
  std::vector<uint16_t> zpack(kPackAlloc, 0);
  e = hipMemcpyAsync(m.params, P.data(), pb, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(m.m1, zeros.data(), pb, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(m.m2, zeros.data(), pb, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(m.pack, zpack.data(), (size_t)kPackAlloc * 2, hipMemcpyHostToDevice, ctx->stream);
  if (e != hipSuccess) {
    free_mlp(m);
    return hip_err(ctx, e, "re_mlp_create upload");
  }
  re_status st = repack(ctx, m);
  if (st == RE_OK) {
    e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) st = hip_err(ctx, e, "re_mlp_create sync");
  }
  if (st) {
    free_mlp(m);
    return st;
  }
  int32_t id = ctx->next_id++;
  ctx->mlps[id] = m;
  *mlp_id = id;
  return RE_OK;
}

RE_API re_status re_mlp_release(re_ctx* ctx, int32_t mlp_id) {
  if (!ctx) return fail(RE_ERR_INVALID, "re_mlp_release: null ctx");
  auto it = ctx->mlps.find(mlp_id);
  if (it == ctx->mlps.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "re_mlp_release: unknown network");
  (void)hipStreamSynchronize(ctx->stream);
  free_mlp(it->second);
  ctx->mlps.erase(it);
  return RE_OK;
}

RE_API re_status re_mlp_set_params(re_ctx* ctx, int32_t mlp_id, const float* params) {
  if (!ctx || !params) return fail(RE_ERR_INVALID, "re_mlp_set_params: null argument");
  DevMlp* m;
  re_status st = get_mlp(ctx, mlp_id, &m);
  if (st) return st;
  HIP_TRY(ctx, hipMemcpyAsync(m->params, params, (size_t)kParams * 4, hipMemcpyHostToDevice, ctx->stream), "set params");
  std::vector<float> zeros(kParams, 0.0f);
  HIP_TRY(ctx, hipMemcpyAsync(m->m1, zeros.data(), (size_t)kParams * 4, hipMemcpyHostToDevice, ctx->stream), "reset m");
  HIP_TRY(ctx, hipMemcpyAsync(m->m2, zeros.data(), (size_t)kParams * 4, hipMemcpyHostToDevice, ctx->stream), "reset v");
  m->step = 0;
  m->cursor = -1;
  st = repack(ctx, *m);
  if (st) return st;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "set params sync");
  return RE_OK;
}

RE_API re_status re_mlp_get_params(re_ctx* ctx, int32_t mlp_id, float* params, float* adam_m, float* adam_v,
                                   int32_t* step) {
  if (!ctx || !params) return fail(RE_ERR_INVALID, "re_mlp_get_params: null argument");
  DevMlp* m;
  re_status st = get_mlp(ctx, mlp_id, &m);
  if (st) return st;
  HIP_TRY(ctx, hipMemcpyAsync(params, m->params, (size_t)kParams * 4, hipMemcpyDeviceToHost, ctx->stream), "get params");
  if (adam_m) HIP_TRY(ctx, hipMemcpyAsync(adam_m, m->m1, (size_t)kParams * 4, hipMemcpyDeviceToHost, ctx->stream), "get m");
  if (adam_v) HIP_TRY(ctx, hipMemcpyAsync(adam_v, m->m2, (size_t)kParams * 4, hipMemcpyDeviceToHost, ctx->stream), "get v");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "get params sync");
  if (step) *step = m->step;
  return RE_OK;
}

RE_API re_status re_mlp_forward(re_ctx* ctx, int32_t mlp_id, const float* pts, int64_t n, float* out,
                                uint32_t flags) {
  if (!ctx || n < 0 || (n > 0 && (!pts || !out))) return fail(RE_ERR_INVALID, "re_mlp_forward: bad arguments");
  DevMlp* m;
  re_status st = get_mlp(ctx, mlp_id, &m);
  if (st || n == 0) return st;
  const bool dev = (flags & RE_DEVICE_PTRS) != 0;
  const float* d_pts = pts;
  float* d_out = out;
  if (!dev) {
    st = ensure_stage(ctx, stage_size({(size_t)n * 12, (size_t)n * 4}));
    if (st) return st;
    StageCursor c{(char*)ctx->stage};
    float* sp = c.take<float>(3 * n);
    d_out = c.take<float>(n);
    HIP_TRY(ctx, hipMemcpyAsync(sp, pts, (size_t)n * 12, hipMemcpyHostToDevice, ctx->stream), "copy pts");
    d_pts = sp;
  }
  HIP_TRY(ctx, ctx->launch_begin(), "event");
  HIP_TRY(ctx, launch_mlp_infer(infer_pack(*m), infer_small(*m), d_pts, n, d_out, ctx->stream), "mlp forward");
  HIP_TRY(ctx, ctx->launch_end(), "event");
  if (!dev) HIP_TRY(ctx, hipMemcpyAsync(out, d_out, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream), "copy out");
  if (!(flags & RE_ASYNC) || !dev) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "mlp forward sync");
  return RE_OK;
}

// Gradient of the batch loss over samples (x, y, z, target) at the current
// parameters; with adam != 0 also takes one Adam step.  n must be a positive
// multiple of 64.
RE_API re_status re_mlp_batch_step(re_ctx* ctx, int32_t mlp_id, const float* samples, int64_t n,
                                   float over_weight, int32_t adam, float lr, float* grad, float* loss) {
  if (!ctx || !samples || n <= 0 || n % 64 != 0 || n > (1 << 26))
    return fail(RE_ERR_INVALID, "re_mlp_batch_step: n must be a positive multiple of 64");
  DevMlp* m;
  re_status st = get_mlp(ctx, mlp_id, &m);
  if (st) return st;
  st = ensure_stage(ctx, stage_size({(size_t)n * 16}));
  if (st) return st;
  HIP_TRY(ctx, hipMemcpyAsync(ctx->stage, samples, (size_t)n * 16, hipMemcpyHostToDevice, ctx->stream), "copy samples");
  HIP_TRY(ctx, launch_mlp_train(m->pack, m->small, (const float4*)ctx->stage, n, 0, (int32_t)n, over_weight,
                                m->partial, m->partial_blocks, ctx->stream), "mlp train");
  if (adam) ++m->step;
  HIP_TRY(ctx, launch_mlp_update(m->params, m->m1, m->m2, m->partial, m->partial_blocks, m->step, lr, 0.9f, 0.999f,
                                 1e-8f, adam ? 1 : 0, m->grad, m->pack, m->small, m->loss, ctx->stream), "mlp update");
  if (grad) HIP_TRY(ctx, hipMemcpyAsync(grad, m->grad, (size_t)kParams * 4, hipMemcpyDeviceToHost, ctx->stream), "copy grad");
  if (loss) HIP_TRY(ctx, hipMemcpyAsync(loss, m->loss, 4, hipMemcpyDeviceToHost, ctx->stream), "copy loss");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "batch step sync");
  return RE_OK;
}

RE_API re_status re_sdf_distance_volume(re_ctx* ctx, int32_t vol_id, int32_t* dist_vol_id) {
  if (!ctx || !dist_vol_id) return fail(RE_ERR_INVALID, "re_sdf_distance_volume: null argument");
  auto it = ctx->volumes.find(vol_id);
  if (it == ctx->volumes.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "re_sdf_distance_volume: unknown volume");
  const DevVolume& f = it->second;
  if (f.dim[0] > 4096 || f.dim[1] > 4096 || f.dim[2] > 4096)
    return set_err(ctx, RE_ERR_INVALID, "re_sdf_distance_volume: dims above 4096");
  DevVolume d = f;
  d.data = nullptr;
  const size_t nvox = (size_t)f.dim[0] * f.dim[1] * f.dim[2];
  float *a = nullptr, *b = nullptr;
  hipError_t e = hipMalloc((void**)&d.data, (size_t)f.bytes);
  if (e == hipSuccess) e = hipMalloc((void**)&a, nvox * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&b, nvox * 4);
  if (e == hipSuccess) e = hipMemsetAsync(d.data, 0, (size_t)f.bytes, ctx->stream);
  if (e == hipSuccess) e = ctx->launch_begin();
  if (e == hipSuccess) e = launch_edt(f, a, b, d.data, ctx->stream);
  if (e == hipSuccess) e = ctx->launch_end();
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (e != hipSuccess) {
    if (d.data) (void)hipFree(d.data);
    return hip_err(ctx, e, "re_sdf_distance_volume");
  }
  int32_t id = ctx->next_id++;
  ctx->volumes[id] = d;
  *dist_vol_id = id;
  return RE_OK;
}

RE_API re_status re_mlp_sample(re_ctx* ctx, int32_t dist_vol_id, int64_t n, uint64_t seed, float near_band,
                               float* out, uint32_t flags) {
  if (!ctx || n <= 0 || !out) return fail(RE_ERR_INVALID, "re_mlp_sample: bad arguments");
  auto it = ctx->volumes.find(dist_vol_id);
  if (it == ctx->volumes.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "re_mlp_sample: unknown volume");
  float4* d_out = (float4*)out;
  const bool dev = (flags & RE_DEVICE_PTRS) != 0;
  if (!dev) {
    re_status st = ensure_stage(ctx, stage_size({(size_t)n * 16}));
    if (st) return st;
    d_out = (float4*)ctx->stage;
  }
  HIP_TRY(ctx, launch_sample(it->second, n, seed, near_band, 256, d_out, ctx->stream), "sampler");
  if (!dev) HIP_TRY(ctx, hipMemcpyAsync(out, d_out, (size_t)n * 16, hipMemcpyDeviceToHost, ctx->stream), "copy samples");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "sampler sync");
  return RE_OK;
}

RE_API void re_hybrid_default_params(re_hybrid_params* p) {
  if (!p) return;
  p->tau_voxels = 2.0f;
  p->margin_voxels = 1.0f;
  p->k_exact = 0.05f;
  p->h_min_voxels = 0.5f;
  p->rejoin_voxels = 4.0f;
  p->n_bisect = 8;
  p->max_iters = 4096;
  p->block = 0;
}

RE_API re_status re_hybrid_trace(re_ctx* ctx, int32_t vol_id, int32_t mlp_id, const re_hybrid_params* params,
                                 const float* ray_o, const float* ray_d, int64_t n, float* t, uint8_t* hit,
                                 int32_t* evals, int32_t* samples, uint32_t flags) {
  if (!ctx) return fail(RE_ERR_INVALID, "re_hybrid_trace: null ctx");
  if (n < 0 || (n > 0 && (!ray_o || !ray_d || !t || !hit))) return set_err(ctx, RE_ERR_INVALID, "re_hybrid_trace: bad arguments");
  auto it = ctx->volumes.find(vol_id);
  if (it == ctx->volumes.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "re_hybrid_trace: unknown volume");
  DevMlp* m = nullptr;
  if (mlp_id >= 0) {
    re_status st = get_mlp(ctx, mlp_id, &m);
    if (st) return st;
  }
  re_hybrid_params p;
  re_hybrid_default_params(&p);
  if (params) p = *params;
  if (p.max_iters < 1 || p.max_iters > (1 << 20) || p.n_bisect < 0 || p.n_bisect > 32 || !(p.h_min_voxels > 0.0f) ||
      !(p.k_exact >= 0.0f))
    return set_err(ctx, RE_ERR_INVALID, "re_hybrid_trace: max_iters in [1,2^20], n_bisect in [0,32], h_min > 0");
  // A surrogate step is s - margin with s > tau: tau must exceed the margin
  // or the steps shrink towards zero.
  if (mlp_id >= 0 && !(p.tau_voxels > p.margin_voxels))
    return set_err(ctx, RE_ERR_INVALID, "re_hybrid_trace: tau_voxels must exceed margin_voxels");
  if (n == 0) return RE_OK;
  const DevVolume& v = it->second;
  const float vox = v.spacing[0];
  HybridParams hp{p.tau_voxels * vox, p.margin_voxels * vox, p.k_exact, p.h_min_voxels * vox, p.rejoin_voxels * vox,
                  p.n_bisect, p.max_iters};
  const bool dev = (flags & RE_DEVICE_PTRS) != 0;
  const float *d_o = ray_o, *d_d = ray_d;
  float* d_t = t;
  uint8_t* d_h = hit;
  int32_t *d_e = evals, *d_s = samples;
  if (!dev) {
    const size_t rb = (size_t)n * 12;
    re_status st = ensure_stage(ctx, stage_size({rb, rb, (size_t)n * 4, (size_t)n, (size_t)n * 4, (size_t)n * 4}));
    if (st) return st;
    StageCursor c{(char*)ctx->stage};
    float* so = c.take<float>(3 * n);
    float* sd = c.take<float>(3 * n);
    d_t = c.take<float>(n);
    d_h = c.take<uint8_t>(n);
    int32_t* se = c.take<int32_t>(n);
    int32_t* ss = c.take<int32_t>(n);
    d_e = evals ? se : nullptr;
    d_s = samples ? ss : nullptr;
    HIP_TRY(ctx, hipMemcpyAsync(so, ray_o, rb, hipMemcpyHostToDevice, ctx->stream), "copy ray_o");
    HIP_TRY(ctx, hipMemcpyAsync(sd, ray_d, rb, hipMemcpyHostToDevice, ctx->stream), "copy ray_d");
    d_o = so;
    d_d = sd;
  }
  HIP_TRY(ctx, ctx->launch_begin(), "event");
  HIP_TRY(ctx, launch_hybrid_trace(v, m ? infer_pack(*m) : nullptr, m ? infer_small(*m) : nullptr, d_o, d_d, n, hp, m != nullptr,
                                   d_t, d_h, d_e, d_s, ctx->stream), "hybrid trace");
  HIP_TRY(ctx, ctx->launch_end(), "event");
  if (!dev) {
    HIP_TRY(ctx, hipMemcpyAsync(t, d_t, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream), "copy t");
    HIP_TRY(ctx, hipMemcpyAsync(hit, d_h, (size_t)n, hipMemcpyDeviceToHost, ctx->stream), "copy hit");
    if (evals) HIP_TRY(ctx, hipMemcpyAsync(evals, d_e, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream), "copy evals");
    if (samples) HIP_TRY(ctx, hipMemcpyAsync(samples, d_s, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream), "copy samples");
  }
  if (!(flags & RE_ASYNC) || !dev) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "hybrid sync");
  return RE_OK;
}

RE_API void re_mlp_default_train_params(re_mlp_train_params* p) {
  if (!p) return;
  p->n_samples = (int64_t)1 << 22;
  p->batch = 65536;
  p->steps = 20000;
  p->lr = 1e-3f;
  p->beta1 = 0.9f;
  p->beta2 = 0.999f;
  p->eps = 1e-8f;
  p->over_weight = 4.0f;
  p->near_band_voxels = 2.0f;
  p->seed = 1;
}

// Trains on n_samples points drawn from the distance volume of vol_id:
// samples are generated once, then `steps` Adam steps each take the next
// `batch` consecutive samples (wrapping).  loss_history (nullable) gets the
// loss of every step.
RE_API re_status re_mlp_train(re_ctx* ctx, int32_t mlp_id, int32_t dist_vol_id, const re_mlp_train_params* params,
                              float* loss_history) {
  if (!ctx) return fail(RE_ERR_INVALID, "re_mlp_train: null ctx");
  re_mlp_train_params p;
  re_mlp_default_train_params(&p);
  if (params) p = *params;
  if (p.batch <= 0 || p.batch % 64 != 0 || p.n_samples < p.batch || p.steps < 0 || p.n_samples > ((int64_t)1 << 28))
    return set_err(ctx, RE_ERR_INVALID, "re_mlp_train: batch a multiple of 64, batch <= n_samples <= 2^28");
  DevMlp* m;
  re_status st = get_mlp(ctx, mlp_id, &m);
  if (st) return st;
  auto it = ctx->volumes.find(dist_vol_id);
  if (it == ctx->volumes.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "re_mlp_train: unknown volume");
  const DevVolume& dv = it->second;
  if (m->loss_cap < p.steps) {
    if (m->loss) (void)hipFree(m->loss);
    m->loss = nullptr;
    HIP_TRY(ctx, hipMalloc((void**)&m->loss, (size_t)(p.steps > 0 ? p.steps : 1) * 4), "loss history");
    m->loss_cap = p.steps;
  }
  // The pool is a pure function of (volume, seed, n_samples, band): a call
  // with the key of the last one continues at its batch cursor.
  uint32_t band_bits;
  std::memcpy(&band_bits, &p.near_band_voxels, 4);
  const bool resume = m->cursor >= 0 && m->pool_seed == p.seed && m->pool_n == p.n_samples &&
                      m->pool_batch == p.batch && m->pool_band_bits == band_bits;
  if (!resume) {
    m->pool_seed = p.seed;
    m->pool_n = p.n_samples;
    m->pool_batch = p.batch;
    m->pool_band_bits = band_bits;
    m->cursor = 0;
  }
  const int64_t first = m->cursor;
  float4* samples = nullptr;
  HIP_TRY(ctx, hipMalloc((void**)&samples, (size_t)p.n_samples * 16), "sample pool");
  hipError_t e = launch_sample(dv, p.n_samples, p.seed, p.near_band_voxels * dv.spacing[0], 256, samples, ctx->stream);
  for (int32_t s = 0; e == hipSuccess && s < p.steps; ++s) {
    const int64_t start = (int64_t)(((unsigned __int128)(first + s) * (uint64_t)p.batch) % (uint64_t)p.n_samples);
    e = launch_mlp_train(m->pack, m->small, samples, p.n_samples, start, p.batch, p.over_weight, m->partial,
                         m->partial_blocks, ctx->stream);
    if (e == hipSuccess) {
      ++m->step;
      e = launch_mlp_update(m->params, m->m1, m->m2, m->partial, m->partial_blocks, m->step, p.lr, p.beta1, p.beta2,
                            p.eps, 1, nullptr, m->pack, m->small, m->loss + s, ctx->stream);
    }
  }
  if (e == hipSuccess && loss_history && p.steps > 0)
    e = hipMemcpyAsync(loss_history, m->loss, (size_t)p.steps * 4, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(samples);
  if (e != hipSuccess) {
    m->cursor = -1;  // the state is not a clean continuation any more
    return hip_err(ctx, e, "re_mlp_train");
  }
  m->cursor = first + p.steps;
  return RE_OK;
}

// Checkpoint file: a 64-byte header, then params, Adam m and Adam v as
// little-endian float32 (param_count each).
struct MlpCheckpointHeader {
  char magic[8];        // "RESURR\0\1"
  int32_t n_params;
  int32_t step;
  uint64_t pool_seed;
  int64_t pool_n;
  int64_t cursor;
  int32_t pool_batch;
  uint32_t pool_band_bits;
  uint8_t reserved[16];
};
static_assert(sizeof(MlpCheckpointHeader) == 64, "checkpoint header");
static const char kMlpMagic[8] = {'R', 'E', 'S', 'U', 'R', 'R', 0, 1};

RE_API re_status re_mlp_save(re_ctx* ctx, int32_t mlp_id, const char* path) {
  if (!ctx || !path) return fail(RE_ERR_INVALID, "re_mlp_save: null argument");
  DevMlp* m;
  re_status st = get_mlp(ctx, mlp_id, &m);
  if (st) return st;
  return guarded(ctx, "re_mlp_save", [&]() -> re_status {
    std::vector<float> buf((size_t)kParams * 3);
    HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    HIP_TRY(ctx, hipMemcpyAsync(buf.data(), m->params, (size_t)kParams * 4, hipMemcpyDeviceToHost, ctx->stream),
            "read params");
    HIP_TRY(ctx, hipMemcpyAsync(buf.data() + kParams, m->m1, (size_t)kParams * 4, hipMemcpyDeviceToHost, ctx->stream),
            "read m");
    HIP_TRY(ctx,
            hipMemcpyAsync(buf.data() + 2 * kParams, m->m2, (size_t)kParams * 4, hipMemcpyDeviceToHost, ctx->stream),
            "read v");
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "read sync");
    MlpCheckpointHeader h{};
    std::memcpy(h.magic, kMlpMagic, 8);
    h.n_params = kParams;
    h.step = m->step;
    h.pool_seed = m->pool_seed;
    h.pool_n = m->pool_n;
    h.cursor = m->cursor;
    h.pool_batch = m->pool_batch;
    h.pool_band_bits = m->pool_band_bits;
    FILE* f = std::fopen(path, "wb");
    if (!f) return set_err(ctx, RE_ERR_IO, std::string("re_mlp_save: cannot open ") + path);
    bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 && std::fwrite(buf.data(), 4, buf.size(), f) == buf.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return set_err(ctx, RE_ERR_IO, std::string("re_mlp_save: write failed: ") + path);
    return RE_OK;
  });
}

RE_API re_status re_mlp_load(re_ctx* ctx, const char* path, int32_t* mlp_id) {
  if (!ctx || !path || !mlp_id) return fail(RE_ERR_INVALID, "re_mlp_load: null argument");
  return guarded(ctx, "re_mlp_load", [&]() -> re_status {
    FILE* f = std::fopen(path, "rb");
    if (!f) return set_err(ctx, RE_ERR_IO, std::string("re_mlp_load: cannot open ") + path);
    MlpCheckpointHeader h{};
    std::vector<float> buf((size_t)kParams * 3);
    bool ok = std::fread(&h, sizeof(h), 1, f) == 1 && std::memcmp(h.magic, kMlpMagic, 8) == 0 &&
              h.n_params == kParams && h.step >= 0 && std::fread(buf.data(), 4, buf.size(), f) == buf.size();
    ok = ok && std::fgetc(f) == EOF;  // nothing after the three vectors
    std::fclose(f);
    if (!ok)
      return set_err(ctx, RE_ERR_IO,
                     std::string("re_mlp_load: not a checkpoint of this network (header, size or length): ") + path);
    int32_t id = -1;
    re_status st = re_mlp_create(ctx, 0, &id);
    if (st) return st;
    DevMlp& m = ctx->mlps[id];
    hipError_t e = hipMemcpyAsync(m.params, buf.data(), (size_t)kParams * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(m.m1, buf.data() + kParams, (size_t)kParams * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(m.m2, buf.data() + 2 * kParams, (size_t)kParams * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) {
      (void)re_mlp_release(ctx, id);
      return hip_err(ctx, e, "re_mlp_load");
    }
    m.step = h.step;
    m.pool_seed = h.pool_seed;
    m.pool_n = h.pool_n;
    m.cursor = h.cursor;
    m.pool_batch = h.pool_batch;
    m.pool_band_bits = h.pool_band_bits;
    st = repack(ctx, m);
    if (st == RE_OK) {
      e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess) st = hip_err(ctx, e, "re_mlp_load sync");
    }
    if (st) {
      (void)re_mlp_release(ctx, id);
      return st;
    }
    *mlp_id = id;
    return RE_OK;
  });
}

}  // extern "C"
