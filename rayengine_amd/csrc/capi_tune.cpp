// C-ABI autotuner of the single-launch closest-hit query: kd-tree leaf size,
// work-group size and kernel variant, chosen by timing the caller's rays on
// the context's device.  No counterpart in the reference, whose launch shape
// was fixed at compile time (renderer.cu:389-410).
#include <hip/hip_runtime.h>

#include <vector>

#include "capi_util.hpp"

using namespace re;

namespace {

constexpr int32_t kLeafDefault[] = {4, 8, 12, 16};
constexpr int32_t kBlockDefault[] = {64, 128, 256};
constexpr int32_t kVariantDefault[] = {3, 4, 5};
constexpr int32_t kMaxReps = 64;  // below the context's ring of launch events

struct Candidates {
  std::vector<int32_t> leaf, block, variant;
  int32_t reps;
};

template <size_t N>
std::vector<int32_t> pick(const int32_t* p, int32_t n, const int32_t (&dflt)[N]) {
  if (p && n > 0) return std::vector<int32_t>(p, p + n);
  return std::vector<int32_t>(dflt, dflt + N);
}

// Device buffers and the tree/scene of one leaf-size round, released on every
// path out of the tuner.
struct TuneState {
  re_ctx* ctx;
  void* rays = nullptr;
  void* out = nullptr;
  re_kdtree* tree = nullptr;
  int32_t scene = -1;
  void drop_scene() {
    if (scene >= 0) (void)re_scene_release(ctx, scene);
    scene = -1;
    if (tree) re_kd_free(tree);
    tree = nullptr;
  }
  ~TuneState() {
    drop_scene();
    if (rays) (void)hipFree(rays);
    if (out) (void)hipFree(out);
  }
};

}  // namespace

extern "C" {

RE_API re_status re_autotune_trace(re_ctx* ctx, const float* tri_v, int64_t n_tris, const re_kd_build_params* build,
                                   const float* ray_o, const float* ray_d, int64_t n, const re_trace_params* base,
                                   const re_tune_space* space, uint32_t flags, re_tune_result* best, float* ms_all) {
  if (!ctx) return fail(RE_ERR_INVALID, "re_autotune_trace: null ctx");
  if (!tri_v || n_tris <= 0 || !ray_o || !ray_d || n <= 0 || !best)
    return set_err(ctx, RE_ERR_INVALID, "re_autotune_trace: bad arguments");
  return guarded(ctx, "re_autotune_trace", [&]() -> re_status {
    Candidates c;
    c.leaf = pick(space ? space->leaf_sizes : nullptr, space ? space->n_leaf_sizes : 0, kLeafDefault);
    c.block = pick(space ? space->blocks : nullptr, space ? space->n_blocks : 0, kBlockDefault);
    c.variant = pick(space ? space->variants : nullptr, space ? space->n_variants : 0, kVariantDefault);
    c.reps = space && space->reps > 0 ? space->reps : 10;
    if (c.reps > kMaxReps) return set_err(ctx, RE_ERR_INVALID, "re_autotune_trace: reps must be at most 64");
    for (int32_t v : c.leaf)
      if (v < 1) return set_err(ctx, RE_ERR_INVALID, "re_autotune_trace: leaf sizes must be >= 1");
    for (int32_t v : c.block)
      if (v != 64 && v != 128 && v != 256) return set_err(ctx, RE_ERR_INVALID, "re_autotune_trace: blocks are 64, 128 or 256");
    for (int32_t v : c.variant)
      if (v < 0 || v > 5) return set_err(ctx, RE_ERR_INVALID, "re_autotune_trace: variants are 0..5");

    re_trace_params p0;
    re_trace_default_params(&p0);
    if (base) p0 = *base;
    re_kd_build_params b0;
    re_kd_default_params(&b0);
    if (build) b0 = *build;

    HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    TuneState S{ctx};
    const size_t ray_bytes = (size_t)n * 12;
    const float* d_o = ray_o;
    const float* d_d = ray_d;
    if (!(flags & RE_DEVICE_PTRS)) {
      HIP_TRY(ctx, hipMalloc(&S.rays, 2 * ray_bytes), "re_autotune_trace: ray buffer");
      HIP_TRY(ctx, hipMemcpyAsync(S.rays, ray_o, ray_bytes, hipMemcpyHostToDevice, ctx->stream), "copy ray_o");
      HIP_TRY(ctx, hipMemcpyAsync((char*)S.rays + ray_bytes, ray_d, ray_bytes, hipMemcpyHostToDevice, ctx->stream),
              "copy ray_d");
      d_o = (const float*)S.rays;
      d_d = (const float*)((char*)S.rays + ray_bytes);
    }
    HIP_TRY(ctx, hipMalloc(&S.out, (size_t)n * 8), "re_autotune_trace: result buffer");
    int32_t* d_prim = (int32_t*)S.out;
    float* d_t = (float*)((char*)S.out + (size_t)n * 4);

    std::vector<float> times((size_t)c.reps);
    bool have = false;
    size_t k = 0;
    for (int32_t leaf : c.leaf) {
      re_kd_build_params bp = b0;
      bp.max_leaf_tris = leaf;
      re_status st = re_kd_build_gpu(ctx, tri_v, n_tris, &bp, &S.tree, nullptr);
      if (st) return st;
      re_object ob{};
      ob.tree = S.tree;
      ob.cos_rot_y = 1.0f;
      st = re_scene_upload(ctx, &ob, 1, &S.scene);
      if (st) return st;
      for (int32_t blk : c.block) {
        for (int32_t var : c.variant) {
          re_trace_params p = p0;
          p.block = blk;
          p.variant = var;
          // one untimed launch, then reps launches back to back on the stream
          st = re_trace_closest(ctx, S.scene, d_o, d_d, n, &p, d_prim, d_t, nullptr, RE_DEVICE_PTRS);
          if (st) return st;
          for (int32_t r = 0; r < c.reps; ++r) {
            st = re_trace_closest(ctx, S.scene, d_o, d_d, n, &p, d_prim, d_t, nullptr, RE_DEVICE_PTRS | RE_ASYNC);
            if (st) return st;
          }
          HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "re_autotune_trace: sync");
          int32_t got = 0;
          st = re_kernel_times(ctx, times.data(), c.reps, &got);
          if (st) return st;
          if (got != c.reps) return set_err(ctx, RE_ERR_HIP, "re_autotune_trace: launch timings missing");
          double sum = 0.0;
          for (float v : times) sum += v;
          const float ms = (float)(sum / c.reps);
          if (ms_all) ms_all[k] = ms;
          ++k;
          if (!have || ms < best->ms) {  // the first minimum wins
            best->max_leaf_tris = leaf;
            best->block = blk;
            best->variant = var;
            best->ms = ms;
            have = true;
          }
        }
      }
      S.drop_scene();
    }
    return RE_OK;
  });
}

}  // extern "C"
