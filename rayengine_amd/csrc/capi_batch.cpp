// C-ABI entry points of the batched mixed workload: one device work queue
// over many triangle scenes and terrain tiles, its launch sequence captured
// once into a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "capi_util.hpp"

using namespace re;

namespace {

// Hybrid / exact-march distances are given in voxels of each tile and scaled
// by that tile's own voxel size, as re_hybrid_trace does for one volume.
HybridParams tile_params(const re_hybrid_params& p, const DevVolume& v) {
  const float vox = v.spacing[0];
  return HybridParams{p.tau_voxels * vox, p.margin_voxels * vox, p.k_exact, p.h_min_voxels * vox,
                      p.rejoin_voxels * vox, p.n_bisect, p.max_iters};
}

void free_batch(DevBatch& b) { release_batch(b); }

}  // namespace

void re::release_batch(DevBatch& b) {
  if (b.exec) (void)hipGraphExecDestroy(b.exec);
  if (b.graph) (void)hipGraphDestroy(b.graph);
  if (b.ev_fork) (void)hipEventDestroy(b.ev_fork);
  if (b.ev_join) (void)hipEventDestroy(b.ev_join);
  if (b.side) (void)hipStreamDestroy(b.side);
  for (void* p : {(void*)b.arena.nodes, (void*)b.arena.leaves, (void*)b.arena.tris, (void*)b.arena.objs,
                  (void*)b.tiles, (void*)b.ray_o, (void*)b.ray_d, (void*)b.prim, (void*)b.t, (void*)b.hit_scratch,
                  (void*)b.items, (void*)b.ctl, (void*)b.remaining, (void*)b.remaining_init, (void*)b.done_ts})
    if (p) (void)hipFree(p);
  b = DevBatch{};
}

extern "C" {

RE_API void re_batch_default_config(re_batch_config* c) {
  if (!c) return;
  c->item_rays = 256;
  c->block = 256;
  c->mlp_id = -1;
  re_trace_default_params(&c->trace);
  re_sdf_default_params(&c->sdf);
  re_hybrid_default_params(&c->hybrid);
  re_hybrid_default_params(&c->march);
  c->side_share = 0.0f;
}

RE_API re_status re_batch_create(re_ctx* ctx, const re_batch_scene* scenes, int32_t n_scenes,
                                 const re_batch_tile* tiles, int32_t n_tiles, const float* ray_o, const float* ray_d,
                                 int64_t n_rays, const re_batch_config* config, int32_t* batch_id) {
  return guarded(ctx, "re_batch_create", [&]() -> re_status {
  if (!ctx || !batch_id || n_scenes < 0 || n_tiles < 0 || n_scenes + n_tiles <= 0 || n_rays <= 0 || !ray_o || !ray_d ||
      (n_scenes > 0 && !scenes) || (n_tiles > 0 && !tiles))
    return fail(RE_ERR_INVALID, "re_batch_create: bad arguments");
  re_batch_config cfg;
  re_batch_default_config(&cfg);
  if (config) cfg = *config;
  if (cfg.item_rays < 1 || cfg.item_rays > (1 << 16) || !(cfg.block == 64 || cfg.block == 128 || cfg.block == 256) ||
      !(cfg.side_share >= 0.0f && cfg.side_share <= 0.9f))
    return set_err(ctx, RE_ERR_INVALID,
                   "re_batch_create: item_rays in [1,65536], block in {64,128,256}, side_share in [0,0.9]");
  for (int32_t k = 0; k < n_scenes; ++k)
    if (!scenes[k].tree || scenes[k].ray_begin < 0 || scenes[k].ray_count < 0 ||
        scenes[k].ray_count > n_rays - scenes[k].ray_begin)
      return set_err(ctx, RE_ERR_INVALID, "re_batch_create: scene ray range outside the ray arrays");
  bool any_surrogate = false;
  for (int32_t k = 0; k < n_tiles; ++k) {
    const re_batch_tile& tl = tiles[k];
    if (tl.ray_begin < 0 || tl.ray_count < 0 || tl.ray_count > n_rays - tl.ray_begin || tl.mode < 0 || tl.mode > 2)
      return set_err(ctx, RE_ERR_INVALID, "re_batch_create: tile ray range or mode invalid");
    if (ctx->volumes.find(tl.vol_id) == ctx->volumes.end())
      return set_err(ctx, RE_ERR_NOT_FOUND, "re_batch_create: unknown tile volume");
    if (tl.mode == 2) {
      any_surrogate = true;
      const int32_t mid = tl.mlp_id >= 0 ? tl.mlp_id : cfg.mlp_id;
      if (ctx->mlps.find(mid) == ctx->mlps.end())
        return set_err(ctx, RE_ERR_NOT_FOUND, "re_batch_create: a surrogate tile needs a valid mlp_id");
    }
  }
  // The per-ray loops are bounded by these parameters: the same limits as the
  // single-workload entry points (re_trace_closest, re_sdf_trace, re_hybrid_trace).
  const re_sdf_params& sp = cfg.sdf;
  const re_hybrid_params& hy = cfg.hybrid;
  const re_hybrid_params& mp = cfg.march;
  if (sp.max_steps < 1 || sp.max_steps > (1 << 20) || !(sp.step_rule == 0 || sp.step_rule == 1))
    return set_err(ctx, RE_ERR_INVALID, "re_batch_create: sdf max_steps in [1,2^20], step_rule 0/1");
  for (const re_hybrid_params* q : {&hy, &mp})
    if (q->max_iters < 1 || q->max_iters > (1 << 20) || q->n_bisect < 0 || q->n_bisect > 32 ||
        !(q->h_min_voxels > 0.0f) || !(q->k_exact >= 0.0f))
      return set_err(ctx, RE_ERR_INVALID,
                     "re_batch_create: hybrid / march max_iters in [1,2^20], n_bisect in [0,32], h_min > 0");
  if (any_surrogate && !(hy.tau_voxels > hy.margin_voxels))
    return set_err(ctx, RE_ERR_INVALID, "re_batch_create: tau_voxels must exceed margin_voxels");
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");

  DevBatch b;
  b.n_scenes = n_scenes;
  b.n_tiles = n_tiles;
  b.n_slots = n_scenes + n_tiles;
  b.n_rays = n_rays;
  b.mlp_id = cfg.mlp_id;
  // Scenes: one arena with one object per scene (identity transform).
  if (n_scenes > 0) {
    std::vector<re_object> objs(n_scenes);
    for (int32_t k = 0; k < n_scenes; ++k) {
      re_object& o = objs[k];
      o = re_object{};
      o.tree = scenes[k].tree;
      o.cos_rot_y = 1.0f;
    }
    int32_t sid = 0;
    re_status st = re_scene_upload(ctx, objs.data(), n_scenes, &sid);
    if (st) return st;
    b.arena = ctx->scenes[sid];
    ctx->scenes.erase(sid);
  }
  const size_t rb = (size_t)n_rays * 12;
  std::vector<BatchTile> htiles(n_tiles);
  for (int32_t k = 0; k < n_tiles; ++k) {
    const DevVolume& v = ctx->volumes[tiles[k].vol_id];
    htiles[k].vol = v.view();
    for (int a = 0; a < 3; ++a) {
      htiles[k].box_min[a] = v.box_min[a];
      htiles[k].box_max[a] = v.box_max[a];
    }
    htiles[k].hp = tile_params(mp, v);
    b.tile_vol.push_back(tiles[k].vol_id);
    b.tile_mode.push_back(tiles[k].mode);
  }
  // Work items, targets in ascending ray count (shortest first).
  BatchPlan plan;
  {
    const re_status st = plan_batch(scenes, n_scenes, tiles, n_tiles, n_rays, cfg.item_rays, plan);
    if (st) {
      free_batch(b);
      const std::string msg = std::string("re_batch_create: ") + g_last_error;  // copied before set_err rewrites it
      return set_err(ctx, st, msg);
    }
  }
  for (int32_t s : plan.surrogate_slots) {
    const re_batch_tile& tl = tiles[s - n_scenes];
    b.surrogate.push_back(SurrogateTile{s - n_scenes, s, tl.ray_begin, tl.ray_count, tl.mlp_id >= 0 ? tl.mlp_id : cfg.mlp_id});
  }
  const std::vector<BatchItem>& items = plan.items;
  const std::vector<int32_t>& rem = plan.remaining;
  b.n_items = (int64_t)items.size();
  hipError_t e = hipSuccess;
  auto alloc = [&](void** p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(p, bytes > 0 ? bytes : 64);
  };
  alloc((void**)&b.tiles, htiles.size() * sizeof(BatchTile));
  alloc((void**)&b.ray_o, rb);
  alloc((void**)&b.ray_d, rb);
  alloc((void**)&b.prim, (size_t)n_rays * 4);
  alloc((void**)&b.t, (size_t)n_rays * 4);
  alloc((void**)&b.hit_scratch, (size_t)n_rays);
  alloc((void**)&b.items, items.size() * sizeof(BatchItem));
  alloc((void**)&b.ctl, 64);
  alloc((void**)&b.remaining, (size_t)b.n_slots * 4);
  alloc((void**)&b.remaining_init, (size_t)b.n_slots * 4);
  alloc((void**)&b.done_ts, (size_t)b.n_slots * 8);
  if (e == hipSuccess && !htiles.empty())
    e = hipMemcpyAsync(b.tiles, htiles.data(), htiles.size() * sizeof(BatchTile), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(b.ray_o, ray_o, rb, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(b.ray_d, ray_d, rb, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess && !items.empty())
    e = hipMemcpyAsync(b.items, items.data(), items.size() * sizeof(BatchItem), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(b.remaining_init, rem.data(), (size_t)b.n_slots * 4, hipMemcpyHostToDevice, ctx->stream);
  // Rays no item covers keep the miss result.
  std::vector<int32_t> miss_p((size_t)n_rays, -1);
  std::vector<float> miss_t((size_t)n_rays, kInf);
  if (e == hipSuccess) e = hipMemcpyAsync(b.prim, miss_p.data(), (size_t)n_rays * 4, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(b.t, miss_t.data(), (size_t)n_rays * 4, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    free_batch(b);
    return hip_err(ctx, e, "re_batch_create");
  }
  int rate_khz = 0;
  if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ctx->device) == hipSuccess && rate_khz > 0)
    b.wall_clock_khz = (double)rate_khz;

  BatchLaunch& L = b.launch;
  L.nodes = b.arena.nodes;
  L.leaves = b.arena.leaves;
  L.tris = b.arena.tris;
  L.objs = b.arena.objs;
  L.tiles = b.tiles;
  L.items = b.items;
  L.n_items = b.n_items;
  L.ray_o = b.ray_o;
  L.ray_d = b.ray_d;
  L.prim = b.prim;
  L.t = b.t;
  L.ctl = b.ctl;
  L.remaining = b.remaining;
  L.remaining_init = b.remaining_init;
  L.done_ts = b.done_ts;
  L.n_slots = b.n_slots;
  L.cap = (cfg.trace.max_leaf_visits > 0 && cfg.trace.max_leaf_visits < kMaxLeafVisits) ? cfg.trace.max_leaf_visits
                                                                                          : kMaxLeafVisits;
  L.one_sided = cfg.trace.one_sided;
  L.secondary = cfg.trace.secondary;
  L.sdf = BatchSdf{cfg.sdf.max_steps, cfg.sdf.eps, cfg.sdf.step_scale, cfg.sdf.step_rule};
  L.block = cfg.block;
  L.grid_blocks = batch_grid(cfg.block);
  if (!b.surrogate.empty() && cfg.side_share > 0.0f) {
    // leave side_share of the slots to the surrogate branch; the persistent
    // kernel's items are taken from a counter, so fewer blocks only changes
    // who runs them
    const int64_t keep = (int64_t)((double)L.grid_blocks * (1.0 - (double)cfg.side_share));
    L.grid_blocks = keep > 0 ? keep : 1;
  }

  // Capture once: the reset, then two branches - the surrogate tiles' hybrid
  // launches (largest first) on a side stream forked from the capture, and
  // the persistent kernel on the context's stream - joined at the end.  The
  // surrogate branch is captured first so its blocks are issued first; the
  // persistent kernel fills the rest of the device and, its items being
  // taken from a counter, absorbs whatever share the hybrid launches leave.
  const bool fork = !b.surrogate.empty();
  if (fork) {
    e = hipStreamCreateWithFlags(&b.side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b.ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b.ev_join, hipEventDisableTiming);
    std::stable_sort(b.surrogate.begin(), b.surrogate.end(),
                     [](const SurrogateTile& x, const SurrogateTile& y) { return x.count > y.count; });
  }
  if (e == hipSuccess) e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) e = launch_batch_reset(L, ctx->stream);
  if (e == hipSuccess && fork) e = hipEventRecord(b.ev_fork, ctx->stream);
  if (e == hipSuccess && fork) e = hipStreamWaitEvent(b.side, b.ev_fork, 0);
  for (const SurrogateTile& st : b.surrogate) {
    if (e != hipSuccess) break;
    const DevVolume& v = ctx->volumes[b.tile_vol[st.tile]];
    const DevMlp& m = ctx->mlps[st.mlp_id];
    const HybridParams hp = tile_params(hy, v);
    e = launch_hybrid_trace(v, infer_pack(m), infer_small(m), b.ray_o + 3 * st.begin, b.ray_d + 3 * st.begin, st.count,
                            hp, true, b.t + st.begin, b.hit_scratch + st.begin, nullptr, nullptr, b.side);
    if (e == hipSuccess) e = launch_hit_to_prim(b.hit_scratch + st.begin, b.prim + st.begin, st.count, b.side);
    if (e == hipSuccess) e = launch_batch_stamp(b.done_ts, st.slot, b.remaining, b.side);
  }
  if (e == hipSuccess && fork) e = hipEventRecord(b.ev_join, b.side);
  if (e == hipSuccess) e = launch_batch(L, ctx->stream);
  if (e == hipSuccess && fork) e = hipStreamWaitEvent(ctx->stream, b.ev_join, 0);
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
  if (e == hipSuccess) e = e2;
  b.graph = g;
  if (e == hipSuccess) e = hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    free_batch(b);
    return hip_err(ctx, e, "re_batch_create graph capture");
  }
  int32_t id = ctx->next_id++;
  ctx->batches[id] = b;
  *batch_id = id;
  return RE_OK;
  });
}

RE_API re_status re_batch_run(re_ctx* ctx, int32_t batch_id, uint32_t flags) {
  if (!ctx) return fail(RE_ERR_INVALID, "re_batch_run: null ctx");
  auto it = ctx->batches.find(batch_id);
  if (it == ctx->batches.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "re_batch_run: unknown batch");
  HIP_TRY(ctx, ctx->launch_begin(), "event");
  HIP_TRY(ctx, hipGraphLaunch(it->second.exec, ctx->stream), "graph launch");
  HIP_TRY(ctx, ctx->launch_end(), "event");
  if (!(flags & RE_ASYNC)) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "batch sync");
  return RE_OK;
}

RE_API re_status re_batch_info(re_ctx* ctx, int32_t batch_id, int64_t* n_rays, int64_t* n_items, int32_t* n_slots) {
  if (!ctx) return fail(RE_ERR_INVALID, "re_batch_info: null ctx");
  auto it = ctx->batches.find(batch_id);
  if (it == ctx->batches.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "re_batch_info: unknown batch");
  if (n_rays) *n_rays = it->second.n_rays;
  if (n_items) *n_items = it->second.n_items;
  if (n_slots) *n_slots = it->second.n_slots;
  return RE_OK;
}

RE_API re_status re_batch_results(re_ctx* ctx, int32_t batch_id, int32_t* prim, float* t) {
  if (!ctx || !prim || !t) return fail(RE_ERR_INVALID, "re_batch_results: null argument");
  auto it = ctx->batches.find(batch_id);
  if (it == ctx->batches.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "re_batch_results: unknown batch");
  const DevBatch& b = it->second;
  HIP_TRY(ctx, hipMemcpyAsync(prim, b.prim, (size_t)b.n_rays * 4, hipMemcpyDeviceToHost, ctx->stream), "copy prim");
  HIP_TRY(ctx, hipMemcpyAsync(t, b.t, (size_t)b.n_rays * 4, hipMemcpyDeviceToHost, ctx->stream), "copy t");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "results sync");
  return RE_OK;
}

RE_API re_status re_batch_completion_ms(re_ctx* ctx, int32_t batch_id, float* ms) {
  if (!ctx || !ms) return fail(RE_ERR_INVALID, "re_batch_completion_ms: null argument");
  auto it = ctx->batches.find(batch_id);
  if (it == ctx->batches.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "re_batch_completion_ms: unknown batch");
  const DevBatch& b = it->second;
  std::vector<unsigned long long> ts((size_t)b.n_slots), ctl(8);
  HIP_TRY(ctx, hipMemcpyAsync(ts.data(), b.done_ts, (size_t)b.n_slots * 8, hipMemcpyDeviceToHost, ctx->stream), "copy ts");
  HIP_TRY(ctx, hipMemcpyAsync(ctl.data(), b.ctl, 64, hipMemcpyDeviceToHost, ctx->stream), "copy ctl");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "completion sync");
  for (int32_t s = 0; s < b.n_slots; ++s)
    ms[s] = ts[s] >= ctl[1] ? (float)((double)(ts[s] - ctl[1]) / b.wall_clock_khz) : 0.0f;
  return RE_OK;
}

RE_API re_status re_batch_release(re_ctx* ctx, int32_t batch_id) {
  if (!ctx) return fail(RE_ERR_INVALID, "re_batch_release: null ctx");
  auto it = ctx->batches.find(batch_id);
  if (it == ctx->batches.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "re_batch_release: unknown batch");
  (void)hipStreamSynchronize(ctx->stream);
  free_batch(it->second);
  ctx->batches.erase(it);
  return RE_OK;
}

}  // extern "C"
