// C-ABI implementation: context, device residency of scenes, and the
// closest-hit entry point.  All device work is ordered on the context's
// stream; errors come back as re_status with a message, never as exit().
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>

#include "capi_util.hpp"

using namespace re;

namespace {

void free_scene(DevScene& s) {
  (void)hipFree(s.nodes);
  (void)hipFree(s.leaves);
  (void)hipFree(s.tris);
  (void)hipFree(s.objs);
  if (s.attr_geom) (void)hipFree(s.attr_geom);
  if (s.attr_nrm) (void)hipFree(s.attr_nrm);
  if (s.attr_uv) (void)hipFree(s.attr_uv);
  s = DevScene{};
}

}  // namespace

extern "C" {

RE_API const char* re_version(void) { return "rayengine-mi355x 0.1 (gfx950)"; }

RE_API re_status re_create(int device, re_ctx** out) {
  if (!out) return fail(RE_ERR_INVALID, "re_create: out is null");
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) return fail(RE_ERR_UNSUPPORTED, "re_create: no HIP device (%s)", hipGetErrorString(e));
  if (device < 0 || device >= n) return fail(RE_ERR_INVALID, "re_create: device %d of %d", device, n);
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) return fail(RE_ERR_HIP, "re_create: %s", hipGetErrorString(e));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(RE_ERR_UNSUPPORTED, "re_create: device is %s, library is built for gfx950", prop.gcnArchName);
  re_ctx* ctx = new (std::nothrow) re_ctx();
  if (!ctx) return fail(RE_ERR_OOM, "re_create: out of host memory");
  ctx->device = device;
  if ((e = hipSetDevice(device)) == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  for (int i = 0; e == hipSuccess && i < 2 * re_ctx::kRing; ++i) e = hipEventCreate(&ctx->ring[i]);
  if (e == hipSuccess) e = hipMalloc(&ctx->counts_buf, kCountsBytes);
  if (e != hipSuccess) {
    re_status st = fail(RE_ERR_HIP, "re_create: %s", hipGetErrorString(e));
    re_destroy(ctx);
    return st;
  }
  // the device address of the fixed area of re_trace_closest_noalloc on this device
  if (re_status st = noalloc_setup(ctx)) {
    re_destroy(ctx);
    return st;
  }
  *out = ctx;
  return RE_OK;
}

RE_API void re_destroy(re_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto& kv : ctx->scenes) free_scene(kv.second);
  for (auto& kv : ctx->volumes) {
    (void)hipFree(kv.second.data);
    if (kv.second.mat) (void)hipFree(kv.second.mat);
  }
  for (auto& kv : ctx->batches) release_batch(kv.second);
  for (auto& kv : ctx->mlps)
    for (void* p : {(void*)kv.second.params, (void*)kv.second.m1, (void*)kv.second.m2, kv.second.pack,
                    (void*)kv.second.small, (void*)kv.second.partial, (void*)kv.second.grad, (void*)kv.second.loss})
      if (p) (void)hipFree(p);
  if (ctx->stage) (void)hipFree(ctx->stage);
  if (ctx->counts_buf) (void)hipFree(ctx->counts_buf);
  for (int i = 0; i < 2 * re_ctx::kRing; ++i)
    if (ctx->ring[i]) (void)hipEventDestroy(ctx->ring[i]);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

RE_API const char* re_last_error(const re_ctx* ctx) {
  if (ctx && !ctx->err.empty()) return ctx->err.c_str();
  return g_last_error;
}

RE_API re_status re_scene_upload(re_ctx* ctx, const re_object* objs, int32_t n_obj, int32_t* scene_id) {
  return guarded(ctx, "re_scene_upload", [&]() -> re_status {
  if (!ctx || !objs || n_obj <= 0 || !scene_id) return set_err(ctx, RE_ERR_INVALID, "re_scene_upload: bad arguments");
  for (int32_t k = 0; k < n_obj; ++k)
    if (!objs[k].tree) return set_err(ctx, RE_ERR_INVALID, "re_scene_upload: object without a tree");
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  // Concatenate, offsetting entry indices, leaf ids, triangle indices and
  // prim ids per object (RayEngine/src/World/RTScene.cpp:263-279).
  std::vector<NodeEntry> nodes;
  std::vector<LeafRec> leaves;
  std::vector<TriRec> tris;
  std::vector<ObjXform> xf(n_obj);
  std::vector<size_t> leaf_end(n_obj);  // one past each object's last leaf
  for (int32_t k = 0; k < n_obj; ++k) {
    const re_object& ob = objs[k];
    const auto& tn = kd_nodes(ob.tree);
    const auto& tl = kd_leaves(ob.tree);
    const auto& tt = kd_tris(ob.tree);
    re_kd_info info;
    re_kd_get_info(ob.tree, &info);
    const uint32_t nb = (uint32_t)nodes.size(), lb = (uint32_t)leaves.size(), tb = (uint32_t)tris.size();
    if ((uint64_t)nb + tn.size() >= 0x3fffffffull || (uint64_t)lb + tl.size() >= 0x3fffffffull)
      return set_err(ctx, RE_ERR_INVALID, "re_scene_upload: scene too large for 30-bit node indices");
    for (NodeEntry e : tn) {
      if ((e.meta >> 30) == 3u) e.meta = kLeafMeta | ((e.meta & 0x3fffffffu) + lb);
      else e.meta = (e.meta & 0xc0000000u) | ((e.meta & 0x3fffffffu) + nb);
      nodes.push_back(e);
    }
    for (LeafRec l : tl) {
      for (int f = 0; f < 6; ++f) {
        int32_t r = l.rope[f];
        if (r == kNoRope) continue;
        l.rope[f] = ref_is_leaf(r) ? leaf_ref(ref_leaf_id(r) + lb) : r + (int32_t)nb;
      }
      l.first += tb;
      leaves.push_back(l);
    }
    leaf_end[k] = leaves.size();
    for (TriRec t : tt) {
      t.prim += ob.prim_offset;
      tris.push_back(t);
    }
    ObjXform& x = xf[k];
    for (int a = 0; a < 3; ++a) {
      x.loc[a] = ob.location[a];
      x.root_min[a] = info.root_min[a];
      x.root_max[a] = info.root_max[a];
    }
    x.sin_y = ob.sin_rot_y;
    x.cos_y = ob.cos_rot_y;
    x.is_character = ob.is_character;
    x.root = (int32_t)nb;
  }
  DevScene s;
  s.n_obj = n_obj;
  s.n_nodes = (int64_t)nodes.size();
  s.n_leaves = (int64_t)leaves.size();
  s.n_trirecs = (int64_t)tris.size();
  // Trailing pad entries (leaf entries of leaf 0) for window reads at the end.
  nodes.resize(nodes.size() + kNodePad, NodeEntry{0.0f, kLeafMeta});
  std::vector<DevLeaf> dleaves(leaves.size());
  // pad[0] of a device leaf record holds its object's index (the walk-start
  // hint check reads it)
  for (int32_t k = 0, i = 0; k < n_obj; ++k)
    for (; (size_t)i < leaf_end[k]; ++i) {
      dleaves[i] = dev_leaf(leaves[i]);
      dleaves[i].pad[0] = (uint32_t)k;
    }
  auto up =[&](void** dst, const void* src, size_t bytes, const char* what) -> re_status {
    if (bytes == 0) bytes = 64;  // keep every pointer valid
    hipError_t e = hipMalloc(dst, bytes);
    if (e != hipSuccess) return hip_err(ctx, e, what);
    if (src) {
      e = hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, ctx->stream);
      if (e != hipSuccess) return hip_err(ctx, e, what);
    }
    s.bytes += (int64_t)bytes;
    return RE_OK;
  };
  re_status st;
  if ((st = up((void**)&s.nodes, nodes.empty() ? nullptr : nodes.data(), nodes.size() * sizeof(NodeEntry), "upload nodes")) ||
      (st = up((void**)&s.leaves, dleaves.empty() ? nullptr : dleaves.data(), dleaves.size() * sizeof(DevLeaf), "upload leaves")) ||
      (st = up((void**)&s.tris, tris.empty() ? nullptr : tris.data(), tris.size() * sizeof(TriRec), "upload tris")) ||
      (st = up((void**)&s.objs, xf.data(), xf.size() * sizeof(ObjXform), "upload objects"))) {
    (void)hipStreamSynchronize(ctx->stream);
    free_scene(s);
    return st;
  }
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    free_scene(s);
    return hip_err(ctx, e, "re_scene_upload sync");
  }
  s.h_leaves = std::move(leaves);
  int32_t id = ctx->next_id++;
  ctx->scenes[id] = std::move(s);
  *scene_id = id;
  return RE_OK;
  });
}

RE_API re_status re_scene_update_objects(re_ctx* ctx, int32_t scene_id, const re_object* objs, int32_t n_obj) {
  if (!ctx || !objs) return fail(RE_ERR_INVALID, "re_scene_update_objects: null argument");
  auto it = ctx->scenes.find(scene_id);
  if (it == ctx->scenes.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "re_scene_update_objects: unknown scene");
  DevScene& s = it->second;
  if (n_obj != s.n_obj) return set_err(ctx, RE_ERR_INVALID, "re_scene_update_objects: object count differs from the upload");
  std::vector<ObjXform> xf((size_t)n_obj);
  HIP_TRY(ctx, hipMemcpyAsync(xf.data(), s.objs, xf.size() * sizeof(ObjXform), hipMemcpyDeviceToHost, ctx->stream), "read objects");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "read objects sync");
  for (int32_t k = 0; k < n_obj; ++k) {
    for (int a = 0; a < 3; ++a) xf[k].loc[a] = objs[k].location[a];
    xf[k].sin_y = objs[k].sin_rot_y;
    xf[k].cos_y = objs[k].cos_rot_y;
    xf[k].is_character = objs[k].is_character;
  }
  HIP_TRY(ctx, hipMemcpyAsync(s.objs, xf.data(), xf.size() * sizeof(ObjXform), hipMemcpyHostToDevice, ctx->stream), "write objects");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "write objects sync");
  return RE_OK;
}

RE_API re_status re_scene_release(re_ctx* ctx, int32_t scene_id) {
  if (!ctx) return fail(RE_ERR_INVALID, "re_scene_release: null ctx");
  auto it = ctx->scenes.find(scene_id);
  if (it == ctx->scenes.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "re_scene_release: unknown scene");
  (void)hipStreamSynchronize(ctx->stream);
  free_scene(it->second);
  ctx->scenes.erase(it);
  return RE_OK;
}

RE_API re_status re_scene_bytes(re_ctx* ctx, int32_t scene_id, int64_t* bytes) {
  if (!ctx || !bytes) return fail(RE_ERR_INVALID, "re_scene_bytes: null argument");
  auto it = ctx->scenes.find(scene_id);
  if (it == ctx->scenes.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "re_scene_bytes: unknown scene");
  *bytes = it->second.bytes;
  return RE_OK;
}

RE_API re_status re_scene_info(re_ctx* ctx, int32_t scene_id, int64_t sizes[4]) {
  if (!ctx || !sizes) return fail(RE_ERR_INVALID, "re_scene_info: null argument");
  auto it = ctx->scenes.find(scene_id);
  if (it == ctx->scenes.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "re_scene_info: unknown scene");
  sizes[0] = it->second.n_nodes;
  sizes[1] = it->second.n_leaves;
  sizes[2] = it->second.n_trirecs;
  sizes[3] = it->second.n_obj;
  return RE_OK;
}

RE_API re_status re_scene_export(re_ctx* ctx, int32_t scene_id, void* nodes, void* leaves, void* tris, void* objs) {
  if (!ctx) return fail(RE_ERR_INVALID, "re_scene_export: null ctx");
  auto it = ctx->scenes.find(scene_id);
  if (it == ctx->scenes.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "re_scene_export: unknown scene");
  const DevScene& s = it->second;
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  if (nodes && s.n_nodes)
    HIP_TRY(ctx, hipMemcpyAsync(nodes, s.nodes, s.n_nodes * sizeof(NodeEntry), hipMemcpyDeviceToHost, ctx->stream), "export nodes");
  // Leaf records in the flat field order (host copy; the device holds them as DevLeaf).
  if (leaves && s.n_leaves) std::memcpy(leaves, s.h_leaves.data(), s.n_leaves * sizeof(LeafRec));
  if (tris && s.n_trirecs)
    HIP_TRY(ctx, hipMemcpyAsync(tris, s.tris, s.n_trirecs * sizeof(TriRec), hipMemcpyDeviceToHost, ctx->stream), "export tris");
  if (objs)
    HIP_TRY(ctx, hipMemcpyAsync(objs, s.objs, s.n_obj * sizeof(ObjXform), hipMemcpyDeviceToHost, ctx->stream), "export objs");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "export sync");
  return RE_OK;
}

RE_API void re_trace_default_params(re_trace_params* p) {
  if (!p) return;
  p->max_leaf_visits = 50;
  p->one_sided = 1;
  p->secondary = 0;
  p->block = 0;
  p->variant = 4;
}

RE_API re_status re_kernel_times(re_ctx* ctx, float* ms, int32_t max, int32_t* count) {
  if (!ctx || !count || max < 0 || (max > 0 && !ms)) return fail(RE_ERR_INVALID, "re_kernel_times: bad arguments");
  int64_t avail = ctx->n_launch < re_ctx::kRing ? ctx->n_launch : re_ctx::kRing;
  int32_t k = (int32_t)(avail < max ? avail : max);
  for (int32_t i = 0; i < k; ++i) {
    int64_t idx = (ctx->n_launch - k + i) % re_ctx::kRing;
    HIP_TRY(ctx, hipEventSynchronize(ctx->ring[2 * idx + 1]), "re_kernel_times sync");
    HIP_TRY(ctx, hipEventElapsedTime(&ms[i], ctx->ring[2 * idx], ctx->ring[2 * idx + 1]), "re_kernel_times");
  }
  *count = k;
  return RE_OK;
}

RE_API re_status re_launch_end_times(re_ctx* ctx, float* ms, int32_t max, int32_t* count) {
  if (!ctx || !count || max < 0 || (max > 0 && !ms)) return fail(RE_ERR_INVALID, "re_launch_end_times: bad arguments");
  int64_t avail = ctx->n_launch < re_ctx::kRing ? ctx->n_launch : re_ctx::kRing;
  int32_t k = (int32_t)(avail < max ? avail : max);
  if (k > 0) {
    const int64_t first = (ctx->n_launch - k) % re_ctx::kRing;
    for (int32_t i = 0; i < k; ++i) {
      int64_t idx = (ctx->n_launch - k + i) % re_ctx::kRing;
      HIP_TRY(ctx, hipEventSynchronize(ctx->ring[2 * idx + 1]), "re_launch_end_times sync");
      HIP_TRY(ctx, hipEventElapsedTime(&ms[i], ctx->ring[2 * first], ctx->ring[2 * idx + 1]), "re_launch_end_times");
    }
  }
  *count = k;
  return RE_OK;
}

RE_API re_status re_last_kernel_ms(re_ctx* ctx, float* ms) {
  if (!ctx || !ms) return fail(RE_ERR_INVALID, "re_last_kernel_ms: null argument");
  if (ctx->n_launch == 0) return set_err(ctx, RE_ERR_INVALID, "re_last_kernel_ms: no launch yet");
  int32_t k = 0;
  return re_kernel_times(ctx, ms, 1, &k);
}

}  // extern "C"

namespace re {

// Shared body of the closest-hit entry points.  counts (host pointer, may be
// null) receives the summed work counters.  With sa given, the work counters,
// the persistent variant's ray counter and the staging of host-pointer calls
// live in that fixed device area (cleared by a fill kernel), and nothing is
// allocated; otherwise they use the context's buffers (the staging buffer
// grows on demand).
re_status trace_common(re_ctx* ctx, int32_t scene_id, const float* ray_o, const float* ray_d, int64_t n,
                       const re_trace_params* params, int32_t* prim_id, float* t, int32_t* obj_id,
                       uint32_t flags, TraceCounts* counts, const ScratchArea* sa, const int32_t* start_leaf,
                       int32_t* hit_leaf) {
  if (!ctx) return fail(RE_ERR_INVALID, "re_trace_closest: null ctx");
  if (n < 0 || (n > 0 && (!ray_o || !ray_d || !prim_id || !t)))
    return set_err(ctx, RE_ERR_INVALID, "re_trace_closest: bad arguments");
  auto it = ctx->scenes.find(scene_id);
  if (it == ctx->scenes.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "re_trace_closest: unknown scene");
  re_trace_params p;
  re_trace_default_params(&p);
  if (params) p = *params;
  if (!(p.block == 0 || p.block == 64 || p.block == 128 || p.block == 256))
    return set_err(ctx, RE_ERR_INVALID, "re_trace_closest: block must be 0, 64, 128 or 256");
  if (p.variant < 0 || p.variant > 5) return set_err(ctx, RE_ERR_INVALID, "re_trace_closest: variant must be 0..5");
  if ((start_leaf || hit_leaf) && ((p.variant != 4 && p.variant != 5) || !(flags & RE_DEVICE_PTRS)))
    return set_err(ctx, RE_ERR_INVALID, "re_trace_closest_hinted: variants 4 and 5 with RE_DEVICE_PTRS only");
  if (n == 0) return RE_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  TraceLaunch L;
  L.scene = &it->second;
  L.n = n;
  L.cap = p.max_leaf_visits;
  L.one_sided = p.one_sided;
  L.secondary = p.secondary;
  L.block = p.block;
  L.counts = nullptr;
  L.variant = p.variant;
  L.start_leaf = start_leaf;
  L.hit_leaf = hit_leaf;
  char* const header = sa ? (char*)sa->base : (char*)ctx->counts_buf;
  L.queue = (unsigned long long*)(header + 64);
  {
    // the block size launch_trace_closest launches with (its default: 256
    // threads for variant 1, one wave otherwise)
    const int blk = trace_launch_block(p.block, p.variant);
    const int slot = blk == 64 ? 0 : (blk == 128 ? 1 : 2);
    int64_t& g = p.variant == 5 ? ctx->coop_refill_grid[slot] : ctx->persistent_grid[slot];
    if (g <= 0) g = p.variant == 5 ? trace_coop_refill_grid(blk) : trace_persistent_grid(blk);
    L.grid_blocks = g;
  }
  const bool dev = (flags & RE_DEVICE_PTRS) != 0;
  const size_t ray_bytes = (size_t)n * 12;
  char* base = nullptr;
  if (!dev) {
    size_t need = 2 * ray_bytes + (size_t)n * 12 + 256 + 64;
    if (sa) {
      if (need > sa->bytes - kScratchHeader)
        return set_err(ctx, RE_ERR_INVALID, "host-pointer rays exceed the fixed staging area");
      base = (char*)sa->base + kScratchHeader;
    } else {
      re_status st = ensure_stage(ctx, need);
      if (st) return st;
      base = (char*)ctx->stage;
    }
    float* d_o = (float*)base;
    float* d_d = (float*)(base + ray_bytes);
    int32_t* d_p = (int32_t*)(base + 2 * ray_bytes);
    float* d_t = (float*)(base + 2 * ray_bytes + (size_t)n * 4);
    int32_t* d_ob = (int32_t*)(base + 2 * ray_bytes + (size_t)n * 8);
    HIP_TRY(ctx, hipMemcpyAsync(d_o, ray_o, ray_bytes, hipMemcpyHostToDevice, ctx->stream), "copy ray_o");
    HIP_TRY(ctx, hipMemcpyAsync(d_d, ray_d, ray_bytes, hipMemcpyHostToDevice, ctx->stream), "copy ray_d");
    L.ray_o = d_o;
    L.ray_d = d_d;
    L.prim = d_p;
    L.t = d_t;
    L.obj = obj_id ? d_ob : nullptr;
  } else {
    L.ray_o = ray_o;
    L.ray_d = ray_d;
    L.prim = prim_id;
    L.t = t;
    L.obj = obj_id;
  }
  TraceCounts* d_counts = nullptr;
  if (counts) {
    d_counts = (TraceCounts*)header;
    L.counts = d_counts;
  }
  if (sa) {
    // work counters (bytes 0-31) and ray counters (from byte 64) of the fixed
    // area start at zero
    HIP_TRY(ctx, launch_zero_words((uint32_t*)header, 16 + kQueueWords, ctx->stream), "scratch clear");
  } else {
    // cleared by a fill kernel, not a memset node, so that callers may
    // capture the call in a graph
    if (counts) HIP_TRY(ctx, launch_zero_words((uint32_t*)d_counts, (int32_t)(sizeof(TraceCounts) / 4), ctx->stream),
                        "counts clear");
    if (L.variant == 1 || L.variant == 5)
      HIP_TRY(ctx, launch_zero_words((uint32_t*)L.queue, kQueueWords, ctx->stream), "queue clear");
  }
  HIP_TRY(ctx, ctx->launch_begin(), "event");
  HIP_TRY(ctx, launch_trace_closest(L, ctx->stream), "trace_closest launch");
  HIP_TRY(ctx, ctx->launch_end(), "event");
  if (!dev) {
    HIP_TRY(ctx, hipMemcpyAsync(prim_id, L.prim, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream), "copy prim");
    HIP_TRY(ctx, hipMemcpyAsync(t, L.t, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream), "copy t");
    if (obj_id) HIP_TRY(ctx, hipMemcpyAsync(obj_id, L.obj, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream), "copy obj");
  }
  if (counts) {
    HIP_TRY(ctx, hipMemcpyAsync(counts, d_counts, sizeof(TraceCounts), hipMemcpyDeviceToHost, ctx->stream), "copy counts");
  }
  if (!(flags & RE_ASYNC) || !dev || counts) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "trace sync");
  return RE_OK;
}

}  // namespace re

extern "C" {

RE_API re_status re_trace_closest(re_ctx* ctx, int32_t scene_id, const float* ray_o, const float* ray_d, int64_t n,
                                  const re_trace_params* params, int32_t* prim_id, float* t, int32_t* obj_id,
                                  uint32_t flags) {
  return trace_common(ctx, scene_id, ray_o, ray_d, n, params, prim_id, t, obj_id, flags, nullptr, nullptr);
}

RE_API re_status re_trace_counts(re_ctx* ctx, int32_t scene_id, const float* ray_o, const float* ray_d, int64_t n,
                                 const re_trace_params* params, int32_t* prim_id, float* t, uint64_t counts[4],
                                 uint32_t flags) {
  if (!counts) return set_err(ctx, RE_ERR_INVALID, "re_trace_counts: counts is null");
  TraceCounts c{0, 0, 0, 0};
  re_status st = trace_common(ctx, scene_id, ray_o, ray_d, n, params, prim_id, t, nullptr, flags, &c, nullptr);
  counts[0] = c.entries;
  counts[1] = c.leaves;
  counts[2] = c.tris;
  counts[3] = c.rays_hit;
  return st;
}

RE_API re_status re_trace_closest_hinted(re_ctx* ctx, int32_t scene_id, const float* ray_o, const float* ray_d,
                                         int64_t n, const re_trace_params* params, const int32_t* start_leaf,
                                         int32_t* prim_id, float* t, int32_t* obj_id, int32_t* hit_leaf,
                                         uint64_t counts[4], uint32_t flags) {
  if (!ctx) return fail(RE_ERR_INVALID, "re_trace_closest_hinted: null ctx");
  if (!(flags & RE_DEVICE_PTRS)) return set_err(ctx, RE_ERR_INVALID, "re_trace_closest_hinted: needs RE_DEVICE_PTRS");
  TraceCounts c{0, 0, 0, 0};
  re_status st = trace_common(ctx, scene_id, ray_o, ray_d, n, params, prim_id, t, obj_id, flags,
                              counts ? &c : nullptr, nullptr, start_leaf, hit_leaf);
  if (counts) {
    counts[0] = c.entries;
    counts[1] = c.leaves;
    counts[2] = c.tris;
    counts[3] = c.rays_hit;
  }
  return st;
}

RE_API re_status re_trace_occluded(re_ctx* ctx, int32_t scene_id, const float* ray_o, const float* ray_d, int64_t n,
                                   const re_trace_params* params, uint8_t* occluded, uint32_t flags) {
  if (!ctx) return fail(RE_ERR_INVALID, "re_trace_occluded: null ctx");
  if (n < 0 || (n > 0 && (!ray_o || !ray_d || !occluded)))
    return set_err(ctx, RE_ERR_INVALID, "re_trace_occluded: bad arguments");
  auto it = ctx->scenes.find(scene_id);
  if (it == ctx->scenes.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "re_trace_occluded: unknown scene");
  re_trace_params p;
  re_trace_default_params(&p);
  p.secondary = 1;  // trace_shadow traverses with is_secondary = true
  if (params) p = *params;
  if (!(p.block == 0 || p.block == 64 || p.block == 128 || p.block == 256))
    return set_err(ctx, RE_ERR_INVALID, "re_trace_occluded: block must be 0, 64, 128 or 256");
  if (n == 0) return RE_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  TraceLaunch L{};
  L.scene = &it->second;
  L.n = n;
  L.cap = p.max_leaf_visits;
  L.one_sided = p.one_sided;
  L.secondary = p.secondary;
  L.block = p.block;
  const bool dev = (flags & RE_DEVICE_PTRS) != 0;
  const size_t ray_bytes = (size_t)n * 12;
  if (!dev) {
    re_status st = ensure_stage(ctx, stage_size({ray_bytes, ray_bytes, (size_t)n}));
    if (st) return st;
    StageCursor c{(char*)ctx->stage};
    float* d_o = c.take<float>(3 * n);
    float* d_d = c.take<float>(3 * n);
    L.occluded = c.take<uint8_t>(n);
    HIP_TRY(ctx, hipMemcpyAsync(d_o, ray_o, ray_bytes, hipMemcpyHostToDevice, ctx->stream), "copy ray_o");
    HIP_TRY(ctx, hipMemcpyAsync(d_d, ray_d, ray_bytes, hipMemcpyHostToDevice, ctx->stream), "copy ray_d");
    L.ray_o = d_o;
    L.ray_d = d_d;
  } else {
    L.ray_o = ray_o;
    L.ray_d = ray_d;
    L.occluded = occluded;
  }
  HIP_TRY(ctx, ctx->launch_begin(), "event");
  HIP_TRY(ctx, launch_trace_occluded(L, ctx->stream), "trace_occluded launch");
  HIP_TRY(ctx, ctx->launch_end(), "event");
  if (!dev)
    HIP_TRY(ctx, hipMemcpyAsync(occluded, L.occluded, (size_t)n, hipMemcpyDeviceToHost, ctx->stream), "copy occluded");
  if (!(flags & RE_ASYNC) || !dev) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "trace_occluded sync");
  return RE_OK;
}


RE_API re_status re_scene_set_attributes(re_ctx* ctx, int32_t scene_id, const float* normals, const float* uvs,
                                         int64_t n_prims) {
  if (!ctx) return fail(RE_ERR_INVALID, "re_scene_set_attributes: null ctx");
  if (!normals || n_prims <= 0) return set_err(ctx, RE_ERR_INVALID, "re_scene_set_attributes: bad arguments");
  auto it = ctx->scenes.find(scene_id);
  if (it == ctx->scenes.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "re_scene_set_attributes: unknown scene");
  DevScene& S = it->second;
  return guarded(ctx, "re_scene_set_attributes", [&]() -> re_status {
    HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    // every prim's geometry from the uploaded triangle records (each prim id
    // appears in one record or, clipped into several leaves, in identical
    // copies)
    std::vector<TriRec> recs((size_t)S.n_trirecs);
    if (S.n_trirecs)
      HIP_TRY(ctx, hipMemcpyAsync(recs.data(), S.tris, recs.size() * sizeof(TriRec), hipMemcpyDeviceToHost, ctx->stream),
              "read triangle records");
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "read sync");
    std::vector<float> geom((size_t)n_prims * 9, 0.0f);
    std::vector<uint8_t> seen((size_t)n_prims, 0);
    for (const TriRec& r : recs) {
      if (r.prim < 0 || r.prim >= n_prims)
        return set_err(ctx, RE_ERR_INVALID, "re_scene_set_attributes: n_prims is below a prim id of the scene");
      float* dst = &geom[(size_t)r.prim * 9];
      if (seen[(size_t)r.prim] && std::memcmp(dst, r.v0, 9 * sizeof(float)) != 0)
        return set_err(ctx, RE_ERR_INVALID,
                       "re_scene_set_attributes: two objects give one prim id different triangles (set prim offsets)");
      std::memcpy(dst, r.v0, 9 * sizeof(float));
      seen[(size_t)r.prim] = 1;
    }
    float *g = nullptr, *nr = nullptr, *uv = nullptr;
    hipError_t e = hipMalloc((void**)&g, geom.size() * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&nr, (size_t)n_prims * 36);
    if (e == hipSuccess && uvs) e = hipMalloc((void**)&uv, (size_t)n_prims * 24);
    if (e == hipSuccess) e = hipMemcpyAsync(g, geom.data(), geom.size() * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(nr, normals, (size_t)n_prims * 36, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && uvs) e = hipMemcpyAsync(uv, uvs, (size_t)n_prims * 24, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
      for (void* p : {(void*)g, (void*)nr, (void*)uv})
        if (p) (void)hipFree(p);
      return hip_err(ctx, e, "re_scene_set_attributes");
    }
    for (void* p : {(void*)S.attr_geom, (void*)S.attr_nrm, (void*)S.attr_uv})
      if (p) (void)hipFree(p);
    S.attr_geom = g;
    S.attr_nrm = nr;
    S.attr_uv = uv;
    S.n_attr = n_prims;
    return RE_OK;
  });
}

RE_API re_status re_debug_bounds_errors(re_ctx* ctx, uint32_t* mask) {
  if (!ctx || !mask) return fail(RE_ERR_INVALID, "re_debug_bounds_errors: null argument");
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "debug bounds sync");
  const hipError_t e = debug_bounds_take(mask);
  if (e == hipErrorNotSupported)
    return set_err(ctx, RE_ERR_UNSUPPORTED, "re_debug_bounds_errors: library built without RE_DEBUG_BOUNDS");
  if (e != hipSuccess) return hip_err(ctx, e, "re_debug_bounds_errors");
  return RE_OK;
}

RE_API re_status re_trace_status(re_ctx* ctx, int32_t scene_id, const float* ray_o, const float* ray_d, int64_t n,
                                 const re_trace_params* params, int32_t* prim_id, float* t, uint8_t* status,
                                 uint32_t flags) {
  if (!ctx) return fail(RE_ERR_INVALID, "re_trace_status: null ctx");
  if (n < 0 || (n > 0 && (!ray_o || !ray_d || !prim_id || !t || !status)))
    return set_err(ctx, RE_ERR_INVALID, "re_trace_status: bad arguments");
  auto it = ctx->scenes.find(scene_id);
  if (it == ctx->scenes.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "re_trace_status: unknown scene");
  re_trace_params p;
  re_trace_default_params(&p);
  if (params) p = *params;
  if (n == 0) return RE_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  StatusLaunch L{};
  L.scene = &it->second;
  L.n = n;
  L.cap = p.max_leaf_visits;
  L.one_sided = p.one_sided;
  L.secondary = p.secondary;
  const bool dev = (flags & RE_DEVICE_PTRS) != 0;
  if (!dev) {
    const size_t rb = (size_t)n * 12;
    re_status st = ensure_stage(ctx, stage_size({rb, rb, (size_t)n * 4, (size_t)n * 4, (size_t)n}));
    if (st) return st;
    StageCursor c{(char*)ctx->stage};
    float* so = c.take<float>(3 * n);
    float* sd = c.take<float>(3 * n);
    L.prim = c.take<int32_t>(n);
    L.t = c.take<float>(n);
    L.status = c.take<uint8_t>(n);
    HIP_TRY(ctx, hipMemcpyAsync(so, ray_o, rb, hipMemcpyHostToDevice, ctx->stream), "copy ray_o");
    HIP_TRY(ctx, hipMemcpyAsync(sd, ray_d, rb, hipMemcpyHostToDevice, ctx->stream), "copy ray_d");
    L.ray_o = so;
    L.ray_d = sd;
  } else {
    L.ray_o = ray_o;
    L.ray_d = ray_d;
    L.prim = prim_id;
    L.t = t;
    L.status = status;
  }
  HIP_TRY(ctx, ctx->launch_begin(), "event");
  HIP_TRY(ctx, launch_trace_status(L, ctx->stream), "trace status");
  HIP_TRY(ctx, ctx->launch_end(), "event");
  if (!dev) {
    HIP_TRY(ctx, hipMemcpyAsync(prim_id, L.prim, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream), "copy prim");
    HIP_TRY(ctx, hipMemcpyAsync(t, L.t, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream), "copy t");
    HIP_TRY(ctx, hipMemcpyAsync(status, L.status, (size_t)n, hipMemcpyDeviceToHost, ctx->stream), "copy status");
  }
  if (!(flags & RE_ASYNC) || !dev) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "status sync");
  return RE_OK;
}

RE_API re_status re_hit_attributes(re_ctx* ctx, int32_t scene_id, const float* ray_o, const float* ray_d, int64_t n,
                                   const re_trace_params* params, const int32_t* prim_id, const int32_t* obj_id,
                                   float* normal, float* uv, float* bary, float* hit_point, uint32_t flags) {
  if (!ctx) return fail(RE_ERR_INVALID, "re_hit_attributes: null ctx");
  if (n < 0 || (n > 0 && (!ray_o || !ray_d || !prim_id)))
    return set_err(ctx, RE_ERR_INVALID, "re_hit_attributes: bad arguments");
  auto it = ctx->scenes.find(scene_id);
  if (it == ctx->scenes.end()) return set_err(ctx, RE_ERR_NOT_FOUND, "re_hit_attributes: unknown scene");
  const DevScene& S = it->second;
  if (!S.attr_geom) return set_err(ctx, RE_ERR_INVALID, "re_hit_attributes: no re_scene_set_attributes for this scene");
  if (S.n_obj > 1 && !obj_id) return set_err(ctx, RE_ERR_INVALID, "re_hit_attributes: obj_id needed with several objects");
  re_trace_params p;
  re_trace_default_params(&p);
  if (params) p = *params;
  if (n == 0) return RE_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  AttrLaunch L{};
  L.scene = &S;
  L.n = n;
  L.one_sided = p.one_sided;
  L.secondary = p.secondary;
  const bool dev = (flags & RE_DEVICE_PTRS) != 0;
  if (!dev) {
    const size_t rb = (size_t)n * 12;
    re_status st = ensure_stage(ctx, stage_size({rb, rb, (size_t)n * 4, (size_t)n * 4, rb, (size_t)n * 8,
                                                 (size_t)n * 8, rb}));
    if (st) return st;
    StageCursor c{(char*)ctx->stage};
    float* so = c.take<float>(3 * n);
    float* sd = c.take<float>(3 * n);
    int32_t* sp = c.take<int32_t>(n);
    int32_t* sob = c.take<int32_t>(n);
    float* snr = c.take<float>(3 * n);
    float* suv = c.take<float>(2 * n);
    float* sba = c.take<float>(2 * n);
    float* shp = c.take<float>(3 * n);
    HIP_TRY(ctx, hipMemcpyAsync(so, ray_o, rb, hipMemcpyHostToDevice, ctx->stream), "copy ray_o");
    HIP_TRY(ctx, hipMemcpyAsync(sd, ray_d, rb, hipMemcpyHostToDevice, ctx->stream), "copy ray_d");
    HIP_TRY(ctx, hipMemcpyAsync(sp, prim_id, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream), "copy prim");
    if (obj_id) HIP_TRY(ctx, hipMemcpyAsync(sob, obj_id, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream), "copy obj");
    L.ray_o = so;
    L.ray_d = sd;
    L.prim = sp;
    L.obj = obj_id ? sob : nullptr;
    L.normal = normal ? snr : nullptr;
    L.uv = uv ? suv : nullptr;
    L.bary = bary ? sba : nullptr;
    L.hit_point = hit_point ? shp : nullptr;
  } else {
    L.ray_o = ray_o;
    L.ray_d = ray_d;
    L.prim = prim_id;
    L.obj = obj_id;
    L.normal = normal;
    L.uv = uv;
    L.bary = bary;
    L.hit_point = hit_point;
  }
  HIP_TRY(ctx, ctx->launch_begin(), "event");
  HIP_TRY(ctx, launch_hit_attributes(L, ctx->stream), "hit attributes");
  HIP_TRY(ctx, ctx->launch_end(), "event");
  if (!dev) {
    if (normal) HIP_TRY(ctx, hipMemcpyAsync(normal, L.normal, (size_t)n * 12, hipMemcpyDeviceToHost, ctx->stream), "copy normal");
    if (uv) HIP_TRY(ctx, hipMemcpyAsync(uv, L.uv, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream), "copy uv");
    if (bary) HIP_TRY(ctx, hipMemcpyAsync(bary, L.bary, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream), "copy bary");
    if (hit_point)
      HIP_TRY(ctx, hipMemcpyAsync(hit_point, L.hit_point, (size_t)n * 12, hipMemcpyDeviceToHost, ctx->stream), "copy point");
  }
  if (!(flags & RE_ASYNC) || !dev) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "attributes sync");
  return RE_OK;
}

}  // extern "C"
