// C-ABI entry point of the device kd-tree build (kernels/kdbuild.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "capi_util.hpp"
#include "kdbuild.hpp"

using namespace re;

namespace re {
re_kdtree* kd_from_arrays(std::vector<NodeEntry>&& nodes, std::vector<LeafRec>&& leaves, std::vector<TriRec>&& tris,
                          const re_kd_info& info);
namespace kdb {
hipError_t build_device(const float* tri_v, int64_t n, const float* root_min, const float* root_max,
                        const Params& P, hipStream_t s, std::vector<NodeEntry>& out_nodes,
                        std::vector<LeafRec>& out_leaves, std::vector<TriRec>& out_tris, int32_t& depth,
                        int32_t& max_leaf, float& device_ms);
}  // namespace kdb
}  // namespace re

extern "C" {

RE_API re_status re_kd_build_gpu(re_ctx* ctx, const float* tri_v, int64_t n_tris, const re_kd_build_params* params,
                                 re_kdtree** out, float* device_ms) {
  if (!ctx) return fail(RE_ERR_INVALID, "re_kd_build_gpu: null ctx");
  if (!out || n_tris < 0 || (n_tris > 0 && !tri_v)) return set_err(ctx, RE_ERR_INVALID, "re_kd_build_gpu: bad arguments");
  if (n_tris >= (int64_t)0x3fffffff) return set_err(ctx, RE_ERR_INVALID, "re_kd_build_gpu: too many triangles");
  re_kd_build_params p;
  re_kd_default_params(&p);
  if (params) p = *params;
  if (p.max_leaf_tris < 1 || p.max_depth < 1 || p.max_depth > 64 || p.n_bins > kdb::kMaxBins)
    return set_err(ctx, RE_ERR_INVALID, "re_kd_build_gpu: max_leaf_tris >= 1, max_depth in [1,64], n_bins <= 64");
  if (p.split_rule != 0)
    return set_err(ctx, RE_ERR_UNSUPPORTED, "re_kd_build_gpu: the spatial-median rule is built on the host (re_kd_build)");
  if (device_ms) *device_ms = 0.0f;
  // Root box and input check on the host, as re_kd_build does them.
  float rmin[3] = {kInf, kInf, kInf}, rmax[3] = {-kInf, -kInf, -kInf};
  for (int64_t i = 0; i < n_tris; ++i) {
    const float* t = tri_v + 9 * i;
    for (int k = 0; k < 9; ++k)
      if (!std::isfinite(t[k])) return set_err(ctx, RE_ERR_INVALID, "re_kd_build_gpu: non-finite vertex");
    for (int a = 0; a < 3; ++a) {
      rmin[a] = std::min(rmin[a], std::min(std::min(t[a], t[3 + a]), t[6 + a]));
      rmax[a] = std::max(rmax[a], std::max(std::max(t[a], t[3 + a]), t[6 + a]));
    }
  }
  if (n_tris == 0) {  // a single empty leaf: nothing for the device to do
    re_status st = re_kd_build(tri_v, 0, &p, out);
    return st ? set_err(ctx, st, "re_kd_build_gpu: empty build failed") : RE_OK;
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  kdb::Params P{p.max_leaf_tris, p.max_depth, p.n_bins, p.clip ? 1 : 0, p.cost_traversal, p.cost_intersect,
                p.empty_bonus};
  std::vector<NodeEntry> nodes;
  std::vector<LeafRec> leaves;
  std::vector<TriRec> tris;
  int32_t depth = 0, max_leaf = 0;
  float ms = 0.0f;
  try {
    hipError_t e = kdb::build_device(tri_v, n_tris, rmin, rmax, P, ctx->stream, nodes, leaves, tris, depth, max_leaf, ms);
    if (e != hipSuccess) return hip_err(ctx, e, "re_kd_build_gpu");
  } catch (const std::bad_alloc&) {
    return set_err(ctx, RE_ERR_OOM, "re_kd_build_gpu: out of host memory");
  }
  re_kd_info in;
  in.n_tris = n_tris;
  in.n_nodes = (int64_t)nodes.size();
  in.n_leaves = (int64_t)leaves.size();
  in.n_trirecs = (int64_t)tris.size();
  in.depth = depth;
  in.max_leaf = max_leaf;
  for (int a = 0; a < 3; ++a) {
    in.root_min[a] = rmin[a];
    in.root_max[a] = rmax[a];
  }
  re_kdtree* t = kd_from_arrays(std::move(nodes), std::move(leaves), std::move(tris), in);
  if (!t) return set_err(ctx, RE_ERR_OOM, "re_kd_build_gpu: out of host memory");
  *out = t;
  if (device_ms) *device_ms = ms;
  return RE_OK;
}

}  // extern "C"
