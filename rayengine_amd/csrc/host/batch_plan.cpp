// Work-item plan of a batched launch (see batch_plan.hpp) and its C-ABI
// entry point re_batch_plan, which needs no device.
#include "batch_plan.hpp"

#include <algorithm>
#include <cstring>
#include <new>
#include <numeric>

#include "errors.hpp"

namespace re {

re_status plan_batch(const re_batch_scene* scenes, int32_t n_scenes, const re_batch_tile* tiles, int32_t n_tiles,
                     int64_t n_rays, int32_t item_rays, BatchPlan& out) {
  out = BatchPlan{};
  if (n_scenes < 0 || n_tiles < 0 || n_scenes + n_tiles <= 0 || n_rays <= 0 || (n_scenes > 0 && !scenes) ||
      (n_tiles > 0 && !tiles) || item_rays < 1 || item_rays > (1 << 16))
    return fail(RE_ERR_INVALID, "batch plan: bad arguments");
  for (int32_t k = 0; k < n_scenes; ++k)
    if (scenes[k].ray_begin < 0 || scenes[k].ray_count < 0 || scenes[k].ray_count > n_rays - scenes[k].ray_begin)
      return fail(RE_ERR_INVALID, "batch plan: scene ray range outside the ray arrays");
  for (int32_t k = 0; k < n_tiles; ++k)
    if (tiles[k].ray_begin < 0 || tiles[k].ray_count < 0 || tiles[k].ray_count > n_rays - tiles[k].ray_begin ||
        tiles[k].mode < 0 || tiles[k].mode > 2)
      return fail(RE_ERR_INVALID, "batch plan: tile ray range or mode invalid");
  const int32_t n_slots = n_scenes + n_tiles;
  auto count_of = [&](int32_t s) { return s < n_scenes ? scenes[s].ray_count : tiles[s - n_scenes].ray_count; };
  int64_t total = 0;
  for (int32_t s = 0; s < n_slots; ++s) {
    const bool surrogate = s >= n_scenes && tiles[s - n_scenes].mode == 2;
    if (!surrogate) total += (count_of(s) + item_rays - 1) / item_rays;
  }
  if (total > kMaxBatchItems) return fail(RE_ERR_INVALID, "batch plan: more than 2^30 work items (raise item_rays)");
  std::vector<int32_t> order(n_slots);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t c) { return count_of(a) < count_of(c); });
  out.items.reserve((size_t)total);
  out.remaining.assign(n_slots, 0);
  for (int32_t s : order) {
    int32_t kind, target;
    int64_t begin, count;
    if (s < n_scenes) {
      kind = kItemTri;
      target = s;
      begin = scenes[s].ray_begin;
      count = scenes[s].ray_count;
    } else {
      const re_batch_tile& tl = tiles[s - n_scenes];
      target = s - n_scenes;
      begin = tl.ray_begin;
      count = tl.ray_count;
      if (tl.mode == 2) {
        out.surrogate_slots.push_back(s);
        out.remaining[s] = 1;
        continue;
      }
      kind = tl.mode == 0 ? kItemExact : kItemPlain;
    }
    for (int64_t off = 0; off < count; off += item_rays) {
      const int32_t c = (int32_t)std::min<int64_t>(item_rays, count - off);
      out.items.push_back(BatchItem{kind, target, begin + off, c, s});
      ++out.remaining[s];
    }
  }
  return RE_OK;
}

}  // namespace re

extern "C" {

RE_API re_status re_batch_plan(const re_batch_scene* scenes, int32_t n_scenes, const re_batch_tile* tiles,
                               int32_t n_tiles, int64_t n_rays, int32_t item_rays, void* items, int64_t max_items,
                               int64_t* n_items, int32_t* remaining) {
  using re::fail;
  if (!n_items || max_items < 0 || (max_items > 0 && !items)) return fail(RE_ERR_INVALID, "re_batch_plan: bad arguments");
  try {
    re::BatchPlan plan;
    const re_status st = re::plan_batch(scenes, n_scenes, tiles, n_tiles, n_rays, item_rays, plan);
    if (st) return st;
    *n_items = (int64_t)plan.items.size();
    if (items) {
      if ((int64_t)plan.items.size() > max_items) return fail(RE_ERR_INVALID, "re_batch_plan: max_items too small");
      std::memcpy(items, plan.items.data(), plan.items.size() * sizeof(re::BatchItem));
    }
    if (remaining) std::memcpy(remaining, plan.remaining.data(), plan.remaining.size() * sizeof(int32_t));
    return RE_OK;
  } catch (const std::bad_alloc&) {
    return fail(RE_ERR_OOM, "re_batch_plan: host allocation failed");
  }
}

}  // extern "C"
