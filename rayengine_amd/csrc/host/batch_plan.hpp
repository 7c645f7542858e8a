// Work-item plan of a batched launch: how the targets' ray ranges are cut into
// the items the persistent kernel takes from its device counter.  Host code,
// shared by re_batch_create and re_batch_plan; the item record is also the
// device record the kernel reads.
#pragma once

#include <cstdint>
#include <vector>

#include "../../../include/rayengine.h"

namespace re {

enum BatchKind : int32_t { kItemTri = 0, kItemExact = 1, kItemPlain = 2 };

struct BatchItem {
  int32_t kind;    // BatchKind
  int32_t target;  // scene object index (triangles) or tile index (terrain)
  int64_t begin;   // first ray
  int32_t count;   // rays in the item
  int32_t slot;    // completion slot (scenes first, then tiles)
};
static_assert(sizeof(BatchItem) == 24, "24-byte item records");

// Work items of one batch are counted by a 32-bit device counter, which each
// wave also advances once past the end.
constexpr int64_t kMaxBatchItems = (int64_t)1 << 30;

struct BatchPlan {
  std::vector<BatchItem> items;
  std::vector<int32_t> remaining;        // per slot: its items (1 for a surrogate tile)
  std::vector<int32_t> surrogate_slots;  // surrogate tiles' slots, in plan order
};

// Targets (slots: scenes 0..n_scenes-1, then tiles) in ascending ray count,
// ties in slot order; each target's range cut into items of item_rays rays
// (the last one shorter).  Surrogate tiles (mode 2) get no items: they run as
// their own launches, and their slot counts one completion.  RE_ERR_INVALID
// for a range outside [0, n_rays), a bad mode or item_rays outside
// [1, 65536]; more than kMaxBatchItems items is RE_ERR_INVALID too.
re_status plan_batch(const re_batch_scene* scenes, int32_t n_scenes, const re_batch_tile* tiles, int32_t n_tiles,
                     int64_t n_rays, int32_t item_rays, BatchPlan& out);

}  // namespace re
