// Device kd-tree build: the records one level of the breadth-first build
// works on.  The build makes the same decisions as the host builder
// (host/kdtree.cpp) with the same floating-point operations, so the tree it
// produces is identical, entry for entry.
#pragma once

#include "re_math.hpp"

namespace re {
namespace kdb {

// A triangle reference inside a node: its id and the box of the part of the
// triangle inside the node.
struct Ref {
  int32_t prim;
  float lo[3], hi[3];
};

// Node modes of a level.
enum : int32_t { kModeLeaf = 0, kModeExact = 1, kModeBinned = 2 };

// An active node of the current level.
struct LNode {
  float lo[3], hi[3];
  int64_t rbegin;    // first reference in the level's reference array
  int32_t n;         // references
  int32_t depth;     // root = 1
  int32_t gid;       // breadth-first node id
  int32_t mode;      // kMode*
  int32_t axis;      // chosen split axis, -1 none
  float pos;         // chosen split position
  int32_t nl, nr;    // counts the split was costed with
  int32_t cl, cr;    // references that actually go below / above (after clipping)
  int32_t split;     // 1 if the node becomes an inner node
  int32_t child;     // index of its below child in the next level (above = child + 1)
  int64_t crbegin;   // first reference of its below child in the next level's array
  int64_t pool;      // leaves: first slot in the leaf-reference pool
};

// A node of the finished tree, in breadth-first order.
struct GNode {
  float lo[3], hi[3];
  int32_t axis;      // -1 for a leaf
  float split;
  int32_t child0, child1;
  int32_t depth;
  int32_t pool_n;    // leaves: references
  int64_t pool;      // leaves: first slot in the leaf-reference pool
};

struct Params {
  int32_t max_leaf, max_depth, n_bins, clip;
  float ct, ci, eb;
};

constexpr int kExactMax = 64;  // nodes with at most this many references test every candidate plane
constexpr int kMaxBins = 64;

}  // namespace kdb
}  // namespace re
