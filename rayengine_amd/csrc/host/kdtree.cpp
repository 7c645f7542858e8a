// Rope kd-tree construction (host).
//
// Replaces RKDTreeCPU (RayEngine/src/Primitives/KDTree.cpp:117-834) and the
// flattening in RKDThreeGPU (RayEngine/src/Primitives/KDThreeGPU.cpp:155-194)
// with a surface-area-heuristic build.  The node semantics the traversal
// relies on are the reference's:
//   * a triangle goes below a split when its bounding-box minimum is < split
//     and above when its maximum is >= split (KDTree.cpp:581-600), so
//     straddlers are duplicated;
//   * leaves carry their split box and six ropes in LEFT, FRONT, RIGHT, BACK,
//     TOP, BOTTOM order; a missing neighbour is -1 (KDTree.cpp:636-692);
//   * the root box is the bounding box of all vertices (KDTree.cpp:165-168).
// Ropes are pushed down to the deepest node that still covers the whole leaf
// face (Popov et al., "Stackless KD-Tree Traversal for High Performance GPU
// Ray Tracing", 2007), the intent of optimizeRopes (KDTree.cpp:694-834).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "errors.hpp"
#include "../re_math.hpp"

using namespace re;

struct re_kdtree {
  std::vector<NodeEntry> nodes;
  std::vector<LeafRec> leaves;
  std::vector<TriRec> tris;
  re_kd_info info;
};

namespace {

struct BNode {
  float bmin[3], bmax[3];
  int axis = -1;
  float split = 0.0f;
  int child[2] = {-1, -1};
  std::vector<int32_t> prims;
  int rope[6] = {-1, -1, -1, -1, -1, -1};
  int depth = 1;
  bool leaf() const { return axis < 0; }
};

inline float area(const float* lo, const float* hi) {
  float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
  return 2.0f * (dx * dy + dx * dz + dy * dz);
}

// Face -> (axis, is_high_side).
const int kFaceAxis[6] = {0, 2, 0, 2, 1, 1};
const bool kFaceHigh[6] = {false, true, true, false, true, false};

class Builder {
 public:
  Builder(const float* tv, int64_t n, const re_kd_build_params& p) : tv_(tv), n_(n), p_(p) {
    tb_.resize((size_t)n * 6);
    for (int64_t i = 0; i < n; ++i) {
      const float* t = tv + 9 * i;
      for (int a = 0; a < 3; ++a) {
        float lo = t[a], hi = t[a];
        lo = std::min(lo, t[3 + a]); hi = std::max(hi, t[3 + a]);
        lo = std::min(lo, t[6 + a]); hi = std::max(hi, t[6 + a]);
        tb_[6 * i + a] = lo;
        tb_[6 * i + 3 + a] = hi;
      }
    }
  }

  std::vector<BNode> nodes;
  float root_min[3], root_max[3];

  void run() {
    for (int a = 0; a < 3; ++a) {
      root_min[a] = kInf;
      root_max[a] = -kInf;
    }
    for (int64_t i = 0; i < n_; ++i)
      for (int a = 0; a < 3; ++a) {
        root_min[a] = std::min(root_min[a], tb_[6 * i + a]);
        root_max[a] = std::max(root_max[a], tb_[6 * i + 3 + a]);
      }
    std::vector<int32_t> all((size_t)n_);
    for (int64_t i = 0; i < n_; ++i) all[i] = (int32_t)i;
    nodes.reserve((size_t)std::max<int64_t>(16, n_ / 2));
    build(root_min, root_max, std::move(all), 1);
    int ropes[6] = {-1, -1, -1, -1, -1, -1};
    link(0, ropes);
  }

 private:
  const float* tv_;
  int64_t n_;
  re_kd_build_params p_;
  std::vector<float> tb_;

  struct Split {
    int axis = -1;
    float pos = 0.0f;
    float cost = 0.0f;
    int64_t nl = 0, nr = 0;
  };

  float cost_of(const float* lo, const float* hi, int axis, float pos, int64_t nl, int64_t nr,
                float inv_sa) const {
    float llo[3] = {lo[0], lo[1], lo[2]}, lhi[3] = {hi[0], hi[1], hi[2]};
    float rlo[3] = {lo[0], lo[1], lo[2]}, rhi[3] = {hi[0], hi[1], hi[2]};
    lhi[axis] = pos;
    rlo[axis] = pos;
    float c = p_.cost_traversal +
              p_.cost_intersect * (area(llo, lhi) * (float)nl + area(rlo, rhi) * (float)nr) * inv_sa;
    if (nl == 0 || nr == 0) c *= p_.empty_bonus;
    return c;
  }

  Split find_split(const float* lo, const float* hi, const std::vector<int32_t>& prims) const {
    Split best;
    best.cost = INFINITY;
    float sa = area(lo, hi);
    float inv_sa = sa > 0.0f ? 1.0f / sa : 0.0f;
    const int64_t n = (int64_t)prims.size();
    for (int a = 0; a < 3; ++a) {
      float ext = hi[a] - lo[a];
      if (!(ext > 0.0f)) continue;
      if (n > 64) {
        const int nb = std::max(4, p_.n_bins);
        std::vector<int64_t> cmin(nb, 0), cmax(nb, 0);
        float scale = (float)nb / ext;
        for (int32_t pi : prims) {
          float pl = std::min(std::max(tb_[6 * pi + a], lo[a]), hi[a]);
          float ph = std::min(std::max(tb_[6 * pi + 3 + a], lo[a]), hi[a]);
          int bl = std::min(nb - 1, std::max(0, (int)((pl - lo[a]) * scale)));
          int bh = std::min(nb - 1, std::max(0, (int)((ph - lo[a]) * scale)));
          cmin[bl]++;
          cmax[bh]++;
        }
        int64_t nl = 0, nr = n;
        for (int k = 1; k < nb; ++k) {
          nl += cmin[k - 1];
          nr -= cmax[k - 1];
          float pos = lo[a] + ext * (float)k / (float)nb;
          if (!(pos > lo[a] && pos < hi[a])) continue;
          float c = cost_of(lo, hi, a, pos, nl, nr, inv_sa);
          if (c < best.cost) best = Split{a, pos, c, nl, nr};
        }
      } else {
        // Exact candidates: every primitive's minimum (it goes above) and the
        // float just past every maximum (it goes below).
        std::vector<float> cand;
        cand.reserve(2 * n);
        for (int32_t pi : prims) {
          cand.push_back(tb_[6 * pi + a]);
          cand.push_back(std::nextafter(tb_[6 * pi + 3 + a], INFINITY));
        }
        std::sort(cand.begin(), cand.end());
        cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
        for (float pos : cand) {
          if (!(pos > lo[a] && pos < hi[a])) continue;
          int64_t nl = 0, nr = 0;
          for (int32_t pi : prims) {
            if (tb_[6 * pi + a] < pos) ++nl;
            if (tb_[6 * pi + 3 + a] >= pos) ++nr;
          }
          float c = cost_of(lo, hi, a, pos, nl, nr, inv_sa);
          if (c < best.cost) best = Split{a, pos, c, nl, nr};
        }
      }
    }
    return best;
  }

  int build(const float* lo, const float* hi, std::vector<int32_t>&& prims, int depth) {
    int id = (int)nodes.size();
    nodes.emplace_back();
    {
      BNode& b = nodes[id];
      std::memcpy(b.bmin, lo, sizeof(b.bmin));
      std::memcpy(b.bmax, hi, sizeof(b.bmax));
      b.depth = depth;
    }
    const int64_t n = (int64_t)prims.size();
    if (n <= p_.max_leaf_tris || depth >= p_.max_depth) {
      nodes[id].prims = std::move(prims);
      return id;
    }
    Split s = find_split(lo, hi, prims);
    if (s.axis < 0 || (s.nl == n && s.nr == n)) {
      nodes[id].prims = std::move(prims);
      return id;
    }
    std::vector<int32_t> L, R;
    L.reserve(s.nl);
    R.reserve(s.nr);
    for (int32_t pi : prims) {
      if (tb_[6 * pi + s.axis] < s.pos) L.push_back(pi);
      if (tb_[6 * pi + 3 + s.axis] >= s.pos) R.push_back(pi);
    }
    if ((int64_t)L.size() == n && (int64_t)R.size() == n) {
      nodes[id].prims = std::move(prims);
      return id;
    }
    std::vector<int32_t>().swap(prims);
    float lhi[3] = {hi[0], hi[1], hi[2]};
    float rlo[3] = {lo[0], lo[1], lo[2]};
    lhi[s.axis] = s.pos;
    rlo[s.axis] = s.pos;
    nodes[id].axis = s.axis;
    nodes[id].split = s.pos;
    int c0 = build(lo, lhi, std::move(L), depth + 1);
    int c1 = build(rlo, hi, std::move(R), depth + 1);
    nodes[id].child[0] = c0;
    nodes[id].child[1] = c1;
    return id;
  }

  // Rope propagation (KDTree.cpp:636-692): the below child's high-side face
  // on the split axis points at the above child and vice versa.
  void link(int id, const int* ropes) {
    BNode& b = nodes[id];
    if (b.leaf()) {
      for (int f = 0; f < 6; ++f) b.rope[f] = push_down(ropes[f], f, b.bmin, b.bmax);
      return;
    }
    int sl, sr;
    if (b.axis == 0) { sl = LEFT; sr = RIGHT; }
    else if (b.axis == 1) { sl = BOTTOM; sr = TOP; }
    else { sl = BACK; sr = FRONT; }
    int rl[6], rr[6];
    for (int f = 0; f < 6; ++f) rl[f] = rr[f] = ropes[f];
    rl[sr] = b.child[1];
    rr[sl] = b.child[0];
    int c0 = b.child[0], c1 = b.child[1];
    link(c0, rl);
    link(c1, rr);
  }

  int push_down(int r, int face, const float* lo, const float* hi) const {
    while (r >= 0 && !nodes[r].leaf()) {
      const BNode& t = nodes[r];
      if (t.axis == kFaceAxis[face]) {
        r = kFaceHigh[face] ? t.child[0] : t.child[1];
      } else if (t.split <= lo[t.axis]) {
        r = t.child[1];
      } else if (t.split >= hi[t.axis]) {
        r = t.child[0];
      } else {
        break;
      }
    }
    return r;
  }
};

}  // namespace

extern "C" {

RE_API void re_kd_default_params(re_kd_build_params* p) {
  if (!p) return;
  p->max_leaf_tris = 8;
  p->max_depth = 24;
  p->cost_traversal = 1.0f;
  p->cost_intersect = 1.5f;
  p->empty_bonus = 0.8f;
  p->n_bins = 32;
}

RE_API re_status re_kd_build(const float* tri_v, int64_t n_tris, const re_kd_build_params* params,
                             re_kdtree** out) {
  if (!out || n_tris < 0 || (n_tris > 0 && !tri_v)) return fail(RE_ERR_INVALID, "re_kd_build: bad arguments");
  if (n_tris >= (int64_t)0x3fffffff) return fail(RE_ERR_INVALID, "re_kd_build: too many triangles");
  for (int64_t i = 0; i < 9 * n_tris; ++i)
    if (!std::isfinite(tri_v[i])) return fail(RE_ERR_INVALID, "re_kd_build: non-finite vertex at float %lld", (long long)i);
  re_kd_build_params p;
  re_kd_default_params(&p);
  if (params) p = *params;
  if (p.max_leaf_tris < 1 || p.max_depth < 1 || p.max_depth > 64)
    return fail(RE_ERR_INVALID, "re_kd_build: max_leaf_tris >= 1 and max_depth in [1,64]");
  re_kdtree* tree = new (std::nothrow) re_kdtree();
  if (!tree) return fail(RE_ERR_OOM, "re_kd_build: out of host memory");
  try {
    Builder b(tri_v, n_tris, p);
    b.run();
    const std::vector<BNode>& bn = b.nodes;
    // Preorder flattening: below child follows its parent, above child index
    // in meta; leaves numbered in the same order.
    std::vector<int> entry(bn.size(), -1), leaf_id(bn.size(), -1);
    tree->nodes.resize(bn.size());
    int next_entry = 0, next_leaf = 0;
    int depth = 0, max_leaf = 0;
    std::vector<int> stack;
    stack.push_back(0);
    std::vector<std::pair<int, int>> fix;  // (entry of parent, bnode of above child)
    // Iterative preorder: push above child first so below is visited next.
    while (!stack.empty()) {
      int id = stack.back();
      stack.pop_back();
      int e = next_entry++;
      entry[id] = e;
      const BNode& nd = bn[id];
      if (nd.leaf()) {
        int lid = next_leaf++;
        leaf_id[id] = lid;
        tree->nodes[e].split = 0.0f;
        tree->nodes[e].meta = kLeafMeta | (uint32_t)lid;
        depth = std::max(depth, nd.depth);
        max_leaf = std::max(max_leaf, (int)nd.prims.size());
      } else {
        tree->nodes[e].split = nd.split;
        tree->nodes[e].meta = (uint32_t)nd.axis << 30;
        fix.emplace_back(e, nd.child[1]);
        stack.push_back(nd.child[1]);
        stack.push_back(nd.child[0]);
      }
    }
    for (auto& f : fix) tree->nodes[f.first].meta |= (uint32_t)entry[f.second];
    tree->leaves.resize(next_leaf);
    // Leaves and triangle records in leaf-id order.
    std::vector<int> by_leaf(next_leaf);
    for (size_t i = 0; i < bn.size(); ++i)
      if (leaf_id[i] >= 0) by_leaf[leaf_id[i]] = (int)i;
    size_t total = 0;
    for (int l = 0; l < next_leaf; ++l) total += bn[by_leaf[l]].prims.size();
    tree->tris.resize(total);
    size_t cur = 0;
    for (int l = 0; l < next_leaf; ++l) {
      const BNode& nd = bn[by_leaf[l]];
      LeafRec& lr = tree->leaves[l];
      std::memset(&lr, 0, sizeof(lr));
      for (int a = 0; a < 3; ++a) {
        lr.bmin[a] = nd.bmin[a];
        lr.bmax[a] = nd.bmax[a];
      }
      for (int f = 0; f < 6; ++f) {
        int r = nd.rope[f];
        lr.rope[f] = r < 0 ? kNoRope : (bn[r].leaf() ? leaf_ref((uint32_t)leaf_id[r]) : entry[r]);
      }
      lr.first = (uint32_t)cur;
      lr.count = (uint32_t)nd.prims.size();
      for (int32_t pi : nd.prims) {
        const float* t = tri_v + 9 * (int64_t)pi;
        TriRec& tr = tree->tris[cur++];
        for (int a = 0; a < 3; ++a) {
          tr.v0[a] = t[a];
          tr.e1[a] = t[3 + a] - t[a];  // RayEngine/src/World/RTScene.cpp:228-248
          tr.e2[a] = t[6 + a] - t[a];
        }
        tr.prim = pi;
      }
    }
    re_kd_info& in = tree->info;
    in.n_tris = n_tris;
    in.n_nodes = (int64_t)tree->nodes.size();
    in.n_leaves = (int64_t)tree->leaves.size();
    in.n_trirecs = (int64_t)tree->tris.size();
    in.depth = depth;
    in.max_leaf = max_leaf;
    for (int a = 0; a < 3; ++a) {
      in.root_min[a] = b.root_min[a];
      in.root_max[a] = b.root_max[a];
    }
  } catch (const std::bad_alloc&) {
    delete tree;
    return fail(RE_ERR_OOM, "re_kd_build: out of host memory");
  }
  *out = tree;
  return RE_OK;
}

RE_API void re_kd_free(re_kdtree* tree) { delete tree; }

RE_API re_status re_kd_get_info(const re_kdtree* tree, re_kd_info* info) {
  if (!tree || !info) return fail(RE_ERR_INVALID, "re_kd_get_info: null argument");
  *info = tree->info;
  return RE_OK;
}

RE_API re_status re_kd_export(const re_kdtree* tree, void* nodes, void* leaves, void* tris) {
  if (!tree) return fail(RE_ERR_INVALID, "re_kd_export: null tree");
  if (nodes) std::memcpy(nodes, tree->nodes.data(), tree->nodes.size() * sizeof(NodeEntry));
  if (leaves) std::memcpy(leaves, tree->leaves.data(), tree->leaves.size() * sizeof(LeafRec));
  if (tris) std::memcpy(tris, tree->tris.data(), tree->tris.size() * sizeof(TriRec));
  return RE_OK;
}

}  // extern "C"

// Internal accessors for the device upload (capi.cpp).
namespace re {
const std::vector<NodeEntry>& kd_nodes(const re_kdtree* t) { return t->nodes; }
const std::vector<LeafRec>& kd_leaves(const re_kdtree* t) { return t->leaves; }
const std::vector<TriRec>& kd_tris(const re_kdtree* t) { return t->tris; }
}  // namespace re
