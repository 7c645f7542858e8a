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

// A triangle reference inside a node: its id and the bounding box of the
// part of the triangle that lies in the node (the whole triangle's box when
// clipping is off).
struct Ref {
  int32_t prim;
  float lo[3], hi[3];
};

class Builder {
 public:
  Builder(const float* tv, int64_t n, const re_kd_build_params& p) : tv_(tv), n_(n), p_(p) {}

  std::vector<BNode> nodes;
  float root_min[3], root_max[3];

  void run() {
    for (int a = 0; a < 3; ++a) {
      root_min[a] = kInf;
      root_max[a] = -kInf;
    }
    std::vector<Ref> all((size_t)n_);
    for (int64_t i = 0; i < n_; ++i) {
      const float* t = tv_ + 9 * i;
      Ref& r = all[i];
      r.prim = (int32_t)i;
      for (int a = 0; a < 3; ++a) {
        r.lo[a] = std::min(std::min(t[a], t[3 + a]), t[6 + a]);
        r.hi[a] = std::max(std::max(t[a], t[3 + a]), t[6 + a]);
        root_min[a] = std::min(root_min[a], r.lo[a]);
        root_max[a] = std::max(root_max[a], r.hi[a]);
      }
    }
    nodes.reserve((size_t)std::max<int64_t>(16, n_ / 2));
    build(root_min, root_max, std::move(all), 1);
    int ropes[6] = {-1, -1, -1, -1, -1, -1};
    link(0, ropes);
  }

 private:
  const float* tv_;
  int64_t n_;
  re_kd_build_params p_;

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

  // The reference's spatial median (KDTree.cpp:192-199,368-389): the longest
  // side of the node box (x on a strict win over y and z, else y on a strict
  // win over z, else z) split at lo + (hi - lo) / 2.
  Split median_split(const float* lo, const float* hi, const std::vector<Ref>& refs) const {
    Split s;
    const float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
    const int a = (ex > ey && ex > ez) ? 0 : (ey > ez ? 1 : 2);
    const float pos = lo[a] + ((hi[a] - lo[a]) / 2.0f);
    if (!(pos > lo[a] && pos < hi[a])) return s;  // no room for a split
    s.axis = a;
    s.pos = pos;
    for (const Ref& r : refs) {
      if (r.lo[a] < pos) ++s.nl;
      if (r.hi[a] >= pos) ++s.nr;
    }
    return s;
  }

  Split find_split(const float* lo, const float* hi, const std::vector<Ref>& refs) const {
    if (p_.split_rule == 1) return median_split(lo, hi, refs);
    Split best;
    best.cost = INFINITY;
    float sa = area(lo, hi);
    float inv_sa = sa > 0.0f ? 1.0f / sa : 0.0f;
    const int64_t n = (int64_t)refs.size();
    for (int a = 0; a < 3; ++a) {
      float ext = hi[a] - lo[a];
      if (!(ext > 0.0f)) continue;
      if (n > 64) {
        const int nb = std::max(4, p_.n_bins);
        std::vector<int64_t> cmin(nb, 0), cmax(nb, 0);
        float scale = (float)nb / ext;
        for (const Ref& r : refs) {
          float pl = std::min(std::max(r.lo[a], lo[a]), hi[a]);
          float ph = std::min(std::max(r.hi[a], lo[a]), hi[a]);
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
        // Exact candidates: every reference's minimum (it goes above) and the
        // float just past every maximum (it goes below).
        std::vector<float> cand;
        cand.reserve(2 * n);
        for (const Ref& r : refs) {
          cand.push_back(r.lo[a]);
          cand.push_back(std::nextafter(r.hi[a], INFINITY));
        }
        std::sort(cand.begin(), cand.end());
        cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
        for (float pos : cand) {
          if (!(pos > lo[a] && pos < hi[a])) continue;
          int64_t nl = 0, nr = 0;
          for (const Ref& r : refs) {
            if (r.lo[a] < pos) ++nl;
            if (r.hi[a] >= pos) ++nr;
          }
          float c = cost_of(lo, hi, a, pos, nl, nr, inv_sa);
          if (c < best.cost) best = Split{a, pos, c, nl, nr};
        }
      }
    }
    return best;
  }

  // Bounding box of the triangle clipped to the box [lo, hi] grown by a
  // relative tolerance (Sutherland-Hodgman in double precision), intersected
  // with [lo, hi].  Returns false when nothing of the triangle lies in it.
  bool clip_box(int32_t prim, const float* lo, const float* hi, Ref& out) const {
    double poly[16][3], tmp[16][3];
    int np = 3;
    const float* t = tv_ + 9 * (int64_t)prim;
    for (int v = 0; v < 3; ++v)
      for (int a = 0; a < 3; ++a) poly[v][a] = t[3 * v + a];
    for (int a = 0; a < 3 && np > 0; ++a) {
      const double tol = 1e-6 * (std::fabs((double)lo[a]) + std::fabs((double)hi[a]) + 1e-3);
      for (int side = 0; side < 2 && np > 0; ++side) {
        // keep x_a >= lo - tol (side 0) or x_a <= hi + tol (side 1)
        const double plane = side == 0 ? (double)lo[a] - tol : (double)hi[a] + tol;
        const double sgn = side == 0 ? 1.0 : -1.0;
        int nq = 0;
        for (int v = 0; v < np; ++v) {
          const double* P = poly[v];
          const double* Q = poly[(v + 1) % np];
          const double dp = sgn * (P[a] - plane), dq = sgn * (Q[a] - plane);
          if (dp >= 0.0) {
            for (int k = 0; k < 3; ++k) tmp[nq][k] = P[k];
            ++nq;
          }
          if ((dp >= 0.0) != (dq >= 0.0)) {
            const double s = dp / (dp - dq);
            for (int k = 0; k < 3; ++k) tmp[nq][k] = P[k] + s * (Q[k] - P[k]);
            ++nq;
          }
        }
        np = nq;
        for (int v = 0; v < np; ++v)
          for (int k = 0; k < 3; ++k) poly[v][k] = tmp[v][k];
      }
    }
    if (np == 0) return false;
    out.prim = prim;
    for (int a = 0; a < 3; ++a) {
      double mn = poly[0][a], mx = poly[0][a];
      for (int v = 1; v < np; ++v) {
        mn = std::min(mn, poly[v][a]);
        mx = std::max(mx, poly[v][a]);
      }
      // Round outward so the float box still contains the clipped polygon.
      float fl = std::nextafter((float)mn, -INFINITY), fh = std::nextafter((float)mx, INFINITY);
      out.lo[a] = std::max(fl, lo[a]);
      out.hi[a] = std::min(fh, hi[a]);
      if (out.lo[a] > out.hi[a]) out.lo[a] = out.hi[a] = std::min(std::max((float)mn, lo[a]), hi[a]);
    }
    return true;
  }

  static std::vector<int32_t> ids(const std::vector<Ref>& refs) {
    std::vector<int32_t> v(refs.size());
    for (size_t i = 0; i < refs.size(); ++i) v[i] = refs[i].prim;
    return v;
  }

  int build(const float* lo, const float* hi, std::vector<Ref>&& refs, int depth) {
    int id = (int)nodes.size();
    nodes.emplace_back();
    {
      BNode& b = nodes[id];
      std::memcpy(b.bmin, lo, sizeof(b.bmin));
      std::memcpy(b.bmax, hi, sizeof(b.bmax));
      b.depth = depth;
    }
    const int64_t n = (int64_t)refs.size();
    if (n <= p_.max_leaf_tris || depth >= p_.max_depth) {
      nodes[id].prims = ids(refs);
      return id;
    }
    Split s = find_split(lo, hi, refs);
    if (s.axis < 0 || (s.nl == n && s.nr == n)) {
      nodes[id].prims = ids(refs);
      return id;
    }
    float lhi[3] = {hi[0], hi[1], hi[2]};
    float rlo[3] = {lo[0], lo[1], lo[2]};
    lhi[s.axis] = s.pos;
    rlo[s.axis] = s.pos;
    std::vector<Ref> L, R;
    L.reserve(s.nl);
    R.reserve(s.nr);
    for (const Ref& r : refs) {
      const bool to_l = r.lo[s.axis] < s.pos, to_r = r.hi[s.axis] >= s.pos;
      if (to_l && to_r && p_.clip) {
        // Straddler: keep it on a side only if the triangle itself reaches
        // into that child, with the box of the part that does.
        Ref c;
        if (clip_box(r.prim, lo, lhi, c)) L.push_back(c);
        if (clip_box(r.prim, rlo, hi, c)) R.push_back(c);
        continue;
      }
      if (to_l) L.push_back(r);
      if (to_r) R.push_back(r);
    }
    if ((int64_t)L.size() == n && (int64_t)R.size() == n) {
      nodes[id].prims = ids(refs);
      return id;
    }
    std::vector<Ref>().swap(refs);
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
  p->clip = 1;
  p->split_rule = 0;
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
  if (p.max_leaf_tris < 1 || p.max_depth < 1 || p.max_depth > 64 || p.split_rule < 0 || p.split_rule > 1)
    return fail(RE_ERR_INVALID, "re_kd_build: max_leaf_tris >= 1, max_depth in [1,64], split_rule 0 or 1");
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
re_kdtree* kd_from_arrays(std::vector<NodeEntry>&& nodes, std::vector<LeafRec>&& leaves, std::vector<TriRec>&& tris,
                          const re_kd_info& info) {
  re_kdtree* t = new (std::nothrow) re_kdtree();
  if (!t) return nullptr;
  t->nodes = std::move(nodes);
  t->leaves = std::move(leaves);
  t->tris = std::move(tris);
  t->info = info;
  return t;
}
const std::vector<NodeEntry>& kd_nodes(const re_kdtree* t) { return t->nodes; }
const std::vector<LeafRec>& kd_leaves(const re_kdtree* t) { return t->leaves; }
const std::vector<TriRec>& kd_tris(const re_kdtree* t) { return t->tris; }
}  // namespace re
