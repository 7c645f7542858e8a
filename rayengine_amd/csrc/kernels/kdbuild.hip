// Rope kd-tree construction on the device, breadth first: every level's
// nodes choose their split planes in parallel (one wave per node), their
// triangle references are partitioned (straddlers clipped in double
// precision) and scattered, in parent order, into the next level; then the
// preorder numbering, the ropes and the flat arrays are computed from the
// finished node set.  Each decision is the host builder's (host/kdtree.cpp)
// with the same floating-point operations in the same order, so the result
// is the host tree, entry for entry; the host build stays the reference the
// device build is tested against.  Replaces the host build of RKDTreeCPU
// (RayEngine/src/Primitives/KDTree.cpp:335-834) and its flattening
// (KDThreeGPU.cpp:155-194), which the reference runs on the CPU before every
// upload (timed at RTScene.cpp:186-198).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../capi_util.hpp"
#include "../kdbuild.hpp"

namespace re {
namespace kdb {
namespace {

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ float area_of(float dx, float dy, float dz) { return 2.0f * (dx * dy + dx * dz + dy * dz); }

// Surface-area cost of splitting box [lo, hi] on `axis` at `pos` with nl
// references below and nr above (host: Builder::cost_of).
__device__ __forceinline__ float split_cost(const Params& P, const float* lo, const float* hi, int axis, float pos,
                                            int64_t nl, int64_t nr, float inv_sa) {
  const float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
  float al, ar;
  if (axis == 0) {
    al = area_of(pos - lo[0], ey, ez);
    ar = area_of(hi[0] - pos, ey, ez);
  } else if (axis == 1) {
    al = area_of(ex, pos - lo[1], ez);
    ar = area_of(ex, hi[1] - pos, ez);
  } else {
    al = area_of(ex, ey, pos - lo[2]);
    ar = area_of(ex, ey, hi[2] - pos);
  }
  float c = P.ct + P.ci * (al * (float)nl + ar * (float)nr) * inv_sa;
  if (nl == 0 || nr == 0) c *= P.eb;
  return c;
}

// nextafter(x, +inf) and nextafter(x, -inf) for finite or infinite x.
__device__ __forceinline__ float next_up(float x) {
  if (x != x || x == kInf) return x;
  if (x == 0.0f) return __uint_as_float(1u);
  const uint32_t u = __float_as_uint(x);
  return __uint_as_float(x > 0.0f ? u + 1u : u - 1u);
}
__device__ __forceinline__ float next_down(float x) {
  if (x != x || x == -kInf) return x;
  if (x == 0.0f) return __uint_as_float(0x80000001u);
  const uint32_t u = __float_as_uint(x);
  return __uint_as_float(x > 0.0f ? u - 1u : u + 1u);
}

__device__ __forceinline__ float get3(const float* v, int a) { return a == 0 ? v[0] : (a == 1 ? v[1] : v[2]); }

// std::min / std::max as the host builder uses them (the first argument
// unless the second compares strictly smaller / larger), so equal values of
// opposite zero sign are picked as on the host.
template <typename T>
__device__ __forceinline__ T smin(T a, T b) { return b < a ? b : a; }
template <typename T>
__device__ __forceinline__ T smax(T a, T b) { return a < b ? b : a; }

// Is (c, a, p) a better split than (bc, ba, bp)?  The host scans axes in
// order and candidates in ascending position and keeps a candidate only if
// its cost is strictly lower, so the winner is the lowest cost, then the
// lowest axis, then the lowest position; costs that are not below +inf
// (including NaN) never win.
__device__ __forceinline__ bool better(float c, int a, float p, float bc, int ba, float bp) {
  if (!(c < kInf)) return false;
  if (c < bc) return true;
  return c == bc && (a < ba || (a == ba && p < bp));
}

// Bounding box of triangle `prim` clipped to [lo, hi] grown by a relative
// tolerance (Sutherland-Hodgman in double precision), intersected with
// [lo, hi] (host: Builder::clip_box).  A triangle clipped by six planes has
// at most 9 vertices.
__device__ bool clip_box(const float* __restrict__ tv, int32_t prim, const float* lo, const float* hi, Ref& out) {
  constexpr int kMaxV = 10;
  double poly[kMaxV][3], tmp[kMaxV][3];
  int np = 3;
  const float* t = tv + 9 * (int64_t)prim;
  for (int v = 0; v < 3; ++v)
    for (int a = 0; a < 3; ++a) poly[v][a] = t[3 * v + a];
  for (int a = 0; a < 3 && np > 0; ++a) {
    const double tol = 1e-6 * (fabs((double)lo[a]) + fabs((double)hi[a]) + 1e-3);
    for (int side = 0; side < 2 && np > 0; ++side) {
      const double plane = side == 0 ? (double)lo[a] - tol : (double)hi[a] + tol;
      const double sgn = side == 0 ? 1.0 : -1.0;
      int nq = 0;
      for (int v = 0; v < np; ++v) {
        const double* Pv = poly[v];
        const double* Qv = poly[(v + 1) % np];
        const double dp = sgn * (Pv[a] - plane), dq = sgn * (Qv[a] - plane);
        if (dp >= 0.0 && nq < kMaxV) {
          for (int k = 0; k < 3; ++k) tmp[nq][k] = Pv[k];
          ++nq;
        }
        if ((dp >= 0.0) != (dq >= 0.0) && nq < kMaxV) {
          const double s = dp / (dp - dq);
          for (int k = 0; k < 3; ++k) tmp[nq][k] = Pv[k] + s * (Qv[k] - Pv[k]);
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
      mn = smin(mn, poly[v][a]);
      mx = smax(mx, poly[v][a]);
    }
    const float fl = next_down((float)mn), fh = next_up((float)mx);
    out.lo[a] = smax(fl, lo[a]);
    out.hi[a] = smin(fh, hi[a]);
    if (out.lo[a] > out.hi[a]) out.lo[a] = out.hi[a] = smin(smax((float)mn, lo[a]), hi[a]);
  }
  return true;
}

// ---------------------------------------------------------------- per level

__global__ void init_refs_kernel(const float* __restrict__ tv, int64_t n, Ref* __restrict__ refs,
                                 int32_t* __restrict__ ref_node) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* t = tv + 9 * i;
  Ref r;
  r.prim = (int32_t)i;
  for (int a = 0; a < 3; ++a) {
    r.lo[a] = smin(smin(t[a], t[3 + a]), t[6 + a]);
    r.hi[a] = smax(smax(t[a], t[3 + a]), t[6 + a]);
  }
  refs[i] = r;
  ref_node[i] = 0;
}

__global__ void classify_kernel(LNode* __restrict__ nodes, int32_t nn, Params P) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  LNode& nd = nodes[i];
  nd.mode = (nd.n <= P.max_leaf || nd.depth >= P.max_depth) ? kModeLeaf
                                                             : (nd.n <= kExactMax ? kModeExact : kModeBinned);
  nd.axis = -1;
  nd.pos = 0.0f;
  nd.nl = nd.nr = 0;
  nd.cl = nd.cr = 0;
  nd.split = 0;
}

// One wave per node with at most 64 references: every reference's minimum
// and the float just past its maximum are candidate planes (host:
// find_split, exact branch).
__global__ void __launch_bounds__(256) exact_split_kernel(LNode* __restrict__ nodes, int32_t nn,
                                                          const Ref* __restrict__ refs, Params P) {
  const int lane = threadIdx.x & 63;
  const int w = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (w >= nn) return;
  LNode& nd = nodes[w];
  if (nd.mode != kModeExact) return;
  const int n = nd.n;
  const bool have = lane < n;
  Ref r{};
  if (have) r = refs[nd.rbegin + lane];
  const float sa = area_of(nd.hi[0] - nd.lo[0], nd.hi[1] - nd.lo[1], nd.hi[2] - nd.lo[2]);
  const float inv_sa = sa > 0.0f ? 1.0f / sa : 0.0f;
  float bc = kInf, bp = 0.0f;
  int ba = -1;
  int64_t bnl = 0, bnr = 0;
  for (int a = 0; a < 3; ++a) {
    const float lo = get3(nd.lo, a), hi = get3(nd.hi, a);
    const float ext = hi - lo;
    if (!(ext > 0.0f)) continue;
    const float rl = get3(r.lo, a), rh = get3(r.hi, a);
    const float c0 = rl, c1 = next_up(rh);
    for (int j = 0; j < 2 * n; ++j) {
      const float mine = (j & 1) ? c1 : c0;
      const float pos = __shfl(mine, j >> 1);
      if (!(pos > lo && pos < hi)) continue;
      const int64_t nl = __popcll(__ballot(have && rl < pos));
      const int64_t nr = __popcll(__ballot(have && rh >= pos));
      const float c = split_cost(P, nd.lo, nd.hi, a, pos, nl, nr, inv_sa);
      if (better(c, a, pos, bc, ba, bp)) {
        bc = c;
        ba = a;
        bp = pos;
        bnl = nl;
        bnr = nr;
      }
    }
  }
  if (lane == 0) {
    nd.axis = ba;
    nd.pos = bp;
    nd.nl = (int32_t)bnl;
    nd.nr = (int32_t)bnr;
  }
}

// Bin counts of the nodes with more than 64 references: per axis, the bin
// of every reference's clamped minimum and of its clamped maximum (host:
// find_split, binned branch).  A block whose chunk of references belongs to
// one node counts in LDS and adds its histogram once.
constexpr int kBinChunk = 1024;

__device__ __forceinline__ int bin_of(float v, float lo, float hi, float scale, int nb) {
  const float pc = smin(smax(v, lo), hi);
  return min(nb - 1, max(0, (int)((pc - lo) * scale)));
}

__global__ void __launch_bounds__(256) bin_count_kernel(const Ref* __restrict__ refs,
                                                        const int32_t* __restrict__ ref_node, int64_t nref,
                                                        const LNode* __restrict__ nodes, int32_t* __restrict__ hist,
                                                        Params P) {
  __shared__ int32_t sh[3 * 2 * kMaxBins];
  const int nb = max(4, P.n_bins);
  const int64_t c0 = (int64_t)blockIdx.x * kBinChunk;
  if (c0 >= nref) return;
  const int64_t c1 = min(c0 + (int64_t)kBinChunk, nref);
  const int32_t first = ref_node[c0], last = ref_node[c1 - 1];
  const bool one = first == last;
  if (one && nodes[first].mode != kModeBinned) return;
  if (one)
    for (int k = threadIdx.x; k < 6 * nb; k += blockDim.x) sh[k] = 0;
  __syncthreads();
  for (int64_t i = c0 + threadIdx.x; i < c1; i += blockDim.x) {
    const int32_t node = one ? first : ref_node[i];
    const LNode& nd = nodes[node];
    if (nd.mode != kModeBinned) continue;
    const Ref r = refs[i];
    int32_t* g = hist + (int64_t)node * 6 * nb;
    for (int a = 0; a < 3; ++a) {
      const float lo = get3(nd.lo, a), hi = get3(nd.hi, a), ext = hi - lo;
      if (!(ext > 0.0f)) continue;
      const float scale = (float)nb / ext;
      const int bl = bin_of(get3(r.lo, a), lo, hi, scale, nb), bh = bin_of(get3(r.hi, a), lo, hi, scale, nb);
      if (one) {
        atomicAdd(&sh[(2 * a) * nb + bl], 1);
        atomicAdd(&sh[(2 * a + 1) * nb + bh], 1);
      } else {
        atomicAdd(&g[(2 * a) * nb + bl], 1);
        atomicAdd(&g[(2 * a + 1) * nb + bh], 1);
      }
    }
  }
  if (!one) return;
  __syncthreads();
  int32_t* g = hist + (int64_t)first * 6 * nb;
  for (int k = threadIdx.x; k < 6 * nb; k += blockDim.x)
    if (sh[k]) atomicAdd(&g[k], sh[k]);
}

// One wave per binned node: lane k costs the plane between bins k-1 and k.
__global__ void __launch_bounds__(256) binned_split_kernel(LNode* __restrict__ nodes, int32_t nn,
                                                           const int32_t* __restrict__ hist, Params P) {
  const int lane = threadIdx.x & 63;
  const int w = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (w >= nn) return;
  LNode& nd = nodes[w];
  if (nd.mode != kModeBinned) return;
  const int nb = max(4, P.n_bins);
  const int64_t n = nd.n;
  const int32_t* g = hist + (int64_t)w * 6 * nb;
  const float sa = area_of(nd.hi[0] - nd.lo[0], nd.hi[1] - nd.lo[1], nd.hi[2] - nd.lo[2]);
  const float inv_sa = sa > 0.0f ? 1.0f / sa : 0.0f;
  float bc = kInf, bp = 0.0f;
  int ba = -1;
  int64_t bnl = 0, bnr = 0;
  for (int a = 0; a < 3; ++a) {
    const float lo = get3(nd.lo, a), hi = get3(nd.hi, a);
    const float ext = hi - lo;
    if (!(ext > 0.0f)) continue;
    const int k = lane;
    float c = kInf, pos = 0.0f;
    int64_t nl = 0, nr = n;
    if (k >= 1 && k < nb) {
      for (int j = 0; j < k; ++j) {
        nl += g[(2 * a) * nb + j];
        nr -= g[(2 * a + 1) * nb + j];
      }
      pos = lo + ext * (float)k / (float)nb;
      if (pos > lo && pos < hi) c = split_cost(P, nd.lo, nd.hi, a, pos, nl, nr, inv_sa);
      if (!(c < kInf)) c = kInf;  // NaN never wins; keep the reduction's order total
    }
    // lowest cost, then lowest k, across the wave
    float wc = c;
    int wk = (c < kInf) ? k : 1 << 20;
    for (int off = 32; off >= 1; off >>= 1) {
      const float oc = __shfl_xor(wc, off);
      const int ok = __shfl_xor(wk, off);
      if (oc < wc || (oc == wc && ok < wk)) {
        wc = oc;
        wk = ok;
      }
    }
    if (wk < (1 << 20)) {
      const float wpos = __shfl(pos, wk);
      const int64_t wnl = __shfl(nl, wk), wnr = __shfl(nr, wk);
      if (better(wc, a, wpos, bc, ba, bp)) {
        bc = wc;
        ba = a;
        bp = wpos;
        bnl = wnl;
        bnr = wnr;
      }
    }
  }
  if (lane == 0) {
    nd.axis = ba;
    nd.pos = bp;
    nd.nl = (int32_t)bnl;
    nd.nr = (int32_t)bnr;
  }
}

__device__ __forceinline__ bool will_try(const LNode& nd) {
  return nd.mode != kModeLeaf && nd.axis >= 0 && !(nd.nl == nd.n && nd.nr == nd.n);
}

// Adds, per node, the below/above counts vl, vr (0 or 1) of the wave's
// lanes; key is the lane's node (-1 for none).  The lanes of a wave usually
// share one node (a node's references are contiguous), so one atomic per
// wave replaces up to 64 on the same address.
__device__ __forceinline__ void wave_count(LNode* __restrict__ nodes, int key, int vl, int vr) {
  const int lane = threadIdx.x & 63;
  const unsigned long long act = __ballot(1);
  const int leader = __ffsll((unsigned long long)act) - 1;
  const int lk = __shfl(key, leader);
  const bool same = key == lk;
  const unsigned long long ml = __ballot(same && vl), mr = __ballot(same && vr);
  if (lane == leader && lk >= 0) {
    if (ml) atomicAdd(&nodes[lk].cl, __popcll(ml));
    if (mr) atomicAdd(&nodes[lk].cr, __popcll(mr));
  }
  if (!same && key >= 0) {
    if (vl) atomicAdd(&nodes[key].cl, 1);
    if (vr) atomicAdd(&nodes[key].cr, 1);
  }
}

// Which side(s) each reference goes to (host: build, partition loop).
// References that straddle the plane while clipping is on are listed for
// part_clip_kernel, which clips them with full waves.
__global__ void __launch_bounds__(256) part_flags_kernel(const Ref* __restrict__ refs,
                                                         const int32_t* __restrict__ ref_node, int64_t nref,
                                                         LNode* __restrict__ nodes, Params P,
                                                         int32_t* __restrict__ fl, int32_t* __restrict__ fr,
                                                         Ref* __restrict__ box_l, Ref* __restrict__ box_r,
                                                         int32_t* __restrict__ strad, int32_t* __restrict__ n_strad) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool in = i < nref;
  const int key = in ? ref_node[i] : -1;
  int l = 0, rr = 0;
  bool st = false;
  if (in && will_try(nodes[key])) {
    const LNode& nd = nodes[key];
    const Ref r = refs[i];
    const bool to_l = get3(r.lo, nd.axis) < nd.pos, to_r = get3(r.hi, nd.axis) >= nd.pos;
    if (to_l && to_r && P.clip) {
      st = true;
    } else {
      l = to_l ? 1 : 0;
      rr = to_r ? 1 : 0;
      box_l[i] = r;
      box_r[i] = r;
    }
  }
  if (in) {
    fl[i] = l;
    fr[i] = rr;
  }
  const unsigned long long sm = __ballot(st);
  if (sm) {
    const int leader = __ffsll(sm) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(n_strad, __popcll(sm));
    base = __shfl(base, leader);
    if (st) strad[base + __popcll(sm & ((1ull << lane) - 1ull))] = (int32_t)i;
  }
  wave_count(nodes, key, l, rr);
}

// Thread 2s + side clips straddler s against the below (side 0) or above
// (side 1) child box (host: Builder::clip_box on both children).
__global__ void __launch_bounds__(256) part_clip_kernel(const float* __restrict__ tv, const Ref* __restrict__ refs,
                                                        const int32_t* __restrict__ ref_node, LNode* __restrict__ nodes,
                                                        const int32_t* __restrict__ strad,
                                                        const int32_t* __restrict__ n_strad, int32_t* __restrict__ fl,
                                                        int32_t* __restrict__ fr, Ref* __restrict__ box_l,
                                                        Ref* __restrict__ box_r) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t ns = *n_strad;
  if ((t - (threadIdx.x & 63)) >= 2 * ns) return;  // whole waves past the list leave together
  const bool in = t < 2 * ns;
  int key = -1, ok = 0;
  const int side = (int)(t & 1);
  if (in) {
    const int64_t i = strad[t >> 1];
    key = ref_node[i];
    const LNode& nd = nodes[key];
    float lo[3] = {nd.lo[0], nd.lo[1], nd.lo[2]}, hi[3] = {nd.hi[0], nd.hi[1], nd.hi[2]};
    if (side == 0) {
      if (nd.axis == 0) hi[0] = nd.pos;
      else if (nd.axis == 1) hi[1] = nd.pos;
      else hi[2] = nd.pos;
    } else {
      if (nd.axis == 0) lo[0] = nd.pos;
      else if (nd.axis == 1) lo[1] = nd.pos;
      else lo[2] = nd.pos;
    }
    Ref c;
    ok = clip_box(tv, refs[i].prim, lo, hi, c) ? 1 : 0;
    if (side == 0) {
      fl[i] = ok;
      if (ok) box_l[i] = c;
    } else {
      fr[i] = ok;
      if (ok) box_r[i] = c;
    }
  }
  wave_count(nodes, key, side == 0 ? ok : 0, side == 1 ? ok : 0);
}

// Totals of the three exclusive scans over a level's nodes.
__global__ void totals_kernel(const int32_t* kids, const int32_t* kx, const int32_t* crefs, const int32_t* cx,
                              const int32_t* leafn, const int32_t* lx, int64_t nn, int64_t* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    out[0] = (int64_t)kx[nn - 1] + kids[nn - 1];
    out[1] = (int64_t)cx[nn - 1] + crefs[nn - 1];
    out[2] = (int64_t)lx[nn - 1] + leafn[nn - 1];
  }
}

__global__ void decide_kernel(LNode* __restrict__ nodes, int32_t nn, int32_t* __restrict__ kids,
                              int32_t* __restrict__ crefs, int32_t* __restrict__ leafn) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  LNode& nd = nodes[i];
  nd.split = (will_try(nd) && !(nd.cl == nd.n && nd.cr == nd.n)) ? 1 : 0;
  kids[i] = nd.split ? 2 : 0;
  crefs[i] = nd.split ? nd.cl + nd.cr : 0;
  leafn[i] = nd.split ? 0 : nd.n;
}

__global__ void emit_kernel(LNode* __restrict__ nodes, int32_t nn, const int32_t* __restrict__ kids_x,
                            const int32_t* __restrict__ crefs_x, const int32_t* __restrict__ leafn_x,
                            LNode* __restrict__ next, int32_t next_gid, int64_t pool_base, GNode* __restrict__ gn) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  LNode& nd = nodes[i];
  GNode g;
  for (int a = 0; a < 3; ++a) {
    g.lo[a] = nd.lo[a];
    g.hi[a] = nd.hi[a];
  }
  g.depth = nd.depth;
  if (nd.split) {
    const int c = kids_x[i];
    nd.child = c;
    nd.crbegin = crefs_x[i];
    LNode b{}, u{};
    for (int a = 0; a < 3; ++a) {
      b.lo[a] = u.lo[a] = nd.lo[a];
      b.hi[a] = u.hi[a] = nd.hi[a];
    }
    if (nd.axis == 0) { b.hi[0] = nd.pos; u.lo[0] = nd.pos; }
    else if (nd.axis == 1) { b.hi[1] = nd.pos; u.lo[1] = nd.pos; }
    else { b.hi[2] = nd.pos; u.lo[2] = nd.pos; }
    b.rbegin = nd.crbegin;
    b.n = nd.cl;
    u.rbegin = nd.crbegin + nd.cl;
    u.n = nd.cr;
    b.depth = u.depth = nd.depth + 1;
    b.gid = next_gid + c;
    u.gid = next_gid + c + 1;
    next[c] = b;
    next[c + 1] = u;
    g.axis = nd.axis;
    g.split = nd.pos;
    g.child0 = b.gid;
    g.child1 = u.gid;
    g.pool_n = 0;
    g.pool = 0;
  } else {
    nd.pool = pool_base + leafn_x[i];
    g.axis = -1;
    g.split = 0.0f;
    g.child0 = g.child1 = -1;
    g.pool_n = nd.n;
    g.pool = nd.pool;
  }
  gn[nd.gid] = g;
}

__global__ void scatter_kernel(const Ref* __restrict__ refs, const int32_t* __restrict__ ref_node, int64_t nref,
                               const LNode* __restrict__ nodes, const int32_t* __restrict__ fl,
                               const int32_t* __restrict__ fr, const int32_t* __restrict__ fl_x,
                               const int32_t* __restrict__ fr_x, const Ref* __restrict__ box_l,
                               const Ref* __restrict__ box_r, Ref* __restrict__ next_refs,
                               int32_t* __restrict__ next_node, int32_t* __restrict__ pool) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nref) return;
  const LNode& nd = nodes[ref_node[i]];
  if (nd.split) {
    if (fl[i]) {
      const int64_t k = nd.crbegin + (fl_x[i] - fl_x[nd.rbegin]);
      next_refs[k] = box_l[i];
      next_node[k] = nd.child;
    }
    if (fr[i]) {
      const int64_t k = nd.crbegin + nd.cl + (fr_x[i] - fr_x[nd.rbegin]);
      next_refs[k] = box_r[i];
      next_node[k] = nd.child + 1;
    }
  } else {
    pool[nd.pool + (i - nd.rbegin)] = refs[i].prim;
  }
}

// ---------------------------------------------------------------- finishing

__global__ void sizes_kernel(const GNode* __restrict__ gn, int32_t g0, int32_t cnt, int32_t* __restrict__ size,
                             int32_t* __restrict__ nleaf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  const int g = g0 + i;
  const GNode& n = gn[g];
  if (n.axis < 0) {
    size[g] = 1;
    nleaf[g] = 1;
  } else {
    size[g] = 1 + size[n.child0] + size[n.child1];
    nleaf[g] = nleaf[n.child0] + nleaf[n.child1];
  }
}

// Preorder entry index and leaf id (host: the iterative preorder of
// re_kd_build), and the ropes handed down (host: Builder::link).
__global__ void preorder_kernel(const GNode* __restrict__ gn, int32_t g0, int32_t cnt,
                                const int32_t* __restrict__ size, const int32_t* __restrict__ nleaf,
                                int32_t* __restrict__ entry, int32_t* __restrict__ lbase,
                                int32_t* __restrict__ ropes) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  const int g = g0 + i;
  const GNode& n = gn[g];
  if (n.axis < 0) return;
  const int c0 = n.child0, c1 = n.child1;
  entry[c0] = entry[g] + 1;
  entry[c1] = entry[g] + 1 + size[c0];
  lbase[c0] = lbase[g];
  lbase[c1] = lbase[g] + nleaf[c0];
  // faces: LEFT 0, FRONT 1, RIGHT 2, BACK 3, TOP 4, BOTTOM 5
  int sl, sr;
  if (n.axis == 0) { sl = 0; sr = 2; }
  else if (n.axis == 1) { sl = 5; sr = 4; }
  else { sl = 3; sr = 1; }
  for (int f = 0; f < 6; ++f) {
    const int r = ropes[6 * g + f];
    ropes[6 * c0 + f] = f == sr ? c1 : r;
    ropes[6 * c1 + f] = f == sl ? c0 : r;
  }
}

__constant__ int kFaceAxisD[6] = {0, 2, 0, 2, 1, 1};
__constant__ int kFaceHighD[6] = {0, 1, 1, 0, 1, 0};

// Leaf records: box, ropes pushed down to the deepest node covering the whole
// face (host: Builder::push_down), and the per-leaf reference counts.
__global__ void leaves_kernel(const GNode* __restrict__ gn, int32_t total, const int32_t* __restrict__ entry,
                              const int32_t* __restrict__ lbase, const int32_t* __restrict__ ropes,
                              LeafRec* __restrict__ leaves, int32_t* __restrict__ leaf_cnt,
                              int32_t* __restrict__ leaf_gid, int32_t* __restrict__ stats) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const GNode& n = gn[g];
  if (n.axis >= 0) return;
  const int lid = lbase[g];
  LeafRec lr;
  for (int a = 0; a < 3; ++a) {
    lr.bmin[a] = n.lo[a];
    lr.bmax[a] = n.hi[a];
  }
  for (int f = 0; f < 6; ++f) {
    int r = ropes[6 * g + f];
    while (r >= 0 && gn[r].axis >= 0) {
      const GNode& t = gn[r];
      if (t.axis == kFaceAxisD[f]) {
        r = kFaceHighD[f] ? t.child0 : t.child1;
      } else if (t.split <= get3(n.lo, t.axis)) {
        r = t.child1;
      } else if (t.split >= get3(n.hi, t.axis)) {
        r = t.child0;
      } else {
        break;
      }
    }
    lr.rope[f] = r < 0 ? kNoRope : (gn[r].axis < 0 ? leaf_ref((uint32_t)lbase[r]) : entry[r]);
  }
  lr.first = 0;
  lr.count = (uint32_t)n.pool_n;
  lr.pad[0] = lr.pad[1] = 0;
  leaves[lid] = lr;
  leaf_cnt[lid] = n.pool_n;
  leaf_gid[lid] = g;
  atomicMax(&stats[0], n.depth);
  atomicMax(&stats[1], n.pool_n);
}

__global__ void entries_kernel(const GNode* __restrict__ gn, int32_t total, const int32_t* __restrict__ entry,
                               const int32_t* __restrict__ lbase, NodeEntry* __restrict__ out) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const GNode& n = gn[g];
  NodeEntry e;
  if (n.axis < 0) {
    e.split = 0.0f;
    e.meta = kLeafMeta | (uint32_t)lbase[g];
  } else {
    e.split = n.split;
    e.meta = ((uint32_t)n.axis << 30) | (uint32_t)entry[n.child1];
  }
  out[entry[g]] = e;
}

// Triangle records in leaf order: v0, e1 = v1 - v0, e2 = v2 - v0, prim
// (RayEngine/src/World/RTScene.cpp:228-248).
__global__ void tris_kernel(const GNode* __restrict__ gn, int32_t n_leaves, const int32_t* __restrict__ leaf_gid,
                            const int32_t* __restrict__ first, const int32_t* __restrict__ pool,
                            const float* __restrict__ tv, LeafRec* __restrict__ leaves, TriRec* __restrict__ tris) {
  const int w = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (w >= n_leaves) return;
  const GNode& n = gn[leaf_gid[w]];
  const int64_t f = first[w];
  if (lane == 0) leaves[w].first = (uint32_t)f;
  for (int j = lane; j < n.pool_n; j += 64) {
    const int32_t pi = pool[n.pool + j];
    const float* t = tv + 9 * (int64_t)pi;
    TriRec tr;
    for (int a = 0; a < 3; ++a) {
      tr.v0[a] = t[a];
      tr.e1[a] = t[3 + a] - t[a];
      tr.e2[a] = t[6 + a] - t[a];
    }
    tr.prim = pi;
    tris[f + j] = tr;
  }
}

// ---------------------------------------------------------------- host side

unsigned blocks(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// Device buffers freed on scope exit.  release() only marks a buffer; it is
// freed by flush(), which the caller calls after synchronizing the stream,
// so no buffer is freed while queued work may still use it.
struct Bufs {
  std::vector<void*> ptrs, dead;
  ~Bufs() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  hipError_t alloc(T** p, size_t count) {
    *p = nullptr;
    hipError_t e = hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) ptrs.push_back(*p);
    return e;
  }
  void release(void* p) {
    if (p && std::find(ptrs.begin(), ptrs.end(), p) != ptrs.end()) dead.push_back(p);
  }
  void flush() {
    for (void* p : dead) {
      auto it = std::find(ptrs.begin(), ptrs.end(), p);
      if (it != ptrs.end()) {
        (void)hipFree(p);
        ptrs.erase(it);
      }
    }
    dead.clear();
  }
};

// A device array that only grows; `keep` elements are copied on growth.
template <typename T>
struct Arr {
  T* p = nullptr;
  int64_t cap = 0;
  hipError_t ensure(Bufs& B, int64_t need, hipStream_t s, int64_t keep = 0) {
    if (need <= cap) return hipSuccess;
    const int64_t nc = std::max<int64_t>(need + need / 2, 64);
    T* q = nullptr;
    hipError_t e = B.alloc(&q, (size_t)nc);
    if (e != hipSuccess) return e;
    if (keep > 0) e = hipMemcpyAsync(q, p, (size_t)keep * sizeof(T), hipMemcpyDeviceToDevice, s);
    B.release(p);
    p = q;
    cap = nc;
    return e;
  }
};

// Exclusive prefix sum with a reusable temporary buffer.
struct Scanner {
  Arr<char> tmp;
  template <typename T>
  hipError_t run(Bufs& B, const T* in, T* out, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    size_t bytes = 0;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, s);
    if (e == hipSuccess) e = tmp.ensure(B, (int64_t)bytes, s);
    if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, in, out, (int)n, s);
    return e;
  }
};

}  // namespace

#define KDB_TRY(expr)                     \
  do {                                    \
    hipError_t e_ = (expr);               \
    if (e_ != hipSuccess) return e_;      \
  } while (0)

// Builds the tree of n (> 0) triangles; the outputs are host vectors.
// device_ms gets the device time of the build (events around all of it).
hipError_t build_device(const float* tri_v, int64_t n, const float* root_min, const float* root_max,
                        const Params& P, hipStream_t s, std::vector<NodeEntry>& out_nodes,
                        std::vector<LeafRec>& out_leaves, std::vector<TriRec>& out_tris, int32_t& depth,
                        int32_t& max_leaf, float& device_ms) {
  Bufs B;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  KDB_TRY(hipEventCreate(&ev0));
  KDB_TRY(hipEventCreate(&ev1));
  struct EvGuard {
    hipEvent_t a, b;
    ~EvGuard() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } evg{ev0, ev1};
  const int nb = std::max(4, P.n_bins);
  float* tv = nullptr;
  KDB_TRY(B.alloc(&tv, (size_t)(9 * n)));
  KDB_TRY(hipMemcpyAsync(tv, tri_v, (size_t)(9 * n) * sizeof(float), hipMemcpyHostToDevice, s));
  KDB_TRY(hipEventRecord(ev0, s));

  // Level arrays: current and next references / nodes swap every level.
  Arr<Ref> refs, nrefs, box_l, box_r;
  Arr<int32_t> ref_node, nnode, fl, fr, flx, frx, strad;
  Arr<LNode> lv, next;
  Arr<int32_t> kids, crefs, leafn, kx, cx, lx, hist;
  Arr<GNode> gn;
  Arr<int32_t> pool;
  Scanner scan;
  int64_t* tot = nullptr;
  int32_t* n_strad = nullptr;
  KDB_TRY(B.alloc(&tot, 3));
  KDB_TRY(B.alloc(&n_strad, 1));
  KDB_TRY(refs.ensure(B, n, s));
  KDB_TRY(ref_node.ensure(B, n, s));
  hipLaunchKernelGGL(init_refs_kernel, dim3(blocks(n, 256)), dim3(256), 0, s, tv, n, refs.p, ref_node.p);
  KDB_TRY(hipGetLastError());
  KDB_TRY(lv.ensure(B, 1, s));
  {
    LNode root{};
    for (int a = 0; a < 3; ++a) {
      root.lo[a] = root_min[a];
      root.hi[a] = root_max[a];
    }
    root.rbegin = 0;
    root.n = (int32_t)n;
    root.depth = 1;
    root.gid = 0;
    KDB_TRY(hipMemcpyAsync(lv.p, &root, sizeof(root), hipMemcpyHostToDevice, s));
  }
  int64_t nn = 1, nref = n, gused = 1, pused = 0;
  KDB_TRY(gn.ensure(B, 1024, s));
  KDB_TRY(pool.ensure(B, n + 1024, s));
  std::vector<std::pair<int32_t, int32_t>> levels;  // (first gid, count)
  int32_t gbase = 0;

  while (nn > 0) {
    levels.emplace_back(gbase, (int32_t)nn);
    hipLaunchKernelGGL(classify_kernel, dim3(blocks(nn, 256)), dim3(256), 0, s, lv.p, (int32_t)nn, P);
    hipLaunchKernelGGL(exact_split_kernel, dim3(blocks(nn * 64, 256)), dim3(256), 0, s, lv.p, (int32_t)nn, refs.p,
                       P);
    KDB_TRY(hist.ensure(B, nn * 6 * nb, s));
    KDB_TRY(hipMemsetAsync(hist.p, 0, (size_t)(nn * 6 * nb) * sizeof(int32_t), s));
    hipLaunchKernelGGL(bin_count_kernel, dim3(blocks(nref, kBinChunk)), dim3(256), 0, s, refs.p, ref_node.p, nref,
                       lv.p, hist.p, P);
    hipLaunchKernelGGL(binned_split_kernel, dim3(blocks(nn * 64, 256)), dim3(256), 0, s, lv.p, (int32_t)nn, hist.p,
                       P);
    KDB_TRY(hipGetLastError());

    for (Arr<int32_t>* a : {&fl, &fr, &flx, &frx, &strad}) KDB_TRY(a->ensure(B, nref, s));
    KDB_TRY(box_l.ensure(B, nref, s));
    KDB_TRY(box_r.ensure(B, nref, s));
    KDB_TRY(hipMemsetAsync(n_strad, 0, sizeof(int32_t), s));
    hipLaunchKernelGGL(part_flags_kernel, dim3(blocks(nref, 256)), dim3(256), 0, s, refs.p, ref_node.p, nref, lv.p, P,
                       fl.p, fr.p, box_l.p, box_r.p, strad.p, n_strad);
    if (P.clip)
      hipLaunchKernelGGL(part_clip_kernel, dim3(blocks(2 * nref, 256)), dim3(256), 0, s, tv, refs.p, ref_node.p,
                         lv.p, strad.p, n_strad, fl.p, fr.p, box_l.p, box_r.p);
    KDB_TRY(hipGetLastError());
    KDB_TRY(scan.run(B, fl.p, flx.p, nref, s));
    KDB_TRY(scan.run(B, fr.p, frx.p, nref, s));

    for (Arr<int32_t>* a : {&kids, &crefs, &leafn, &kx, &cx, &lx}) KDB_TRY(a->ensure(B, nn, s));
    hipLaunchKernelGGL(decide_kernel, dim3(blocks(nn, 256)), dim3(256), 0, s, lv.p, (int32_t)nn, kids.p, crefs.p,
                       leafn.p);
    KDB_TRY(hipGetLastError());
    KDB_TRY(scan.run(B, kids.p, kx.p, nn, s));
    KDB_TRY(scan.run(B, crefs.p, cx.p, nn, s));
    KDB_TRY(scan.run(B, leafn.p, lx.p, nn, s));
    hipLaunchKernelGGL(totals_kernel, dim3(1), dim3(64), 0, s, kids.p, kx.p, crefs.p, cx.p, leafn.p, lx.p, nn, tot);
    int64_t t3[3] = {0, 0, 0};
    KDB_TRY(hipMemcpyAsync(t3, tot, sizeof(t3), hipMemcpyDeviceToHost, s));
    KDB_TRY(hipStreamSynchronize(s));
    const int64_t n_next = t3[0], r_next = t3[1], l_now = t3[2];
    if (gbase + nn + n_next > (int64_t)0x3fffffff || r_next > (int64_t)0x7fffffff) return hipErrorInvalidValue;

    KDB_TRY(gn.ensure(B, gused + n_next, s, gused));
    KDB_TRY(pool.ensure(B, pused + l_now, s, pused));
    KDB_TRY(next.ensure(B, n_next, s));
    KDB_TRY(nrefs.ensure(B, r_next, s));
    KDB_TRY(nnode.ensure(B, r_next, s));
    const int32_t next_gid = gbase + (int32_t)nn;
    hipLaunchKernelGGL(emit_kernel, dim3(blocks(nn, 256)), dim3(256), 0, s, lv.p, (int32_t)nn, kx.p, cx.p, lx.p,
                       next.p, next_gid, pused, gn.p);
    hipLaunchKernelGGL(scatter_kernel, dim3(blocks(nref, 256)), dim3(256), 0, s, refs.p, ref_node.p, nref, lv.p,
                       fl.p, fr.p, flx.p, frx.p, box_l.p, box_r.p, nrefs.p, nnode.p, pool.p);
    KDB_TRY(hipGetLastError());
    std::swap(lv, next);
    std::swap(refs, nrefs);
    std::swap(ref_node, nnode);
    gused += n_next;
    pused += l_now;
    gbase = next_gid;
    nn = n_next;
    nref = r_next;
    if (!B.dead.empty()) {
      KDB_TRY(hipStreamSynchronize(s));
      B.flush();
    }
  }

  const int32_t total = (int32_t)gused;
  int32_t *size = nullptr, *nleaf = nullptr, *entry = nullptr, *lbase = nullptr, *ropes = nullptr;
  KDB_TRY(B.alloc(&size, (size_t)total));
  KDB_TRY(B.alloc(&nleaf, (size_t)total));
  KDB_TRY(B.alloc(&entry, (size_t)total));
  KDB_TRY(B.alloc(&lbase, (size_t)total));
  KDB_TRY(B.alloc(&ropes, (size_t)(6 * total)));
  for (auto it = levels.rbegin(); it != levels.rend(); ++it)
    hipLaunchKernelGGL(sizes_kernel, dim3(blocks(it->second, 256)), dim3(256), 0, s, gn.p, it->first, it->second,
                       size, nleaf);
  KDB_TRY(hipGetLastError());
  KDB_TRY(hipMemsetAsync(entry, 0, sizeof(int32_t), s));
  KDB_TRY(hipMemsetAsync(lbase, 0, sizeof(int32_t), s));
  KDB_TRY(hipMemsetAsync(ropes, 0xff, 6 * sizeof(int32_t), s));
  for (auto& lvl : levels)
    hipLaunchKernelGGL(preorder_kernel, dim3(blocks(lvl.second, 256)), dim3(256), 0, s, gn.p, lvl.first, lvl.second,
                       size, nleaf, entry, lbase, ropes);
  KDB_TRY(hipGetLastError());
  int32_t n_leaves = 0;
  KDB_TRY(hipMemcpyAsync(&n_leaves, nleaf, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  KDB_TRY(hipStreamSynchronize(s));

  LeafRec* leaves = nullptr;
  NodeEntry* entries = nullptr;
  int32_t *leaf_cnt = nullptr, *leaf_gid = nullptr, *first = nullptr, *stats = nullptr;
  KDB_TRY(B.alloc(&leaves, (size_t)n_leaves));
  KDB_TRY(B.alloc(&entries, (size_t)total));
  KDB_TRY(B.alloc(&leaf_cnt, (size_t)n_leaves));
  KDB_TRY(B.alloc(&leaf_gid, (size_t)n_leaves));
  KDB_TRY(B.alloc(&first, (size_t)n_leaves + 1));
  KDB_TRY(B.alloc(&stats, 2));
  KDB_TRY(hipMemsetAsync(stats, 0, 2 * sizeof(int32_t), s));
  hipLaunchKernelGGL(leaves_kernel, dim3(blocks(total, 256)), dim3(256), 0, s, gn.p, total, entry, lbase, ropes,
                     leaves, leaf_cnt, leaf_gid, stats);
  hipLaunchKernelGGL(entries_kernel, dim3(blocks(total, 256)), dim3(256), 0, s, gn.p, total, entry, lbase, entries);
  KDB_TRY(hipGetLastError());
  KDB_TRY(scan.run(B, leaf_cnt, first, n_leaves, s));
  int64_t n_tris = pused;  // every leaf reference becomes one triangle record
  TriRec* tris = nullptr;
  KDB_TRY(B.alloc(&tris, (size_t)n_tris));
  hipLaunchKernelGGL(tris_kernel, dim3(blocks((int64_t)n_leaves * 64, 256)), dim3(256), 0, s, gn.p, n_leaves,
                     leaf_gid, first, pool.p, tv, leaves, tris);
  KDB_TRY(hipGetLastError());
  KDB_TRY(hipEventRecord(ev1, s));

  out_nodes.resize((size_t)total);
  out_leaves.resize((size_t)n_leaves);
  out_tris.resize((size_t)n_tris);
  int32_t st[2] = {0, 0};
  KDB_TRY(hipMemcpyAsync(out_nodes.data(), entries, (size_t)total * sizeof(NodeEntry), hipMemcpyDeviceToHost, s));
  KDB_TRY(hipMemcpyAsync(out_leaves.data(), leaves, (size_t)n_leaves * sizeof(LeafRec), hipMemcpyDeviceToHost, s));
  KDB_TRY(hipMemcpyAsync(out_tris.data(), tris, (size_t)n_tris * sizeof(TriRec), hipMemcpyDeviceToHost, s));
  KDB_TRY(hipMemcpyAsync(st, stats, sizeof(st), hipMemcpyDeviceToHost, s));
  KDB_TRY(hipStreamSynchronize(s));
  KDB_TRY(hipEventElapsedTime(&device_ms, ev0, ev1));
  depth = st[0];
  max_leaf = st[1];
  return hipSuccess;
}

}  // namespace kdb
}  // namespace re
