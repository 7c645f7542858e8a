// Per-ray device paths shared by the single-workload kernels and the batched
// work-queue kernel, so a ray gives bit-identical results whichever kernel
// runs it.
#pragma once

#include "internal.hpp"

namespace re {

// Bounds-checked debug build (compiled with RE_DEBUG_BOUNDS, in the
// translation unit that defines RE_DEBUG_BOUNDS_HERE: the single-scene trace
// kernels): every node window, leaf record and triangle run a walk reads is
// checked against the scene's arrays first.  An index outside them sets a bit
// in g_dbg_errors (1 node entry, 2 leaf record, 4 triangle run) and is
// replaced by 0 (an empty run), so the walk goes on without reading outside
// the arrays; its result is then meaningless and the error word says so.
#if defined(RE_DEBUG_BOUNDS) && defined(RE_DEBUG_BOUNDS_HERE)
struct DebugBounds {
  uint32_t n_entries, n_leaves, n_tris, pad;
};
static __constant__ DebugBounds g_dbg_bounds;
static __device__ unsigned int g_dbg_errors;
__device__ __forceinline__ uint32_t dbg_index(uint32_t i, uint32_t n, unsigned bit) {
  if (i < n) return i;
  atomicOr(&g_dbg_errors, bit);
  return 0u;
}
__device__ __forceinline__ void dbg_run(uint32_t& first, uint32_t& count) {
  if ((uint64_t)first + count <= g_dbg_bounds.n_tris) return;
  atomicOr(&g_dbg_errors, 4u);
  first = 0u;
  count = 0u;
}
#define RE_DBG_ENTRY(c) ((int32_t)dbg_index((uint32_t)(c), g_dbg_bounds.n_entries, 1u))
#define RE_DBG_LEAF(l) dbg_index((uint32_t)(l), g_dbg_bounds.n_leaves, 2u)
#define RE_DBG_RUN(f, n) dbg_run((f), (n))
#else
#define RE_DBG_ENTRY(c) (c)
#define RE_DBG_LEAF(l) (l)
#define RE_DBG_RUN(f, n) ((void)0)
#endif

struct TriGeom {
  f3 v0, e1, e2;
};

__device__ __forceinline__ TriGeom load_geom(const TriRec* __restrict__ tp, uint32_t j) {
  const float* f = tp[j].v0;
  return TriGeom{f3{f[0], f[1], f[2]}, f3{f[3], f[4], f[5]}, f3{f[6], f[7], f[8]}};
}

// Tests a leaf's triangles in record order (the reference's in-leaf order,
// ray_functions.cuh:295-324).  The geometry of up to four records is loaded
// before any of them is tested, so a lane has four independent loads in
// flight instead of one; the prim id is read only for an accepted hit.
//
// ANY: stop at the first accepted hit (any-hit query; the reference's
// intersection_detected becomes true there and never reverts).
// BATCH records are loaded per round (4 by default; fewer where registers
// are short).
template <bool ONE_SIDED, bool COUNT, bool ANY = false, int BATCH = 4>
__device__ __forceinline__ void test_leaf(const TriRec* __restrict__ tp, uint32_t count, f3 ro, f3 rd,
                                          float& t_exit, bool& found, float& hit_t, int32_t& hit_prim,
                                          unsigned& c_tri) {
  for (uint32_t j = 0; j < count; j += BATCH) {
    if (ANY && found) return;
    TriGeom g[BATCH];
#pragma unroll
    for (int q = 0; q < BATCH; ++q)
      if (j + q < count) g[q] = load_geom(tp, j + q);
#pragma unroll
    for (int q = 0; q < BATCH; ++q) {
      if (j + q >= count) break;
      if (COUNT) ++c_tri;
      float t;
      const bool h = ONE_SIDED ? ray_tri(ro, rd, g[q].v0, g[q].e1, g[q].e2, t)
                               : ray_tri_2s(ro, rd, g[q].v0, g[q].e1, g[q].e2, t);
      if (h && t < t_exit) {
        found = true;
        hit_t = t;
        hit_prim = tp[j + q].prim;
        t_exit = t;
        if (ANY) break;
      }
    }
  }
}

struct RayCounts {
  unsigned ent = 0, leaf = 0, tri = 0;
};

// The two entries starting at c as one 16-byte load per lane (the entry
// array is padded so a window at the last entry stays in bounds): .x/.y are
// the split bits and meta of entry c, .z/.w those of entry c + 1.
static_assert(kNodeWindow == 2, "the descent reads two-entry windows");
__device__ __forceinline__ uint4 load_window(const NodeEntry* __restrict__ nodes, int32_t c) {
  uint4 q;
  __builtin_memcpy(&q, nodes + RE_DBG_ENTRY(c), sizeof(q));
  return q;
}

// Descends from entry cur, whose window q is loaded, to the leaf containing
// pe (the down-walk of ray_functions.cuh:289-291: below if pe < split, else
// above).  The below child of an entry is the next entry, so a "below" step
// from the window's first entry needs no new load; any other step does.
// Returns the leaf id.  The loop is bounded by the tree: the builders number
// entries in preorder, so both children of an entry lie after it and inside
// the array (checked on the host for every built tree shape), and every step
// goes one level down; a descent ends at a leaf within the tree's depth.
template <bool COUNT>
__device__ __forceinline__ uint32_t descend(const NodeEntry* __restrict__ nodes, int32_t cur, uint4 q, f3 pe,
                                            unsigned& c_ent) {
  for (;;) {
    if (COUNT) ++c_ent;
    uint32_t meta = q.y;
    uint32_t ax = meta >> 30;
    if (ax == 3u) return meta & 0x3fffffffu;
    if (get(pe, (int)ax) < __uint_as_float(q.x)) {
      if (COUNT) ++c_ent;
      meta = q.w;
      ax = meta >> 30;
      if (ax == 3u) return meta & 0x3fffffffu;
      cur = (get(pe, (int)ax) < __uint_as_float(q.z)) ? cur + 2 : (int32_t)(meta & 0x3fffffffu);
    } else {
      cur = (int32_t)(meta & 0x3fffffffu);
    }
    q = load_window(nodes, cur);
  }
}

// The state that decides a walk's next visit is (cur, t_entry, t_exit).
// When it equals the state of one of the last two visits, no hit was
// accepted in between (an accepted hit lowers t_exit), so from here the walk
// repeats the same visits, accepting nothing, until the visit cap: the
// result is final and the walk stops.  This happens when a leaf's exit point
// lies within the face tolerance of two faces and the ropes lead back and
// forth between two leaves at the same t_entry (ray_functions.cuh:341-347);
// the reference walks on to the cap.
struct WalkMemo {
  int32_t c1 = kNoRope, c2 = kNoRope;  // cur is never kNoRope inside a walk
  float e1 = 0.0f, e2 = 0.0f, x1 = 0.0f, x2 = 0.0f;
  __device__ __forceinline__ bool repeats(int32_t cur, float te, float tx) {
    if ((cur == c1 && te == e1 && tx == x1) || (cur == c2 && te == e2 && tx == x2)) return true;
    c2 = c1;
    e2 = e1;
    x2 = x1;
    c1 = cur;
    e1 = te;
    x1 = tx;
    return false;
  }
};

// Closest hit of one ray over objects [k0, k1): stackless_kdtree_traversal and
// trace_scene, RayEngine/src/cuda-src/ray_functions.cuh:230-396.  With ANY the
// walk of an object stops at its first accepted hit and the objects after the
// first that reports one are skipped: best_prim >= 0 then says whether any
// object detects an intersection, which is all trace_shadow
// (ray_functions.cuh:365-374) delivers to its callers.
template <bool COUNT, bool ONE_SIDED, bool ANY = false>
__device__ __forceinline__ void trace_ray(const NodeEntry* __restrict__ nodes, const DevLeaf* __restrict__ leaves,
                                          const TriRec* __restrict__ tris, const ObjXform* __restrict__ objs,
                                          int32_t k0, int32_t k1, f3 o, f3 d, int32_t cap, bool secondary,
                                          float& best_t, int32_t& best_prim, int32_t& best_obj, RayCounts& rc) {
  best_t = kInf;
  best_prim = -1;
  best_obj = -1;
  for (int32_t k = k0; k < k1; ++k) {
    const ObjXform ob = objs[k];
    f3 ro, rd;
    xform_ray(ob, o, d, secondary, ro, rd);
    // the slab reciprocals once per object, not per leaf (bit-identical)
    const f3 fr{1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z};
    float t_entry, t_exit;
    if (!ray_box_inv(ob.root_min, ob.root_max, ro, fr, t_entry, t_exit)) continue;
    int32_t cur = ob.root;
    bool found = false;
    float hit_t = kInf;
    int32_t hit_prim = -1;
    int32_t visits = 0;
    uint4 w = make_uint4(0u, 0u, 0u, 0u);
    bool have_window = false;
    WalkMemo memo;
    while (t_entry < t_exit && visits < cap) {
      if (memo.repeats(cur, t_entry, t_exit)) break;
      ++visits;
      const f3 pe = add(ro, smul(t_entry, rd));
      uint32_t leaf_id;
      if (ref_is_leaf(cur)) {
        leaf_id = ref_leaf_id(cur);
      } else {
        if (!have_window) w = load_window(nodes, cur);
        leaf_id = descend<COUNT>(nodes, cur, w, pe, rc.ent);
      }
      // The box and the triangle run (32 bytes); the exit face and its rope
      // depend on the box alone, so they are found before the triangle tests
      // and only that rope is loaded; the first window of the next descent is
      // fetched while the triangles are.  Results are unchanged: the
      // reference's order (triangles, then exit, :295-347) only matters for
      // what the loop does next, which still waits for the tests.
      const DevLeaf* lp = leaves + RE_DBG_LEAF(leaf_id);
      const float4 h0 = *reinterpret_cast<const float4*>(lp->bmin);
      const float4 h1 = *reinterpret_cast<const float4*>(lp->bmax + 1);
      const float bmin[3] = {h0.x, h0.y, h0.z}, bmax[3] = {h0.w, h1.x, h1.y};
      uint32_t first = __float_as_uint(h1.z), count = __float_as_uint(h1.w);
      RE_DBG_RUN(first, count);
      if (COUNT) ++rc.leaf;
      float tn, tf;
      const bool box_ok = ray_box_inv(bmin, bmax, ro, fr, tn, tf);
      const int face = box_ok ? exit_face(bmin, bmax, add(ro, smul(tf, rd))) : -1;
      const int32_t r = face >= 0 ? lp->rope[face] : kNoRope;
      have_window = r >= 0;
      if (have_window) w = load_window(nodes, r);
      test_leaf<ONE_SIDED, COUNT, ANY>(tris + first, count, ro, rd, t_exit, found, hit_t, hit_prim, rc.tri);
      if (ANY && found) break;
      if (!box_ok) break;
      t_entry = tf;
      if (r == kNoRope) break;
      cur = r;
    }
    if (found && hit_t < best_t) {
      best_t = hit_t;
      best_prim = hit_prim;
      best_obj = k;
    }
    if (ANY && found) return;
  }
}

// ---------------------------------------------------------------------------
// Cooperative variant: the wave stages its lanes' leaf triangles together.
//
// One ray per lane, as trace_closest_kernel, but every leaf visit is done by
// the whole wave at once.  Each lane first walks down to its leaf (its own
// descent, the leaf record, the exit rope), then the triangles of all the
// lanes' leaves are laid end to end (an exclusive scan of the counts) and
// tested one triangle per lane, 64 at a time: lane e loads triangle e of
// that list - so a leaf's records are read by consecutive lanes, one
// coalesced stream instead of one scattered 40-byte fetch per lane and
// triangle - and tests it against the ray of the lane owning that leaf
// (fetched with cross-lane reads).  A hit with t below the owner's t_exit is
// folded into the owner's LDS slot with an atomic minimum of
// (t bits << 32 | record index).  For t > 0 the float bits order like the
// values, so the slot ends at the smallest t and, among equal t, the first
// record: exactly the hit the sequential in-order loop keeps (a hit is taken
// when t < t_exit and then lowers t_exit, so the last hit taken is the
// smallest t, and a later equal t is never taken;
// RayEngine/src/cuda-src/ray_functions.cuh:295-324).  Results are therefore
// bit-identical to the other variants.  The rest of the walk (descent, slab
// exit, rope, visit cap, repeated-state stop) is trace_ray's.
// ---------------------------------------------------------------------------
// Per-wave LDS of the cooperative leaf test: each visiting lane's ray, t_exit
// and record base; the owner lane of every list entry (while the list fits);
// each lane's hit slot.
constexpr int kCoopMap = 512;
// Triangle records one coherent round stages (in the ray area: 2 KB).
constexpr int kCoopStage = 48;
struct CoopWave {
  float4 ray[64][2];             // {ro, t_exit}, {rd, base = first - offset as bits}
  uint8_t owner[kCoopMap];       // list entry -> lane
  unsigned long long slot[64];   // (t bits << 32 | record) of the lane's best hit
};

// Inclusive sum over the 64 lanes of a wave (every lane active), on the
// data-parallel primitives of the vector ALU: Hillis-Steele within each row
// of 16 lanes (shifts by 1, 2, 4 and 8 lanes; a source outside the row gives
// 0), then row 0's total added into row 1 and row 2's into row 3 (broadcast
// of lane 15 of the row before), and rows 0-1's total into rows 2 and 3
// (broadcast of lane 31).  Six vector adds instead of six cross-lane reads
// through the LDS crossbar with their index arithmetic.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111 /* row_shr:1 */, 0xf, 0xf, false);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112 /* row_shr:2 */, 0xf, 0xf, false);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114 /* row_shr:4 */, 0xf, 0xf, false);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118 /* row_shr:8 */, 0xf, 0xf, false);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142 /* row_bcast:15 */, 0xa, 0xf, false);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143 /* row_bcast:31 */, 0xc, 0xf, false);
  return x;
}

template <bool ONE_SIDED>
__device__ __forceinline__ void coop_leaf_test(const TriRec* __restrict__ tris, bool at_leaf, uint32_t first,
                                               uint32_t cnt, f3 ro, f3 rd, float& t_exit, bool& found, float& hit_t,
                                               int32_t& hit_prim, CoopWave& W) {
  const int lane = threadIdx.x & 63;
  const unsigned long long at_mask = __ballot(at_leaf);
  const int lead = __ffsll((long long)at_mask) - 1;  // wave-uniform
  const uint32_t lead_first = (uint32_t)__builtin_amdgcn_readlane((int)first, lead);
  const uint32_t lc = (uint32_t)__builtin_amdgcn_readlane((int)cnt, lead);
  if (__ballot(at_leaf && (first != lead_first || cnt != lc)) == 0ull) {
    // Every visiting lane tests the same run of records (the same leaf; an
    // empty leaf may share its first index with another leaf, so the count is
    // compared too): the run is staged once in LDS (one coalesced load, up to
    // kCoopStage records a round) and each lane tests it in record order
    // against its own ray - the sequential rule itself, read from LDS instead
    // of one scattered fetch per lane.
    TriRec* const stage = reinterpret_cast<TriRec*>(&W.ray[0][0]);
#pragma unroll 1
    for (uint32_t base = 0; base < lc; base += kCoopStage) {
      const uint32_t m = lc - base < (uint32_t)kCoopStage ? lc - base : (uint32_t)kCoopStage;
      if ((uint32_t)lane < m) stage[lane] = tris[lead_first + base + (uint32_t)lane];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (at_leaf) {
        for (uint32_t j = 0; j < m; ++j) {
          const float* f = stage[j].v0;
          float t;
          const f3 v0{f[0], f[1], f[2]}, e1{f[3], f[4], f[5]}, e2{f[6], f[7], f[8]};
          const bool h = ONE_SIDED ? ray_tri(ro, rd, v0, e1, e2, t) : ray_tri_2s(ro, rd, v0, e1, e2, t);
          if (h && t < t_exit) {
            found = true;
            hit_t = t;
            hit_prim = stage[j].prim;
            t_exit = t;
          }
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    return;
  }
  const uint32_t c = at_leaf ? cnt : 0u;
  const uint32_t incl = wave_incl_scan(c);
  const uint32_t excl = incl - c;
  // in a scalar register, so the loops below exit on scalar branches taken by
  // the whole wave (their bodies hold wave barriers)
  const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
  W.slot[lane] = ~0ull;
  if (at_leaf) {
    W.ray[lane][0] = make_float4(ro.x, ro.y, ro.z, t_exit);
    W.ray[lane][1] = make_float4(rd.x, rd.y, rd.z, __uint_as_float(first - excl));
  }
  // The list is taken kCoopMap entries at a time: the owners write their
  // lane into the entries of the window, then the wave tests the window's
  // entries 64 at a time.
#pragma unroll 1
  for (uint32_t wb = 0; wb < total; wb += kCoopMap) {
    const uint32_t we = total - wb < (uint32_t)kCoopMap ? total : wb + (uint32_t)kCoopMap;
    if (at_leaf) {
      const uint32_t k0 = excl > wb ? excl : wb, k1 = excl + c < we ? excl + c : we;
      for (uint32_t k = k0; k < k1; ++k) W.owner[k - wb] = (uint8_t)lane;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll 1
    for (uint32_t base = wb; base < we; base += 64) {
      const uint32_t e = base + (uint32_t)lane;
      if (e < we) {
        const int L = W.owner[e - wb];
        const float4 a = W.ray[L][0], b = W.ray[L][1];
        const uint32_t rec = __float_as_uint(b.w) + e;
        const TriGeom g = load_geom(tris, rec);
        const f3 oro{a.x, a.y, a.z}, ord{b.x, b.y, b.z};
        float t;
        const bool h =
            ONE_SIDED ? ray_tri(oro, ord, g.v0, g.e1, g.e2, t) : ray_tri_2s(oro, ord, g.v0, g.e1, g.e2, t);
        if (h && t < a.w)
          atomicMin(&W.slot[L], ((unsigned long long)__float_as_uint(t) << 32) | (unsigned long long)rec);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  if (at_leaf) {
    const unsigned long long sv = W.slot[lane];
    if (sv != ~0ull) {
      found = true;
      hit_t = __uint_as_float((uint32_t)(sv >> 32));
      hit_prim = tris[(uint32_t)sv].prim;
      t_exit = hit_t;
    }
  }
}

// Walk-start hint (no counterpart in the reference, whose walks all start at
// the root, ray_functions.cuh:273-291).  The first visit of a walk of object k
// descends from the root to the leaf holding pe, the point at the walk's
// t_entry, going below where pe < split and above otherwise.  A leaf's cell
// box is built from its ancestors' splits (the below child's upper bound and
// the above child's lower bound are the split itself, and every split lies in
// its node's cell), so its lower bound on an axis is the largest split the
// path went above at and its upper bound the smallest split it went below at.
// Hence when bmin <= pe < bmax holds on every axis, every comparison on the
// path from the root to that leaf goes the way the path does, and the descent
// ends there: the walk may start at the leaf, and its visits (and so its
// result) are those of the walk from the root.  Otherwise, for a NaN pe, a
// leaf of another object (the object index is kept in the device record's
// pad[0]) or an id outside [0, n_leaves), the walk starts at the root.
__device__ __forceinline__ int32_t hinted_start(const DevLeaf* __restrict__ leaves, int64_t n_leaves, int32_t hint,
                                                int32_t k, f3 pe, int32_t root) {
  if (hint < 0 || (int64_t)hint >= n_leaves) return root;
  const DevLeaf* lp = leaves + hint;
  const float4 h0 = *reinterpret_cast<const float4*>(lp->bmin);
  const float2 h1 = *reinterpret_cast<const float2*>(lp->bmax + 1);
  const bool inside = lp->pad[0] == (uint32_t)k && pe.x >= h0.x && pe.y >= h0.y && pe.z >= h0.z && pe.x < h0.w &&
                      pe.y < h1.x && pe.z < h1.y;
  return inside ? leaf_ref((uint32_t)hint) : root;
}

// The closest hit of every lane's ray over objects [k0, k1) (k0 + 1 with
// ONE_OBJ), the wave walking together: each lane makes its own descent, leaf
// record read, exit and rope step (trace_ray's walk, same visit cap and
// repeated-state stop), and every leaf visit's triangle tests are done by
// coop_leaf_test for all visiting lanes at once.  Lanes with has_ray false take
// part in the wave's work without a ray of their own.  Results are
// bit-identical to trace_ray's.  Every lane of the wave must call it (it
// holds wave-wide operations); the loop ends on a wave-uniform branch.
template <bool COUNT, bool ONE_SIDED, bool ONE_OBJ, int HINT = 0>
__device__ __forceinline__ void coop_trace_wave(const NodeEntry* __restrict__ nodes,
                                                const DevLeaf* __restrict__ leaves, const TriRec* __restrict__ tris,
                                                const ObjXform* __restrict__ objs, int32_t k0, int32_t k1,
                                                bool has_ray, f3 o, f3 d, int32_t cap, bool secondary, float& best_t,
                                                int32_t& best_prim, int32_t& best_obj, RayCounts& rc, CoopWave& W,
                                                int32_t hint = -1, int64_t n_leaves = 0,
                                                int32_t* best_leaf = nullptr) {
  const int32_t k_end = ONE_OBJ ? k0 + 1 : k1;
  best_t = kInf;
  best_prim = -1;
  best_obj = -1;
  int32_t last_leaf = -1;  // HINT == 2: the walk's last visited leaf
  if (HINT == 2) *best_leaf = -1;
  int32_t k_obj = k0 - 1;
  bool live = false, open = false, done = !has_ray;
  f3 ro{0, 0, 0}, rd{0, 0, 0}, fr{0, 0, 0};
  float t_entry = 0.0f, t_exit = 0.0f, hit_t = kInf;
  int32_t hit_prim = -1, cur = 0, visits = 0;
  bool found = false, have_window = false;
  uint4 w = make_uint4(0u, 0u, 0u, 0u);
  WalkMemo memo;
  // starts the next object whose root box the ray enters (trace_ray's object
  // loop); a walk that cannot take a first visit closes at once with no hit
  auto start_next = [&]() {
    while (!live && k_obj + 1 < k_end) {
      ++k_obj;
      const ObjXform ob = objs[k_obj];
      xform_ray(ob, o, d, secondary, ro, rd);
      fr = f3{1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z};
      if (!ray_box_inv(ob.root_min, ob.root_max, ro, fr, t_entry, t_exit)) continue;
      cur = ob.root;
      if (HINT != 0) cur = hinted_start(leaves, n_leaves, hint, k_obj, add(ro, smul(t_entry, rd)), cur);
      found = false;
      hit_t = kInf;
      hit_prim = -1;
      visits = 0;
      have_window = false;
      memo = WalkMemo{};
      live = t_entry < t_exit && visits < cap;
      open = live;
    }
    if (!live) done = true;
  };
  if (!done) start_next();
#pragma unroll 1
  for (;;) {
    // a lane between walks closes the last one and, with more objects,
    // starts the next
    if (!live && !done) {
      if (open && found && hit_t < best_t) {
        best_t = hit_t;
        best_prim = hit_prim;
        best_obj = k_obj;
        if (HINT == 2) *best_leaf = last_leaf;
      }
      open = false;
      if (ONE_OBJ) done = true;
      else start_next();
    }
    if (__ballot(!done) == 0ull) break;
    bool at_leaf = false, box_ok = false;
    uint32_t first = 0, cnt = 0;
    float tf = 0.0f;
    int32_t r = kNoRope;
    if (live) {
      if (memo.repeats(cur, t_entry, t_exit)) {
        live = false;
      } else {
        ++visits;
        const f3 pe = add(ro, smul(t_entry, rd));
        uint32_t leaf_id;
        if (ref_is_leaf(cur)) {
          leaf_id = ref_leaf_id(cur);
        } else {
          if (!have_window) w = load_window(nodes, cur);
          leaf_id = descend<COUNT>(nodes, cur, w, pe, rc.ent);
        }
        if (HINT == 2) last_leaf = (int32_t)leaf_id;
        const DevLeaf* lp = leaves + RE_DBG_LEAF(leaf_id);
        const float4 h0 = *reinterpret_cast<const float4*>(lp->bmin);
        const float4 h1 = *reinterpret_cast<const float4*>(lp->bmax + 1);
        const float bmin[3] = {h0.x, h0.y, h0.z}, bmax[3] = {h0.w, h1.x, h1.y};
        first = __float_as_uint(h1.z);
        cnt = __float_as_uint(h1.w);
        RE_DBG_RUN(first, cnt);
        if (COUNT) ++rc.leaf;
        float tn;
        box_ok = ray_box_inv(bmin, bmax, ro, fr, tn, tf);
        const int face = box_ok ? exit_face(bmin, bmax, add(ro, smul(tf, rd))) : -1;
        r = face >= 0 ? lp->rope[face] : kNoRope;
        have_window = r >= 0;
        if (have_window) w = load_window(nodes, r);
        at_leaf = true;
      }
    }
    if (__ballot(at_leaf) != 0ull) {
      coop_leaf_test<ONE_SIDED>(tris, at_leaf, first, cnt, ro, rd, t_exit, found, hit_t, hit_prim, W);
      if (COUNT && at_leaf) rc.tri += cnt;
    }
    if (at_leaf) {
      if (!box_ok) {
        live = false;
      } else {
        t_entry = tf;
        if (r == kNoRope) {
          live = false;
        } else {
          cur = r;
          live = t_entry < t_exit && visits < cap;
        }
      }
    }
  }
}

// Reference sphere-trace march (single_ray_sphere_trace,
// RayEngine/src/cuda-src/sphere_tracing.cuh:182-216) clipped to the volume
// box: sample, step, then test d <= eps * t on the stepped t.
__device__ __forceinline__ bool plain_march(const VolView& V, const float* bmin, const float* bmax, f3 o, f3 d,
                                            int32_t max_steps, float eps, float step_scale, int32_t step_rule,
                                            float& t_out, int32_t& steps) {
  float t_near, t_far;
  steps = 0;
  t_out = kInf;
  if (!ray_box(bmin, bmax, o, d, t_near, t_far)) return false;
  float t = t_near;
  while (t < t_far && steps < max_steps) {
    const f3 p = add(o, smul(t, d));
    const float dist = trilinear(V, p);
    const float k = step_rule == 0 ? step_scale : fmaxf(0.45f, fminf(t * 0.6f, 0.8f));
    t += dist * k;
    ++steps;
    if (dist <= eps * t) {
      t_out = t;
      return true;
    }
  }
  return false;
}

// Exact-field march: from the volume entry, step max(f * k_exact, h_min)
// while the trilinear field stays positive; bisect the first sign change
// n_bisect times.  A ray entering at f <= 0 hits at the entry.
__device__ __forceinline__ bool exact_march(const VolView& V, const float* bmin, const float* bmax, f3 o, f3 d,
                                            float k_exact, float h_min, int32_t n_bisect, int32_t max_iters,
                                            float& t_out, int32_t& samples) {
  float tn, tf;
  samples = 0;
  t_out = kInf;
  if (!ray_box(bmin, bmax, o, d, tn, tf)) return false;
  float t = tn;
  float f_prev = trilinear(V, add(o, smul(t, d)));
  ++samples;
  if (f_prev <= 0.0f) {
    t_out = t;
    return true;
  }
  for (int it = 0; it < max_iters; ++it) {
    const float h = fmaxf(f_prev * k_exact, h_min);
    const float t_next = t + h;
    const float f_next = trilinear(V, add(o, smul(t_next, d)));
    ++samples;
    if (f_next <= 0.0f) {
      float a = t, b = t_next;
      for (int k = 0; k < n_bisect; ++k) {
        const float mid = 0.5f * (a + b);
        const float fm = trilinear(V, add(o, smul(mid, d)));
        ++samples;
        if (fm > 0.0f) a = mid; else b = mid;
      }
      t_out = 0.5f * (a + b);
      return true;
    }
    t = t_next;
    f_prev = f_next;
    if (t >= tf) return false;
  }
  return false;
}

}  // namespace re
