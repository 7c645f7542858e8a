// Rope kd-tree closest-hit traversal for gfx950.
//
// Semantics: stackless_kdtree_traversal and trace_scene,
// RayEngine/src/cuda-src/ray_functions.cuh:230-396 - per object transform,
// root slab test, then while (t_entry < t_exit && visits < cap): descend from
// the current node to a leaf at p_entry, test the leaf's triangles in order
// (a hit with t < t_exit shrinks t_exit), leave the leaf through its slab
// exit and follow the rope of the first matching face.  Objects keep the
// strictly closest hit.
//
// One ray per lane; the node/leaf/triangle records are laid out for single
// wide loads (8-byte entries, 64-byte leaves, 40-byte triangles in leaf
// order), so each step of the pointer chase is one or a few cache lines.
#define RE_DEBUG_BOUNDS_HERE 1
#include <cstdlib>
#include "../device_paths.hpp"

namespace re {
namespace {

// One ray per lane.  REVERSE hands the last block of rays to the first
// dispatched work-group: when costly rays sit at the end of the array (the
// secondary rays a renderer appends after its primaries) they start first and
// the cheap ones fill in behind them instead of forming the kernel's tail.
// ONE_OBJ: the scene holds one object, so the object loop runs once and the
// untransformed ray and the running best hit need no registers of their own
// (the scenes the benchmarks trace; multi-object scenes take ONE_OBJ false).
template <bool COUNT, bool ONE_SIDED, bool REVERSE, bool ONE_OBJ = false>
__device__ __forceinline__ void closest_body(const NodeEntry* __restrict__ nodes, const DevLeaf* __restrict__ leaves,
                                             const TriRec* __restrict__ tris, const ObjXform* __restrict__ objs,
                                             int32_t n_obj, const float* __restrict__ ray_o,
                                             const float* __restrict__ ray_d, int64_t n, int32_t cap,
                                             int32_t secondary, int32_t* __restrict__ prim_out,
                                             float* __restrict__ t_out, int32_t* __restrict__ obj_out,
                                             TraceCounts* __restrict__ counts) {
  const int64_t blk = REVERSE ? (int64_t)(gridDim.x - 1 - blockIdx.x) : (int64_t)blockIdx.x;
  const int64_t i = blk * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const f3 o{ray_o[3 * i], ray_o[3 * i + 1], ray_o[3 * i + 2]};
  const f3 d{ray_d[3 * i], ray_d[3 * i + 1], ray_d[3 * i + 2]};
  float best_t;
  int32_t best_prim, best_obj;
  RayCounts rc;
  trace_ray<COUNT, ONE_SIDED>(nodes, leaves, tris, objs, 0, ONE_OBJ ? 1 : n_obj, o, d, cap, secondary != 0, best_t,
                              best_prim, best_obj, rc);
  prim_out[i] = best_prim;
  t_out[i] = best_t;
  if (obj_out) obj_out[i] = best_obj;
  if (COUNT) {
    atomicAdd(&counts->entries, (unsigned long long)rc.ent);
    atomicAdd(&counts->leaves, (unsigned long long)rc.leaf);
    atomicAdd(&counts->tris, (unsigned long long)rc.tri);
    if (best_prim >= 0) atomicAdd(&counts->rays_hit, 1ull);
  }
}

#define RE_CLOSEST_PARAMS                                                                                  \
  const NodeEntry *__restrict__ nodes, const DevLeaf *__restrict__ leaves, const TriRec *__restrict__ tris, \
      const ObjXform *__restrict__ objs, int32_t n_obj, const float *__restrict__ ray_o,                    \
      const float *__restrict__ ray_d, int64_t n, int32_t cap, int32_t secondary,                          \
      int32_t *__restrict__ prim_out, float *__restrict__ t_out, int32_t *__restrict__ obj_out,            \
      TraceCounts *__restrict__ counts
#define RE_CLOSEST_ARGS \
  nodes, leaves, tris, objs, n_obj, ray_o, ray_d, n, cap, secondary, prim_out, t_out, obj_out, counts

template <bool COUNT, bool ONE_SIDED, bool REVERSE>
__global__ void __launch_bounds__(256) trace_closest_kernel(RE_CLOSEST_PARAMS) {
  closest_body<COUNT, ONE_SIDED, REVERSE>(RE_CLOSEST_ARGS);
}

// The same with the register budget held to 80 VGPRs (6 waves per SIMD).
template <bool COUNT, bool ONE_SIDED, bool ONE_OBJ>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6)))
trace_closest_o6_kernel(RE_CLOSEST_PARAMS) {
  closest_body<COUNT, ONE_SIDED, true, ONE_OBJ>(RE_CLOSEST_ARGS);
}

// Any-hit query (trace_shadow, RayEngine/src/cuda-src/ray_functions.cuh:365-374):
// one ray per lane, last block first, 6-wave budget; each ray stops at its
// first accepted hit.
template <bool ONE_SIDED>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6)))
trace_occluded_kernel(const NodeEntry* __restrict__ nodes, const DevLeaf* __restrict__ leaves,
                      const TriRec* __restrict__ tris, const ObjXform* __restrict__ objs, int32_t n_obj,
                      const float* __restrict__ ray_o, const float* __restrict__ ray_d, int64_t n, int32_t cap,
                      int32_t secondary, uint8_t* __restrict__ occluded) {
  const int64_t n_blk = (n + blockDim.x - 1) / blockDim.x;
  const int64_t i = (n_blk - 1 - (int64_t)blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const f3 o{ray_o[3 * i], ray_o[3 * i + 1], ray_o[3 * i + 2]};
  const f3 d{ray_d[3 * i], ray_d[3 * i + 1], ray_d[3 * i + 2]};
  float best_t;
  int32_t best_prim, best_obj;
  RayCounts rc;
  trace_ray<false, ONE_SIDED, true>(nodes, leaves, tris, objs, 0, n_obj, o, d, cap, secondary != 0, best_t,
                                    best_prim, best_obj, rc);
  occluded[i] = best_prim >= 0 ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Persistent variant.  A wave keeps 64 rays in flight and refills each lane
// as soon as its ray finishes, from a wave-private chunk of CHUNK
// consecutive ray indices taken from a global counter (one atomic per
// chunk).  One loop iteration advances every live lane by one leaf visit, so
// a ray that needs many visits (the cap, or a crowded leaf) no longer holds
// 63 idle lanes.  Per-ray arithmetic is identical to trace_closest_kernel.
// ---------------------------------------------------------------------------
constexpr int kChunk = 64;

template <bool COUNT, bool ONE_SIDED>
__global__ void __launch_bounds__(256) trace_persistent_kernel(
    const NodeEntry* __restrict__ nodes, const DevLeaf* __restrict__ leaves,
    const TriRec* __restrict__ tris, const ObjXform* __restrict__ objs, int32_t n_obj,
    const float* __restrict__ ray_o, const float* __restrict__ ray_d, int64_t n, int32_t cap,
    int32_t secondary, int32_t* __restrict__ prim_out, float* __restrict__ t_out,
    int32_t* __restrict__ obj_out, TraceCounts* __restrict__ counts,
    unsigned long long* __restrict__ queue) {
  const int lane = (int)(threadIdx.x & 63);
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int64_t pool_next = 0, pool_end = 0;  // wave-uniform
  bool drained = false;                 // wave-uniform

  int64_t ray = -1;
  f3 o{0, 0, 0}, d{0, 0, 0}, ro{0, 0, 0}, rd{0, 0, 0}, fr{0, 0, 0};
  float t_entry = 0.0f, t_exit = 0.0f, hit_t = kInf, best_t = kInf;
  int32_t cur = 0, visits = 0, k_obj = 0, hit_prim = -1, best_prim = -1, best_obj = -1;
  bool live = false, found = false;
  WalkMemo memo;
  unsigned c_ent = 0, c_leaf = 0, c_tri = 0, c_hit = 0;

  for (;;) {
    // ---- refill idle lanes (wave-uniform control flow) ----
    const unsigned long long need = __ballot(ray < 0);
    if (need != 0ull && !drained) {
      if (pool_next >= pool_end) {
        // every lane takes part in the atomic (lane 0 adds the chunk, the
        // others 0) and the result is read back with readfirstlane, so the
        // chunk start is a scalar and no lane-dependent branch precedes the
        // loop's exit (a lane-0-only atomic followed by a shuffle can be
        // structured into an inner loop that lanes 1-63 never leave)
        const unsigned long long v = atomicAdd(queue, lane == 0 ? (unsigned long long)kChunk : 0ull);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
        const unsigned long long b = ((unsigned long long)hi << 32) | lo;
        pool_next = (int64_t)b;
        pool_end = (int64_t)b + kChunk < n ? (int64_t)b + kChunk : n;
        if ((int64_t)b >= n) drained = true;
      }
      if (!drained) {
        const int64_t avail = pool_end - pool_next;
        const int rank = __popcll(need & below);
        if (ray < 0 && rank < avail) {
          ray = pool_next + rank;
          o = f3{ray_o[3 * ray], ray_o[3 * ray + 1], ray_o[3 * ray + 2]};
          d = f3{ray_d[3 * ray], ray_d[3 * ray + 1], ray_d[3 * ray + 2]};
          best_t = kInf;
          best_prim = -1;
          best_obj = -1;
          k_obj = -1;
          live = false;
        }
        const int64_t taken = (int64_t)__popcll(need) < avail ? (int64_t)__popcll(need) : avail;
        pool_next += taken;
      }
    }
    if (!__any(ray >= 0)) {
      if (drained) break;
      continue;
    }
    if (ray < 0) continue;

    // ---- start the next object whose root box the ray hits ----
    while (!live && k_obj + 1 < n_obj) {
      ++k_obj;
      const ObjXform ob = objs[k_obj];
      xform_ray(ob, o, d, secondary != 0, ro, rd);
      fr = f3{1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z};
      if (!ray_box_inv(ob.root_min, ob.root_max, ro, fr, t_entry, t_exit)) continue;
      cur = ob.root;
      visits = 0;
      found = false;
      hit_t = kInf;
      hit_prim = -1;
      memo = WalkMemo{};
      live = (t_entry < t_exit && visits < cap);
    }

    // ---- one leaf visit ----
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
          leaf_id = descend<COUNT>(nodes, cur, load_window(nodes, cur), pe, c_ent);
        }
        const DevLeaf* lp = leaves + RE_DBG_LEAF(leaf_id);
        const DevLeaf L = *lp;
        if (COUNT) ++c_leaf;
        uint32_t lf = L.first, lc = L.count;
        RE_DBG_RUN(lf, lc);
        test_leaf<ONE_SIDED, COUNT>(tris + lf, lc, ro, rd, t_exit, found, hit_t, hit_prim, c_tri);
        float tn, tf;
        if (!ray_box_inv(L.bmin, L.bmax, ro, fr, tn, tf)) {
          live = false;
        } else {
          t_entry = tf;
          const int face = exit_face(L.bmin, L.bmax, add(ro, smul(t_entry, rd)));
          const int32_t r = face >= 0 ? lp->rope[face] : kNoRope;
          if (r == kNoRope) {
            live = false;
          } else {
            cur = r;
            live = (t_entry < t_exit && visits < cap);
          }
        }
      }
      if (!live && found && hit_t < best_t) {
        best_t = hit_t;
        best_prim = hit_prim;
        best_obj = k_obj;
      }
    }

    // ---- retire the ray once every object is done ----
    if (!live && k_obj + 1 >= n_obj) {
      prim_out[ray] = best_prim;
      t_out[ray] = best_t;
      if (obj_out) obj_out[ray] = best_obj;
      if (COUNT && best_prim >= 0) ++c_hit;
      ray = -1;
    }
  }
  if (COUNT) {
    atomicAdd(&counts->entries, (unsigned long long)c_ent);
    atomicAdd(&counts->leaves, (unsigned long long)c_leaf);
    atomicAdd(&counts->tris, (unsigned long long)c_tri);
    atomicAdd(&counts->rays_hit, (unsigned long long)c_hit);
  }
}

// Cooperative variant (see coop_trace_wave): one ray per lane, last block of
// rays first, 6-wave register budget.
// HINT (as in the refill kernel below): 1 takes each ray's walk-start leaf
// from start_leaf (may be null), 2 also writes each ray's hit leaf.
template <bool COUNT, bool ONE_SIDED, bool ONE_OBJ, int HINT = 0>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6)))
trace_coop_kernel(RE_CLOSEST_PARAMS, const int32_t* __restrict__ start_leaf, int32_t* __restrict__ hit_leaf,
                  int64_t n_leaves) {
  // one CoopWave per wave of the block (dynamic LDS sized by the launch)
  extern __shared__ CoopWave waves[];
  CoopWave& wv = waves[threadIdx.x >> 6];
  const int64_t blk = (int64_t)(gridDim.x - 1 - blockIdx.x);  // last block of rays first
  const int64_t i = blk * blockDim.x + threadIdx.x;
  const bool has_ray = i < n;
  const int64_t ic = has_ray ? i : 0;
  const f3 o{ray_o[3 * ic], ray_o[3 * ic + 1], ray_o[3 * ic + 2]};
  const f3 d{ray_d[3 * ic], ray_d[3 * ic + 1], ray_d[3 * ic + 2]};
  float best_t;
  int32_t best_prim, best_obj;
  RayCounts rc;
  const int32_t hint = (HINT != 0 && start_leaf && has_ray) ? start_leaf[i] : -1;
  int32_t best_leaf = -1;
  coop_trace_wave<COUNT, ONE_SIDED, ONE_OBJ, HINT>(nodes, leaves, tris, objs, 0, ONE_OBJ ? 1 : n_obj, has_ray, o, d,
                                                   cap, secondary != 0, best_t, best_prim, best_obj, rc, wv, hint,
                                                   n_leaves, &best_leaf);
  if (has_ray) {
    prim_out[i] = best_prim;
    t_out[i] = best_t;
    if (obj_out) obj_out[i] = best_obj;
    if (HINT == 2) hit_leaf[i] = best_leaf;
    if (COUNT) {
      atomicAdd(&counts->entries, (unsigned long long)rc.ent);
      atomicAdd(&counts->leaves, (unsigned long long)rc.leaf);
      atomicAdd(&counts->tris, (unsigned long long)rc.tri);
      if (best_prim >= 0) atomicAdd(&counts->rays_hit, 1ull);
    }
  }
}

// Cooperative variant with refill (variant 5): persistent waves; once at
// least half of a wave's lanes have finished their rays, those lanes take the
// next rays at the start of the following visit round, from a wave-private
// chunk of 64 consecutive rays taken from a counter of its XCD's part of the
// ray array, so a wave is not
// held by its longest walk with most lanes idle (on incoherent bounce rays
// the longest walk of a wave is about four times the mean).  Waiting for half
// a wave keeps the rays taken together consecutive, so they stay as coherent
// as their neighbours in the array are.  The leaf work is coop_leaf_test's,
// shared by every lane at a leaf whatever ray it holds; the rays are taken
// from the end of the array, so appended bounce rays start first.  Per-ray
// results equal the other variants' (each ray's walk is the same sequence of
// visits).
constexpr int kRefillLanes = 32;  // refill threshold: half a wave

// HINT: 0 none; 1 each ray's walk may start at its start_leaf entry (may be
// null; see hinted_start); 2 as 1, and each ray's hit leaf is written to
// hit_leaf: the last leaf visited by the walk that found the closest hit, -1
// for a miss.  A hit is taken at any t below the walk's t_exit, possibly
// beyond the cell of the leaf whose triangle gave it; the walk then goes on
// until its t_entry passes the hit, so its last leaf is the one whose cell
// holds the hit point (when the walk ends there, not at the visit cap or a
// missing rope), the leaf a ray leaving that point starts in.
template <bool COUNT, bool ONE_SIDED, bool ONE_OBJ, int HINT = 0>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6)))
trace_coop_refill_kernel(RE_CLOSEST_PARAMS, unsigned long long* __restrict__ queue,
                         const int32_t* __restrict__ start_leaf, int32_t* __restrict__ hit_leaf, int64_t n_leaves) {
  extern __shared__ CoopWave waves[];
  CoopWave& W = waves[threadIdx.x >> 6];
  const int lane = (int)(threadIdx.x & 63);
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int32_t k_end = ONE_OBJ ? 1 : n_obj;
  int64_t pool_next = 0, pool_end = 0;  // wave-uniform
  bool drained = false;                 // wave-uniform
  // The rays are split into kQueueSlots equal parts, each with its own
  // counter; a wave takes chunks from the part of the XCD it runs on (its L2
  // then holds the part's tree lines) and, once that part is taken, from the
  // others in turn.
  const int xcd = (int)(__builtin_amdgcn_s_getreg((20 /* HW_REG_XCC_ID */) | (0 << 6) | ((4 - 1) << 11)) &
                        (unsigned)(kQueueSlots - 1));
  int tried = 0;  // wave-uniform: parts found taken
  const int64_t part = ((n + kQueueSlots - 1) / kQueueSlots + 63) / 64 * 64;
  int64_t ray = -1;
  f3 o{0, 0, 0}, d{0, 0, 0};  // kept only with several objects (the next object's transform)
  f3 ro{0, 0, 0}, rd{0, 0, 0}, fr{0, 0, 0};
  float t_entry = 0.0f, t_exit = 0.0f, hit_t = kInf, best_t = kInf;
  int32_t k_obj = -1, cur = 0, visits = 0, hit_prim = -1, best_prim = -1, best_obj = -1;
  int32_t hint = -1, walk_leaf = -1, best_leaf = -1;  // HINT only
  bool live = false, open = false, found = false, have_window = false;
  uint4 w = make_uint4(0u, 0u, 0u, 0u);
  WalkMemo memo;
  RayCounts rc;
  unsigned c_hit = 0;
  // the walk of object k from the ray (o, d): its first visit, or none
  auto start_walk = [&](const ObjXform& ob, int32_t k, f3 oo, f3 dd) {
    xform_ray(ob, oo, dd, secondary != 0, ro, rd);
    fr = f3{1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z};
    if (!ray_box_inv(ob.root_min, ob.root_max, ro, fr, t_entry, t_exit)) return;
    cur = ob.root;
    if (HINT != 0) cur = hinted_start(leaves, n_leaves, hint, k, add(ro, smul(t_entry, rd)), cur);
    found = false;
    hit_t = kInf;
    hit_prim = -1;
    visits = 0;
    have_window = false;
    memo = WalkMemo{};
    live = t_entry < t_exit && visits < cap;
    open = live;
  };
  // the next object whose root box the ray enters (trace_ray's object loop)
  auto start_next = [&]() {
    while (!live && k_obj + 1 < k_end) {
      ++k_obj;
      start_walk(objs[k_obj], k_obj, o, d);
    }
  };
  auto finish = [&]() {
    prim_out[ray] = best_prim;
    t_out[ray] = best_t;
    if (obj_out) obj_out[ray] = best_obj;
    if (HINT == 2) hit_leaf[ray] = best_leaf;
    if (COUNT && best_prim >= 0) ++c_hit;
    ray = -1;
  };
#pragma unroll 1
  for (;;) {
    // close a finished walk: the next object, or the ray's result
    if (ray >= 0 && !live) {
      if (open && found && hit_t < best_t) {
        best_t = hit_t;
        best_prim = hit_prim;
        best_obj = ONE_OBJ ? 0 : k_obj;
        if (HINT == 2) best_leaf = walk_leaf;
      }
      open = false;
      if (!ONE_OBJ) start_next();
      if (!live) finish();
    }
    // refill (wave-uniform), once half the lanes or more have no ray: they
    // take the next rays of the wave's chunk; the chunk start is taken with
    // every lane in the atomic and read back with readfirstlane, so it is a
    // scalar.  While fewer lanes are idle, the others hold live walks, so the
    // wave keeps making progress, and a wave with no ray always refills.
    const unsigned long long need = __ballot(ray < 0);
    if (__popcll(need) >= kRefillLanes && !drained) {
      if (pool_next >= pool_end) {
        // at most one chunk request per round: a part found taken moves the
        // wave on to the next part at the next round (every wave tries all
        // parts before it stops, so every chunk is taken once)
        const int x = (xcd + tried) & (kQueueSlots - 1);
        const int64_t lo_x = part * x < n ? part * x : n;
        const int64_t hi_x = part * (x + 1) < n ? part * (x + 1) : n;
        const unsigned long long v = atomicAdd(queue + kQueueStride * x, lane == 0 ? 64ull : 0ull);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
        const int64_t b = lo_x + (int64_t)(((unsigned long long)hi << 32) | lo);
        if (b < hi_x) {
          pool_next = b;
          pool_end = b + 64 < hi_x ? b + 64 : hi_x;
        } else if (++tried == kQueueSlots) {
          drained = true;
        }
      }
      if (!drained && pool_next < pool_end) {
        const int64_t avail = pool_end - pool_next;
        const int rank = __popcll(need & below);
        if (ray < 0 && rank < avail) {
          ray = n - 1 - (pool_next + rank);
          const f3 no{ray_o[3 * ray], ray_o[3 * ray + 1], ray_o[3 * ray + 2]};
          const f3 nd{ray_d[3 * ray], ray_d[3 * ray + 1], ray_d[3 * ray + 2]};
          best_t = kInf;
          best_prim = -1;
          best_obj = -1;
          if (HINT == 2) best_leaf = -1;
          if (HINT != 0) hint = start_leaf ? start_leaf[ray] : -1;
          live = false;
          open = false;
          if (ONE_OBJ) {
            start_walk(objs[0], 0, no, nd);
          } else {
            o = no;
            d = nd;
            k_obj = -1;
            start_next();
          }
          if (!live) finish();  // enters no object: a miss
        }
        const int64_t taken = (int64_t)__popcll(need) < avail ? (int64_t)__popcll(need) : avail;
        pool_next += taken;
      }
    }
    if (__ballot(ray >= 0) == 0ull) {
      if (drained) break;
      continue;
    }
    // one visit round (as coop_trace_wave)
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
        if (HINT == 2) walk_leaf = (int32_t)leaf_id;
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
  if (COUNT) {
    atomicAdd(&counts->entries, (unsigned long long)rc.ent);
    atomicAdd(&counts->leaves, (unsigned long long)rc.leaf);
    atomicAdd(&counts->tris, (unsigned long long)rc.tri);
    atomicAdd(&counts->rays_hit, (unsigned long long)c_hit);
  }
}

// Closest hit with each ray's status: trace_ray's walk (same visits, same
// cap and repeated-state stop, so the same prim and t) with the way the walk
// of the last object ended recorded as the oracle numbers it, and a flag for
// non-finite input.  A debugging aid: one ray per thread, no specialisation.
template <bool ONE_SIDED>
__global__ void __launch_bounds__(256) trace_status_kernel(const NodeEntry* __restrict__ nodes,
                                                           const DevLeaf* __restrict__ leaves,
                                                           const TriRec* __restrict__ tris,
                                                           const ObjXform* __restrict__ objs, int32_t n_obj,
                                                           const float* __restrict__ ray_o,
                                                           const float* __restrict__ ray_d, int64_t n, int32_t cap,
                                                           int32_t secondary, int32_t* __restrict__ prim_out,
                                                           float* __restrict__ t_out, uint8_t* __restrict__ status) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const f3 o{ray_o[3 * i], ray_o[3 * i + 1], ray_o[3 * i + 2]};
  const f3 d{ray_d[3 * i], ray_d[3 * i + 1], ray_d[3 * i + 2]};
  const bool finite = isfinite(o.x) && isfinite(o.y) && isfinite(o.z) && isfinite(d.x) && isfinite(d.y) &&
                      isfinite(d.z) && (d.x != 0.0f || d.y != 0.0f || d.z != 0.0f);
  float best_t = kInf;
  int32_t best_prim = -1;
  uint8_t reason = 5;
  unsigned c_tri = 0;
  for (int32_t k = 0; k < n_obj; ++k) {
    const ObjXform ob = objs[k];
    f3 ro, rd;
    xform_ray(ob, o, d, secondary != 0, ro, rd);
    const f3 fr{1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z};
    float t_entry, t_exit;
    if (!ray_box_inv(ob.root_min, ob.root_max, ro, fr, t_entry, t_exit)) {
      reason = 5;
      continue;
    }
    int32_t cur = ob.root, hit_prim = -1, visits = 0;
    bool found = false;
    float hit_t = kInf;
    WalkMemo memo;
    uint8_t why = 255;
    while (t_entry < t_exit && visits < cap) {
      if (memo.repeats(cur, t_entry, t_exit)) {
        why = 6;
        break;
      }
      ++visits;
      const f3 pe = add(ro, smul(t_entry, rd));
      unsigned c_ent = 0;
      const uint32_t leaf_id =
          ref_is_leaf(cur) ? ref_leaf_id(cur) : descend<false>(nodes, cur, load_window(nodes, cur), pe, c_ent);
      const DevLeaf L = leaves[RE_DBG_LEAF(leaf_id)];
      uint32_t lf = L.first, lc = L.count;
      RE_DBG_RUN(lf, lc);
      test_leaf<ONE_SIDED, false>(tris + lf, lc, ro, rd, t_exit, found, hit_t, hit_prim, c_tri);
      float tn, tf;
      if (!ray_box_inv(L.bmin, L.bmax, ro, fr, tn, tf)) {
        why = 2;
        break;
      }
      t_entry = tf;
      const int face = exit_face(L.bmin, L.bmax, add(ro, smul(t_entry, rd)));
      const int32_t r = face >= 0 ? L.rope[face] : kNoRope;
      if (r == kNoRope) {
        why = face < 0 ? 3 : 4;
        break;
      }
      cur = r;
    }
    reason = why != 255 ? why : (t_entry < t_exit ? 1 : 0);
    if (found && hit_t < best_t) {
      best_t = hit_t;
      best_prim = hit_prim;
    }
  }
  prim_out[i] = best_prim;
  t_out[i] = best_t;
  status[i] = (uint8_t)(reason | (finite ? 0u : 0x80u));
}

// Hit attributes of traced rays: the barycentrics of the hit triangle (the
// Moller-Trumbore test of the hit repeated on the same transformed ray and
// record, so u, v are the traversal's own), the interpolated vertex normal
// (gpu_get_tri_normal, RayEngine/src/cuda-src/ray_functions.cuh:146-155), the
// interpolated uv (get_uvs, :157-166) or, without vertex uvs, (u, v) itself
// (the reference's material-without-uvs case, :206-213), and the hit point
// ro + t rd in object space (:204).  Misses get zeros.
__global__ void __launch_bounds__(256) hit_attr_kernel(const ObjXform* __restrict__ objs,
                                                       const float* __restrict__ geom, const float* __restrict__ nrm,
                                                       const float* __restrict__ uvs, const float* __restrict__ ray_o,
                                                       const float* __restrict__ ray_d, const int32_t* __restrict__ prim,
                                                       const int32_t* __restrict__ obj, int64_t n, int64_t n_attr,
                                                       int32_t n_obj, int32_t one_sided,
                                                       int32_t secondary, float* __restrict__ normal,
                                                       float* __restrict__ uv, float* __restrict__ bary,
                                                       float* __restrict__ hit_point) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t p = prim[i];
  float nx = 0.0f, ny = 0.0f, nz = 0.0f, uu = 0.0f, vv = 0.0f, tu = 0.0f, tv = 0.0f, hx = 0.0f, hy = 0.0f, hz = 0.0f;
  const int32_t k = obj ? obj[i] : 0;
  if (p >= 0 && (p >= n_attr || k < 0 || k >= n_obj)) {
    // ids outside the scene's prims or objects: not from a traversal of it
    nx = ny = nz = uu = vv = tu = tv = hx = hy = hz = __builtin_nanf("");
  } else if (p >= 0) {
    const ObjXform ob = objs[k];
    const f3 o{ray_o[3 * i], ray_o[3 * i + 1], ray_o[3 * i + 2]};
    const f3 d{ray_d[3 * i], ray_d[3 * i + 1], ray_d[3 * i + 2]};
    f3 ro, rd;
    xform_ray(ob, o, d, secondary != 0, ro, rd);
    const float* g = geom + 9 * (int64_t)p;
    float t, u, v;
    if (ray_tri_uv(ro, rd, f3{g[0], g[1], g[2]}, f3{g[3], g[4], g[5]}, f3{g[6], g[7], g[8]}, one_sided == 0, t, u,
                   v)) {
      const float* q = nrm + 9 * (int64_t)p;
      const float w = 1.0f - (u + v);
      const f3 nn = normalize(add(add(smul(w, f3{q[0], q[1], q[2]}), smul(u, f3{q[3], q[4], q[5]})),
                                  smul(v, f3{q[6], q[7], q[8]})));
      nx = nn.x;
      ny = nn.y;
      nz = nn.z;
      if (uvs) {
        const float* c = uvs + 6 * (int64_t)p;
        const float w2 = 1.0f - u - v;
        uu = (w2 * c[0] + u * c[2]) + v * c[4];
        vv = (w2 * c[1] + u * c[3]) + v * c[5];
      } else {
        uu = u;
        vv = v;
      }
      tu = u;
      tv = v;
      const f3 hp = add(ro, smul(t, rd));
      hx = hp.x;
      hy = hp.y;
      hz = hp.z;
    } else {
      // the ray does not hit its prim: inputs not from one traversal
      nx = ny = nz = uu = vv = tu = tv = hx = hy = hz = __builtin_nanf("");
    }
  }
  if (normal) {
    normal[3 * i] = nx;
    normal[3 * i + 1] = ny;
    normal[3 * i + 2] = nz;
  }
  if (uv) {
    uv[2 * i] = uu;
    uv[2 * i + 1] = vv;
  }
  if (bary) {
    bary[2 * i] = tu;
    bary[2 * i + 1] = tv;
  }
  if (hit_point) {
    hit_point[3 * i] = hx;
    hit_point[3 * i + 1] = hy;
    hit_point[3 * i + 2] = hz;
  }
}

__global__ void zero_words_kernel(uint32_t* __restrict__ p, int32_t n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

}  // namespace

// The scene's array sizes for the bounds checks of the debug build (a no-op
// otherwise); the copy is synchronous, so launches on other streams of the
// process see one scene's sizes at a time.
static hipError_t set_debug_bounds(const DevScene& S, hipStream_t s) {
#ifdef RE_DEBUG_BOUNDS
  hipError_t e = hipStreamSynchronize(s);
  DebugBounds b{(uint32_t)S.n_nodes, (uint32_t)S.n_leaves, (uint32_t)S.n_trirecs, 0u};
  // RE_DEBUG_BOUNDS_LEAF_LIMIT=k checks leaf ids against min(k, leaves) instead:
  // ids at or above k are then flagged and read as leaf 0 (still inside the
  // array; walks end at their leaf-visit cap), which shows the check firing.
  if (const char* lim = std::getenv("RE_DEBUG_BOUNDS_LEAF_LIMIT")) {
    const long k = std::strtol(lim, nullptr, 10);
    if (k > 0 && (uint64_t)k < b.n_leaves) b.n_leaves = (uint32_t)k;
  }
  if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(g_dbg_bounds), &b, sizeof(b), 0, hipMemcpyHostToDevice);
  return e;
#else
  (void)S;
  (void)s;
  return hipSuccess;
#endif
}

hipError_t launch_hit_attributes(const AttrLaunch& L, hipStream_t s) {
  if (L.n <= 0) return hipSuccess;
  const DevScene& S = *L.scene;
  hipLaunchKernelGGL(hit_attr_kernel, dim3((unsigned)((L.n + 255) / 256)), dim3(256), 0, s, S.objs, S.attr_geom,
                     S.attr_nrm, S.attr_uv, L.ray_o, L.ray_d, L.prim, L.obj, L.n, S.n_attr, S.n_obj, L.one_sided,
                     L.secondary, L.normal, L.uv, L.bary, L.hit_point);
  return hipGetLastError();
}

hipError_t launch_trace_status(const StatusLaunch& L, hipStream_t s) {
  if (L.n <= 0) return hipSuccess;
  {
    const hipError_t e = set_debug_bounds(*L.scene, s);
    if (e != hipSuccess) return e;
  }
  const DevScene& S = *L.scene;
  const int32_t cap = (L.cap > 0 && L.cap < kMaxLeafVisits) ? L.cap : kMaxLeafVisits;
  const dim3 grid((unsigned)((L.n + 255) / 256));
  if (L.one_sided)
    hipLaunchKernelGGL(trace_status_kernel<true>, grid, dim3(256), 0, s, S.nodes, S.leaves, S.tris, S.objs, S.n_obj,
                       L.ray_o, L.ray_d, L.n, cap, L.secondary, L.prim, L.t, L.status);
  else
    hipLaunchKernelGGL(trace_status_kernel<false>, grid, dim3(256), 0, s, S.nodes, S.leaves, S.tris, S.objs,
                       S.n_obj, L.ray_o, L.ray_d, L.n, cap, L.secondary, L.prim, L.t, L.status);
  return hipGetLastError();
}

hipError_t launch_zero_words(uint32_t* p, int32_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(zero_words_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

int64_t trace_persistent_grid(int block) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, trace_persistent_kernel<false, true>, block, 0) !=
      hipSuccess)
    return 0;
  return (int64_t)cus * (per_cu > 0 ? per_cu : 1);
}

int64_t trace_coop_refill_grid(int block) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, trace_coop_refill_kernel<false, true, true>, block,
                                                    (size_t)(block / 64) * sizeof(CoopWave)) != hipSuccess)
    return 0;
  return (int64_t)cus * (per_cu > 0 ? per_cu : 1);
}

hipError_t debug_bounds_take(uint32_t* mask) {
#ifdef RE_DEBUG_BOUNDS
  unsigned int v = 0, z = 0;
  hipError_t e = hipMemcpyFromSymbol(&v, HIP_SYMBOL(g_dbg_errors), sizeof(v), 0, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(g_dbg_errors), &z, sizeof(z), 0, hipMemcpyHostToDevice);
  *mask = v;
  return e;
#else
  *mask = 0;
  return hipErrorNotSupported;
#endif
}

hipError_t launch_trace_closest(const TraceLaunch& L, hipStream_t s) {
  if (L.n <= 0) return hipSuccess;
  {
    const hipError_t e = set_debug_bounds(*L.scene, s);
    if (e != hipSuccess) return e;
  }
  // Default work-group: 256 threads for the persistent kernel, one wave for
  // the one-ray-per-thread kernels (finer dispatch granularity at the tail;
  // 64 and 128 measured alike and ahead of 256 on the cluster workload).
  const int block = trace_launch_block(L.block, L.variant);
  if (L.variant == 5) {
    // Persistent cooperative waves: enough blocks to fill every CU at the
    // kernel's occupancy, never more than the rays need.
    const int64_t want = (L.n + block - 1) / block;
    int64_t grid = L.grid_blocks > 0 ? L.grid_blocks : want;
    if (grid > want) grid = want;
    const DevScene& S5 = *L.scene;
    const int32_t cap5 = (L.cap > 0 && L.cap < kMaxLeafVisits) ? L.cap : kMaxLeafVisits;
    const int hint = L.hit_leaf ? 2 : (L.start_leaf ? 1 : 0);
#define RE_LAUNCH_REFILL_H(C, O, ONE, H)                                                                       \
  hipLaunchKernelGGL((trace_coop_refill_kernel<C, O, ONE, H>), dim3((unsigned)grid), dim3(block),              \
                     (block / 64) * sizeof(CoopWave), s, S5.nodes, S5.leaves, S5.tris, S5.objs, S5.n_obj,        \
                     L.ray_o, L.ray_d, L.n, cap5, L.secondary, L.prim, L.t, L.obj, L.counts, L.queue,            \
                     L.start_leaf, L.hit_leaf, S5.n_leaves)
#define RE_LAUNCH_REFILL(C, O, ONE)                                         \
  do {                                                                      \
    if (hint == 2) RE_LAUNCH_REFILL_H(C, O, ONE, 2);                        \
    else if (hint == 1) RE_LAUNCH_REFILL_H(C, O, ONE, 1);                   \
    else RE_LAUNCH_REFILL_H(C, O, ONE, 0);                                  \
  } while (0)
    const bool one = S5.n_obj == 1;
    if (L.counts) {
      if (L.one_sided) { if (one) RE_LAUNCH_REFILL(true, true, true); else RE_LAUNCH_REFILL(true, true, false); }
      else { if (one) RE_LAUNCH_REFILL(true, false, true); else RE_LAUNCH_REFILL(true, false, false); }
    } else {
      if (L.one_sided) { if (one) RE_LAUNCH_REFILL(false, true, true); else RE_LAUNCH_REFILL(false, true, false); }
      else { if (one) RE_LAUNCH_REFILL(false, false, true); else RE_LAUNCH_REFILL(false, false, false); }
    }
#undef RE_LAUNCH_REFILL
#undef RE_LAUNCH_REFILL_H
    return hipGetLastError();
  }
  const DevScene& S = *L.scene;
  // A ray can alternate between two leaves when its exit point lies within
  // the rope tolerance of two faces; the visit cap is what ends it, so
  // "unlimited" is a large bound that every ray reaches.
  int32_t cap = (L.cap > 0 && L.cap < kMaxLeafVisits) ? L.cap : kMaxLeafVisits;
  if (L.variant == 1) {
    // Persistent: enough blocks to fill every CU at the kernel's occupancy,
    // never more than the rays need.
    int64_t want = (L.n + block - 1) / block;
    int64_t grid = L.grid_blocks > 0 ? L.grid_blocks : want;
    if (grid > want) grid = want;
#define RE_LAUNCH(C, O)                                                                                     \
  hipLaunchKernelGGL((trace_persistent_kernel<C, O>), dim3((unsigned)grid), dim3(block), 0, s, S.nodes,     \
                     S.leaves, S.tris, S.objs, S.n_obj, L.ray_o, L.ray_d, L.n, cap, L.secondary, L.prim, L.t, \
                     L.obj, L.counts, L.queue)
    if (L.counts) {
      if (L.one_sided) RE_LAUNCH(true, true); else RE_LAUNCH(true, false);
    } else {
      if (L.one_sided) RE_LAUNCH(false, true); else RE_LAUNCH(false, false);
    }
#undef RE_LAUNCH
    return hipGetLastError();
  }
  const int64_t grid = (L.n + block - 1) / block;
#define RE_LAUNCH(K, ...)                                                                                       \
  hipLaunchKernelGGL((K<__VA_ARGS__>), dim3((unsigned)grid), dim3(block), 0, s, S.nodes, S.leaves, S.tris, S.objs, \
                     S.n_obj, L.ray_o, L.ray_d, L.n, cap, L.secondary, L.prim, L.t, L.obj, L.counts)
#define RE_LAUNCH_ALL(K, ...)                                             \
  do {                                                                    \
    if (L.counts) {                                                       \
      if (L.one_sided) RE_LAUNCH(K, true, true __VA_ARGS__);              \
      else RE_LAUNCH(K, true, false __VA_ARGS__);                         \
    } else {                                                              \
      if (L.one_sided) RE_LAUNCH(K, false, true __VA_ARGS__);             \
      else RE_LAUNCH(K, false, false __VA_ARGS__);                        \
    }                                                                     \
  } while (0)
  if (L.variant == 4) {
    const int hint = L.hit_leaf ? 2 : (L.start_leaf ? 1 : 0);
#define RE_LAUNCH_COOP_H(C, O, ONE, H)                                                                        \
  hipLaunchKernelGGL((trace_coop_kernel<C, O, ONE, H>), dim3((unsigned)grid), dim3(block),                    \
                     (block / 64) * sizeof(CoopWave), s, S.nodes, S.leaves,                                      \
                     S.tris, S.objs, S.n_obj, L.ray_o, L.ray_d, L.n, cap, L.secondary, L.prim, L.t, L.obj,        \
                     L.counts, L.start_leaf, L.hit_leaf, S.n_leaves)
#define RE_LAUNCH_COOP(C, O, ONE)                                         \
  do {                                                                    \
    if (hint == 2) RE_LAUNCH_COOP_H(C, O, ONE, 2);                        \
    else if (hint == 1) RE_LAUNCH_COOP_H(C, O, ONE, 1);                   \
    else RE_LAUNCH_COOP_H(C, O, ONE, 0);                                  \
  } while (0)
    const bool one = S.n_obj == 1;
    if (L.counts) {
      if (L.one_sided) { if (one) RE_LAUNCH_COOP(true, true, true); else RE_LAUNCH_COOP(true, true, false); }
      else { if (one) RE_LAUNCH_COOP(true, false, true); else RE_LAUNCH_COOP(true, false, false); }
    } else {
      if (L.one_sided) { if (one) RE_LAUNCH_COOP(false, true, true); else RE_LAUNCH_COOP(false, true, false); }
      else { if (one) RE_LAUNCH_COOP(false, false, true); else RE_LAUNCH_COOP(false, false, false); }
    }
#undef RE_LAUNCH_COOP
#undef RE_LAUNCH_COOP_H
  } else if (L.variant == 2) RE_LAUNCH_ALL(trace_closest_kernel, , true);
  else if (L.variant == 3 && S.n_obj == 1) RE_LAUNCH_ALL(trace_closest_o6_kernel, , true);
  else if (L.variant == 3) RE_LAUNCH_ALL(trace_closest_o6_kernel, , false);
  else RE_LAUNCH_ALL(trace_closest_kernel, , false);
#undef RE_LAUNCH_ALL
#undef RE_LAUNCH
  return hipGetLastError();
}

hipError_t launch_trace_occluded(const TraceLaunch& L, hipStream_t s) {
  if (L.n <= 0) return hipSuccess;
  {
    const hipError_t e = set_debug_bounds(*L.scene, s);
    if (e != hipSuccess) return e;
  }
  const int block = L.block > 0 ? L.block : 64;
  const DevScene& S = *L.scene;
  const int32_t cap = (L.cap > 0 && L.cap < kMaxLeafVisits) ? L.cap : kMaxLeafVisits;
  const int64_t grid = (L.n + block - 1) / block;
  if (L.one_sided)
    hipLaunchKernelGGL(trace_occluded_kernel<true>, dim3((unsigned)grid), dim3(block), 0, s, S.nodes, S.leaves,
                       S.tris, S.objs, S.n_obj, L.ray_o, L.ray_d, L.n, cap, L.secondary, L.occluded);
  else
    hipLaunchKernelGGL(trace_occluded_kernel<false>, dim3((unsigned)grid), dim3(block), 0, s, S.nodes, S.leaves,
                       S.tris, S.objs, S.n_obj, L.ray_o, L.ray_d, L.n, cap, L.secondary, L.occluded);
  return hipGetLastError();
}

}  // namespace re
