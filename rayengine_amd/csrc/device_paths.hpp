// Per-ray device paths shared by the single-workload kernels and the batched
// work-queue kernel, so a ray gives bit-identical results whichever kernel
// runs it.
#pragma once

#include "internal.hpp"

namespace re {

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
template <bool ONE_SIDED, bool COUNT, bool ANY = false>
__device__ __forceinline__ void test_leaf(const TriRec* __restrict__ tp, uint32_t count, f3 ro, f3 rd,
                                          float& t_exit, bool& found, float& hit_t, int32_t& hit_prim,
                                          unsigned& c_tri) {
  for (uint32_t j = 0; j < count; j += 4) {
    if (ANY && found) return;
    TriGeom g[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (j + q < count) g[q] = load_geom(tp, j + q);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
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
  __builtin_memcpy(&q, nodes + c, sizeof(q));
  return q;
}

// Descends from entry cur, whose window q is loaded, to the leaf containing
// pe (the down-walk of ray_functions.cuh:289-291: below if pe < split, else
// above).  The below child of an entry is the next entry, so a "below" step
// from the window's first entry needs no new load; any other step does.
// Returns the leaf id.
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
      const DevLeaf* lp = leaves + leaf_id;
      const float4 h0 = *reinterpret_cast<const float4*>(lp->bmin);
      const float4 h1 = *reinterpret_cast<const float4*>(lp->bmax + 1);
      const float bmin[3] = {h0.x, h0.y, h0.z}, bmax[3] = {h0.w, h1.x, h1.y};
      const uint32_t first = __float_as_uint(h1.z), count = __float_as_uint(h1.w);
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
