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
#include "../internal.hpp"

namespace re {
namespace {

template <bool COUNT, bool ONE_SIDED>
__global__ void __launch_bounds__(256) trace_closest_kernel(
    const NodeEntry* __restrict__ nodes, const LeafRec* __restrict__ leaves,
    const TriRec* __restrict__ tris, const ObjXform* __restrict__ objs, int32_t n_obj,
    const float* __restrict__ ray_o, const float* __restrict__ ray_d, int64_t n, int32_t cap,
    int32_t secondary, int32_t* __restrict__ prim_out, float* __restrict__ t_out,
    int32_t* __restrict__ obj_out, TraceCounts* __restrict__ counts) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const f3 o{ray_o[3 * i], ray_o[3 * i + 1], ray_o[3 * i + 2]};
  const f3 d{ray_d[3 * i], ray_d[3 * i + 1], ray_d[3 * i + 2]};
  float best_t = kInf;
  int32_t best_prim = -1, best_obj = -1;
  unsigned c_ent = 0, c_leaf = 0, c_tri = 0;

  for (int32_t k = 0; k < n_obj; ++k) {
    const ObjXform ob = objs[k];
    f3 ro, rd;
    xform_ray(ob, o, d, secondary != 0, ro, rd);
    float t_entry, t_exit;
    if (!ray_box(ob.root_min, ob.root_max, ro, rd, t_entry, t_exit)) continue;
    int32_t cur = ob.root;
    bool found = false;
    float hit_t = kInf;
    int32_t hit_prim = -1;
    int32_t visits = 0;
    while (t_entry < t_exit && visits < cap) {
      ++visits;
      const f3 pe = add(ro, smul(t_entry, rd));
      while (!ref_is_leaf(cur)) {
        const NodeEntry e = nodes[cur];
        if (COUNT) ++c_ent;
        const uint32_t ax = e.meta >> 30;
        const int32_t other = (int32_t)(e.meta & 0x3fffffffu);
        if (ax == 3u) {
          cur = leaf_ref((uint32_t)other);
          break;
        }
        cur = (get(pe, (int)ax) < e.split) ? cur + 1 : other;
      }
      const LeafRec L = leaves[ref_leaf_id(cur)];
      if (COUNT) ++c_leaf;
      for (uint32_t j = 0; j < L.count; ++j) {
        const TriRec tr = tris[L.first + j];
        if (COUNT) ++c_tri;
        float t;
        bool h = ONE_SIDED ? ray_tri(ro, rd, f3{tr.v0[0], tr.v0[1], tr.v0[2]},
                                     f3{tr.e1[0], tr.e1[1], tr.e1[2]},
                                     f3{tr.e2[0], tr.e2[1], tr.e2[2]}, t)
                           : ray_tri_2s(ro, rd, f3{tr.v0[0], tr.v0[1], tr.v0[2]},
                                        f3{tr.e1[0], tr.e1[1], tr.e1[2]},
                                        f3{tr.e2[0], tr.e2[1], tr.e2[2]}, t);
        if (h && t < t_exit) {
          found = true;
          hit_t = t;
          hit_prim = tr.prim;
          t_exit = t;
        }
      }
      float tn, tf;
      if (!ray_box(L.bmin, L.bmax, ro, rd, tn, tf)) break;
      t_entry = tf;
      const f3 px = add(ro, smul(t_entry, rd));
      const int face = exit_face(L.bmin, L.bmax, px);
      if (face < 0) break;
      const int32_t r = L.rope[face];
      if (r == kNoRope) break;
      cur = r;
    }
    if (found && hit_t < best_t) {
      best_t = hit_t;
      best_prim = hit_prim;
      best_obj = k;
    }
  }
  prim_out[i] = best_prim;
  t_out[i] = best_t;
  if (obj_out) obj_out[i] = best_obj;
  if (COUNT) {
    atomicAdd(&counts->entries, (unsigned long long)c_ent);
    atomicAdd(&counts->leaves, (unsigned long long)c_leaf);
    atomicAdd(&counts->tris, (unsigned long long)c_tri);
    if (best_prim >= 0) atomicAdd(&counts->rays_hit, 1ull);
  }
}

}  // namespace

hipError_t launch_trace_closest(const TraceLaunch& L, hipStream_t s) {
  if (L.n <= 0) return hipSuccess;
  const int block = L.block > 0 ? L.block : 256;
  const int64_t grid = (L.n + block - 1) / block;
  const DevScene& S = *L.scene;
  // A ray can alternate between two leaves when its exit point lies within
  // the rope tolerance of two faces; the visit cap is what ends it, so
  // "unlimited" is a large bound that every ray reaches.
  int32_t cap = (L.cap > 0 && L.cap < kMaxLeafVisits) ? L.cap : kMaxLeafVisits;
#define RE_LAUNCH(C, O)                                                                      \
  hipLaunchKernelGGL((trace_closest_kernel<C, O>), dim3((unsigned)grid), dim3(block), 0, s,  \
                     S.nodes, S.leaves, S.tris, S.objs, S.n_obj, L.ray_o, L.ray_d, L.n, cap, \
                     L.secondary, L.prim, L.t, L.obj, L.counts)
  if (L.counts) {
    if (L.one_sided) RE_LAUNCH(true, true); else RE_LAUNCH(true, false);
  } else {
    if (L.one_sided) RE_LAUNCH(false, true); else RE_LAUNCH(false, false);
  }
#undef RE_LAUNCH
  return hipGetLastError();
}

}  // namespace re
