// CPU oracle for the RayEngine ray-query path.  TEST INFRASTRUCTURE ONLY:
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it as
// the checker; the product library never links or calls it.
//
// Each function restates a reference routine (hadalhw17/RayEngine, paths
// relative to its root) in plain C++ with its own vector code, compiled
// -O2 -ffp-contract=off -fno-fast-math, so FP32 results are the IEEE results
// of the reference's operation order.  What pins each function (closed forms,
// known-answer vectors, brute force) is listed in its header comment; the
// tests under tests/ run those checks.
//
// Flattened tree layout consumed here (produced by the library's builder):
//   node entry  8 B  {float split; uint32 meta}  meta>>30 = axis, 3 = leaf;
//                    inner: below child = entry+1, above = meta & 0x3fffffff;
//                    leaf: leaf id = meta & 0x3fffffff
//   leaf       64 B  {bmin[3], bmax[3], int32 rope[6], uint32 first, count, pad[2]}
//   tri        40 B  {v0[3], e1[3], e2[3], int32 prim}
//   rope/ref   int32: >= 0 node entry; -1 none; other negatives: leaf id | 0x80000000
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <random>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API extern "C" __attribute__((visibility("default")))

namespace {

const float K_EPSILON = 1e-4f;   // RayEngine/src/RayEngine/RayEngine.h:24
const float K_INFINITY = 1e20f;  // RayEngine/src/RayEngine/RayEngine.h:25

template <typename T>
struct V3 {
  T x, y, z;
};
template <typename T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename T> inline T dotp(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> inline V3<T> crossp(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// normalize(v) = v * rsqrt(dot(v,v)) with rsqrt = 1/sqrt (the host form of
// RayEngine/src/helper_math.h:66-69,1313-1317).
template <typename T> inline V3<T> unit(V3<T> v) {
  T inv = T(1) / std::sqrt(dotp(v, v));
  return {v.x * inv, v.y * inv, v.z * inv};
}
template <typename T> inline T comp(V3<T> v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

struct Node { float split; uint32_t meta; };
struct Leaf { float bmin[3], bmax[3]; int32_t rope[6]; uint32_t first, count, pad[2]; };
struct Tri { float v0[3], e1[3], e2[3]; int32_t prim; };
struct Obj {
  float loc[3];
  float sin_y, cos_y;
  int32_t is_character;
  int32_t root;
  float root_min[3], root_max[3];
};
static_assert(sizeof(Node) == 8 && sizeof(Leaf) == 64 && sizeof(Tri) == 40 && sizeof(Obj) == 52, "layout");

// gpu_ray_box_intersect, RayEngine/src/cuda-src/ray_functions.cuh:61-92.
// fminf/fmaxf are the CUDA device functions (IEEE minNum/maxNum); fsel
// (helper_math.h:35-39) clamps t_near to >= 0.
template <typename T>
bool box_test(const T bmin[3], const T bmax[3], V3<T> o, V3<T> d, T& t_near, T& t_far) {
  V3<T> fr{T(1) / d.x, T(1) / d.y, T(1) / d.z};
  T t1 = (bmin[0] - o.x) * fr.x, t2 = (bmax[0] - o.x) * fr.x;
  T t3 = (bmin[1] - o.y) * fr.y, t4 = (bmax[1] - o.y) * fr.y;
  T t5 = (bmin[2] - o.z) * fr.z, t6 = (bmax[2] - o.z) * fr.z;
  T tmin = std::fmax(std::fmax(std::fmin(t1, t2), std::fmin(t3, t4)), std::fmin(t5, t6));
  T tmax = std::fmin(std::fmin(std::fmax(t1, t2), std::fmax(t3, t4)), std::fmax(t5, t6));
  if (tmax < 0) return false;
  if (tmin > tmax) return false;
  t_near = tmin;
  t_far = tmax;
  t_near = (t_near >= 0) ? t_near : T(0);
  return true;
}

// gpu_ray_tri_intersect, RayEngine/src/cuda-src/ray_functions.cuh:173-223
// (one-sided: a < K_EPSILON misses).  two_sided selects the |a| < K_EPSILON
// rejection of the host triIntersect (RayEngine/src/Primitives/KDTree.cpp:909-949)
// with the device routine's remaining steps.
template <typename T>
bool tri_test(V3<T> o, V3<T> d, V3<T> v0, V3<T> e1, V3<T> e2, bool two_sided, T& t_out, T* uv) {
  const T eps = (T)K_EPSILON;
  V3<T> h = crossp(d, e2);
  T a = dotp(e1, h);
  if (two_sided ? (a > -eps && a < eps) : (a < eps)) return false;
  T f = T(1) / a;
  V3<T> s = o - v0;
  T u = f * dotp(s, h);
  if (u < T(0) || u > T(1)) return false;
  V3<T> q = crossp(s, e1);
  T v = f * dotp(d, q);
  if (v < T(0) || u + v > T(1)) return false;
  T t = f * dotp(e2, q);
  if (t > eps) {
    t_out = t;
    if (uv) { uv[0] = u; uv[1] = v; }
    return true;
  }
  return false;
}

// is_point_to_the_left_of_split, RayEngine/src/cuda-src/kd_tree_functions.cuh:9-25.
inline bool left_of_split(int axis, float split, V3<float> p) {
  if (axis == 0) return p.x < split;
  if (axis == 1) return p.y < split;
  if (axis == 2) return p.z < split;
  return false;
}

// get_neighboring_node_index, RayEngine/src/cuda-src/kd_tree_functions.cuh:31-62.
inline int32_t neighbour(const Leaf& L, V3<float> p) {
  if (std::fabs(p.x - L.bmin[0]) < K_EPSILON) return L.rope[0];  // LEFT
  if (std::fabs(p.z - L.bmax[2]) < K_EPSILON) return L.rope[1];  // FRONT
  if (std::fabs(p.x - L.bmax[0]) < K_EPSILON) return L.rope[2];  // RIGHT
  if (std::fabs(p.z - L.bmin[2]) < K_EPSILON) return L.rope[3];  // BACK
  if (std::fabs(p.y - L.bmax[1]) < K_EPSILON) return L.rope[4];  // TOP
  if (std::fabs(p.y - L.bmin[1]) < K_EPSILON) return L.rope[5];  // BOTTOM
  return -1;
}

// Object transform, RayEngine/src/cuda-src/ray_functions.cuh:238-267.
template <typename T>
void transform(const Obj& ob, V3<T> o, V3<T> d, bool secondary, V3<T>& no, V3<T>& nd) {
  T s = (T)ob.sin_y, c = (T)ob.cos_y;
  nd = {d.x * c + d.z * s, d.y, -d.x * s + d.z * c};
  no = {T(0), T(0), T(0)};
  if (!ob.is_character) {
    no = {o.x * c + o.z * s, o.y, -o.x * s + o.z * c};
    no = no + V3<T>{(T)ob.loc[0], (T)ob.loc[1], (T)ob.loc[2]};
  }
  if (ob.is_character && secondary) {
    T s2 = -s, c2 = c;  // sinf(-a) = -sinf(a), cosf(-a) = cosf(a)
    nd = {d.x * c2 + d.z * s2, d.y, -d.x * s2 + d.z * c2};
    no = {(T)ob.loc[0] * c2 + (T)ob.loc[2] * s2, -(T)ob.loc[1], -(T)ob.loc[0] * s2 + (T)ob.loc[2] * c2};
  }
  nd = unit(nd);
}

// Work counters of a ray and how its last traversal loop ended:
// 0 t_entry >= t_exit, 1 visit cap, 2 leaf slab miss, 3 no face within
// K_EPSILON, 4 rope -1 (left the root box), 5 root box missed, 6 the walk
// state repeated (only with cut).
struct Counts { uint64_t entries, leaves, tris, reason; };

// stackless_kdtree_traversal, RayEngine/src/cuda-src/ray_functions.cuh:230-360,
// for one object.  Returns true with (t, prim) of the accepted hit.
//
// cut: stop when (cur, t_entry, t_exit) at the top of a visit equals its value
// at one of the two visits before.  Nothing is accepted between two equal
// states (an accepted hit lowers t_exit) and the next visit depends on that
// state alone, so the uncut walk repeats them to the cap with the same result;
// cut only changes the counters, to the work of the shortened walk the
// library runs.  Equality of the results with and without cut is tested.
bool traverse(const Node* nodes, const Leaf* leaves, const Tri* tris, const Obj& ob, V3<float> o, V3<float> d,
              int cap, bool two_sided, bool secondary, float& t_hit, int32_t& prim, Counts& cnt,
              bool cut = false) {
  V3<float> no, nd;
  transform(ob, o, d, secondary, no, nd);
  float t_entry, t_exit;
  if (!box_test(ob.root_min, ob.root_max, no, nd, t_entry, t_exit)) {
    cnt.reason = 5;
    return false;
  }
  bool detected = false;
  int limit = 0;
  int32_t cur = ob.root;  // current node as a ref
  int32_t seen_c[2] = {-1, -1};  // a walk's cur is never -1
  float seen_e[2] = {0.0f, 0.0f}, seen_x[2] = {0.0f, 0.0f};
  while (t_entry < t_exit && limit < cap) {
    if (cut) {
      for (int q = 0; q < 2; ++q)
        if (seen_c[q] == cur && seen_e[q] == t_entry && seen_x[q] == t_exit) {
          cnt.reason = 6;
          return detected;
        }
      seen_c[1] = seen_c[0];
      seen_e[1] = seen_e[0];
      seen_x[1] = seen_x[0];
      seen_c[0] = cur;
      seen_e[0] = t_entry;
      seen_x[0] = t_exit;
    }
    ++limit;
    V3<float> p_entry = no + (t_entry * nd);
    // Down traversal to a leaf (:289-291).
    while (cur >= 0) {
      const Node& e = nodes[cur];
      ++cnt.entries;
      uint32_t axis = e.meta >> 30;
      if (axis == 3u) { cur = (int32_t)(0x80000000u | (e.meta & 0x3fffffffu)); break; }
      cur = left_of_split((int)axis, e.split, p_entry) ? cur + 1 : (int32_t)(e.meta & 0x3fffffffu);
    }
    const Leaf& L = leaves[(uint32_t)cur & 0x7fffffffu];
    ++cnt.leaves;
    // Leaf triangles in order (:295-324).
    for (uint32_t i = L.first; i < L.first + L.count; ++i) {
      const Tri& tr = tris[i];
      ++cnt.tris;
      float t;
      if (tri_test<float>(no, nd, {tr.v0[0], tr.v0[1], tr.v0[2]}, {tr.e1[0], tr.e1[1], tr.e1[2]},
                          {tr.e2[0], tr.e2[1], tr.e2[2]}, two_sided, t, nullptr)) {
        if (t < t_exit) {
          detected = true;
          t_hit = t;
          prim = tr.prim;
          t_exit = t;
        }
      }
    }
    float tn, tf;
    if (!box_test(L.bmin, L.bmax, no, nd, tn, tf)) { cnt.reason = 2; return detected; }  // :328-337
    t_entry = tf;
    V3<float> p_exit = no + (t_entry * nd);
    int32_t next = neighbour(L, p_exit);  // :341-347
    if (next == -1) {
      int face = -1;
      for (int f = 0; f < 6 && face < 0; ++f) {
        const float* pl = (f == 1 || f == 2 || f == 4) ? L.bmax : L.bmin;
        const int ax = (f == 0 || f == 2) ? 0 : ((f == 1 || f == 3) ? 2 : 1);
        const float pc = ax == 0 ? p_exit.x : (ax == 1 ? p_exit.y : p_exit.z);
        if (std::fabs(pc - pl[ax]) < K_EPSILON) face = f;
      }
      cnt.reason = face < 0 ? 3 : 4;
      return detected;
    }
    cur = next;
  }
  cnt.reason = (t_entry < t_exit) ? 1 : 0;
  return detected;
}

}  // namespace

// ---------------------------------------------------------------- traversal

// trace_scene (RayEngine/src/cuda-src/ray_functions.cuh:376-396) over every
// object, for n rays, OpenMP over rays with schedule(dynamic, 64) as in
// RayEngine/src/Engine/RayTracer.cpp:154.  counts (nullable) gets 4 per ray:
// entries, leaves, triangles tested and the end reason of the last loop.
// cap <= 0 means the ceiling 2^20 (a ray whose exit point lies within
// K_EPSILON of two faces can alternate between two leaves until the cap).  Pinned by brute-force equality at cap = inf
// (orc_brute), the icosphere closed form and the primitive KATs.  cut: see
// traverse.
ORC_API void orc_trace(const void* nodes, const void* leaves, const void* tris, const void* objs, int n_obj,
                       const float* ray_o, const float* ray_d, int64_t n, int cap, int two_sided, int secondary,
                       int32_t* prim_out, float* t_out, int32_t* obj_out, uint64_t* counts, int threads, int cut) {
  const Node* N = (const Node*)nodes;
  const Leaf* L = (const Leaf*)leaves;
  const Tri* T = (const Tri*)tris;
  const Obj* O = (const Obj*)objs;
  if (cap <= 0 || cap > (1 << 20)) cap = 1 << 20;  // same ceiling as the library
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
#endif
  for (int64_t i = 0; i < n; ++i) {
    V3<float> o{ray_o[3 * i], ray_o[3 * i + 1], ray_o[3 * i + 2]};
    V3<float> d{ray_d[3 * i], ray_d[3 * i + 1], ray_d[3 * i + 2]};
    float best = K_INFINITY;
    int32_t bp = -1, bo = -1;
    Counts c{0, 0, 0, 0};
    for (int k = 0; k < n_obj; ++k) {
      float t = K_INFINITY;
      int32_t p = -1;
      if (traverse(N, L, T, O[k], o, d, cap, two_sided != 0, secondary != 0, t, p, c, cut != 0) && t < best) {
        best = t;
        bp = p;
        bo = k;
      }
    }
    prim_out[i] = bp;
    t_out[i] = best;
    if (obj_out) obj_out[i] = bo;
    if (counts) {
      counts[4 * i] = c.entries;
      counts[4 * i + 1] = c.leaves;
      counts[4 * i + 2] = c.tris;
      counts[4 * i + 3] = c.reason;
    }
  }
}

// trace_shadow (RayEngine/src/cuda-src/ray_functions.cuh:365-374): the
// closest-hit walk of every object (stackless_kdtree_traversal, is_secondary
// as given; the reference passes true), each overwriting the shared
// HitResult when it detects an intersection.  Its callers
// (material_functions.cuh:68-119) read only `hits`, so the result is 1 when
// some object's walk detects an intersection.  Pinned by its relation to
// orc_trace (per object, detected <=> a closest hit) and by known answers.
ORC_API void orc_occluded(const void* nodes, const void* leaves, const void* tris, const void* objs, int n_obj,
                          const float* ray_o, const float* ray_d, int64_t n, int cap, int two_sided, int secondary,
                          uint8_t* out, int threads) {
  const Node* N = (const Node*)nodes;
  const Leaf* L = (const Leaf*)leaves;
  const Tri* T = (const Tri*)tris;
  const Obj* O = (const Obj*)objs;
  if (cap <= 0 || cap > (1 << 20)) cap = 1 << 20;
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
#endif
  for (int64_t i = 0; i < n; ++i) {
    V3<float> o{ray_o[3 * i], ray_o[3 * i + 1], ray_o[3 * i + 2]};
    V3<float> d{ray_d[3 * i], ray_d[3 * i + 1], ray_d[3 * i + 2]};
    bool hits = false;
    for (int k = 0; k < n_obj; ++k) {
      float t = K_INFINITY;
      int32_t p = -1;
      Counts c{0, 0, 0, 0};
      if (traverse(N, L, T, O[k], o, d, cap, two_sided != 0, secondary != 0, t, p, c)) hits = true;
    }
    out[i] = hits ? 1 : 0;
  }
}

// Brute force over a triangle soup (9 floats per triangle), the check the
// reference's gpu_bruteforce_ray_cast performs (RayEngine/src/cuda-src/renderer.cu:328-382):
// root-box gate, then every triangle in index order with a strict t < best.
// Edges are formed in FP32 as in RTScene::init_triangles
// (RayEngine/src/World/RTScene.cpp:228-248).  One object at identity
// transform with the ray direction renormalised as the traversal does.
ORC_API void orc_brute(const float* tri_v, int64_t n_tris, const float* root_min, const float* root_max,
                       const float* ray_o, const float* ray_d, int64_t n, int two_sided, int32_t* prim_out,
                       float* t_out, int threads) {
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
#endif
  for (int64_t i = 0; i < n; ++i) {
    V3<float> o{ray_o[3 * i], ray_o[3 * i + 1], ray_o[3 * i + 2]};
    V3<float> d{ray_d[3 * i], ray_d[3 * i + 1], ray_d[3 * i + 2]};
    Obj id{};
    id.cos_y = 1.0f;
    V3<float> no, nd;
    transform(id, o, d, false, no, nd);
    float best = K_INFINITY;
    int32_t bp = -1;
    float tn, tf;
    if (box_test(root_min, root_max, no, nd, tn, tf)) {
      for (int64_t k = 0; k < n_tris; ++k) {
        const float* v = tri_v + 9 * k;
        V3<float> v0{v[0], v[1], v[2]};
        V3<float> e1{v[3] - v[0], v[4] - v[1], v[5] - v[2]};
        V3<float> e2{v[6] - v[0], v[7] - v[1], v[8] - v[2]};
        float t;
        if (tri_test<float>(no, nd, v0, e1, e2, two_sided != 0, t, nullptr) && t < best) {
          best = t;
          bp = (int32_t)k;
        }
      }
    }
    prim_out[i] = bp;
    t_out[i] = best;
  }
}

// FP64 twin of orc_brute: vertices, edges and the ray in double (the ray
// direction is normalised in double).  Gives the rounding error of the FP32
// t (pinned against the textbook Moller-Trumbore KATs in the tests).
ORC_API void orc_brute_f64(const float* tri_v, int64_t n_tris, const float* ray_o, const float* ray_d, int64_t n,
                           int two_sided, int32_t* prim_out, double* t_out, int threads) {
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
#endif
  for (int64_t i = 0; i < n; ++i) {
    V3<double> o{ray_o[3 * i], ray_o[3 * i + 1], ray_o[3 * i + 2]};
    V3<double> d = unit(V3<double>{ray_d[3 * i], ray_d[3 * i + 1], ray_d[3 * i + 2]});
    double best = (double)K_INFINITY;
    int32_t bp = -1;
    for (int64_t k = 0; k < n_tris; ++k) {
      const float* v = tri_v + 9 * k;
      V3<double> v0{v[0], v[1], v[2]};
      V3<double> e1{(double)v[3] - v[0], (double)v[4] - v[1], (double)v[5] - v[2]};
      V3<double> e2{(double)v[6] - v[0], (double)v[7] - v[1], (double)v[8] - v[2]};
      double t;
      if (tri_test<double>(o, d, v0, e1, e2, two_sided != 0, t, nullptr) && t < best) {
        best = t;
        bp = (int32_t)k;
      }
    }
    prim_out[i] = bp;
    t_out[i] = best;
  }
}

// Single primitive tests for known-answer checks.
ORC_API int orc_ray_tri(const float* o, const float* d, const float* v0, const float* e1, const float* e2,
                        int two_sided, float* t, float* uv) {
  return tri_test<float>({o[0], o[1], o[2]}, {d[0], d[1], d[2]}, {v0[0], v0[1], v0[2]}, {e1[0], e1[1], e1[2]},
                         {e2[0], e2[1], e2[2]}, two_sided != 0, *t, uv)
             ? 1
             : 0;
}

ORC_API int orc_ray_box(const float* bmin, const float* bmax, const float* o, const float* d, float* tn, float* tf) {
  return box_test<float>(bmin, bmax, {o[0], o[1], o[2]}, {d[0], d[1], d[2]}, *tn, *tf) ? 1 : 0;
}

ORC_API int orc_neighbour_face(const float* bmin, const float* bmax, const float* p) {
  Leaf L;
  std::memset(&L, 0, sizeof(L));
  for (int a = 0; a < 3; ++a) { L.bmin[a] = bmin[a]; L.bmax[a] = bmax[a]; }
  for (int f = 0; f < 6; ++f) L.rope[f] = f;
  return neighbour(L, {p[0], p[1], p[2]});
}

// ------------------------------------------------------------ hit attributes

// Hit attributes of a traced ray (prim, obj as the traversal returned them):
// the Moller-Trumbore test of the hit repeated on the object-space ray gives
// the barycentrics (u, v) and t; then
//   normal = normalize(w n0 + u n1 + v n2), w = 1 - (u + v)
//     (gpu_get_tri_normal, RayEngine/src/cuda-src/ray_functions.cuh:146-155),
//   uv = w' uv0 + u uv1 + v uv2, w' = 1 - u - v (get_uvs, :157-166), or (u, v)
//     when there are no vertex uvs (:206-213),
//   hit point = ro + t rd in object space (:204).
// Misses (prim < 0) give zeros.  Pinned by closed forms: flat vertex normals
// give the face normal, a ray aimed at v0 + a e1 + b e2 gives barycentrics
// (a, b) and the linear interpolation of the uvs and normals there, and the
// hit point lies on the triangle (tests/test_attr_oracle.py).
ORC_API void orc_hit_attributes(const float* geom, const float* nrm, const float* uvs, const void* objs,
                                const float* ray_o, const float* ray_d, int64_t n, const int32_t* prim,
                                const int32_t* obj, int two_sided, int secondary, float* normal, float* uv,
                                float* bary, float* hit_point) {
  const Obj* O = static_cast<const Obj*>(objs);
  for (int64_t i = 0; i < n; ++i) {
    float out[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};  // normal, uv, bary, hit point
    const int32_t p = prim[i];
    if (p >= 0) {
      const Obj& ob = O[obj ? obj[i] : 0];
      V3<float> no, nd;
      transform(ob, V3<float>{ray_o[3 * i], ray_o[3 * i + 1], ray_o[3 * i + 2]},
                V3<float>{ray_d[3 * i], ray_d[3 * i + 1], ray_d[3 * i + 2]}, secondary != 0, no, nd);
      const float* g = geom + 9 * p;
      float t = 0.0f, b[2] = {0.0f, 0.0f};
      if (tri_test<float>(no, nd, {g[0], g[1], g[2]}, {g[3], g[4], g[5]}, {g[6], g[7], g[8]}, two_sided != 0, t, b)) {
        const float u = b[0], v = b[1];
        const float* q = nrm + 9 * p;
        const float w = 1.0f - (u + v);
        const V3<float> nn = unit((w * V3<float>{q[0], q[1], q[2]} + u * V3<float>{q[3], q[4], q[5]}) +
                                  v * V3<float>{q[6], q[7], q[8]});
        out[0] = nn.x;
        out[1] = nn.y;
        out[2] = nn.z;
        if (uvs) {
          const float* c = uvs + 6 * p;
          const float w2 = 1.0f - u - v;
          out[3] = (w2 * c[0] + u * c[2]) + v * c[4];
          out[4] = (w2 * c[1] + u * c[3]) + v * c[5];
        } else {
          out[3] = u;
          out[4] = v;
        }
        out[5] = u;
        out[6] = v;
        const V3<float> hp = no + t * nd;
        out[7] = hp.x;
        out[8] = hp.y;
        out[9] = hp.z;
      } else {
        for (float& x : out) x = std::nanf("");
      }
    }
    if (normal) std::memcpy(normal + 3 * i, out, 12);
    if (uv) std::memcpy(uv + 2 * i, out + 3, 8);
    if (bary) std::memcpy(bary + 2 * i, out + 5, 8);
    if (hit_point) std::memcpy(hit_point + 3 * i, out + 7, 12);
  }
}

// ------------------------------------------------------------ Perlin table

// RPerlinNoise constructor, RayEngine/src/World/PerlinNoise.h:16-52, with the
// standard library's own engine and distributions.  std::bind stores a copy
// of the engine, so the gradient loop and the swap loop each draw from their
// own copy of the freshly seeded engine.  perm gets the doubled 512-entry
// table, grads (nullable) the 256 gradients.
ORC_API void orc_perlin_table(uint32_t seed, uint32_t* perm, float* grads) {
  const unsigned tableSize = 256, tableSizeMask = 255;
  std::mt19937 generator(seed);
  std::uniform_real_distribution<float> distribution;
  auto dice = std::bind(distribution, generator);
  uint32_t table[512];
  for (unsigned i = 0; i < tableSize; ++i) {
    double theta = std::acos(2 * dice() - 1);
    double phi = 2.f * dice() * 3.14159265358979323846;
    if (grads) {
      grads[3 * i] = (float)(std::cos(phi) * std::sin(theta));
      grads[3 * i + 1] = (float)(std::sin(phi) * std::sin(theta));
      grads[3 * i + 2] = (float)std::cos(theta);
    }
    table[i] = i;
  }
  std::uniform_int_distribution<unsigned> distributionInt;
  auto diceInt = std::bind(distributionInt, generator);
  for (unsigned i = 0; i < tableSize; ++i) std::swap(table[i], table[diceInt() & tableSizeMask]);
  for (unsigned i = 0; i < tableSize; ++i) table[tableSize + i] = table[i];
  std::memcpy(perm, table, sizeof(table));
}

// k-th (1-based) output of std::mt19937 seeded with seed; the C++ standard
// fixes the 10000th output of the default seed 5489 at 4123659995.
ORC_API uint32_t orc_mt19937_nth(uint32_t seed, int64_t k) {
  std::mt19937 g(seed);
  uint32_t v = 0;
  for (int64_t i = 0; i < k; ++i) v = g();
  return v;
}
