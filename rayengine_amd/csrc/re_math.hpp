// Shared FP32 vector math and the ray-query primitives of the rope kd-tree path.
//
// Every routine here is compiled twice: by g++ for the host library (scene
// preparation) and by hipcc for gfx950 kernels.  Both builds use
// -ffp-contract=off and IEEE division/sqrt, and every expression below fixes
// its evaluation order, so host and device produce bit-identical floats.
//
// Semantics follow the reference routines cited on each function; the vector
// operators are written from their mathematical definitions (dot is summed
// left to right, cross is the standard right-handed product, normalize is
// v * (1 / sqrt(dot(v, v)))).
#pragma once

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RE_HD __host__ __device__ __forceinline__
#else
#define RE_HD inline
#endif

namespace re {

// RayEngine/src/RayEngine/RayEngine.h:24-25
constexpr float kEps = 1e-4f;
constexpr float kInf = 1e20f;

struct f3 {
  float x, y, z;
};

RE_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
RE_HD f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
RE_HD f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
RE_HD f3 mul(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
RE_HD f3 smul(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
RE_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RE_HD f3 cross(f3 a, f3 b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
RE_HD f3 normalize(f3 v) {
  float inv = 1.0f / sqrtf(dot(v, v));
  return mul(v, inv);
}
RE_HD float get(f3 v, int axis) { return axis == 0 ? v.x : (axis == 1 ? v.y : v.z); }

// IEEE minNum/maxNum (a NaN operand yields the other operand), the behaviour
// of fminf/fmaxf that the slab test relies on for axis-parallel rays.
RE_HD float fmin_(float a, float b) { return fminf(a, b); }
RE_HD float fmax_(float a, float b) { return fmaxf(a, b); }

// Slab test.  RayEngine/src/cuda-src/ray_functions.cuh:61-92 (fsel:
// RayEngine/src/helper_math.h:35-39).  Returns false when the box is behind
// the ray or missed; t_near is clamped to >= 0.
RE_HD bool ray_box(const float* bmin, const float* bmax, f3 o, f3 d, float& t_near,
                   float& t_far) {
  float fx = 1.0f / d.x, fy = 1.0f / d.y, fz = 1.0f / d.z;
  float t1 = (bmin[0] - o.x) * fx;
  float t2 = (bmax[0] - o.x) * fx;
  float t3 = (bmin[1] - o.y) * fy;
  float t4 = (bmax[1] - o.y) * fy;
  float t5 = (bmin[2] - o.z) * fz;
  float t6 = (bmax[2] - o.z) * fz;
  float tmin = fmax_(fmax_(fmin_(t1, t2), fmin_(t3, t4)), fmin_(t5, t6));
  float tmax = fmin_(fmin_(fmax_(t1, t2), fmax_(t3, t4)), fmax_(t5, t6));
  if (tmax < 0.0f) return false;
  if (tmin > tmax) return false;
  t_near = tmin >= 0.0f ? tmin : 0.0f;
  t_far = tmax;
  return true;
}

// The same slab test with the direction reciprocals computed once per ray
// (fr = 1/d per component, bit-identical to forming them per call).
RE_HD bool ray_box_inv(const float* bmin, const float* bmax, f3 o, f3 fr, float& t_near, float& t_far) {
  float t1 = (bmin[0] - o.x) * fr.x;
  float t2 = (bmax[0] - o.x) * fr.x;
  float t3 = (bmin[1] - o.y) * fr.y;
  float t4 = (bmax[1] - o.y) * fr.y;
  float t5 = (bmin[2] - o.z) * fr.z;
  float t6 = (bmax[2] - o.z) * fr.z;
  float tmin = fmax_(fmax_(fmin_(t1, t2), fmin_(t3, t4)), fmin_(t5, t6));
  float tmax = fmin_(fmin_(fmax_(t1, t2), fmax_(t3, t4)), fmax_(t5, t6));
  if (tmax < 0.0f) return false;
  if (tmin > tmax) return false;
  t_near = tmin >= 0.0f ? tmin : 0.0f;
  t_far = tmax;
  return true;
}

// One-sided Moller-Trumbore.  RayEngine/src/cuda-src/ray_functions.cuh:173-223.
// det < kEps rejects back faces and near-parallel rays; a hit needs t > kEps.
RE_HD bool ray_tri(f3 o, f3 d, f3 v0, f3 e1, f3 e2, float& t_out) {
  f3 h = cross(d, e2);
  float a = dot(e1, h);
  if (a < kEps) return false;
  float f = 1.0f / a;
  f3 s = sub(o, v0);
  float u = f * dot(s, h);
  if (u < 0.0f || u > 1.0f) return false;
  f3 q = cross(s, e1);
  float v = f * dot(d, q);
  if (v < 0.0f || u + v > 1.0f) return false;
  float t = f * dot(e2, q);
  if (t > kEps) {
    t_out = t;
    return true;
  }
  return false;
}

// ray_tri / ray_tri_2s that also return the barycentrics (u, v) of the hit,
// the same operations in the same order (RayEngine/src/cuda-src/ray_functions.cuh:173-223).
RE_HD bool ray_tri_uv(f3 o, f3 d, f3 v0, f3 e1, f3 e2, bool two_sided, float& t_out, float& u_out, float& v_out) {
  f3 h = cross(d, e2);
  float a = dot(e1, h);
  if (two_sided ? (a > -kEps && a < kEps) : (a < kEps)) return false;
  float f = 1.0f / a;
  f3 s = sub(o, v0);
  float u = f * dot(s, h);
  if (u < 0.0f || u > 1.0f) return false;
  f3 q = cross(s, e1);
  float v = f * dot(d, q);
  if (v < 0.0f || u + v > 1.0f) return false;
  float t = f * dot(e2, q);
  if (t > kEps) {
    t_out = t;
    u_out = u;
    v_out = v;
    return true;
  }
  return false;
}

// Two-sided variant (|det| < kEps rejects), used when the one_sided
// trace parameter is 0.  Same operation order as ray_tri.
RE_HD bool ray_tri_2s(f3 o, f3 d, f3 v0, f3 e1, f3 e2, float& t_out) {
  f3 h = cross(d, e2);
  float a = dot(e1, h);
  if (a > -kEps && a < kEps) return false;
  float f = 1.0f / a;
  f3 s = sub(o, v0);
  float u = f * dot(s, h);
  if (u < 0.0f || u > 1.0f) return false;
  f3 q = cross(s, e1);
  float v = f * dot(d, q);
  if (v < 0.0f || u + v > 1.0f) return false;
  float t = f * dot(e2, q);
  if (t > kEps) {
    t_out = t;
    return true;
  }
  return false;
}

// Box faces in rope order.  RayEngine/src/RayEngine/RayEngine.h:134-142.
enum Face { LEFT = 0, FRONT = 1, RIGHT = 2, BACK = 3, TOP = 4, BOTTOM = 5 };

// Rope selection: first face (LEFT, FRONT, RIGHT, BACK, TOP, BOTTOM) whose
// plane lies within kEps of the exit point.
// RayEngine/src/cuda-src/kd_tree_functions.cuh:31-62.  Returns -1 for none.
RE_HD int exit_face(const float* bmin, const float* bmax, f3 p) {
  if (fabsf(p.x - bmin[0]) < kEps) return LEFT;
  if (fabsf(p.z - bmax[2]) < kEps) return FRONT;
  if (fabsf(p.x - bmax[0]) < kEps) return RIGHT;
  if (fabsf(p.z - bmin[2]) < kEps) return BACK;
  if (fabsf(p.y - bmax[1]) < kEps) return TOP;
  if (fabsf(p.y - bmin[1]) < kEps) return BOTTOM;
  return -1;
}

// ---------------------------------------------------------------------------
// Flattened rope kd-tree layout (shared by the host builder and the kernels).
//
// node[]  : 8-byte entries {float split; uint32 meta}.  meta>>30 is the split
//           axis (0,1,2) for an inner node, whose "below" child is the next
//           entry and whose "above" child is entry (meta & 0x3fffffff); 3 marks
//           a leaf entry whose low 30 bits are the leaf id.
// leaf[]  : 64-byte records {bmin[3], bmax[3], rope[6], first, count, pad[2]}.
//           rope values are node refs (below) and first indexes tri[].
// tri[]   : 40-byte records {v0[3], e1[3], e2[3], prim_id}, stored in leaf
//           order (triangles straddling a split appear in every leaf they
//           overlap), so a leaf's triangles are contiguous.
// refs    : int32.  ref >= 0 is a node entry index; (ref & 0x80000000) with
//           ref != -1 is a leaf id in the low 31 bits; -1 is "no neighbour".
// ---------------------------------------------------------------------------
constexpr uint32_t kLeafMeta = 3u << 30;
// Ceiling of the per-ray leaf-visit cap (a cap <= 0 means this ceiling).
constexpr int32_t kMaxLeafVisits = 1 << 20;
constexpr int32_t kNoRope = -1;
// Node entries are read kNodeWindow at a time (an inner node's below child is
// the next entry), so the uploaded entry array carries kNodeWindow - 1 or more
// trailing pad entries that a window starting at the last entry may read.
// Two entries (one 16-byte load per lane) measured fastest: windows of 1, 4
// and 8 entries took 2 %, 7 % and 18 % longer on the cluster workload.
constexpr int kNodeWindow = 2;
constexpr int kNodePad = 2;

struct NodeEntry {
  float split;
  uint32_t meta;
};

struct LeafRec {
  float bmin[3];
  float bmax[3];
  int32_t rope[6];
  uint32_t first;
  uint32_t count;
  uint32_t pad[2];
};

struct TriRec {
  float v0[3];
  float e1[3];
  float e2[3];
  int32_t prim;
};

// Device copy of a leaf record, fields reordered so that the box and the
// triangle run come first: a visit loads those 32 bytes, finds the exit face,
// and then loads the one rope it follows (4 bytes) instead of all 64 bytes.
// The vector-memory data path (64 bytes per clock per CU) is what the
// traversal saturates, so the bytes a wave loads set its speed.
struct DevLeaf {
  float bmin[3];
  float bmax[3];
  uint32_t first;
  uint32_t count;
  int32_t rope[6];
  uint32_t pad[2];  // pad[0]: the leaf's object index in its scene (set at upload)
};

RE_HD DevLeaf dev_leaf(const LeafRec& l) {
  DevLeaf d;
  for (int a = 0; a < 3; ++a) {
    d.bmin[a] = l.bmin[a];
    d.bmax[a] = l.bmax[a];
  }
  d.first = l.first;
  d.count = l.count;
  for (int f = 0; f < 6; ++f) d.rope[f] = l.rope[f];
  d.pad[0] = d.pad[1] = 0;
  return d;
}

static_assert(sizeof(NodeEntry) == 8, "node entry is 8 bytes");
static_assert(sizeof(LeafRec) == 64, "leaf record is 64 bytes");
static_assert(sizeof(DevLeaf) == 64, "device leaf record is 64 bytes");
static_assert(sizeof(TriRec) == 40, "triangle record is 40 bytes");

RE_HD bool ref_is_leaf(int32_t r) { return r < -1; }
RE_HD int32_t leaf_ref(uint32_t leaf_id) { return (int32_t)(0x80000000u | leaf_id); }
RE_HD uint32_t ref_leaf_id(int32_t r) { return (uint32_t)r & 0x7fffffffu; }

// Per-object transform applied before traversal.
// RayEngine/src/cuda-src/ray_functions.cuh:238-267: rotate about Y by the
// object's rotation.y (sin/cos supplied by the host), translate the origin by
// the object's location unless it is a character, then renormalise d.
struct ObjXform {
  float loc[3];
  float sin_y, cos_y;
  int32_t is_character;
  int32_t root;  // node entry index of the object's root
  float root_min[3];
  float root_max[3];
};

RE_HD void xform_ray(const ObjXform& ob, f3 o, f3 d, bool secondary, f3& no, f3& nd) {
  float s = ob.sin_y, c = ob.cos_y;
  nd.x = d.x * c + d.z * s;
  nd.y = d.y;
  nd.z = (-d.x) * s + d.z * c;
  no = f3{0.0f, 0.0f, 0.0f};
  if (!ob.is_character) {
    no.x = o.x * c + o.z * s;
    no.y = o.y;
    no.z = (-o.x) * s + o.z * c;
    no = add(no, f3{ob.loc[0], ob.loc[1], ob.loc[2]});
  }
  if (ob.is_character && secondary) {
    // The reference recomputes sin/cos of -rotation.y here; sin(-a) = -sin(a)
    // and cos(-a) = cos(a) hold exactly for IEEE sinf/cosf.
    float s2 = -s, c2 = c;
    nd.x = d.x * c2 + d.z * s2;
    nd.y = d.y;
    nd.z = (-d.x) * s2 + d.z * c2;
    no.x = ob.loc[0] * c2 + ob.loc[2] * s2;
    no.y = -ob.loc[1];
    no.z = (-ob.loc[0]) * s2 + ob.loc[2] * c2;
  }
  nd = normalize(nd);
}

}  // namespace re
