// rayengine_compat.hpp - reference-named host API over the C ABI.
//
// Header-only C++17.  Host code written against hadalhw17/RayEngine's scene,
// kd-tree and render entry points keeps its call sites: the names, argument
// meaning and object model below follow the reference (file:line on each),
// and every call lands in librayengine.so (include/rayengine.h).  Device
// state lives in one process-wide context on device 0, as the reference's
// TU-global state did (RayEngine/src/cuda-src/CUDARayTracing.cuh:15-85);
// unlike the reference, failures throw rayengine_compat::Error instead of
// calling exit() (RayEngine/src/RayEngine/RayEngine.h:56-63).
#pragma once

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "rayengine.h"

namespace rayengine_compat {

struct Error : std::runtime_error {
  re_status status;
  Error(re_status st, const std::string& m) : std::runtime_error(m), status(st) {}
};

inline void check(re_status st, const re_ctx* ctx = nullptr) {
  if (st != RE_OK) throw Error(st, re_last_error(ctx));
}

// Host vector types of the reference's interface (CUDA vector_types.h).
struct float2 { float x, y; };
struct float3 { float x, y, z; };
struct uint3 { uint32_t x, y, z; };

// RCamera, RayEngine/src/Primitives/Camera.h:7-16 (fields the rays use).
struct RCamera {
  float3 campos{0, 0, 0};
  float3 view{0, 0, -1};
  float3 camdown{0, -1, 0};
  float2 fov{60, 60};
  float focial_distance = 4.0f;
};

// GPUSceneObject, RayEngine/src/cuda-src/gpu_structs.h:69-90 (transform
// fields; the material stays with the caller).
struct GPUSceneObject {
  size_t index_of_first_prim = 0;
  size_t num_prims = 0;
  float3 location{0, 0, 0};
  float3 rotation{0, 0, 0};
  bool is_character = false;
};

// HitResult, RayEngine/src/cuda-src/gpu_structs.h:92-111, with the
// primitive id the reference does not return.
struct HitResult {
  float t = 1e20f;
  int prim_id = -1;
  int obj_index = -1;
  bool hits = false;
};

// RKDTreeCPU, RayEngine/src/Primitives/KDTree.h:65: built from an indexed
// mesh.  The tree is the SAH rope kd-tree of re_kd_build.
class RKDTreeCPU {
 public:
  RKDTreeCPU(const float3* verts, const uint3* faces, const float3* /*norms*/, const float2* /*uvs*/, int num_verts,
             int num_faces, int /*num_norms*/, int /*num_uvs*/, const re_kd_build_params* params = nullptr) {
    soup_.resize((size_t)num_faces * 9);
    for (int f = 0; f < num_faces; ++f) {
      const uint32_t idx[3] = {faces[f].x, faces[f].y, faces[f].z};
      for (int k = 0; k < 3; ++k) {
        if ((int)idx[k] >= num_verts) throw Error(RE_ERR_INVALID, "RKDTreeCPU: face index out of range");
        const float3& v = verts[idx[k]];
        soup_[9 * f + 3 * k] = v.x;
        soup_[9 * f + 3 * k + 1] = v.y;
        soup_[9 * f + 3 * k + 2] = v.z;
      }
    }
    check(re_kd_build(soup_.data(), num_faces, params, &tree_));
  }
  RKDTreeCPU(const RKDTreeCPU&) = delete;
  RKDTreeCPU& operator=(const RKDTreeCPU&) = delete;
  ~RKDTreeCPU() { re_kd_free(tree_); }
  const re_kdtree* handle() const { return tree_; }
  re_kd_info info() const {
    re_kd_info i;
    check(re_kd_get_info(tree_, &i));
    return i;
  }
  int num_faces() const { return (int)(soup_.size() / 9); }

 private:
  std::vector<float> soup_;
  re_kdtree* tree_ = nullptr;
};

// RKDThreeGPU (RayEngine/src/Primitives/KDThreeGPU.h:9) flattened the CPU
// tree for upload; re_kd_build already produces the flattened arrays.
using RKDThreeGPU = RKDTreeCPU;

namespace detail {
struct State {
  re_ctx* ctx = nullptr;
  int32_t scene = -1;
  int32_t volume = -1;
  std::vector<const RKDTreeCPU*> trees;
  std::vector<int32_t> prim_offsets;
  RCamera camera;
  ~State() {
    if (ctx) re_destroy(ctx);
  }
};
inline State& state() {
  static State s;
  return s;
}
inline re_ctx* ctx() {
  State& s = state();
  if (!s.ctx) check(re_create(0, &s.ctx));
  return s.ctx;
}
inline std::vector<re_object> objects(const std::vector<GPUSceneObject>& objs) {
  State& s = state();
  if (objs.size() != s.trees.size()) throw Error(RE_ERR_INVALID, "one GPUSceneObject per tree");
  std::vector<re_object> o(objs.size());
  for (size_t k = 0; k < objs.size(); ++k) {
    o[k] = re_object{};
    o[k].tree = s.trees[k]->handle();
    o[k].location[0] = objs[k].location.x;
    o[k].location[1] = objs[k].location.y;
    o[k].location[2] = objs[k].location.z;
    // sin/cos in double, rounded once (the Python host mirror does the same).
    o[k].sin_rot_y = (float)std::sin((double)objs[k].rotation.y);
    o[k].cos_rot_y = (float)std::cos((double)objs[k].rotation.y);
    o[k].is_character = objs[k].is_character ? 1 : 0;
    o[k].prim_offset = s.prim_offsets[k];
  }
  return o;
}
}  // namespace detail

// copy_memory, RayEngine/src/cuda-src/cuda_memory_functions.cuh:58-211:
// uploads every object's tree (triangle ids offset by the preceding objects'
// triangle counts, RTScene.cpp:263-279) and keeps the camera.
inline void copy_memory(const std::vector<RKDThreeGPU*>& trees, const RCamera& cam,
                        const std::vector<GPUSceneObject>& objs) {
  detail::State& s = detail::state();
  re_ctx* c = detail::ctx();
  if (s.scene >= 0) check(re_scene_release(c, s.scene), c);
  s.scene = -1;
  s.trees.assign(trees.begin(), trees.end());
  s.prim_offsets.clear();
  int32_t off = 0;
  for (const RKDThreeGPU* t : trees) {
    s.prim_offsets.push_back(off);
    off += t->num_faces();
  }
  std::vector<re_object> o = detail::objects(objs);
  check(re_scene_upload(c, o.data(), (int32_t)o.size(), &s.scene), c);
  s.camera = cam;
}

// update_objects, RayEngine/src/cuda-src/cuda_memory_functions.cuh:50-56.
inline void update_objects(const std::vector<GPUSceneObject>& objs) {
  detail::State& s = detail::state();
  if (s.scene < 0) throw Error(RE_ERR_INVALID, "update_objects before copy_memory");
  std::vector<re_object> o = detail::objects(objs);
  check(re_scene_update_objects(s.ctx, s.scene, o.data(), (int32_t)o.size()), s.ctx);
}

// trace_primary_rays, RayEngine/src/cuda-src/renderer.cu:389-410: one
// closest-hit query per pixel of the camera, pixel (x, y) at y * width + x.
inline void trace_primary_rays(HitResult* hit_results, const RCamera& cam, uint32_t width, uint32_t height,
                               int max_leaf_visits = 50) {
  detail::State& s = detail::state();
  if (s.scene < 0) throw Error(RE_ERR_INVALID, "trace_primary_rays before copy_memory");
  re_camera rc;
  rc.pos[0] = cam.campos.x; rc.pos[1] = cam.campos.y; rc.pos[2] = cam.campos.z;
  rc.view[0] = cam.view.x; rc.view[1] = cam.view.y; rc.view[2] = cam.view.z;
  rc.camdown[0] = cam.camdown.x; rc.camdown[1] = cam.camdown.y; rc.camdown[2] = cam.camdown.z;
  rc.fov[0] = cam.fov.x; rc.fov[1] = cam.fov.y;
  rc.focal_distance = cam.focial_distance;
  const size_t n = (size_t)width * height;
  std::vector<float> o(3 * n), d(3 * n), t(n);
  std::vector<int32_t> p(n), ob(n);
  check(re_gen_camera_rays(&rc, (int)width, (int)height, o.data(), d.data()));
  re_trace_params tp;
  re_trace_default_params(&tp);
  tp.max_leaf_visits = max_leaf_visits;
  check(re_trace_closest(s.ctx, s.scene, o.data(), d.data(), (int64_t)n, &tp, p.data(), t.data(), ob.data(), 0),
        s.ctx);
  for (size_t i = 0; i < n; ++i) hit_results[i] = HitResult{t[i], p[i], ob[i], p[i] >= 0};
}

// generate_noise / cuda_update_chunk_gen, RayEngine/src/cuda-src/renderer.cu:812-836,982-1005:
// (re)generates the terrain volume of the given resolution and permutation
// seed (RPerlinNoise's default 2016).
inline void generate_noise(int dim = 256, uint32_t seed = 2016) {
  detail::State& s = detail::state();
  re_ctx* c = detail::ctx();
  if (s.volume >= 0) check(re_volume_release(c, s.volume), c);
  s.volume = -1;
  re_noise_params p;
  re_noise_default_params(&p);
  for (int a = 0; a < 3; ++a) {
    p.dim[a] = dim;
    p.spacing[a] = 1.0f / (float)dim;
  }
  p.seed = seed;
  check(re_noise_generate(c, &p, &s.volume), c);
}

// The trace part of render_sphere_trace, RayEngine/src/cuda-src/renderer.cu:142-235.
inline void sphere_trace_primary_rays(HitResult* hit_results, const RCamera& cam, uint32_t width, uint32_t height) {
  detail::State& s = detail::state();
  if (s.volume < 0) throw Error(RE_ERR_INVALID, "sphere trace before generate_noise");
  re_camera rc;
  rc.pos[0] = cam.campos.x; rc.pos[1] = cam.campos.y; rc.pos[2] = cam.campos.z;
  rc.view[0] = cam.view.x; rc.view[1] = cam.view.y; rc.view[2] = cam.view.z;
  rc.camdown[0] = cam.camdown.x; rc.camdown[1] = cam.camdown.y; rc.camdown[2] = cam.camdown.z;
  rc.fov[0] = cam.fov.x; rc.fov[1] = cam.fov.y;
  rc.focal_distance = cam.focial_distance;
  const size_t n = (size_t)width * height;
  std::vector<float> o(3 * n), d(3 * n), t(n);
  std::vector<uint8_t> h(n);
  check(re_gen_camera_rays(&rc, (int)width, (int)height, o.data(), d.data()));
  check(re_sdf_trace(s.ctx, s.volume, nullptr, o.data(), d.data(), (int64_t)n, t.data(), h.data(), nullptr, nullptr, 0),
        s.ctx);
  for (size_t i = 0; i < n; ++i) hit_results[i] = HitResult{t[i], h[i] ? 0 : -1, h[i] ? 0 : -1, h[i] != 0};
}

// trace_shadow, RayEngine/src/cuda-src/ray_functions.cuh:365-374, for a
// batch of rays against the uploaded scene: occluded[i] = 1 when some object
// blocks ray i (traversed with is_secondary = true, as trace_shadow does).
inline void trace_shadow_rays(uint8_t* occluded, const float3* origins, const float3* dirs, size_t n,
                              int max_leaf_visits = 50) {
  detail::State& s = detail::state();
  if (s.scene < 0) throw Error(RE_ERR_INVALID, "trace_shadow_rays before copy_memory");
  std::vector<float> o(3 * n), d(3 * n);
  for (size_t i = 0; i < n; ++i) {
    o[3 * i] = origins[i].x; o[3 * i + 1] = origins[i].y; o[3 * i + 2] = origins[i].z;
    d[3 * i] = dirs[i].x; d[3 * i + 1] = dirs[i].y; d[3 * i + 2] = dirs[i].z;
  }
  re_trace_params tp;
  re_trace_default_params(&tp);
  tp.max_leaf_visits = max_leaf_visits;
  tp.secondary = 1;
  check(re_trace_occluded(s.ctx, s.scene, o.data(), d.data(), (int64_t)n, &tp, occluded, 0), s.ctx);
}

// TerrainBrushType / TerrainBrush, RayEngine/src/RayEngine/RayEngine.h:77-90.
enum TerrainBrushType { SPHERE_ADD = 0, SPHERE_SUBTRACT = 1, CUBE_ADD = 2, CUBE_SUBTRACT = 3 };
struct int3 { int x, y, z; };
struct TerrainBrush {
  TerrainBrushType brush_type;
  int3 brush_extent;
  int material_index;
  bool snap_to_grid;
};

// insert_sphere_to_texture, RayEngine/src/cuda-src/renderer.cu:30-92: the
// brush ray is sphere traced (the reference's clamp step rule and 1e-5
// threshold) and, on a hit, the brush is applied around the hit point.
// Returns whether the ray hit.
inline bool insert_sphere_to_texture(const TerrainBrush& brush, float3 ray_o, float3 ray_dir) {
  detail::State& s = detail::state();
  if (s.volume < 0) throw Error(RE_ERR_INVALID, "brush before generate_noise / load_map");
  const float o[3] = {ray_o.x, ray_o.y, ray_o.z}, d[3] = {ray_dir.x, ray_dir.y, ray_dir.z};
  re_sdf_params sp;
  re_sdf_default_params(&sp);
  sp.step_rule = 1;
  sp.eps = 10e-6f;
  float t = 0.0f;
  uint8_t hit = 0;
  check(re_sdf_trace(s.ctx, s.volume, &sp, o, d, 1, &t, &hit, nullptr, nullptr, 0), s.ctx);
  if (!hit) return false;
  re_brush b;
  re_brush_default(&b);
  b.type = (int32_t)brush.brush_type;
  for (int a = 0; a < 3; ++a) b.center[a] = o[a] + t * d[a];
  b.extent[0] = (float)brush.brush_extent.x;
  b.extent[1] = (float)brush.brush_extent.y;
  b.extent[2] = (float)brush.brush_extent.z;
  b.material = (float)brush.material_index;
  b.snap_to_grid = brush.snap_to_grid ? 1 : 0;
  check(re_volume_edit(s.ctx, s.volume, &b), s.ctx);
  return true;
}

// save_map / load_map, RayEngine/src/cuda-src/renderer.cu:838-947, with the
// file name given by the caller (the reference derives it from the chunk).
inline void save_map(const std::string& path) {
  detail::State& s = detail::state();
  if (s.volume < 0) throw Error(RE_ERR_INVALID, "save_map without a volume");
  check(re_volume_save_rsdf(s.ctx, s.volume, path.c_str()), s.ctx);
}
inline void load_map(const std::string& path) {
  detail::State& s = detail::state();
  re_ctx* c = detail::ctx();
  int32_t v = -1;
  check(re_volume_load_rsdf(c, path.c_str(), &v), c);
  if (s.volume >= 0) check(re_volume_release(c, s.volume), c);
  s.volume = v;
}

// The camera cuda_render_frame renders (the reference keeps it in device
// state set by initialize_volume_render, sphere_tracing.cuh:617-755).
inline void set_camera(const RCamera& cam) {
  (void)detail::ctx();
  detail::state().camera = cam;
}

// cuda_render_frame, RayEngine/src/cuda-src/renderer.cu:701-738, headless:
// sphere traces the volume for the camera (with normals) and colours the
// frame (sky for misses, the normal material for hits, RGBA8 with R in the
// low byte) into output, width * height pixels.
inline void cuda_render_frame(uint32_t* output, const uint32_t& w, const uint32_t& h) {
  detail::State& s = detail::state();
  if (s.volume < 0) throw Error(RE_ERR_INVALID, "cuda_render_frame before generate_noise / load_map");
  const RCamera& cam = s.camera;
  re_camera rc;
  rc.pos[0] = cam.campos.x; rc.pos[1] = cam.campos.y; rc.pos[2] = cam.campos.z;
  rc.view[0] = cam.view.x; rc.view[1] = cam.view.y; rc.view[2] = cam.view.z;
  rc.camdown[0] = cam.camdown.x; rc.camdown[1] = cam.camdown.y; rc.camdown[2] = cam.camdown.z;
  rc.fov[0] = cam.fov.x; rc.fov[1] = cam.fov.y;
  rc.focal_distance = cam.focial_distance;
  const size_t n = (size_t)w * h;
  std::vector<float> o(3 * n), d(3 * n), t(n), nrm(3 * n);
  std::vector<uint8_t> hit(n);
  check(re_gen_camera_rays(&rc, (int)w, (int)h, o.data(), d.data()));
  check(re_sdf_trace(s.ctx, s.volume, nullptr, o.data(), d.data(), (int64_t)n, t.data(), hit.data(), nullptr,
                     nrm.data(), 0), s.ctx);
  check(re_frame_shade(s.ctx, d.data(), hit.data(), nrm.data(), (int32_t)w, (int32_t)h, nullptr, output, 0), s.ctx);
}

// free_memory, RayEngine/src/cuda-src/sphere_tracing.cuh:757-767.
inline void free_memory() {
  detail::State& s = detail::state();
  if (s.ctx) re_destroy(s.ctx);
  s.ctx = nullptr;
  s.scene = -1;
  s.volume = -1;
  s.trees.clear();
}

}  // namespace rayengine_compat
