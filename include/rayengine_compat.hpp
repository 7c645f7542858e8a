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
#include <cstring>
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
struct float4 { float x, y, z, w; };
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
// primitive id the reference does not return and without its shading colour.
// normal, uv and hit_point are filled when copy_memory was given vertex
// normals (ray_functions.cuh:146-166,204-213): hit_point in the hit object's
// space, as the reference computes it; ray_o / ray_dir are the world ray.
struct HitResult {
  float t = 1e20f;
  float3 normal{0, 0, 0};
  float3 hit_point{0, 0, 0};
  float3 ray_dir{0, 0, 0};
  float3 ray_o{0, 0, 0};
  float2 uv{0, 0};
  bool hits = false;
  int obj_index = -1;
  int prim_id = -1;
};

// RenderingSettings / SceneSettings, RayEngine/src/RayEngine/RayEngine.h:93-124.
enum RenderQuality { LOW = 0, MEDIUM = 1, HIGH = 2 };
struct RenderingSettings {
  float texture_scale = 1;
  RenderQuality quality = MEDIUM;
  bool gamma = false;
  bool vignetting = false;
  float vignetting_k = 0.25f;
  bool gravity = false;
  bool god_mode = false;
};
struct SceneSettings {
  float3 light_pos{0, 25, 0};
  int soft_shadow_k = 32;
  float light_intensity = 1;
  bool enable_fog = false;
  float fog_deisity = 0;
  float noise_freuency = 0;
  float noise_amplitude = 0;
  float noise_redistrebution = 0;
  int terracing = 1;
  bool should_terrace = false;
  uint3 volume_resolution{256, 256, 256};
  float3 world_size{1, 1, 1};
};

// VoxelMaterial, RayEngine/src/Engine/Material.h:5: textures of the shading
// path; initialize_volume_render accepts and ignores them.
struct VoxelMaterial {};

// GridNode / Grid, RayEngine/src/World/Grid.h:11-60 (the fields the volume
// upload reads): voxels x fastest, index x + dim.x * (y + dim.y * z).
struct GridNode {
  float3 point{0, 0, 0};
  float distance = 9999999999.f;
  float material = 0;
};
struct Grid {
  std::vector<GridNode> voxels;
  float3 spacing{0, 0, 0};
  uint3 sdf_dim{0, 0, 0};
  float3 box_max{0, 0, 0};
};

namespace RayEngine {
// RChunk, RayEngine/src/World/Chunk.h:8-25: a grid and its world location.
class RChunk {
 public:
  RChunk() = default;
  explicit RChunk(const Grid& g) : distance_field(g) {}
  const Grid& get_sdf() const { return distance_field; }
  const float3& get_location() const { return location; }
  void set_location(float3 l) { location = l; }
  void update_sdf_box(const float3& new_box_max) { distance_field.box_max = new_box_max; }

 private:
  Grid distance_field;
  float3 location{0, 0, 0};
};

// RPerlinNoise, RayEngine/src/World/PerlinNoise.h:13-61: the doubled
// 512-entry permutation table of a seed, from the library's portable
// mt19937 shuffle (re_perlin_table).  The reference's gradient table is not
// read by its noise primitive (sdf_functions.cuh:83-135) and is left out.
class RPerlinNoise {
 public:
  static const unsigned tableSize = 256;
  static const unsigned tableSizeMask = tableSize - 1;
  explicit RPerlinNoise(const unsigned& seed = 2016) : seed_(seed) { check(re_perlin_table(seed, permutationTable)); }
  virtual ~RPerlinNoise() = default;
  unsigned seed() const { return seed_; }
  uint32_t permutationTable[tableSize * 2];

 private:
  unsigned seed_;
};
}  // namespace RayEngine

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

// RKDThreeGPU, RayEngine/src/Primitives/KDThreeGPU.h:9: the reference
// flattened the CPU tree into upload arrays here (KDThreeGPU.cpp:155-194).
// re_kd_build already produced them, so this view keeps the CPU tree (which
// must outlive it) and answers the reference's getters from its arrays.
class RKDThreeGPU {
 public:
  explicit RKDThreeGPU(RKDTreeCPU* cpu) : cpu_(cpu) {
    if (!cpu) throw Error(RE_ERR_INVALID, "RKDThreeGPU: null tree");
  }
  const RKDTreeCPU* cpu() const { return cpu_; }
  int num_faces() const { return cpu_->num_faces(); }
  size_t get_num_faces() const { return (size_t)cpu_->num_faces(); }
  size_t GetNumNodes() const { return (size_t)cpu_->info().n_nodes; }
  size_t get_root_index() const { return 0; }  // the root is entry 0 of the preorder array
  // The triangle ids in leaf order (the reference's obj_index_list): one per
  // triangle record, straddler copies included.
  std::vector<int> GetIndexList() const {
    const re_kd_info i = cpu_->info();
    std::vector<unsigned char> recs((size_t)i.n_trirecs * 40);
    check(re_kd_export(cpu_->handle(), nullptr, nullptr, recs.data()));
    std::vector<int> ids((size_t)i.n_trirecs);
    for (size_t k = 0; k < ids.size(); ++k) std::memcpy(&ids[k], recs.data() + 40 * k + 36, sizeof(int));
    return ids;
  }

 private:
  RKDTreeCPU* cpu_;
};

namespace detail {
struct State {
  re_ctx* ctx = nullptr;
  int32_t scene = -1;
  int32_t volume = -1;
  std::vector<const RKDTreeCPU*> trees;
  std::vector<int32_t> prim_offsets;
  int32_t n_prims = 0;
  bool attributes = false;  // copy_memory was given vertex normals
  RCamera camera;
  RenderingSettings render_settings;
  SceneSettings scene_settings;
  float3 chunk_location{0, 0, 0};  // the volume instance's location
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
inline re_camera camera(const RCamera& cam) {
  re_camera rc;
  rc.pos[0] = cam.campos.x; rc.pos[1] = cam.campos.y; rc.pos[2] = cam.campos.z;
  rc.view[0] = cam.view.x; rc.view[1] = cam.view.y; rc.view[2] = cam.view.z;
  rc.camdown[0] = cam.camdown.x; rc.camdown[1] = cam.camdown.y; rc.camdown[2] = cam.camdown.z;
  rc.fov[0] = cam.fov.x; rc.fov[1] = cam.fov.y;
  rc.focal_distance = cam.focial_distance;
  return rc;
}
// Camera rays of a width x height frame taken into the volume's frame: the
// origin minus the chunk location (hit_result.ray_o -= instances[0].location,
// RayEngine/src/cuda-src/renderer.cu:185).
inline void volume_rays(const RCamera& cam, uint32_t w, uint32_t h, std::vector<float>& o, std::vector<float>& d) {
  const size_t n = (size_t)w * h;
  o.resize(3 * n);
  d.resize(3 * n);
  re_camera rc = camera(cam);
  check(re_gen_camera_rays(&rc, (int)w, (int)h, o.data(), d.data()));
  const float3 L = state().chunk_location;
  for (size_t i = 0; i < n; ++i) {
    o[3 * i] -= L.x;
    o[3 * i + 1] -= L.y;
    o[3 * i + 2] -= L.z;
  }
}
inline void replace_volume(int32_t v) {
  State& s = state();
  if (s.volume >= 0) check(re_volume_release(s.ctx, s.volume), s.ctx);
  s.volume = v;
}
// The chunk's grid as one volume with its materials (cuda_update_chunk's
// upload, RayEngine/src/cuda-src/sphere_tracing.cuh:565-615).
inline void upload_chunk(const RayEngine::RChunk& chunk) {
  re_ctx* c = ctx();
  const Grid& g = chunk.get_sdf();
  const int32_t dim[3] = {(int32_t)g.sdf_dim.x, (int32_t)g.sdf_dim.y, (int32_t)g.sdf_dim.z};
  const float sp[3] = {g.spacing.x, g.spacing.y, g.spacing.z};
  const size_t n = (size_t)g.sdf_dim.x * g.sdf_dim.y * g.sdf_dim.z;
  if (g.voxels.size() != n) throw Error(RE_ERR_INVALID, "chunk grid: voxels.size() != sdf_dim.x * sdf_dim.y * sdf_dim.z");
  std::vector<float> pairs(2 * n);
  for (size_t i = 0; i < n; ++i) {
    pairs[2 * i] = g.voxels[i].distance;
    pairs[2 * i + 1] = g.voxels[i].material;
  }
  int32_t v = -1;
  check(re_volume_upload_pairs(c, dim, sp, pairs.data(), &v), c);
  replace_volume(v);
  state().chunk_location = chunk.get_location();
}
// The terrain volume of the scene settings at a field origin: resolution
// volume_resolution, spacing world_size / volume_resolution per axis
// (renderer.cu:815-822).  The field is the library's 3D fBm terrain
// (re_noise_generate); the reference's noise_freuency / noise_amplitude /
// noise_redistrebution / terracing shape its 2D heightfield, which that
// field replaces, and are not read.
inline void generate_volume(const float3& origin, uint32_t seed) {
  State& s = state();
  re_ctx* c = ctx();
  re_noise_params p;
  re_noise_default_params(&p);
  const uint32_t r[3] = {s.scene_settings.volume_resolution.x, s.scene_settings.volume_resolution.y,
                         s.scene_settings.volume_resolution.z};
  const float ws[3] = {s.scene_settings.world_size.x, s.scene_settings.world_size.y, s.scene_settings.world_size.z};
  for (int a = 0; a < 3; ++a) {
    p.dim[a] = (int32_t)r[a];
    p.spacing[a] = ws[a] / (float)r[a];
  }
  p.origin[0] = origin.x;
  p.origin[1] = origin.y;
  p.origin[2] = origin.z;
  p.seed = seed;
  int32_t v = -1;
  check(re_noise_generate(c, &p, &v), c);
  replace_volume(v);
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
  s.attributes = false;
  s.trees.clear();
  for (const RKDThreeGPU* t : trees) s.trees.push_back(t->cpu());
  s.prim_offsets.clear();
  int32_t off = 0;
  for (const RKDThreeGPU* t : trees) {
    s.prim_offsets.push_back(off);
    off += t->num_faces();
  }
  s.n_prims = off;
  std::vector<re_object> o = detail::objects(objs);
  check(re_scene_upload(c, o.data(), (int32_t)o.size(), &s.scene), c);
  s.camera = cam;
}

// copy_memory in the reference's full call shape
// (RayEngine/src/cuda-src/cuda_memory_functions.cuh:58-62, packed by
// RTScene::init_triangles, RTScene.cpp:228-279): three entries per triangle
// in scene prim order.  The trees carry their triangle records, so
// triangles (v0, e1, e2) is checked for its size only; normals (n0, n1, n2)
// and uvs (uv0, uv1, uv2, or empty) become the hit attributes that
// trace_primary_rays returns.  textures feed the shading path and are not
// read.  bruteforce = true asked for the reference's per-pixel loop over all
// triangles (gpu_bruteforce_ray_cast, renderer.cu:326-386), a debugging path
// outside the ray queries: it throws RE_ERR_UNSUPPORTED.
inline void copy_memory(std::vector<RKDThreeGPU*> tree, RCamera sceneCam, std::vector<float4> triangles,
                        std::vector<float4> normals, std::vector<float2> uvs, std::vector<GPUSceneObject> objs,
                        std::vector<float3> textures, bool bruteforce = false) {
  (void)textures;
  if (bruteforce) throw Error(RE_ERR_UNSUPPORTED, "copy_memory: the brute-force ray caster is not provided");
  int64_t total = 0;
  for (const RKDThreeGPU* t : tree) total += t->num_faces();
  if ((int64_t)triangles.size() != 3 * total)
    throw Error(RE_ERR_INVALID, "copy_memory: triangles must hold 3 entries (v0, e1, e2) per triangle");
  if (!normals.empty() && (int64_t)normals.size() != 3 * total)
    throw Error(RE_ERR_INVALID, "copy_memory: normals must hold 3 entries per triangle");
  if (!uvs.empty() && (int64_t)uvs.size() != 3 * total)
    throw Error(RE_ERR_INVALID, "copy_memory: uvs must hold 3 entries per triangle or none");
  copy_memory(tree, sceneCam, objs);
  if (normals.empty()) return;
  detail::State& s = detail::state();
  std::vector<float> nr((size_t)total * 9), uv(uvs.empty() ? 0 : (size_t)total * 6);
  for (int64_t i = 0; i < 3 * total; ++i) {
    nr[3 * i] = normals[i].x;
    nr[3 * i + 1] = normals[i].y;
    nr[3 * i + 2] = normals[i].z;
    if (!uvs.empty()) {
      uv[2 * i] = uvs[i].x;
      uv[2 * i + 1] = uvs[i].y;
    }
  }
  check(re_scene_set_attributes(s.ctx, s.scene, nr.data(), uvs.empty() ? nullptr : uv.data(), total), s.ctx);
  s.attributes = true;
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
  re_camera rc = detail::camera(cam);
  const size_t n = (size_t)width * height;
  std::vector<float> o(3 * n), d(3 * n), t(n);
  std::vector<int32_t> p(n), ob(n);
  check(re_gen_camera_rays(&rc, (int)width, (int)height, o.data(), d.data()));
  re_trace_params tp;
  re_trace_default_params(&tp);
  tp.max_leaf_visits = max_leaf_visits;
  check(re_trace_closest(s.ctx, s.scene, o.data(), d.data(), (int64_t)n, &tp, p.data(), t.data(), ob.data(), 0),
        s.ctx);
  std::vector<float> nrm, uv, hp;
  if (s.attributes) {
    nrm.resize(3 * n);
    uv.resize(2 * n);
    hp.resize(3 * n);
    check(re_hit_attributes(s.ctx, s.scene, o.data(), d.data(), (int64_t)n, &tp, p.data(), ob.data(), nrm.data(),
                            uv.data(), nullptr, hp.data(), 0),
          s.ctx);
  }
  for (size_t i = 0; i < n; ++i) {
    HitResult r;
    r.t = t[i];
    r.prim_id = p[i];
    r.obj_index = ob[i];
    r.hits = p[i] >= 0;
    r.ray_o = float3{o[3 * i], o[3 * i + 1], o[3 * i + 2]};
    r.ray_dir = float3{d[3 * i], d[3 * i + 1], d[3 * i + 2]};
    if (s.attributes) {
      r.normal = float3{nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]};
      r.uv = float2{uv[2 * i], uv[2 * i + 1]};
      r.hit_point = float3{hp[3 * i], hp[3 * i + 1], hp[3 * i + 2]};
    }
    hit_results[i] = r;
  }
}

// initialize_volume_render, RayEngine/src/cuda-src/sphere_tracing.cuh:617-755
// (caller SDFScene::init_cuda_res, SDFScene.cpp:42-46): keeps the camera and
// both settings and uploads the chunk's grid with its materials, placed at
// the chunk's location.  The reference reads only the chunk's one grid, so
// num_sdf must be 1.  The materials' textures belong to the shading path and
// are not read; the noise table is the one re_noise_generate derives from
// the seed when a chunk is generated.
inline void initialize_volume_render(RCamera& sceneCam, const RayEngine::RChunk& world_chunk, const int& num_sdf,
                                     const std::vector<VoxelMaterial>& materials,
                                     const RenderingSettings& render_settings, const SceneSettings& scene_settings,
                                     const RayEngine::RPerlinNoise& noise) {
  (void)materials;
  (void)noise;
  if (num_sdf != 1) throw Error(RE_ERR_INVALID, "initialize_volume_render: num_sdf must be 1");
  detail::State& s = detail::state();
  (void)detail::ctx();
  s.camera = sceneCam;
  s.render_settings = render_settings;
  s.scene_settings = scene_settings;
  detail::upload_chunk(world_chunk);
}

// cuda_update_chunk, RayEngine/src/cuda-src/sphere_tracing.cuh:565-615
// (SDFScene::update_chunk, SDFScene.cpp:48-53): replaces the volume with the
// chunk's grid at the chunk's location.
inline void cuda_update_chunk(const RayEngine::RChunk& world_chunk, const RayEngine::RPerlinNoise& noise) {
  (void)noise;
  detail::upload_chunk(world_chunk);
}

// cuda_update_chunk_gen, RayEngine/src/cuda-src/renderer.cu:982-1005
// (SDFScene::generate_chunk, SDFScene.cpp:55-59): generates the chunk's
// terrain with the noise's permutation seed, the field sampled at the chunk's
// location plus each voxel's offset, so neighbouring chunks continue one
// another; the volume is placed at that location.
inline void cuda_update_chunk_gen(const RayEngine::RChunk& world_chunk, const RayEngine::RPerlinNoise& noise) {
  detail::generate_volume(world_chunk.get_location(), noise.seed());
  detail::state().chunk_location = world_chunk.get_location();
}

// generate_noise(pos), RayEngine/src/cuda-src/renderer.cu:812-836 (the
// generator sliders, UILayer.cpp:262-318): regenerates the volume of the
// scene settings at field origin 0 and places it at the origin.  As in the
// reference, pos is not read (its kernel is launched with make_float3(0),
// renderer.cu:829-830).  The permutation seed is RPerlinNoise's default 2016.
inline void generate_noise(const float3& pos) {
  (void)pos;
  detail::generate_volume(float3{0, 0, 0}, 2016);
  detail::state().chunk_location = float3{0, 0, 0};
}

// A cube of dim^3 voxels over [0, 1]^3 with the given permutation seed: the
// scene-settings form above with volume_resolution dim and world_size 1.
inline void generate_noise(int dim = 256, uint32_t seed = 2016) {
  detail::State& s = detail::state();
  const SceneSettings keep = s.scene_settings;
  s.scene_settings.volume_resolution = uint3{(uint32_t)dim, (uint32_t)dim, (uint32_t)dim};
  s.scene_settings.world_size = float3{1, 1, 1};
  try {
    detail::generate_volume(float3{0, 0, 0}, seed);
  } catch (...) {
    s.scene_settings = keep;
    throw;
  }
  s.scene_settings = keep;
  s.chunk_location = float3{0, 0, 0};
}

// update_render_settings, RayEngine/src/cuda-src/sphere_tracing.cuh:769-774:
// the settings the next generate_noise / cuda_update_chunk_gen
// (volume_resolution, world_size) and cuda_render_frame (gamma, vignetting,
// vignetting_k) use.
inline void update_render_settings(const RenderingSettings& render_settings, const SceneSettings& scene_settings) {
  detail::State& s = detail::state();
  s.render_settings = render_settings;
  s.scene_settings = scene_settings;
}

// The trace part of render_sphere_trace, RayEngine/src/cuda-src/renderer.cu:142-235.
inline void sphere_trace_primary_rays(HitResult* hit_results, const RCamera& cam, uint32_t width, uint32_t height) {
  detail::State& s = detail::state();
  if (s.volume < 0) throw Error(RE_ERR_INVALID, "sphere trace before generate_noise");
  const size_t n = (size_t)width * height;
  std::vector<float> o, d, t(n);
  std::vector<uint8_t> h(n);
  detail::volume_rays(cam, width, height, o, d);
  check(re_sdf_trace(s.ctx, s.volume, nullptr, o.data(), d.data(), (int64_t)n, t.data(), h.data(), nullptr, nullptr, 0),
        s.ctx);
  for (size_t i = 0; i < n; ++i) {
    HitResult r;
    r.t = t[i];
    r.hits = h[i] != 0;
    r.prim_id = r.obj_index = h[i] ? 0 : -1;
    r.ray_o = float3{o[3 * i], o[3 * i + 1], o[3 * i + 2]};
    r.ray_dir = float3{d[3 * i], d[3 * i + 1], d[3 * i + 2]};
    hit_results[i] = r;
  }
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
  // the ray in the volume's frame (hit_result.ray_o -= instances[0].location, renderer.cu:40)
  const float3 L = s.chunk_location;
  const float o[3] = {ray_o.x - L.x, ray_o.y - L.y, ray_o.z - L.z}, d[3] = {ray_dir.x, ray_dir.y, ray_dir.z};
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
  const size_t n = (size_t)w * h;
  std::vector<float> o, d, t(n), nrm(3 * n);
  std::vector<uint8_t> hit(n);
  detail::volume_rays(s.camera, w, h, o, d);
  check(re_sdf_trace(s.ctx, s.volume, nullptr, o.data(), d.data(), (int64_t)n, t.data(), hit.data(), nullptr,
                     nrm.data(), 0), s.ctx);
  // RenderingSettings' gamma and vignetting (post-processing of the frame)
  re_frame_params fp;
  re_frame_default_params(&fp);
  fp.gamma = s.render_settings.gamma ? 1 : 0;
  fp.vignetting = s.render_settings.vignetting ? 1 : 0;
  fp.vignetting_k = s.render_settings.vignetting_k;
  check(re_frame_shade(s.ctx, d.data(), hit.data(), nrm.data(), (int32_t)w, (int32_t)h, &fp, output, 0), s.ctx);
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
