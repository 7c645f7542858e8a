// Reference-style host code against include/rayengine_compat.hpp: builds a
// kd-tree from an indexed mesh (RKDTreeCPU), uploads it with its packed
// triangles, vertex normals and uvs (copy_memory, packed as
// RTScene::init_triangles packs them), moves the object (update_objects) and
// traces the camera's primary rays (trace_primary_rays).  Mode "cpu" stops
// after the tree build; mode "gpu" writes prim ids and t bits, then normals,
// uvs and hit points, for comparison.  Mode "volume" drives the terrain entry
// points in the order SDFScene uses them (initialize_volume_render,
// cuda_update_chunk, cuda_update_chunk_gen, generate_noise,
// update_render_settings, cuda_render_frame) and dumps each volume.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "rayengine_compat.hpp"

using namespace rayengine_compat;

namespace {

void dump_volume(const std::string& path) {
  detail::State& s = detail::state();
  int32_t dim[3];
  float sp[3];
  int64_t bytes = 0;
  check(re_volume_info(s.ctx, s.volume, dim, sp, &bytes), s.ctx);
  std::vector<float> g((size_t)dim[0] * dim[1] * dim[2]), m(g.size());
  check(re_volume_read(s.ctx, s.volume, g.data()), s.ctx);
  check(re_volume_read_material(s.ctx, s.volume, m.data()), s.ctx);
  FILE* f = std::fopen(path.c_str(), "wb");
  std::fwrite(dim, 4, 3, f);
  std::fwrite(sp, 4, 3, f);
  std::fwrite(g.data(), 4, g.size(), f);
  std::fwrite(m.data(), 4, m.size(), f);
  std::fclose(f);
}

void dump_frame(const std::string& path, uint32_t w, uint32_t h) {
  std::vector<uint32_t> px(w * h);
  cuda_render_frame(px.data(), w, h);
  FILE* g = std::fopen(path.c_str(), "wb");
  std::fwrite(px.data(), 4, px.size(), g);
  std::fclose(g);
}

// SDFScene's terrain life cycle against the compat entry points.
int volume_mode(const std::string& dir) {
  // a 24^3 chunk grid of a sphere of radius 0.3 around (0.5, 0.5, 0.5),
  // material 1 inside, 2 outside
  Grid grid;
  const uint32_t n = 24;
  grid.sdf_dim = uint3{n, n, n};
  grid.spacing = float3{1.0f / 24, 1.0f / 24, 1.0f / 24};
  grid.box_max = float3{1, 1, 1};
  grid.voxels.resize((size_t)n * n * n);
  for (uint32_t z = 0; z < n; ++z)
    for (uint32_t y = 0; y < n; ++y)
      for (uint32_t x = 0; x < n; ++x) {
        GridNode& g = grid.voxels[x + n * (y + n * z)];
        g.point = float3{x * grid.spacing.x, y * grid.spacing.y, z * grid.spacing.z};
        const float dx = g.point.x - 0.5f, dy = g.point.y - 0.5f, dz = g.point.z - 0.5f;
        g.distance = std::sqrt(dx * dx + dy * dy + dz * dz) - 0.3f;
        g.material = g.distance < 0 ? 1.0f : 2.0f;
      }
  RayEngine::RChunk chunk(grid);
  chunk.set_location(float3{0.25f, 0.0f, -0.5f});
  RayEngine::RPerlinNoise noise(2016);
  RCamera cam;
  cam.campos = float3{0.75f, 0.5f, 2.5f};
  cam.view = float3{0.0f, 0.0f, -1.0f};
  std::vector<VoxelMaterial> materials(2);
  RenderingSettings rs;
  SceneSettings ss;
  ss.volume_resolution = uint3{40, 24, 33};
  ss.world_size = float3{2.0f, 1.0f, 1.5f};
  initialize_volume_render(cam, chunk, 1, materials, rs, ss, noise);
  dump_volume(dir + "/init.bin");
  dump_frame(dir + "/init_frame.bin", 64, 48);
  // the chunk edited on the host and uploaded again
  grid.voxels[0].material = 7.0f;
  RayEngine::RChunk edited(grid);
  edited.set_location(float3{0.0f, 0.0f, 0.0f});
  cuda_update_chunk(edited, noise);
  dump_volume(dir + "/update.bin");
  // a generated chunk at a world location, with another seed
  RayEngine::RChunk far_chunk;
  far_chunk.set_location(float3{3.0f, 0.0f, -2.0f});
  RayEngine::RPerlinNoise noise7(7);
  cuda_update_chunk_gen(far_chunk, noise7);
  dump_volume(dir + "/gen.bin");
  // generator sliders: generate_noise(pos) with new settings
  ss.volume_resolution = uint3{32, 16, 20};
  ss.world_size = float3{1.0f, 0.5f, 0.75f};
  rs.gamma = true;
  rs.vignetting = true;
  rs.vignetting_k = 0.4f;
  update_render_settings(rs, ss);
  generate_noise(float3{5.0f, 1.0f, 5.0f});
  dump_volume(dir + "/noise.bin");
  RCamera top;
  top.campos = float3{0.5f, 0.9f, -0.3f};
  top.view = float3{0.0f, -0.55f, 1.0f};
  set_camera(top);
  dump_frame(dir + "/noise_frame.bin", 64, 48);
  std::printf("volume ok\n");
  free_memory();
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc > 2 && std::strcmp(argv[1], "volume") == 0) return volume_mode(argv[2]);
  const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
  // Icosphere subdivided once, as an indexed mesh with duplicated vertices.
  int64_t n_tris = 0;
  re_gen_icosphere(2, 1.0f, nullptr, &n_tris);
  std::vector<float> soup(9 * n_tris);
  re_gen_icosphere(2, 1.0f, soup.data(), &n_tris);
  std::vector<float3> verts(3 * n_tris);
  std::vector<uint3> faces(n_tris);
  for (int64_t i = 0; i < 3 * n_tris; ++i) verts[i] = float3{soup[3 * i], soup[3 * i + 1], soup[3 * i + 2]};
  for (int64_t f = 0; f < n_tris; ++f) faces[f] = uint3{(uint32_t)(3 * f), (uint32_t)(3 * f + 1), (uint32_t)(3 * f + 2)};
  RKDTreeCPU tree(verts.data(), faces.data(), nullptr, nullptr, (int)verts.size(), (int)faces.size(), 0, 0);
  re_kd_info info = tree.info();
  RKDThreeGPU gpu_tree(&tree);  // RTScene::build_scene, RTScene.cpp:48
  const std::vector<int> index_list = gpu_tree.GetIndexList();
  std::printf("{\"n_tris\": %lld, \"n_leaves\": %lld, \"n_trirecs\": %lld, \"depth\": %d, \"index_list\": %zu, "
              "\"index_max\": %d, \"n_nodes\": %zu}\n",
              (long long)info.n_tris, (long long)info.n_leaves, (long long)info.n_trirecs, info.depth, index_list.size(),
              index_list.empty() ? -1 : *std::max_element(index_list.begin(), index_list.end()), gpu_tree.GetNumNodes());
  if (!gpu) return 0;
  RCamera cam;
  cam.campos = float3{0, 0, 5};
  std::vector<RKDThreeGPU*> trees = {&gpu_tree};
  std::vector<GPUSceneObject> objs(1);
  objs[0].num_prims = (size_t)n_tris;
  // RTScene::init_triangles packing: v0, v1 - v0, v2 - v0 per triangle; the
  // vertex normals (here the unit vertex positions of the sphere) and uvs
  // (here spherical angles) per corner
  std::vector<float4> packed, normals;
  std::vector<float2> uvs;
  for (int64_t f = 0; f < n_tris; ++f) {
    const float3 v[3] = {verts[3 * f], verts[3 * f + 1], verts[3 * f + 2]};
    packed.push_back(float4{v[0].x, v[0].y, v[0].z, 0});
    packed.push_back(float4{v[1].x - v[0].x, v[1].y - v[0].y, v[1].z - v[0].z, 0});
    packed.push_back(float4{v[2].x - v[0].x, v[2].y - v[0].y, v[2].z - v[0].z, 0});
    for (int k = 0; k < 3; ++k) {
      const float r = std::sqrt(v[k].x * v[k].x + v[k].y * v[k].y + v[k].z * v[k].z);
      normals.push_back(float4{v[k].x / r, v[k].y / r, v[k].z / r, 0});
      uvs.push_back(float2{std::atan2(v[k].z, v[k].x), std::acos(v[k].y / r)});
    }
  }
  copy_memory(trees, cam, packed, normals, uvs, objs, std::vector<float3>{}, false);
  bool refused = false;
  try {
    copy_memory(trees, cam, packed, normals, uvs, objs, std::vector<float3>{}, true);
  } catch (const Error& e) {
    refused = e.status == RE_ERR_UNSUPPORTED;
  }
  std::printf("bruteforce refused %d\n", refused ? 1 : 0);
  {
    // the attribute arrays as the C ABI takes them, for the comparison
    FILE* fa = std::fopen((std::string(argv[2]) + ".attr_in").c_str(), "wb");
    for (const float4& q : normals) std::fwrite(&q, 4, 3, fa);
    for (const float2& q : uvs) std::fwrite(&q, 4, 2, fa);
    std::fclose(fa);
  }
  objs[0].location = float3{0.25f, -0.1f, 0.0f};
  objs[0].rotation = float3{0.0f, 0.3f, 0.0f};
  update_objects(objs);
  const uint32_t w = 64, h = 48;
  std::vector<HitResult> hits(w * h);
  trace_primary_rays(hits.data(), cam, w, h);
  FILE* f = std::fopen(argv[2], "wb");
  for (const HitResult& r : hits) {
    std::fwrite(&r.prim_id, 4, 1, f);
    std::fwrite(&r.t, 4, 1, f);
  }
  std::fclose(f);
  FILE* fh = std::fopen((std::string(argv[2]) + ".attr").c_str(), "wb");
  for (const HitResult& r : hits) {
    std::fwrite(&r.normal, 4, 3, fh);
    std::fwrite(&r.uv, 4, 2, fh);
    std::fwrite(&r.hit_point, 4, 3, fh);
    std::fwrite(&r.ray_o, 4, 3, fh);
    std::fwrite(&r.ray_dir, 4, 3, fh);
  }
  std::fclose(fh);
  int n_hit = 0;
  for (const HitResult& r : hits) n_hit += r.hits;
  std::printf("hits %d of %u\n", n_hit, w * h);
  if (argc > 3) {
    // trace_shadow: eight rays from outside the sphere, four towards it, four away
    std::vector<float3> so(8), sd(8);
    for (int i = 0; i < 8; ++i) {
      so[i] = float3{0.1f * i, 0.05f * i, 4.0f};
      sd[i] = float3{0.0f, 0.0f, i < 4 ? -1.0f : 1.0f};
    }
    std::vector<uint8_t> occ(8);
    trace_shadow_rays(occ.data(), so.data(), sd.data(), 8);
    std::printf("shadow");
    for (uint8_t v : occ) std::printf(" %d", v);
    std::printf("\n");
    // terrain: generate, brush from above, save and reload, render a frame
    const std::string dir = argv[3];
    generate_noise(64, 2016);
    TerrainBrush brush{SPHERE_ADD, int3{1, 1, 1}, 3, false};
    const bool bhit = insert_sphere_to_texture(brush, float3{0.5f, 2.0f, 0.5f}, float3{0.0f, -1.0f, 0.0f});
    std::printf("brush %d\n", bhit ? 1 : 0);
    save_map(dir + "/map.rsdf");
    load_map(dir + "/map.rsdf");
    RCamera tc;
    tc.campos = float3{0.5f, 0.95f, -0.35f};
    tc.view = float3{0.0f, -0.55f, 1.0f};
    set_camera(tc);
    const uint32_t fw = 96, fh = 64;
    std::vector<uint32_t> px(fw * fh);
    cuda_render_frame(px.data(), fw, fh);
    FILE* g = std::fopen((dir + "/frame.bin").c_str(), "wb");
    std::fwrite(px.data(), 4, px.size(), g);
    std::fclose(g);
    check(re_write_ppm((dir + "/frame.ppm").c_str(), px.data(), (int32_t)fw, (int32_t)fh));
  }
  free_memory();
  return 0;
}
