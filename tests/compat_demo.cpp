// Reference-style host code against include/rayengine_compat.hpp: builds a
// kd-tree from an indexed mesh (RKDTreeCPU), uploads it (copy_memory),
// moves the object (update_objects) and traces the camera's primary rays
// (trace_primary_rays).  Mode "cpu" stops after the tree build; mode "gpu"
// writes prim ids and t bits to the output file for comparison.
#include <cstdio>
#include <cstring>
#include <vector>

#include "rayengine_compat.hpp"

using namespace rayengine_compat;

int main(int argc, char** argv) {
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
  std::printf("{\"n_tris\": %lld, \"n_leaves\": %lld, \"n_trirecs\": %lld, \"depth\": %d}\n", (long long)info.n_tris,
              (long long)info.n_leaves, (long long)info.n_trirecs, info.depth);
  if (!gpu) return 0;
  RCamera cam;
  cam.campos = float3{0, 0, 5};
  std::vector<RKDThreeGPU*> trees = {&tree};
  std::vector<GPUSceneObject> objs(1);
  copy_memory(trees, cam, objs);
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
