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
  free_memory();
  return 0;
}
