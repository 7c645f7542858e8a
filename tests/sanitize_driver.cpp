// Host library code (scene generators, kd-tree build and export, camera and
// bounce rays, Perlin table, .rsdf header checks) and the CPU oracle driven
// through their C entry points, for a build under AddressSanitizer and
// UndefinedBehaviorSanitizer.  Every call uses small inputs, including the
// edge cases the library must refuse or pass through (empty soups,
// degenerate triangles, NaN and zero directions, truncated files).  Exits
// non-zero on a failed check; the sanitizers abort on their own findings.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "rayengine.h"

extern "C" {
void orc_trace(const void* nodes, const void* leaves, const void* tris, const void* objs, int n_obj, const float* o,
               const float* d, int64_t n, int cap, int two_sided, int secondary, int32_t* prim, float* t,
               int32_t* obj, uint64_t* counts, int threads, int cut);
void orc_occluded(const void* nodes, const void* leaves, const void* tris, const void* objs, int n_obj, const float* o,
                  const float* d, int64_t n, int cap, int two_sided, int secondary, uint8_t* out, int threads);
void orc_brute(const float* tri_v, int64_t n_tris, const float* root_min, const float* root_max, const float* o,
               const float* d, int64_t n, int two_sided, int32_t* prim, float* t, int threads);
void orc_hit_attributes(const float* geom, const float* nrm, const float* uvs, const void* objs, const float* o,
                        const float* d, int64_t n, const int32_t* prim, const int32_t* obj, int two_sided,
                        int secondary, float* normal, float* uv, float* bary, float* hit_point);
void orc_perlin_table(uint32_t seed, uint32_t* perm, float* grads);
void orc_noise_fill(const uint32_t* perm, const int32_t* dim, const float* spacing, const float* origin, int octaves,
                    float lacunarity, float gain, float base_freq, float y_offset, float* out, int threads);
void orc_noise_fill_f64(const uint32_t* perm, const int32_t* dim, const float* spacing, const float* origin,
                        int octaves, float lacunarity, float gain, float base_freq, float y_offset, double* out,
                        int threads);
void orc_sdf_trace(const float* grid, const int32_t* dim, const float* spacing, const float* o, const float* d,
                   int64_t n, int max_steps, float eps, float scale, int rule, float* t, uint8_t* hit, int32_t* steps,
                   float* nrm, int threads);
void orc_truth_trace(const double* grid, const int32_t* dim, const float* spacing, const float* o, const float* d,
                     int64_t n, double h_voxels, double* t, uint8_t* hit, int threads);
void orc_volume_edit(float* dist, float* mat, const int32_t* dim, const float* spacing, int type, const float* center,
                     const float* extent, float material, int snap, float k);
int orc_mlp_n_params(void);
void orc_mlp_forward(const float* params, const float* pts, int64_t n, double* out);
}

namespace {

int g_failed = 0;

void expect(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s\n", what);
    ++g_failed;
  }
}

struct ObjRec {
  float loc[3];
  float sin_y, cos_y;
  int32_t is_character, root;
  float root_min[3], root_max[3];
};
static_assert(sizeof(ObjRec) == 52, "object record layout");

void trace_tree(const std::vector<float>& soup, int64_t n_tris, const re_kd_build_params* p) {
  re_kdtree* tree = nullptr;
  expect(re_kd_build(soup.data(), n_tris, p, &tree) == RE_OK, "re_kd_build");
  re_kd_info info;
  expect(re_kd_get_info(tree, &info) == RE_OK, "re_kd_get_info");
  std::vector<unsigned char> nodes((size_t)info.n_nodes * 8 + 8), leaves((size_t)info.n_leaves * 64 + 64),
      tris((size_t)info.n_trirecs * 40 + 40);
  expect(re_kd_export(tree, nodes.data(), leaves.data(), tris.data()) == RE_OK, "re_kd_export");
  ObjRec ob{};
  ob.cos_y = 1.0f;
  for (int a = 0; a < 3; ++a) {
    ob.root_min[a] = info.root_min[a];
    ob.root_max[a] = info.root_max[a];
  }
  re_camera cam;
  cam.pos[0] = 0.2f; cam.pos[1] = 0.1f; cam.pos[2] = 4.0f;
  cam.view[0] = 0; cam.view[1] = 0; cam.view[2] = -1;
  cam.camdown[0] = 0; cam.camdown[1] = -1; cam.camdown[2] = 0;
  cam.fov[0] = 60; cam.fov[1] = 60;
  cam.focal_distance = 4;
  const int w = 24, h = 16;
  const int64_t n = (int64_t)w * h + 4;
  std::vector<float> o(3 * n), d(3 * n);
  expect(re_gen_camera_rays(&cam, w, h, o.data(), d.data()) == RE_OK, "re_gen_camera_rays");
  // four degenerate rays: NaN, zero, axis-aligned, origin inside the box
  const float nan = std::numeric_limits<float>::quiet_NaN();
  const float extra[4][6] = {{0, 0, 4, nan, 0, -1}, {0, 0, 4, 0, 0, 0}, {0, 0, 4, 0, 0, -1}, {0, 0, 0, 1, 0, 0}};
  for (int k = 0; k < 4; ++k)
    for (int c = 0; c < 3; ++c) {
      o[3 * (w * h + k) + c] = extra[k][c];
      d[3 * (w * h + k) + c] = extra[k][3 + c];
    }
  std::vector<int32_t> prim(n), obj(n), prim_b(n);
  std::vector<float> t(n), t_b(n);
  std::vector<uint64_t> counts(4 * n);
  for (int cut = 0; cut < 2; ++cut)
    for (int cap : {1, 50}) {
      orc_trace(nodes.data(), leaves.data(), tris.data(), &ob, 1, o.data(), d.data(), n, cap, 0, 0, prim.data(),
                t.data(), obj.data(), counts.data(), 2, cut);
    }
  std::vector<uint8_t> occ(n);
  orc_occluded(nodes.data(), leaves.data(), tris.data(), &ob, 1, o.data(), d.data(), n, 50, 1, 1, occ.data(), 2);
  orc_brute(soup.data(), n_tris, info.root_min, info.root_max, o.data(), d.data(), n, 0, prim_b.data(), t_b.data(), 2);
  // attributes of the traced hits: flat per-corner normals and uvs
  std::vector<float> geom((size_t)n_tris * 9), nrm((size_t)n_tris * 9), uvs((size_t)n_tris * 6);
  for (int64_t i = 0; i < n_tris; ++i) {
    const float* v = &soup[9 * i];
    for (int c = 0; c < 3; ++c) {
      geom[9 * i + c] = v[c];
      geom[9 * i + 3 + c] = v[3 + c] - v[c];
      geom[9 * i + 6 + c] = v[6 + c] - v[c];
    }
    for (int k = 0; k < 9; ++k) nrm[9 * i + k] = (k % 3 == 2) ? 1.0f : 0.0f;
    for (int k = 0; k < 6; ++k) uvs[6 * i + k] = 0.1f * k;
  }
  std::vector<float> an(3 * n), auv(2 * n), ab(2 * n), ahp(3 * n);
  orc_trace(nodes.data(), leaves.data(), tris.data(), &ob, 1, o.data(), d.data(), n, 50, 0, 0, prim.data(), t.data(),
            obj.data(), nullptr, 2, 0);
  orc_hit_attributes(geom.data(), nrm.data(), uvs.data(), &ob, o.data(), d.data(), n, prim.data(), obj.data(), 0, 0,
                     an.data(), auv.data(), ab.data(), ahp.data());
  // bounce rays off the hits
  std::vector<float> bo(3 * n), bd(3 * n);
  int64_t nb = 0;
  expect(re_gen_bounce_rays(soup.data(), n_tris, o.data(), d.data(), prim.data(), t.data(), n, 7, bo.data(), bd.data(),
                            &nb) == RE_OK,
         "re_gen_bounce_rays");
  expect(nb >= 0 && nb <= n, "bounce count");
  re_kd_free(tree);
}

void write_file(const std::string& path, const void* data, size_t bytes) {
  FILE* f = std::fopen(path.c_str(), "wb");
  std::fwrite(data, 1, bytes, f);
  std::fclose(f);
}

}  // namespace

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  // scenes
  int64_t n_ico = 0;
  expect(re_gen_icosphere(2, 1.0f, nullptr, &n_ico) == RE_OK && n_ico == 320, "icosphere size");
  std::vector<float> ico((size_t)n_ico * 9);
  expect(re_gen_icosphere(2, 1.0f, ico.data(), &n_ico) == RE_OK, "icosphere");
  std::vector<float> rnd(9 * 500), clu(9 * 2000);
  expect(re_gen_random_tris(5, 500, -1.5f, 1.5f, 0.05f, 0.3f, rnd.data()) == RE_OK, "random tris");
  expect(re_gen_cluster_soup(3, 2000, 16, -2.0f, 2.0f, 0.3f, -3.0f, 0.5f, clu.data()) == RE_OK, "cluster soup");
  re_kd_build_params p;
  re_kd_default_params(&p);
  trace_tree(ico, n_ico, nullptr);
  trace_tree(rnd, 500, &p);
  p.max_leaf_tris = 1;
  p.max_depth = 40;
  trace_tree(clu, 2000, &p);
  p.max_depth = 2;
  trace_tree(clu, 2000, &p);
  // the reference's spatial-median rule with bounding-box assignment
  re_kd_default_params(&p);
  p.split_rule = 1;
  p.clip = 0;
  p.max_leaf_tris = 20;
  trace_tree(rnd, 500, &p);
  // degenerate triangles (zero area, repeated vertices) next to good ones
  std::vector<float> degen(ico);
  for (int k = 0; k < 9; ++k) degen[k] = 0.25f;
  for (int k = 9; k < 18; ++k) degen[k] = (k % 3) * 0.5f;
  trace_tree(degen, n_ico, nullptr);
  // refused input
  re_kdtree* none = nullptr;
  expect(re_kd_build(ico.data(), -1, nullptr, &none) != RE_OK, "negative triangle count refused");
  // an empty soup builds an empty tree or is refused, without touching memory
  if (re_kd_build(ico.data(), 0, nullptr, &none) == RE_OK) {
    re_kd_info ei;
    expect(re_kd_get_info(none, &ei) == RE_OK && ei.n_tris == 0, "empty tree info");
    re_kd_free(none);
  }
  expect(re_gen_camera_rays(nullptr, 4, 4, nullptr, nullptr) != RE_OK, "null camera refused");
  // Perlin tables: the library's portable shuffle is a permutation, doubled
  uint32_t perm[512], operm[512];
  float grads[768];
  expect(re_perlin_table(2016, perm) == RE_OK, "perlin table");
  orc_perlin_table(2016, operm, grads);
  bool seen[256] = {};
  for (int i = 0; i < 256; ++i) {
    expect(perm[i] < 256 && perm[i + 256] == perm[i], "perm doubled");
    seen[perm[i] & 255] = true;
  }
  for (bool s : seen) expect(s, "perm is a permutation");
  // terrain: noise fill, traces, edits
  const int32_t dim[3] = {12, 10, 14};
  const float sp[3] = {0.1f, 0.1f, 0.08f}, origin[3] = {0.5f, -0.25f, 2.0f};
  const size_t nv = (size_t)dim[0] * dim[1] * dim[2];
  std::vector<float> grid(nv), mat(nv, 0.0f);
  std::vector<double> grid64(nv);
  orc_noise_fill(perm, dim, sp, nullptr, 5, 2.0f, 0.5f, 4.0f, 0.4f, grid.data(), 2);
  orc_noise_fill(perm, dim, sp, origin, 3, 2.0f, 0.5f, 4.0f, 0.4f, grid.data(), 1);
  orc_noise_fill_f64(perm, dim, sp, nullptr, 5, 2.0f, 0.5f, 4.0f, 0.4f, grid64.data(), 2);
  const int64_t nr = 64;
  std::vector<float> ro(3 * nr), rd(3 * nr), rt(nr), rn(3 * nr);
  std::vector<uint8_t> rh(nr);
  std::vector<int32_t> rs(nr);
  std::vector<double> rt64(nr);
  for (int64_t i = 0; i < nr; ++i) {
    ro[3 * i] = 0.05f + 0.015f * (float)i;
    ro[3 * i + 1] = 1.5f;
    ro[3 * i + 2] = 0.5f;
    rd[3 * i] = 0.0f;
    rd[3 * i + 1] = i % 7 == 3 ? 0.0f : -1.0f;
    rd[3 * i + 2] = i % 5 == 1 ? 1.0f : 0.0f;
  }
  orc_sdf_trace(grid.data(), dim, sp, ro.data(), rd.data(), nr, 256, 1e-4f, 1.0f, 0, rt.data(), rh.data(), rs.data(),
                rn.data(), 2);
  orc_sdf_trace(grid.data(), dim, sp, ro.data(), rd.data(), nr, 3, 1e-5f, 0.9f, 1, rt.data(), rh.data(), rs.data(),
                nullptr, 1);
  orc_truth_trace(grid64.data(), dim, sp, ro.data(), rd.data(), nr, 0.05, rt64.data(), rh.data(), 2);
  const float c[3] = {0.6f, 0.4f, 0.5f}, ext[3] = {2.0f, 1.0f, 3.0f};
  for (int type = 0; type < 4; ++type)
    for (int snap = 0; snap < 2; ++snap) orc_volume_edit(grid.data(), mat.data(), dim, sp, type, c, ext, 3.0f, snap, 0.1f);
  // .rsdf files: a good one, a truncated one, a bad header, a missing one
  std::vector<float> file(6 + 2 * nv, 0.0f);
  std::memcpy(file.data(), dim, 12);
  std::memcpy(file.data() + 3, sp, 12);
  write_file(dir + "/good.rsdf", file.data(), file.size() * 4);
  write_file(dir + "/short.rsdf", file.data(), file.size() * 4 - 8);
  const int32_t bad_dim[3] = {1 << 30, 1 << 30, 2};
  std::memcpy(file.data(), bad_dim, 12);
  write_file(dir + "/huge.rsdf", file.data(), 24);
  int32_t fd[3];
  float fs[3];
  expect(re_rsdf_info((dir + "/good.rsdf").c_str(), fd, fs) == RE_OK && fd[2] == 14, "good rsdf");
  expect(re_rsdf_info((dir + "/short.rsdf").c_str(), fd, fs) == RE_ERR_IO, "short rsdf refused");
  expect(re_rsdf_info((dir + "/huge.rsdf").c_str(), fd, fs) == RE_ERR_IO, "huge header refused");
  expect(re_rsdf_info((dir + "/missing.rsdf").c_str(), fd, fs) == RE_ERR_IO, "missing rsdf refused");
  // surrogate network forward pass
  const int np = orc_mlp_n_params();
  std::vector<float> params((size_t)np);
  for (int i = 0; i < np; ++i) params[i] = 0.01f * (float)((i * 37) % 101 - 50);
  std::vector<double> mo(16);
  orc_mlp_forward(params.data(), ro.data(), 16, mo.data());
  std::printf("sanitized run done, %d failed checks\n", g_failed);
  return g_failed == 0 ? 0 : 1;
}
