// Synthetic scene and ray generation (host).  These produce the workloads the
// ray-query path is measured on; they replace the reference's mesh loaders
// (RStaticMesh, RayEngine/src/Objects/RStaticMesh.cpp:68-128) with seeded
// generators, and restate its camera ray generator.
#include <map>
#include <utility>
#include <vector>

#include "errors.hpp"
#include "rng.hpp"

using namespace re;

namespace re {
thread_local char g_last_error[512] = {0};
}

namespace {

void put_tri(float* out, int64_t i, f3 a, f3 b, f3 c) {
  float* p = out + 9 * i;
  p[0] = a.x; p[1] = a.y; p[2] = a.z;
  p[3] = b.x; p[4] = b.y; p[5] = b.z;
  p[6] = c.x; p[7] = c.y; p[8] = c.z;
}

// Triangle with centroid c and edges e1 = l1*u1, e2 = l2*u2.
void centred_tri(float* out, int64_t i, f3 c, f3 e1, f3 e2) {
  f3 s = add(e1, e2);
  f3 v0 = sub(c, mul(s, 1.0f / 3.0f));
  put_tri(out, i, v0, add(v0, e1), add(v0, e2));
}

}  // namespace

extern "C" {

RE_API re_status re_gen_icosphere(int subdiv, float radius, float* out, int64_t* n_tris) {
  if (subdiv < 0 || subdiv > 8 || !n_tris) return fail(RE_ERR_INVALID, "re_gen_icosphere: subdiv in [0,8]");
  int64_t n = 20;
  for (int i = 0; i < subdiv; ++i) n *= 4;
  *n_tris = n;
  if (!out) return RE_OK;
  const float X = 0.525731112119133606f, Z = 0.850650808352039932f;
  std::vector<f3> v = {{-X, 0, Z}, {X, 0, Z},  {-X, 0, -Z}, {X, 0, -Z}, {0, Z, X},  {0, Z, -X},
                       {0, -Z, X}, {0, -Z, -X}, {Z, X, 0},  {-Z, X, 0}, {Z, -X, 0}, {-Z, -X, 0}};
  for (auto& p : v) p = normalize(p);
  std::vector<int> f = {0, 4, 1, 0, 9, 4, 9, 5, 4, 4, 5, 8, 4, 8, 1,  8, 10, 1, 8, 3, 10, 5, 3, 8, 5, 2, 3, 2, 7, 3,
                        7, 10, 3, 7, 6, 10, 7, 11, 6, 11, 0, 6, 0, 1, 6, 6, 1, 10, 9, 0, 11, 9, 11, 2, 9, 2, 5, 7, 2, 11};
  // Orient every face outward (counter-clockwise seen from outside).
  for (size_t k = 0; k < f.size(); k += 3) {
    f3 a = v[f[k]], b = v[f[k + 1]], c = v[f[k + 2]];
    f3 nrm = cross(sub(b, a), sub(c, a));
    if (dot(nrm, add(add(a, b), c)) < 0.0f) std::swap(f[k + 1], f[k + 2]);
  }
  for (int s = 0; s < subdiv; ++s) {
    std::map<std::pair<int, int>, int> mid;
    auto midpoint = [&](int a, int b) {
      auto key = std::make_pair(a < b ? a : b, a < b ? b : a);
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      f3 m = normalize(mul(add(v[a], v[b]), 0.5f));
      v.push_back(m);
      int id = (int)v.size() - 1;
      mid[key] = id;
      return id;
    };
    std::vector<int> g;
    g.reserve(f.size() * 4);
    for (size_t k = 0; k < f.size(); k += 3) {
      int a = f[k], b = f[k + 1], c = f[k + 2];
      int ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
      int t[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
      g.insert(g.end(), t, t + 12);
    }
    f.swap(g);
  }
  for (int64_t i = 0; i < n; ++i)
    put_tri(out, i, mul(v[f[3 * i]], radius), mul(v[f[3 * i + 1]], radius), mul(v[f[3 * i + 2]], radius));
  return RE_OK;
}

RE_API re_status re_gen_random_tris(uint64_t seed, int64_t n, float lo, float hi, float len_lo, float len_hi,
                                    float* out) {
  if (n < 0 || !out) return fail(RE_ERR_INVALID, "re_gen_random_tris: bad arguments");
  Pcg32 rng(seed, 0x52415952ull);  // stream id "RAYR"
  for (int64_t i = 0; i < n; ++i) {
    f3 c{rng.uniform(lo, hi), rng.uniform(lo, hi), rng.uniform(lo, hi)};
    float l1 = rng.uniform(len_lo, len_hi), l2 = rng.uniform(len_lo, len_hi);
    f3 u1 = rng.unit_vector(), u2 = rng.unit_vector();
    centred_tri(out, i, c, mul(u1, l1), mul(u2, l2));
  }
  return RE_OK;
}

RE_API re_status re_gen_cluster_soup(uint64_t seed, int64_t n, int n_clusters, float lo, float hi, float sigma,
                                     float ln_mu, float ln_sigma, float* out) {
  if (n < 0 || n_clusters <= 0 || !out) return fail(RE_ERR_INVALID, "re_gen_cluster_soup: bad arguments");
  Pcg32 rng(seed, 0x434c5553ull);  // stream id "CLUS"
  std::vector<f3> centres(n_clusters);
  for (auto& c : centres) c = f3{rng.uniform(lo, hi), rng.uniform(lo, hi), rng.uniform(lo, hi)};
  for (int64_t i = 0; i < n; ++i) {
    const f3& cc = centres[rng.next() % (uint32_t)n_clusters];
    f3 c{(float)(cc.x + sigma * rng.normal()), (float)(cc.y + sigma * rng.normal()),
         (float)(cc.z + sigma * rng.normal())};
    float l1 = (float)exp(ln_mu + ln_sigma * rng.normal());
    float l2 = (float)exp(ln_mu + ln_sigma * rng.normal());
    f3 u1 = rng.unit_vector(), u2 = rng.unit_vector();
    centred_tri(out, i, c, mul(u1, l1), mul(u2, l2));
  }
  return RE_OK;
}

// generate_ray, RayEngine/src/cuda-src/ray_functions.cuh:13-55.  The
// half-angle product is formed in double as in the reference
// (fov * 0.5 * (M_PI / 180)) and rounded to float before tanf.
RE_API re_status re_gen_camera_rays(const re_camera* cam, int width, int height, float* ray_o, float* ray_d) {
  if (!cam || width < 2 || height < 2 || !ray_o || !ray_d)
    return fail(RE_ERR_INVALID, "re_gen_camera_rays: need width, height >= 2");
  const double kPi = 3.14159265358979323846;
  f3 campos{cam->pos[0], cam->pos[1], cam->pos[2]};
  f3 view = normalize(f3{cam->view[0], cam->view[1], cam->view[2]});
  f3 up = normalize(f3{cam->camdown[0], cam->camdown[1], cam->camdown[2]});
  f3 hax = normalize(cross(view, up));
  f3 vax = normalize(cross(hax, view));
  f3 middle = add(campos, view);
  f3 horizontal = mul(hax, tanf((float)(cam->fov[0] * 0.5 * (kPi / 180))));
  f3 vertical = mul(vax, tanf((float)(cam->fov[1] * 0.5 * (kPi / 180))));
  for (int y = 0; y < height; ++y) {
    for (int x = 0; x < width; ++x) {
      float sx = (float)x / ((float)width - 1.0f);
      float sy = 1.0f - ((float)y / ((float)height - 1.0f));
      f3 p1 = add(add(middle, mul(horizontal, (2 * sx) - 1)), mul(vertical, (2 * sy) - 1));
      f3 pimg = add(campos, mul(sub(p1, campos), cam->focal_distance));
      f3 d = normalize(normalize(sub(pimg, campos)));
      int64_t i = (int64_t)y * width + x;
      ray_o[3 * i] = campos.x; ray_o[3 * i + 1] = campos.y; ray_o[3 * i + 2] = campos.z;
      ray_d[3 * i] = d.x; ray_d[3 * i + 1] = d.y; ray_d[3 * i + 2] = d.z;
    }
  }
  return RE_OK;
}

// Cosine-weighted bounce (Malley's method: uniform disk sample lifted to the
// hemisphere) about the geometric normal facing the incoming ray.  The
// offset 1e-4 * n follows the reference's secondary-ray origin offset
// (RayEngine/src/cuda-src/renderer.cu:579).  The tangent frame is the
// branchless basis of Duff et al., "Building an Orthonormal Basis,
// Revisited" (JCGT 2017).
RE_API re_status re_gen_bounce_rays(const float* tri_v, int64_t n_tris, const float* ray_o, const float* ray_d,
                                    const int32_t* prim_id, const float* t, int64_t n, uint64_t seed, float* out_o,
                                    float* out_d, int64_t* n_out) {
  if (!tri_v || !ray_o || !ray_d || !prim_id || !t || !out_o || !out_d || !n_out || n < 0)
    return fail(RE_ERR_INVALID, "re_gen_bounce_rays: bad arguments");
  int64_t m = 0;
  for (int64_t i = 0; i < n; ++i) {
    int32_t p = prim_id[i];
    if (p < 0) continue;
    if (p >= n_tris) return fail(RE_ERR_INVALID, "re_gen_bounce_rays: prim id %d out of range", p);
    const float* tv = tri_v + 9 * (int64_t)p;
    f3 v0{tv[0], tv[1], tv[2]}, v1{tv[3], tv[4], tv[5]}, v2{tv[6], tv[7], tv[8]};
    f3 nrm = normalize(cross(sub(v1, v0), sub(v2, v0)));
    f3 o{ray_o[3 * i], ray_o[3 * i + 1], ray_o[3 * i + 2]};
    f3 d{ray_d[3 * i], ray_d[3 * i + 1], ray_d[3 * i + 2]};
    if (dot(nrm, d) > 0.0f) nrm = mul(nrm, -1.0f);
    f3 hit = add(o, mul(d, t[i]));
    f3 org = add(hit, mul(nrm, 1e-4f));
    Pcg32 rng(seed, (uint64_t)i);
    float sx, sy, r2;
    do {
      sx = 2.0f * rng.uniform() - 1.0f;
      sy = 2.0f * rng.uniform() - 1.0f;
      r2 = sx * sx + sy * sy;
    } while (r2 >= 1.0f);
    float sz = sqrtf(1.0f - r2);
    float sign = nrm.z >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (sign + nrm.z);
    float b = nrm.x * nrm.y * a;
    f3 tx{1.0f + sign * nrm.x * nrm.x * a, sign * b, -sign * nrm.x};
    f3 ty{b, sign + nrm.y * nrm.y * a, -nrm.y};
    f3 dir = normalize(add(add(mul(tx, sx), mul(ty, sy)), mul(nrm, sz)));
    out_o[3 * m] = org.x; out_o[3 * m + 1] = org.y; out_o[3 * m + 2] = org.z;
    out_d[3 * m] = dir.x; out_d[3 * m + 1] = dir.y; out_d[3 * m + 2] = dir.z;
    ++m;
  }
  *n_out = m;
  return RE_OK;
}

}  // extern "C"
