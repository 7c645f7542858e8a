// CPU oracle, terrain part.  TEST INFRASTRUCTURE ONLY (see oracle.cpp).
//
// Restates the noise primitive, the fBm terrain field, software trilinear
// sampling and the sphere-trace march in FP32 with the reference's
// operation order, plus an FP64 twin of the whole chain (field generated,
// sampled and marched in double) that measures the FP32 hit-depth error.
// Grids here are linear, x fastest: index x + dx*(y + dy*z).
#include <cmath>
#include <cstdint>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API extern "C" __attribute__((visibility("default")))

namespace {

// gradientDotV, RayEngine/src/cuda-src/sdf_functions.cuh:25-48.
template <typename T>
T gradient_dot(uint32_t perm, T x, T y, T z) {
  switch (perm & 15) {
    case 0: return x + y;
    case 1: return -x + y;
    case 2: return x - y;
    case 3: return -x - y;
    case 4: return x + z;
    case 5: return -x + z;
    case 6: return x - z;
    case 7: return -x - z;
    case 8: return y + z;
    case 9: return -y + z;
    case 10: return y - z;
    case 11: return -y - z;
    case 12: return y + x;
    case 13: return -x + y;
    case 14: return -y + z;
    case 15: return -y - z;
  }
  return 0;
}

// quintic, sdf_functions.cuh:56-60.
template <typename T>
T quintic(T t) { return t * t * t * (t * (t * 6 - 15) + 10); }

// hash: P[P[P[x] + y] + z] over the doubled table (the build's definition of
// the reference's texture hash, whose texture contents are undefined).
inline uint32_t hash3(const uint32_t* P, int x, int y, int z) { return P[P[P[x] + y] + z]; }

// eval, sdf_functions.cuh:83-135 (value only).
template <typename T>
T eval(const uint32_t* P, T px, T py, T pz) {
  int xi0 = ((int)std::floor(px)) & 255;
  int yi0 = ((int)std::floor(py)) & 255;
  int zi0 = ((int)std::floor(pz)) & 255;
  int xi1 = (xi0 + 1) & 255, yi1 = (yi0 + 1) & 255, zi1 = (zi0 + 1) & 255;
  T tx = px - (T)((int)std::floor(px));
  T ty = py - (T)((int)std::floor(py));
  T tz = pz - (T)((int)std::floor(pz));
  T u = quintic(tx), v = quintic(ty), w = quintic(tz);
  T x0 = tx, x1 = tx - 1, y0 = ty, y1 = ty - 1, z0 = tz, z1 = tz - 1;
  T a = gradient_dot<T>(hash3(P, xi0, yi0, zi0), x0, y0, z0);
  T b = gradient_dot<T>(hash3(P, xi1, yi0, zi0), x1, y0, z0);
  T c = gradient_dot<T>(hash3(P, xi0, yi1, zi0), x0, y1, z0);
  T d = gradient_dot<T>(hash3(P, xi1, yi1, zi0), x1, y1, z0);
  T e = gradient_dot<T>(hash3(P, xi0, yi0, zi1), x0, y0, z1);
  T f = gradient_dot<T>(hash3(P, xi1, yi0, zi1), x1, y0, z1);
  T g = gradient_dot<T>(hash3(P, xi0, yi1, zi1), x0, y1, z1);
  T h = gradient_dot<T>(hash3(P, xi1, yi1, zi1), x1, y1, z1);
  T k0 = a;
  T k1 = (b - a);
  T k2 = (c - a);
  T k3 = (e - a);
  T k4 = (a + d - b - c);
  T k5 = (a + f - b - e);
  T k6 = (a + g - c - e);
  T k7 = (b + c + e + h - a - d - f - g);
  return k0 + k1 * u + k2 * v + k3 * w + k4 * u * v + k5 * u * w + k6 * v * w + k7 * u * v * w;
}

struct Fbm {
  int octaves;
  float lacunarity, gain, base_freq, y_offset;
};

// Terrain field: (y - y_offset) - sum_o gain^o * eval(p * freq_o), freq_o =
// base_freq * lacunarity^o accumulated by repeated multiplication.
template <typename T>
T field(const uint32_t* P, const Fbm& fp, T x, T y, T z) {
  T sum = 0, amp = 1, freq = (T)fp.base_freq;
  for (int o = 0; o < fp.octaves; ++o) {
    sum = sum + amp * eval<T>(P, x * freq, y * freq, z * freq);
    freq = freq * (T)fp.lacunarity;
    amp = amp * (T)fp.gain;
  }
  return (y - (T)fp.y_offset) - sum;
}

template <typename T>
struct Grid {
  const T* v;
  int dx, dy, dz;
  T sx, sy, sz;
  T at(int x, int y, int z) const { return v[((int64_t)z * dy + y) * dx + x]; }
};

// interpolate / bilinear (sdf_functions.cuh:209-265) on grid coordinates
// p / spacing, cell index clamped to [0, dim-2].
template <typename T>
T trilerp(const Grid<T>& G, T px, T py, T pz) {
  T gx = px / G.sx, gy = py / G.sy, gz = pz / G.sz;
  int ix = (int)gx, iy = (int)gy, iz = (int)gz;
  ix = ix < 0 ? 0 : (ix > G.dx - 2 ? G.dx - 2 : ix);
  iy = iy < 0 ? 0 : (iy > G.dy - 2 ? G.dy - 2 : iy);
  iz = iz < 0 ? 0 : (iz > G.dz - 2 ? G.dz - 2 : iz);
  T tx = gx - (T)ix, ty = gy - (T)iy, tz = gz - (T)iz;
  T c000 = G.at(ix, iy, iz), c100 = G.at(ix + 1, iy, iz);
  T c010 = G.at(ix, iy + 1, iz), c110 = G.at(ix + 1, iy + 1, iz);
  T c001 = G.at(ix, iy, iz + 1), c101 = G.at(ix + 1, iy, iz + 1);
  T c011 = G.at(ix, iy + 1, iz + 1), c111 = G.at(ix + 1, iy + 1, iz + 1);
  T a0 = c000 * (T(1) - tx) + c100 * tx;
  T b0 = c010 * (T(1) - tx) + c110 * tx;
  T e = a0 * (T(1) - ty) + b0 * ty;
  T a1 = c001 * (T(1) - tx) + c101 * tx;
  T b1 = c011 * (T(1) - tx) + c111 * tx;
  T f = a1 * (T(1) - ty) + b1 * ty;
  return e * (1 - tz) + f * tz;
}

// Slab test of ray_functions.cuh:61-92 (see oracle.cpp).
template <typename T>
bool slab(const T* bmin, const T* bmax, const T* o, const T* d, T& tn, T& tf) {
  T f0 = T(1) / d[0], f1 = T(1) / d[1], f2 = T(1) / d[2];
  T t1 = (bmin[0] - o[0]) * f0, t2 = (bmax[0] - o[0]) * f0;
  T t3 = (bmin[1] - o[1]) * f1, t4 = (bmax[1] - o[1]) * f1;
  T t5 = (bmin[2] - o[2]) * f2, t6 = (bmax[2] - o[2]) * f2;
  T tmin = std::fmax(std::fmax(std::fmin(t1, t2), std::fmin(t3, t4)), std::fmin(t5, t6));
  T tmax = std::fmin(std::fmin(std::fmax(t1, t2), std::fmax(t3, t4)), std::fmax(t5, t6));
  if (tmax < 0) return false;
  if (tmin > tmax) return false;
  tn = tmin >= 0 ? tmin : T(0);
  tf = tmax;
  return true;
}

// single_ray_sphere_trace (RayEngine/src/cuda-src/sphere_tracing.cuh:182-216)
// with a step cap: sample, step, then test on the stepped t.  Normal:
// compute_sdf_normal (sdf_functions.cuh:506-516) on the trilinear field.
template <typename T>
void march(const Grid<T>& G, const T* bmin, const T* bmax, const T* o, const T* d, int max_steps, T eps, T scale,
           int rule, T& t_out, uint8_t& hit_out, int32_t& steps_out, T* nrm) {
  T tn, tf;
  bool hit = false;
  int steps = 0;
  T t = (T)1e20f;
  if (slab<T>(bmin, bmax, o, d, tn, tf)) {
    t = tn;
    while (t < tf && steps < max_steps) {
      T px = o[0] + t * d[0], py = o[1] + t * d[1], pz = o[2] + t * d[2];
      T dist = trilerp<T>(G, px, py, pz);
      T k = rule == 0 ? scale : std::fmax((T)0.45f, std::fmin(t * (T)0.6f, (T)0.8f));
      t += dist * k;
      ++steps;
      if (dist <= eps * t) { hit = true; break; }
    }
  }
  t_out = hit ? t : (T)1e20f;
  hit_out = hit ? 1 : 0;
  steps_out = steps;
  if (nrm) {
    T nx = 0, ny = 0, nz = 0;
    if (hit) {
      T delta = (T)std::fmin((float)(0.001 * (double)t), 0.002f);
      T px = o[0] + t * d[0], py = o[1] + t * d[1], pz = o[2] + t * d[2];
      T c = trilerp<T>(G, px, py, pz);
      nx = trilerp<T>(G, px + delta, py + T(0), pz + T(0)) - c;
      ny = trilerp<T>(G, px + T(0), py + delta, pz + T(0)) - c;
      nz = trilerp<T>(G, px + T(0), py + T(0), pz + delta) - c;
      T inv = T(1) / std::sqrt(nx * nx + ny * ny + nz * nz);
      nx = nx * inv; ny = ny * inv; nz = nz * inv;
    }
    nrm[0] = nx; nrm[1] = ny; nrm[2] = nz;
  }
}

// The library's exact-field march (device_paths.hpp exact_march; the
// reference has no counterpart, its tracer is march above): from the box
// entry step max(f * k_exact, h_min) while the trilinear field stays
// positive, bisect the first sign change n_bisect times, and take the
// midpoint of the last bracket.  A ray entering at f <= 0 hits at the entry.
template <typename T>
bool exact_march(const Grid<T>& G, const T* bmin, const T* bmax, const T* o, const T* d, T k_exact, T h_min,
                 int n_bisect, int max_iters, T& t_out, int32_t& samples) {
  T tn, tf;
  samples = 0;
  t_out = (T)1e20f;
  if (!slab<T>(bmin, bmax, o, d, tn, tf)) return false;
  auto f = [&](T t) { return trilerp<T>(G, o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2]); };
  T t = tn;
  T f_prev = f(t);
  ++samples;
  if (f_prev <= T(0)) {
    t_out = t;
    return true;
  }
  for (int it = 0; it < max_iters; ++it) {
    const T h = std::fmax(f_prev * k_exact, h_min);
    const T t_next = t + h;
    const T f_next = f(t_next);
    ++samples;
    if (f_next <= T(0)) {
      T a = t, b = t_next;
      for (int k = 0; k < n_bisect; ++k) {
        const T mid = T(0.5) * (a + b);
        const T fm = f(mid);
        ++samples;
        if (fm > T(0)) a = mid; else b = mid;
      }
      t_out = T(0.5) * (a + b);
      return true;
    }
    t = t_next;
    f_prev = f_next;
    if (t >= tf) return false;
  }
  return false;
}

}  // namespace

// Surrogate network value at one point in FP32 (oracle_mlp.cpp).
extern "C" float orc_mlp_eval_f32(const float* params, float x, float y, float z);

ORC_API float orc_grad_dot(uint32_t h, float x, float y, float z) { return gradient_dot<float>(h, x, y, z); }

ORC_API float orc_noise_eval(const uint32_t* perm, float x, float y, float z) { return eval<float>(perm, x, y, z); }
ORC_API double orc_noise_eval_f64(const uint32_t* perm, double x, double y, double z) {
  return eval<double>(perm, x, y, z);
}

// Terrain field sampled at every voxel p = origin + (x,y,z) * spacing (the
// generate_noise loop, RayEngine/src/cuda-src/renderer.cu:105-126; origin is
// the chunk location cuda_update_chunk_gen passes, :982-1000; null = 0).
ORC_API void orc_noise_fill(const uint32_t* perm, const int32_t* dim, const float* spacing, const float* origin,
                            int octaves, float lacunarity, float gain, float base_freq, float y_offset, float* out,
                            int threads) {
  const float zero[3] = {0.0f, 0.0f, 0.0f};
  const float* o = origin ? origin : zero;
  Fbm fp{octaves, lacunarity, gain, base_freq, y_offset};
  const int64_t n = (int64_t)dim[0] * dim[1] * dim[2];
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(threads)
#endif
  for (int64_t i = 0; i < n; ++i) {
    int x = (int)(i % dim[0]), y = (int)((i / dim[0]) % dim[1]), z = (int)(i / ((int64_t)dim[0] * dim[1]));
    out[i] = field<float>(perm, fp, o[0] + (float)x * spacing[0], o[1] + (float)y * spacing[1],
                          o[2] + (float)z * spacing[2]);
  }
}

ORC_API void orc_noise_fill_f64(const uint32_t* perm, const int32_t* dim, const float* spacing, const float* origin,
                                int octaves, float lacunarity, float gain, float base_freq, float y_offset, double* out,
                                int threads) {
  const float zero[3] = {0.0f, 0.0f, 0.0f};
  const float* o = origin ? origin : zero;
  Fbm fp{octaves, lacunarity, gain, base_freq, y_offset};
  const int64_t n = (int64_t)dim[0] * dim[1] * dim[2];
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(threads)
#endif
  for (int64_t i = 0; i < n; ++i) {
    int x = (int)(i % dim[0]), y = (int)((i / dim[0]) % dim[1]), z = (int)(i / ((int64_t)dim[0] * dim[1]));
    out[i] = field<double>(perm, fp, (double)o[0] + (double)x * spacing[0], (double)o[1] + (double)y * spacing[1],
                           (double)o[2] + (double)z * spacing[2]);
  }
}

ORC_API float orc_trilinear(const float* grid, const int32_t* dim, const float* spacing, float x, float y, float z) {
  Grid<float> G{grid, dim[0], dim[1], dim[2], spacing[0], spacing[1], spacing[2]};
  return trilerp<float>(G, x, y, z);
}

// FP32 sphere trace over a linear grid, volume box [0, dim * spacing].
ORC_API void orc_sdf_trace(const float* grid, const int32_t* dim, const float* spacing, const float* ray_o,
                           const float* ray_d, int64_t n, int max_steps, float eps, float scale, int rule, float* t,
                           uint8_t* hit, int32_t* steps, float* nrm, int threads) {
  Grid<float> G{grid, dim[0], dim[1], dim[2], spacing[0], spacing[1], spacing[2]};
  float bmin[3] = {0, 0, 0};
  float bmax[3] = {(float)dim[0] * spacing[0], (float)dim[1] * spacing[1], (float)dim[2] * spacing[2]};
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
#endif
  for (int64_t i = 0; i < n; ++i)
    march<float>(G, bmin, bmax, ray_o + 3 * i, ray_d + 3 * i, max_steps, eps, scale, rule, t[i], hit[i], steps[i],
                 nrm ? nrm + 3 * i : nullptr);
}

// FP64 twin: double grid, double ray, double march.
ORC_API void orc_sdf_trace_f64(const double* grid, const int32_t* dim, const float* spacing, const float* ray_o,
                               const float* ray_d, int64_t n, int max_steps, float eps, float scale, int rule,
                               double* t, uint8_t* hit, int32_t* steps, int threads) {
  Grid<double> G{grid, dim[0], dim[1], dim[2], spacing[0], spacing[1], spacing[2]};
  double bmin[3] = {0, 0, 0};
  double bmax[3] = {(double)dim[0] * spacing[0], (double)dim[1] * spacing[1], (double)dim[2] * spacing[2]};
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
#endif
  for (int64_t i = 0; i < n; ++i) {
    double o[3] = {ray_o[3 * i], ray_o[3 * i + 1], ray_o[3 * i + 2]};
    double d[3] = {ray_d[3 * i], ray_d[3 * i + 1], ray_d[3 * i + 2]};
    march<double>(G, bmin, bmax, o, d, max_steps, (double)eps, (double)scale, rule, t[i], hit[i], steps[i], nullptr);
  }
}

// Ground truth for the accuracy of a terrain tracer: the first crossing of
// the zero set along each ray, in FP64 (double grid, double ray), found by
// marching fixed steps of h_voxels voxels from the volume entry and
// bisecting the first sign change 40 times.  A ray that starts inside
// (f <= 0 at the entry) hits at the entry; misses return 1e20.
ORC_API void orc_truth_trace(const double* grid, const int32_t* dim, const float* spacing, const float* ray_o,
                             const float* ray_d, int64_t n, double h_voxels, double* t_out, uint8_t* hit_out,
                             int threads) {
  Grid<double> G{grid, dim[0], dim[1], dim[2], spacing[0], spacing[1], spacing[2]};
  double bmin[3] = {0, 0, 0};
  double bmax[3] = {(double)dim[0] * spacing[0], (double)dim[1] * spacing[1], (double)dim[2] * spacing[2]};
  const double h = h_voxels * (double)spacing[0];
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
#endif
  for (int64_t i = 0; i < n; ++i) {
    double o[3] = {ray_o[3 * i], ray_o[3 * i + 1], ray_o[3 * i + 2]};
    double d[3] = {ray_d[3 * i], ray_d[3 * i + 1], ray_d[3 * i + 2]};
    double tn, tf;
    t_out[i] = 1e20;
    hit_out[i] = 0;
    if (!slab<double>(bmin, bmax, o, d, tn, tf)) continue;
    auto f = [&](double t) { return trilerp<double>(G, o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2]); };
    double t = tn, fp = f(t);
    if (fp <= 0) { t_out[i] = t; hit_out[i] = 1; continue; }
    while (t < tf) {
      double tq = t + h, fq = f(tq);
      if (fq <= 0) {
        double a = t, b = tq;
        for (int k = 0; k < 40; ++k) {
          double m = 0.5 * (a + b);
          if (f(m) > 0) a = m; else b = m;
        }
        t_out[i] = 0.5 * (a + b);
        hit_out[i] = 1;
        break;
      }
      t = tq;
    }
  }
}

// ---------------------------------------------------------------- SDF edit

namespace {

// A (distance, material) pair as the reference's float2.
struct Pair {
  float d, m;
};

// sdf_fminf / sdf_fmaxf select a whole pair by distance; fminf on float2 is
// componentwise (RayEngine/src/cuda-src/sdf_functions.cuh:270-280).
Pair pick_smaller(Pair a, Pair b) { return a.d < b.d ? a : b; }
Pair pick_larger(Pair a, Pair b) { return a.d > b.d ? a : b; }

// sdf_smin (sdf_functions.cuh:283-289): the smaller pair, its distance
// lowered by h^2 k / 4 with h = max(k - |a - b|, 0) / k.
Pair smooth_min(Pair a, Pair b, float k) {
  const float h = std::fmax(k - std::fabs(a.d - b.d), 0.0f) / k;
  const Pair s = pick_smaller(a, b);
  return Pair{s.d - h * h * k * 0.25f, s.m};
}

// sphere_distance and aabb_distance (sdf_functions.cuh:422-437).
float norm3(float x, float y, float z) { return std::sqrt(x * x + y * y + z * z); }
float sphere_dist(float x, float y, float z, float r) { return norm3(x, y, z) - r; }
float aabb_dist(float x, float y, float z, const float* s) {
  const float dx = std::fabs(x) - s[0], dy = std::fabs(y) - s[1], dz = std::fabs(z) - s[2];
  return norm3(std::fmax(dx, 0.0f), std::fmax(dy, 0.0f), std::fmax(dz, 0.0f)) +
         std::fmin(std::fmax(dx, std::fmax(dy, dz)), 0.0f);
}

}  // namespace

// insert_sphere_to_texture (RayEngine/src/cuda-src/renderer.cu:30-92) once the
// brush ray has found its hit point `center`: for every voxel, p = voxel
// point - center (x and z snapped to multiples of extent[0] when snap), then
// SPHERE_ADD smooth-min, SPHERE_SUBTRACT / CUBE_SUBTRACT pair-max against the
// negated brush distance, CUBE_ADD componentwise min.  dist and mat are
// linear x-fastest grids, edited in place.
ORC_API void orc_volume_edit(float* dist, float* mat, const int32_t* dim, const float* spacing, int type,
                             const float* center, const float* extent, float material, int snap, float k) {
  for (int z = 0; z < dim[2]; ++z)
    for (int y = 0; y < dim[1]; ++y)
      for (int x = 0; x < dim[0]; ++x) {
        const int64_t i = ((int64_t)z * dim[1] + y) * dim[0] + x;
        float px = (float)x * spacing[0] - center[0];
        const float py = (float)y * spacing[1] - center[1];
        float pz = (float)z * spacing[2] - center[2];
        if (snap) {
          px = std::floor(px - std::fabs(std::fmod(px, extent[0])));
          pz = std::floor(pz - std::fabs(std::fmod(pz, extent[0])));
        }
        const Pair a{dist[i], mat[i]};
        Pair r;
        switch (type) {
          case 0: r = smooth_min(a, Pair{sphere_dist(px, py, pz, extent[0]), material}, k); break;
          case 1: r = pick_larger(a, Pair{-sphere_dist(px, py, pz, extent[0]), material}); break;
          case 2: r = Pair{std::fmin(a.d, aabb_dist(px, py, pz, extent)), std::fmin(a.m, material)}; break;
          default: r = pick_larger(a, Pair{-aabb_dist(px, py, pz, extent), material}); break;
        }
        dist[i] = r.d;
        mat[i] = r.m;
      }
}

// Exact-field march over a linear FP32 grid, volume box [0, dim * spacing];
// distances in world units.
ORC_API void orc_exact_march(const float* grid, const int32_t* dim, const float* spacing, const float* ray_o,
                             const float* ray_d, int64_t n, float k_exact, float h_min, int n_bisect, int max_iters,
                             float* t, uint8_t* hit, int32_t* samples, int threads) {
  Grid<float> G{grid, dim[0], dim[1], dim[2], spacing[0], spacing[1], spacing[2]};
  const float bmin[3] = {0, 0, 0};
  const float bmax[3] = {(float)dim[0] * spacing[0], (float)dim[1] * spacing[1], (float)dim[2] * spacing[2]};
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
#endif
  for (int64_t i = 0; i < n; ++i) {
    float tt;
    int32_t s;
    const bool h = exact_march<float>(G, bmin, bmax, ray_o + 3 * i, ray_d + 3 * i, k_exact, h_min, n_bisect,
                                      max_iters, tt, s);
    t[i] = tt;
    hit[i] = h ? 1 : 0;
    if (samples) samples[i] = s;
  }
}

// The library's hybrid tracer per ray (surrogate.hip hybrid_trace_kernel,
// whose wave scheduling leaves every ray's steps as written here; no
// reference counterpart), with the network evaluated in FP32 instead of the
// device's BF16.  Each advance: in the surrogate phase one network value s at
// the current point: s > tau steps t += s - margin (t >= t_far ends with a
// miss), else the exact phase starts here (f <= 0 at this point is a hit at
// t); in the exact phase one exact-march step (as exact_march) in the same
// advance; after rejoin of exact travel without a crossing the surrogate
// phase resumes.  At most max_iters advances.  net = NULL gives a network
// whose value is the constant const_s (tests pin the state machine with it).
ORC_API void orc_hybrid_trace(const float* grid, const int32_t* dim, const float* spacing, const float* net,
                              float const_s, const float* ray_o, const float* ray_d, int64_t n, float tau,
                              float margin, float k_exact, float h_min, float rejoin, int n_bisect, int max_iters,
                              float* t_out, uint8_t* hit_out, int32_t* evals_out, int32_t* samples_out,
                              int threads) {
  Grid<float> G{grid, dim[0], dim[1], dim[2], spacing[0], spacing[1], spacing[2]};
  const float bmin[3] = {0, 0, 0};
  const float bmax[3] = {(float)dim[0] * spacing[0], (float)dim[1] * spacing[1], (float)dim[2] * spacing[2]};
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
#endif
  for (int64_t i = 0; i < n; ++i) {
    const float* o = ray_o + 3 * i;
    const float* d = ray_d + 3 * i;
    auto f = [&](float t) { return trilerp<float>(G, o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2]); };
    enum { SURR, EXACT, DONE };
    int phase = DONE;
    bool hit = false;
    float t = 0.0f, t_far = 0.0f, f_prev = 0.0f, travel = 0.0f, tn, tf;
    int32_t evals = 0, samples = 0, advances = 0;
    if (slab<float>(bmin, bmax, o, d, tn, tf)) {
      t = tn;
      t_far = tf;
      phase = SURR;
    }
    while (phase != DONE) {
      if (advances >= max_iters) break;
      if (phase == SURR) {
        const float px = o[0] + t * d[0], py = o[1] + t * d[1], pz = o[2] + t * d[2];
        const float s = net ? orc_mlp_eval_f32(net, px, py, pz) : const_s;
        ++evals;
        if (s > tau) {
          t += s - margin;
          if (t >= t_far) phase = DONE;
        } else {
          phase = EXACT;
          travel = 0.0f;
          f_prev = f(t);
          ++samples;
          if (f_prev <= 0.0f) {
            hit = true;
            phase = DONE;
          }
        }
      }
      if (phase == EXACT) {
        const float h = std::fmax(f_prev * k_exact, h_min);
        const float t_next = t + h;
        const float f_next = f(t_next);
        ++samples;
        if (f_next <= 0.0f) {
          float a = t, b = t_next;
          for (int k = 0; k < n_bisect; ++k) {
            const float mid = 0.5f * (a + b);
            const float fm = f(mid);
            ++samples;
            if (fm > 0.0f) a = mid; else b = mid;
          }
          t = 0.5f * (a + b);
          hit = true;
          phase = DONE;
        } else {
          t = t_next;
          f_prev = f_next;
          travel += h;
          if (t >= t_far) phase = DONE;
          else if (travel > rejoin) phase = SURR;
        }
      }
      ++advances;
    }
    t_out[i] = hit ? t : 1e20f;
    hit_out[i] = hit ? 1 : 0;
    if (evals_out) evals_out[i] = evals;
    if (samples_out) samples_out[i] = samples;
  }
}
