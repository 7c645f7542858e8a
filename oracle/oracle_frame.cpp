// CPU oracle, frame part.  TEST INFRASTRUCTURE ONLY (see oracle.cpp).
//
// Restates the per-pixel tail of render_sphere_trace
// (RayEngine/src/cuda-src/renderer.cu:257-283): the single-scattering sky of
// compute_incident_light (material_functions.cuh:295-344) with the
// Atmosphere constants (Atmosphere.cuh), its tone curve, simple_shade for
// hits (cuda_helper_functions.h:137-140, mix at :71-76), gamma, vignetting,
// clip and rgbaFloatToInt (:60-68).  FP32 with libm's expf / powf where the
// reference uses __expf / pow; an FP64 twin of the sky measures its
// rounding.
#include <cmath>
#include <cstdint>

#define ORC_API extern "C" __attribute__((visibility("default")))

namespace {

// Same layout as re_frame_params.
struct FrameParams {
  float sun[3], er, ar, hr, hm, origin[3], base[3];
  int32_t gamma, vignetting;
  float vignetting_k;
};

template <typename T>
struct V {
  T x, y, z;
};
template <typename T>
V<T> add(V<T> a, V<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T>
V<T> scale(T s, V<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename T>
T dot(V<T> a, V<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// solve_quadratic + gpu_ray_sphere_intersect (ray_functions.cuh:95-126).
template <typename T>
bool sphere_hit(V<T> o, V<T> d, T radius, T& t0, T& t1) {
  const T a = dot(d, d), b = 2 * dot(d, o), c = dot(o, o) - radius * radius;
  const T discr = b * b - (T)4 * a * c;
  if (discr < 0) return false;
  if (discr == 0) {
    t0 = t1 = (T)-0.5 * b / a;
  } else {
    const T q = (b > 0) ? (T)-0.5 * (b + std::sqrt(discr)) : (T)-0.5 * (b - std::sqrt(discr));
    t0 = q / a;
    t1 = c / q;
  }
  if (t0 > t1) std::swap(t0, t1);
  return true;
}

template <typename T>
V<T> sky(const FrameParams& P, const T* betaR, T betaM, V<T> orig, V<T> dir) {
  const V<T> sun{(T)P.sun[0], (T)P.sun[1], (T)P.sun[2]};
  T tmin = 0, tmax = (T)INFINITY, t0 = 0, t1 = 0;
  if (!sphere_hit(orig, dir, (T)P.ar, t0, t1) || t1 < 0) return {1, 1, 1};
  if (t0 > tmin && t0 > 0) tmin = t0;
  if (t1 < tmax) tmax = t1;
  const T seg = (tmax - tmin) / (T)32;
  T tc = tmin, odr = 0, odm = 0;
  V<T> sr{0, 0, 0}, sm{0, 0, 0};
  const T mu = dot(dir, sun);
  const T pr = (T)(3.0 / (16.0 * M_PI) * (double)(1 + mu * mu));
  const T g = (T)0.76f;
  const T pm = (T)(3.0 / (8.0 * M_PI) * (double)((1 - g * g) * (1 + mu * mu)) /
                   (double)((2 + g * g) * std::pow(1 + g * g - 2 * g * mu, (T)1.5)));
  for (int i = 0; i < 32; ++i) {
    const V<T> sp = add(orig, scale(tc + seg * (T)0.5, dir));
    const T h = std::sqrt(dot(sp, sp)) - (T)P.er;
    const T hr = std::exp(-h / (T)P.hr) * seg, hm = std::exp(-h / (T)P.hm) * seg;
    odr += hr;
    odm += hm;
    T l0 = 0, l1 = 0;
    sphere_hit(sp, sun, (T)P.ar, l0, l1);
    const T segl = l1 / (T)8;
    T tl = 0, olr = 0, olm = 0;
    int j;
    for (j = 0; j < 8; ++j) {
      const V<T> spl = add(sp, scale(tl + segl * (T)0.5, sun));
      const T hl = std::sqrt(dot(spl, spl)) - (T)P.er;
      if (hl < 0) break;
      olr += std::exp(-hl / (T)P.hr) * segl;
      olm += std::exp(-hl / (T)P.hm) * segl;
      tl += segl;
    }
    if (j == 8) {
      const T m = betaM * (T)1.1f * (odm + olm);
      const T ax = std::exp(-(betaR[0] * (odr + olr) + m)), ay = std::exp(-(betaR[1] * (odr + olr) + m)),
              az = std::exp(-(betaR[2] * (odr + olr) + m));
      sr = add(sr, V<T>{ax * hr, ay * hr, az * hr});
      sm = add(sm, V<T>{ax * hm, ay * hm, az * hm});
    }
    tc += seg;
  }
  const T km = 100 * betaM * pm;
  return {(sr.x * betaR[0] * pr + sm.x + km) * 20, (sr.y * betaR[1] * pr + sm.y + km) * 20,
          (sr.z * betaR[2] * pr + sm.z + km) * 20};
}

float sat(float x) { return x > 0.0f ? (x < 1.0f ? x : 1.0f) : 0.0f; }
float tone(float c) { return c < 1.413f ? std::pow(c * 0.38317f, 1.0f / 2.2f) : 1.0f - std::exp(-c); }

const float kBetaR[3] = {3.8e-6f, 13.5e-6f, 33.1e-6f};
const float kBetaM = 21e-6f;

}  // namespace

// Sky radiance of n directions before the tone curve, FP32, with the given
// scattering coefficients (the Atmosphere's are 3.8e-6, 13.5e-6, 33.1e-6 and
// 21e-6).
ORC_API void orc_sky(const FrameParams* P, const float* beta_r, float beta_m, const float* dirs, int64_t n,
                     float* out) {
  const V<float> o{P->origin[0], P->origin[1], P->origin[2]};
  for (int64_t i = 0; i < n; ++i) {
    const V<float> c = sky<float>(*P, beta_r, beta_m, o, V<float>{dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]});
    out[3 * i] = c.x;
    out[3 * i + 1] = c.y;
    out[3 * i + 2] = c.z;
  }
}

ORC_API void orc_sky_f64(const FrameParams* P, const float* beta_r, float beta_m, const float* dirs, int64_t n,
                         double* out) {
  const double br[3] = {beta_r[0], beta_r[1], beta_r[2]};
  const V<double> o{P->origin[0], P->origin[1], P->origin[2]};
  for (int64_t i = 0; i < n; ++i) {
    const V<double> c = sky<double>(*P, br, (double)beta_m, o, V<double>{dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]});
    out[3 * i] = c.x;
    out[3 * i + 1] = c.y;
    out[3 * i + 2] = c.z;
  }
}

ORC_API float orc_sky_tone(float c) { return tone(c); }

ORC_API void orc_frame_shade(const float* ray_d, const uint8_t* hit, const float* normal, int32_t width, int32_t height,
                             const FrameParams* P, uint32_t* rgba) {
  const V<float> o{P->origin[0], P->origin[1], P->origin[2]};
  for (int32_t y = 0; y < height; ++y)
    for (int32_t x = 0; x < width; ++x) {
      const int64_t i = (int64_t)y * width + x;
      const float u = x / (float)width, v = y / (float)height;
      const V<float> d{ray_d[3 * i], ray_d[3 * i + 1], ray_d[3 * i + 2]};
      float c[3];
      if (hit[i]) {
        const V<float> nrm{normal[3 * i], normal[3 * i + 1], normal[3 * i + 2]};
        const float f = std::fmax(0.f, dot(nrm, V<float>{-d.x, -d.y, -d.z}) / 2);
        for (int k = 0; k < 3; ++k) c[k] = sat(P->base[k] + 0.5f * f);
      } else {
        const V<float> s = sky<float>(*P, kBetaR, kBetaM, o, d);
        c[0] = tone(s.x);
        c[1] = tone(s.y);
        c[2] = tone(s.z);
      }
      if (P->gamma)
        for (float& q : c) q = std::pow(q, 1.0f / 2.2f);
      if (P->vignetting) {
        const float w = 0.5f + 0.5f * std::pow(16.0f * u * v * (1.0f - u) * (1.0f - v), P->vignetting_k);
        for (float& q : c) q *= w;
      }
      for (float& q : c) q = sat(q);
      rgba[i] = (255u << 24) | ((uint32_t)(c[2] * 255) << 16) | ((uint32_t)(c[1] * 255) << 8) | (uint32_t)(c[0] * 255);
    }
}
