// CPU oracle, SDF surrogate part.  TEST INFRASTRUCTURE ONLY (see oracle.cpp).
//
// The surrogate has no reference code.  This FP64 restatement of its
// definition is the checker for the GPU forward, backward and Adam kernels:
//   features(p) = [sin(2 pi f_k p_d), cos(2 pi f_k p_d)] for d < 3, k < 16,
//                 feature index 32 d + 2 k + {0: sin, 1: cos}, f_k = 2^(k/2);
//                 then p.x, p.y, p.z (indices 96..98)
//   h_0 = silu(W0 x + b0), h_{l} = silu(W_l h_{l-1} + b_l) for l = 1..3,
//   out = W4 . h_3 + b4,   silu(z) = z / (1 + exp(-z))
// Parameters are one flat float vector:
//   W0[64][99] b0[64] W1[64][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64] b4[1]
// Loss over a batch: mean_i w_i (out_i - y_i)^2 with w_i = over_weight when
// out_i > y_i (an over-estimate of the distance) and 1 otherwise.
// Pinned by finite differences of the FP64 loss (tests/test_mlp_oracle.py).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#define ORC_API extern "C" __attribute__((visibility("default")))

namespace {

constexpr int kIn = 99, kH = 64, kFreq = 16;
constexpr int oW0 = 0, ob0 = oW0 + kH * kIn, oW1 = ob0 + kH, ob1 = oW1 + kH * kH, oW2 = ob1 + kH,
              ob2 = oW2 + kH * kH, oW3 = ob2 + kH, ob3 = oW3 + kH * kH, oW4 = ob3 + kH, ob4 = oW4 + kH,
              kParams = ob4 + 1;

void features(const double* p, double* x) {
  for (int d = 0; d < 3; ++d)
    for (int k = 0; k < kFreq; ++k) {
      double f = std::pow(2.0, 0.5 * k);
      double a = 2.0 * M_PI * f * p[d];
      x[32 * d + 2 * k] = std::sin(a);
      x[32 * d + 2 * k + 1] = std::cos(a);
    }
  x[96] = p[0];
  x[97] = p[1];
  x[98] = p[2];
}

inline double sigmoid(double z) { return 1.0 / (1.0 + std::exp(-z)); }
inline double act(double z) { return z * sigmoid(z); }
// d silu / dz = s (1 + z (1 - s)), s = sigmoid(z).
inline double dact(double z) {
  const double s = sigmoid(z);
  return s * (1.0 + z * (1.0 - s));
}

struct Acts {
  double x[kIn];
  double z[4][kH];
  double h[4][kH];
  double out;
};

void forward(const double* P, const double* p, Acts& A) {
  features(p, A.x);
  const int oW[4] = {oW0, oW1, oW2, oW3}, ob[4] = {ob0, ob1, ob2, ob3};
  for (int l = 0; l < 4; ++l) {
    const int nin = l == 0 ? kIn : kH;
    const double* in = l == 0 ? A.x : A.h[l - 1];
    for (int i = 0; i < kH; ++i) {
      double s = P[ob[l] + i];
      for (int j = 0; j < nin; ++j) s += P[oW[l] + i * nin + j] * in[j];
      A.z[l][i] = s;
      A.h[l][i] = act(s);
    }
  }
  double o = P[ob4];
  for (int j = 0; j < kH; ++j) o += P[oW4 + j] * A.h[3][j];
  A.out = o;
}

}  // namespace

ORC_API int orc_mlp_n_params(void) { return kParams; }

ORC_API void orc_mlp_features(const float* pts, int64_t n, double* out) {
  for (int64_t i = 0; i < n; ++i) {
    double p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    features(p, out + kIn * i);
  }
}

// Forward pass in FP64 with the given (FP32-stored) parameters.
ORC_API void orc_mlp_forward(const float* params, const float* pts, int64_t n, double* out) {
  std::vector<double> P(params, params + kParams);
  Acts A;
  for (int64_t i = 0; i < n; ++i) {
    double p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    forward(P.data(), p, A);
    out[i] = A.out;
  }
}

// Loss and its gradient over a batch, FP64.
ORC_API double orc_mlp_loss_grad(const float* params, const float* pts, const float* target, int64_t n,
                                 float over_weight, double* grad) {
  std::vector<double> P(params, params + kParams);
  std::memset(grad, 0, sizeof(double) * kParams);
  Acts A;
  double loss = 0.0;
  const int oW[4] = {oW0, oW1, oW2, oW3}, ob[4] = {ob0, ob1, ob2, ob3};
  for (int64_t i = 0; i < n; ++i) {
    double p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    forward(P.data(), p, A);
    double r = A.out - (double)target[i];
    double w = r > 0.0 ? (double)over_weight : 1.0;
    loss += w * r * r / (double)n;
    double dout = 2.0 * w * r / (double)n;
    grad[ob4] += dout;
    double dh[kH], dz[kH];
    for (int j = 0; j < kH; ++j) {
      grad[oW4 + j] += dout * A.h[3][j];
      dh[j] = dout * P[oW4 + j];
    }
    for (int l = 3; l >= 0; --l) {
      const int nin = l == 0 ? kIn : kH;
      const double* in = l == 0 ? A.x : A.h[l - 1];
      for (int k = 0; k < kH; ++k) {
        dz[k] = dh[k] * dact(A.z[l][k]);
      }
      for (int k = 0; k < kH; ++k) {
        grad[ob[l] + k] += dz[k];
        for (int j = 0; j < nin; ++j) grad[oW[l] + k * nin + j] += dz[k] * in[j];
      }
      if (l > 0) {
        for (int j = 0; j < kH; ++j) {
          double s = 0.0;
          for (int k = 0; k < kH; ++k) s += P[oW[l] + k * kH + j] * dz[k];
          dh[j] = s;
        }
      }
    }
  }
  return loss;
}

// One Adam step (Kingma & Ba 2015, Algorithm 1) on the flat parameters, in
// FP64 from FP32 state; `step` is the 1-based step count.
ORC_API void orc_adam(float* params, float* m, float* v, const float* grad, int n, int step, float lr, float b1,
                      float b2, float eps) {
  const double c1 = 1.0 - std::pow((double)b1, step), c2 = 1.0 - std::pow((double)b2, step);
  for (int i = 0; i < n; ++i) {
    double g = grad[i];
    double mi = (double)b1 * m[i] + (1.0 - (double)b1) * g;
    double vi = (double)b2 * v[i] + (1.0 - (double)b2) * g * g;
    m[i] = (float)mi;
    v[i] = (float)vi;
    params[i] = (float)((double)params[i] - (double)lr * (mi / c1) / (std::sqrt(vi / c2) + (double)eps));
  }
}

// The network in FP32 (float weights, features, accumulation and SiLU):
// the CPU stand-in for the device's BF16 evaluation in the hybrid tracer.
ORC_API float orc_mlp_eval_f32(const float* P, float x, float y, float z) {
  float in[kIn], h[kH], g[kH];
  const float p[3] = {x, y, z};
  const float two_pi = 6.28318530717958647692f;
  for (int d = 0; d < 3; ++d)
    for (int k = 0; k < kFreq; ++k) {
      const float a = two_pi * std::exp2(0.5f * (float)k) * p[d];
      in[32 * d + 2 * k] = std::sin(a);
      in[32 * d + 2 * k + 1] = std::cos(a);
    }
  in[96] = x;
  in[97] = y;
  in[98] = z;
  const int oW[4] = {oW0, oW1, oW2, oW3}, ob[4] = {ob0, ob1, ob2, ob3};
  const float* src = in;
  for (int l = 0; l < 4; ++l) {
    const int nin = l == 0 ? kIn : kH;
    for (int i = 0; i < kH; ++i) {
      const float* w = P + oW[l] + i * nin;
      float s = 0.0f;
      for (int j = 0; j < nin; ++j) s += w[j] * src[j];
      s += P[ob[l] + i];
      g[i] = s / (1.0f + std::exp(-s));
    }
    std::copy(g, g + kH, h);
    src = h;
  }
  float o = 0.0f;
  for (int j = 0; j < kH; ++j) o += P[oW4 + j] * h[j];
  return o + P[ob4];
}
