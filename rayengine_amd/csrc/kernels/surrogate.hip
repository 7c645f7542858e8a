// Learned SDF surrogate of the terrain field: training-data kernels (signed
// Euclidean distance transform of the field's zero set, point sampler), the
// fused MFMA forward/backward training kernel, the Adam + operand-packing
// kernel, and batched inference.  No reference counterpart; the CPU oracle's
// FP64 forward/backward checks these kernels.
//
// MFMA layout and the register chaining of layers: csrc/mlp_layout.hpp.
#include "../device_paths.hpp"
#include "../mlp_layout.hpp"

namespace re {
namespace {

using namespace mlp;

// Work-group size of the kernels that stage the network in LDS (42 KB per
// group): 8 waves share one copy, so LDS allows 3 groups = 24 waves per CU
// and the register budget, not LDS, sets the occupancy.
constexpr int kSurrBlock = 512;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// sin(2 pi x) and cos(2 pi x): v_sin_f32 / v_cos_f32 take the angle in
// revolutions; the argument is reduced to [0, 1) first.
__device__ __forceinline__ float sin_rev(float x) { return __builtin_amdgcn_sinf(__builtin_amdgcn_fractf(x)); }
__device__ __forceinline__ float cos_rev(float x) { return __builtin_amdgcn_cosf(__builtin_amdgcn_fractf(x)); }

// Hidden activation, SiLU: z * sigmoid(z), with sigmoid(z) =
// 1 / (1 + 2^(-z log2 e)) from the hardware exp2 and reciprocal (about 1 ulp
// each; the activations are rounded to BF16 anyway), not the IEEE division
// sequence.
__device__ __forceinline__ float sigmoid_fast(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * -1.44269504089f));
}
__device__ __forceinline__ float silu(float z) { return z * sigmoid_fast(z); }
// The same unit on the scaled pre-activation u = -log2(e) z of the inference
// operands: u / (1 + 2^u) = -log2(e) SiLU(z).
__device__ __forceinline__ float silu_scaled(float u) { return u * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(u)); }

// Its derivative: s (1 + z (1 - s)) with s = sigmoid(z).
__device__ __forceinline__ float dsilu(float z) {
  const float s = sigmoid_fast(z);
  return s * (1.0f + z * (1.0f - s));
}

// Device PCG32 (same generator as the host scene generators).
struct DevPcg {
  uint64_t state, inc;
  __device__ DevPcg(uint64_t initstate, uint64_t initseq) {
    state = 0u;
    inc = (initseq << 1u) | 1u;
    next();
    state += initstate;
    next();
  }
  __device__ uint32_t next() {
    uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
  }
  __device__ float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
};

// ----------------------------------------------------------------- features

// The B operands of layer 0 for this lane's point: K-step c < 3 holds the
// sin/cos pairs of frequencies 4g..4g+3 of coordinate c (f_k = 2^(k/2), so
// f_{4g+2} = 2 f_{4g} and f_{4g+3} = 2 f_{4g+1}: those two pairs come from
// the double-angle identities, which keeps 4 of the 8 transcendentals);
// K-step 3 holds the raw coordinates (g = 0, s < 3) and zero padding.
__device__ __forceinline__ void features(f3 p, int g, bf16x8 xb[4]) {
  const float pc[3] = {p.x, p.y, p.z};
  const float f0 = (float)(1 << (2 * g));
  const float f1 = f0 * 1.41421356237f;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float s0 = sin_rev(f0 * pc[c]), c0 = cos_rev(f0 * pc[c]);
    const float s1 = sin_rev(f1 * pc[c]), c1 = cos_rev(f1 * pc[c]);
    xb[c][0] = (__bf16)s0;
    xb[c][1] = (__bf16)c0;
    xb[c][2] = (__bf16)s1;
    xb[c][3] = (__bf16)c1;
    xb[c][4] = (__bf16)(2.0f * s0 * c0);
    xb[c][5] = (__bf16)((c0 - s0) * (c0 + s0));
    xb[c][6] = (__bf16)(2.0f * s1 * c1);
    xb[c][7] = (__bf16)((c1 - s1) * (c1 + s1));
  }
#pragma unroll
  for (int s = 0; s < 8; ++s) xb[3][s] = (__bf16)((g == 0 && s < 3) ? pc[s] : 0.0f);
}

// ----------------------------------------------------------------- forward

struct TileAct {
  f32x4 z[4][4];    // pre-activations [layer][M-tile], lane holds neurons 16m+4g+r
  bf16x8 hb[4][2];  // bf16 B operands of h0..h3 (hb[3] feeds nothing but the staging)
  float h3[4][4];   // FP32 h3 for the output layer
};

// SCALED: the inference operands (pre-activations carried as -log2(e) z).
template <bool KEEP, bool SCALED = false>
__device__ __forceinline__ float forward_tile(const bf16x8* __restrict__ packF, const float* __restrict__ small,
                                              const bf16x8 xb[4], int lane, TileAct& T) {
  const int g = lane >> 4;
  bf16x8 in[2];
  // layer 0: 4 K-steps of 32 features
  {
    f32x4 acc[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      acc[m] = *reinterpret_cast<const f32x4*>(&small[16 * m + 4 * g]);  // the bias starts the sum
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[m] = mfma16(packF[(oPackF0 >> 3) + (m * 4 + c) * 64 + lane], xb[c], acc[m]);
    }
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float z = acc[m][r];
        if (KEEP) T.z[0][m][r] = z;
        T.hb[0][m >> 1][(m & 1) * 4 + r] = (__bf16)(SCALED ? silu_scaled(z) : silu(z));
      }
  }
  // layers 1..3: 2 K-steps of 32 neurons, chained in registers
#pragma unroll
  for (int L = 1; L < 4; ++L) {
    in[0] = T.hb[L - 1][0];
    in[1] = T.hb[L - 1][1];
    const int off = (oPackF1 + (L - 1) * kPackL) >> 3;
    f32x4 acc[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      acc[m] = *reinterpret_cast<const f32x4*>(&small[64 * L + 16 * m + 4 * g]);
#pragma unroll
      for (int c = 0; c < 2; ++c) acc[m] = mfma16(packF[off + (m * 2 + c) * 64 + lane], in[c], acc[m]);
    }
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float z = acc[m][r];
        if (KEEP) T.z[L][m][r] = z;
        const float h = SCALED ? silu_scaled(z) : silu(z);
        T.hb[L][m >> 1][(m & 1) * 4 + r] = (__bf16)h;
        if (L == 3) T.h3[m][r] = h;
      }
  }
  // output: partial dot over this lane's 16 neurons, summed over the 4 groups
  float o = 0.0f;
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int r = 0; r < 4; ++r) o += small[256 + 16 * m + 4 * g + r] * T.h3[m][r];
  o += __shfl_xor(o, 16);
  o += __shfl_xor(o, 32);
  return o + small[320];
}

// Copies the packed forward operands (and optionally the backward ones) and
// the small parameters into LDS.
__device__ __forceinline__ void load_weights(const bf16x8* __restrict__ gpack, const float* __restrict__ gsmall,
                                             bf16x8* spack, float* ssmall, int n_vec) {
  for (int i = threadIdx.x; i < n_vec; i += blockDim.x) spack[i] = gpack[i];
  for (int i = threadIdx.x; i < kSmall; i += blockDim.x) ssmall[i] = gsmall[i];
  __syncthreads();
}

// ----------------------------------------------------------------- inference

__global__ void __launch_bounds__(512) mlp_infer_kernel(const bf16x8* __restrict__ gpack,
                                                        const float* __restrict__ gsmall,
                                                        const float* __restrict__ pts, int64_t n,
                                                        float* __restrict__ out) {
  __shared__ bf16x8 spack[oPackB1 >> 3];
  __shared__ __align__(16) float ssmall[kSmall];
  load_weights(gpack, gsmall, spack, ssmall, oPackB1 >> 3);
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t n_waves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const int64_t tiles = (n + 15) / 16;
  for (int64_t tile = wave; tile < tiles; tile += n_waves) {
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");  // weight reads stay in the loop
    int64_t i = tile * 16 + (lane & 15);
    const int64_t ic = i < n ? i : n - 1;
    const f3 p{pts[3 * ic], pts[3 * ic + 1], pts[3 * ic + 2]};
    bf16x8 xb[4];
    features(p, lane >> 4, xb);
    TileAct T;
    const float o = forward_tile<false, true>(spack, ssmall, xb, lane, T);
    if (lane < 16 && i < n) out[i] = o;
  }
}

// ----------------------------------------------------------------- training

// Per 64-point chunk (16 per wave) every wave runs forward and backward on its
// tile and stages, transposed as [neuron][point] in LDS, the bf16 dZ of every
// layer and the input activations of layers 0..3.  Wave w then accumulates the
// weight gradient of layer w over the chunk's 64 points with K = points, and
// its bias gradient as the product with a column of ones.  The output layer's
// gradient is summed in registers, per lane over its own points, and reduced
// over lanes and waves once at the end.
constexpr int kChunkPts = 64;
// Staged rows are padded by 16 bytes (72 bf16): the 16 rows of one MFMA operand
// read then cover all 64 LDS banks once, and the four neuron groups of a
// staging write fall in different banks.
constexpr int kRow = kChunkPts + 8;
// Layer-0 inputs staged: the 99 used, padded to 7 column tiles of 16.
constexpr int kInStage = 112;
constexpr int kTrainBlock = 256;

struct TrainLds {
  bf16x8 pack[kPack >> 3];            // 64 KB: forward and backward operands
  float small[kSmall + 3];
  __bf16 dzT[4][kH][kRow];            // 36 KB
  __bf16 xT[kInStage][kRow];          // 15.75 KB
  __bf16 hT[3][kH][kRow];             // 27 KB (h0..h2, the inputs of layers 1..3)
};
static_assert(sizeof(TrainLds) <= 160 * 1024, "one training block must fit the LDS of a CU");
static_assert(sizeof(__bf16) * 4 * kH * kRow >= sizeof(float) * (4 * kH + 8),
              "the final reduction reuses the dZ staging area");
static_assert((kPack >> 3) % kTrainBlock == 0, "the operand pack is copied in whole rounds");

__device__ __forceinline__ int64_t train_sample_index(int64_t batch_start, int ch, int p_local, int64_t n_samples) {
  return (batch_start + (int64_t)ch * kChunkPts + p_local) % n_samples;
}

__global__ void __launch_bounds__(kTrainBlock) mlp_train_kernel(const bf16x8* __restrict__ gpack,
                                                                const float* __restrict__ gsmall,
                                                                const float4* __restrict__ samples, int64_t n_samples,
                                                                int64_t batch_start, int32_t batch, float over_weight,
                                                                float* __restrict__ partial) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  TrainLds& S = *reinterpret_cast<TrainLds*>(lds_raw);
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int g = lane >> 4, j = lane & 15;
  const int p_local = 16 * w + j;
  const float inv_b = 1.0f / (float)batch;
  const int n_chunks = batch / kChunkPts;

  // The first chunk's sample is requested before the operand copy, and the
  // copy issues all of a thread's loads before its first LDS store.
  float4 smp_next = make_float4(0.f, 0.f, 0.f, 0.f);
  if ((int)blockIdx.x < n_chunks) smp_next = samples[train_sample_index(batch_start, blockIdx.x, p_local, n_samples)];
  {
    constexpr int kPer = (kPack >> 3) / kTrainBlock;
    bf16x8 r[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) r[k] = gpack[threadIdx.x + kTrainBlock * k];
    for (int i = threadIdx.x; i < kSmall; i += kTrainBlock) S.small[i] = gsmall[i];
#pragma unroll
    for (int k = 0; k < kPer; ++k) S.pack[threadIdx.x + kTrainBlock * k] = r[k];
  }
  __syncthreads();

  f32x4 acc[4][7], accb[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    accb[a] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int b = 0; b < 7; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  }
  // B operand of the bias product: column 0 all ones, the other columns zero
  bf16x8 ones;
#pragma unroll
  for (int s = 0; s < 8; ++s) ones[s] = (__bf16)(j == 0 ? 1.0f : 0.0f);
  // per lane, over its points: dout * h3 for neurons 16 m + 4 g + r (index
  // 4 m + r), and dout (g == 0 lanes)
  float dw4r[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) dw4r[q] = 0.0f;
  float db4r = 0.0f, loss = 0.0f;

  for (int ch = blockIdx.x; ch < n_chunks; ch += gridDim.x) {
    const float4 smp = smp_next;
    if (ch + (int)gridDim.x < n_chunks)
      smp_next = samples[train_sample_index(batch_start, ch + gridDim.x, p_local, n_samples)];
    const f3 p{smp.x, smp.y, smp.z};
    bf16x8 xb[4];
    features(p, g, xb);
    TileAct T;
    const float o = forward_tile<true>(S.pack, S.small, xb, lane, T);
    const float r = o - smp.w;
    const float wgt = r > 0.0f ? over_weight : 1.0f;
    const float dout = 2.0f * wgt * r * inv_b;
    if (g == 0) {
      loss += wgt * r * r * inv_b;
      db4r += dout;
    }
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int r2 = 0; r2 < 4; ++r2) dw4r[4 * m + r2] += dout * T.h3[m][r2];

    // backward through the hidden layers, chaining dZ in registers
    float dz[4][4];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int r2 = 0; r2 < 4; ++r2) dz[m][r2] = dout * S.small[256 + 16 * m + 4 * g + r2] * dsilu(T.z[3][m][r2]);
#pragma unroll
    for (int L = 3; L >= 0; --L) {
      // stage dZ_L
#pragma unroll
      for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int r2 = 0; r2 < 4; ++r2) S.dzT[L][16 * m + 4 * g + r2][p_local] = (__bf16)dz[m][r2];
      if (L == 0) break;
      bf16x8 dzb[2];
#pragma unroll
      for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int r2 = 0; r2 < 4; ++r2) dzb[m >> 1][(m & 1) * 4 + r2] = (__bf16)dz[m][r2];
      const int off = (oPackB1 + (L - 1) * kPackL) >> 3;
      f32x4 dh[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        dh[m] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int c = 0; c < 2; ++c) dh[m] = mfma16(S.pack[off + (m * 2 + c) * 64 + lane], dzb[c], dh[m]);
      }
#pragma unroll
      for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int r2 = 0; r2 < 4; ++r2) dz[m][r2] = dh[m][r2] * dsilu(T.z[L - 1][m][r2]);
    }
    // stage the inputs of layers 0..3: features (rows 0..111), h0..h2
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (c == 3 && g >= 2) break;  // rows 112..127 are zero padding, never read
#pragma unroll
      for (int s = 0; s < 8; ++s) S.xT[32 * c + 8 * g + s][p_local] = xb[c][s];
    }
#pragma unroll
    for (int L = 0; L < 3; ++L)
#pragma unroll
      for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int r2 = 0; r2 < 4; ++r2) S.hT[L][16 * m + 4 * g + r2][p_local] = T.hb[L][m >> 1][(m & 1) * 4 + r2];
    __syncthreads();

    // weight gradient of layer w: dW = dZ^T (64 x pts) . H (pts x in), K = points
    const int n_tiles = (w == 0) ? 7 : 4;
#pragma unroll
    for (int kc = 0; kc < 2; ++kc) {
#pragma unroll
      for (int mi = 0; mi < 4; ++mi) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(&S.dzT[w][16 * mi + j][32 * kc + 8 * g]);
        accb[mi] = mfma16(a, ones, accb[mi]);
#pragma unroll
        for (int nj = 0; nj < 7; ++nj) {
          if (nj >= n_tiles) break;
          const __bf16* brow = (w == 0) ? &S.xT[16 * nj + j][32 * kc + 8 * g] : &S.hT[w - 1][16 * nj + j][32 * kc + 8 * g];
          acc[mi][nj] = mfma16(a, *reinterpret_cast<const bf16x8*>(brow), acc[mi][nj]);
        }
      }
    }
    __syncthreads();
  }

  // write this block's partial gradient (flat parameter layout) and loss
  float* P = partial + (int64_t)blockIdx.x * (kParams + 1);
  {
    const int nin = (w == 0) ? kIn : kH;
    const int oW = (w == 0) ? oW0 : (w == 1 ? oW1 : (w == 2 ? oW2 : oW3));
    const int ob = (w == 0) ? ob0 : (w == 1 ? ob1 : (w == 2 ? ob2 : ob3));
    const int n_tiles = (w == 0) ? 7 : 4;
    if (j == 0)
#pragma unroll
      for (int mi = 0; mi < 4; ++mi)
#pragma unroll
        for (int r2 = 0; r2 < 4; ++r2) P[ob + 16 * mi + 4 * g + r2] = accb[mi][r2];
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
      for (int nj = 0; nj < 7; ++nj) {
        if (nj >= n_tiles) break;
#pragma unroll
        for (int r2 = 0; r2 < 4; ++r2) {
          const int i = 16 * mi + 4 * g + r2, k = 16 * nj + j;
          if (k < nin) P[oW + i * nin + k] = acc[mi][nj][r2];
        }
      }
  }
  // the output layer: over the 16 points of a lane group, then over the four
  // waves (the loop ended on a barrier: the staging area is free)
  float* red = reinterpret_cast<float*>(&S.dzT[0][0][0]);  // [wave][64], then [wave][2]
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    float v = dw4r[q];
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 4);
    v += __shfl_xor(v, 8);
    if (j == 0) red[w * 64 + 16 * (q >> 2) + 4 * g + (q & 3)] = v;
  }
  float d4 = db4r, l = loss;
  for (int off = 32; off > 0; off >>= 1) {
    d4 += __shfl_xor(d4, off);
    l += __shfl_xor(l, off);
  }
  if (lane == 0) {
    red[4 * 64 + 2 * w] = d4;
    red[4 * 64 + 2 * w + 1] = l;
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    const int n = threadIdx.x;
    P[oW4 + n] = red[n] + red[64 + n] + red[128 + n] + red[192 + n];
    if (n == 0) {
      P[ob4] = red[256] + red[258] + red[260] + red[262];
      P[kParams] = red[257] + red[259] + red[261] + red[263];
    }
  }
}

// Sum of the blocks' partial gradients; optionally an Adam step
// (Kingma & Ba 2015) on the FP32 master parameters; then the bf16 operand
// pack and FP32 small parameters are rewritten from the parameters.
__device__ __forceinline__ void pack_param(int p, float wv, __bf16* pack, float* small) {
  if (p < ob0) {  // W0
    const int i = p / kIn, k = p % kIn;
    const int m = i >> 4, c = k >> 5, gg = (k & 31) >> 3, s = k & 7;
    pack[oPackF0 + pack_index(m, c, 4, 16 * gg + (i & 15), s)] = (__bf16)wv;
    pack[oPackInfer + oPackF0 + pack_index(m, c, 4, 16 * gg + (i & 15), s)] = (__bf16)(wv * kNegLog2e);
    return;
  }
  const int oWs[3] = {oW1, oW2, oW3}, obs[4] = {ob0, ob1, ob2, ob3};
  for (int L = 0; L < 4; ++L)
    if (p >= obs[L] && p < obs[L] + kH) {
      small[64 * L + (p - obs[L])] = wv;
      small[oSmallInfer + 64 * L + (p - obs[L])] = wv * kNegLog2e;
      return;
    }
  for (int L = 1; L < 4; ++L)
    if (p >= oWs[L - 1] && p < oWs[L - 1] + kH * kH) {
      const int i = (p - oWs[L - 1]) / kH, k = (p - oWs[L - 1]) % kH;
      int gg, s;
      kperm_inv(k & 31, gg, s);
      pack[oPackF1 + (L - 1) * kPackL + pack_index(i >> 4, k >> 5, 2, 16 * gg + (i & 15), s)] = (__bf16)wv;
      pack[oPackInfer + oPackF1 + (L - 1) * kPackL + pack_index(i >> 4, k >> 5, 2, 16 * gg + (i & 15), s)] = (__bf16)wv;
      kperm_inv(i & 31, gg, s);
      pack[oPackB1 + (L - 1) * kPackL + pack_index(k >> 4, i >> 5, 2, 16 * gg + (k & 15), s)] = (__bf16)wv;
      return;
    }
  if (p >= oW4 && p < oW4 + kH) {
    small[256 + (p - oW4)] = wv;
    small[oSmallInfer + 256 + (p - oW4)] = wv * kNegLn2;
  } else if (p == ob4) {
    small[320] = wv;
    small[oSmallInfer + 320] = wv;
  }
}

// First stage of the sum over blocks: row kFold * s of the partial-gradient
// array receives the sum of rows kFold * s .. kFold * s + kFold - 1, each
// column by one thread (fixed order, so the result is deterministic).
constexpr int kFold = 16;

__global__ void partial_fold_kernel(float* __restrict__ partial, int32_t n_blocks) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > kParams) return;
  const int b0 = blockIdx.y * kFold;
  const int b1 = b0 + kFold < n_blocks ? b0 + kFold : n_blocks;
  float acc = 0.0f;
  for (int b = b0; b < b1; ++b) acc += partial[(int64_t)b * (kParams + 1) + p];
  partial[(int64_t)b0 * (kParams + 1) + p] = acc;
}

__global__ void mlp_update_kernel(float* __restrict__ params, float* __restrict__ m1, float* __restrict__ m2,
                                  const float* __restrict__ partial, int32_t n_blocks, int32_t step, float lr,
                                  float b1, float b2, float eps, int32_t do_adam, float* __restrict__ grad_out,
                                  __bf16* __restrict__ pack, float* __restrict__ small, float* __restrict__ loss_out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > kParams) return;
  float gsum = 0.0f;
  if (partial)
    for (int b = 0; b < n_blocks; b += kFold) gsum += partial[(int64_t)b * (kParams + 1) + p];
  if (p == kParams) {
    if (loss_out) *loss_out = gsum;
    return;
  }
  if (grad_out) grad_out[p] = gsum;
  float wv = params[p];
  if (do_adam) {
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    const float mm = b1 * m1[p] + (1.0f - b1) * gsum;
    const float vv = b2 * m2[p] + (1.0f - b2) * gsum * gsum;
    m1[p] = mm;
    m2[p] = vv;
    wv = wv - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    params[p] = wv;
  }
  pack_param(p, wv, pack, small);
}

// The two stages above in one launch: a block covers 64 parameters with four
// quarters of threads; quarter q sums the kFold-row groups q, q + 4, ... in
// row order, then one thread per parameter adds the group sums in group order
// and takes the Adam step.  The additions are the ones the two stages make,
// in the same order, so the result is the same to the bit.
constexpr int kReduceParams = 64;
constexpr int kMaxFoldGroups = 64;

__global__ void __launch_bounds__(256) grad_reduce_update_kernel(
    float* __restrict__ params, float* __restrict__ m1, float* __restrict__ m2, const float* __restrict__ partial,
    int32_t n_blocks, int32_t step, float lr, float b1, float b2, float eps, int32_t do_adam,
    float* __restrict__ grad_out, __bf16* __restrict__ pack, float* __restrict__ small, float* __restrict__ loss_out) {
  __shared__ float gs[kMaxFoldGroups][kReduceParams];
  const int pl = threadIdx.x & (kReduceParams - 1), q = threadIdx.x / kReduceParams;
  const int p = blockIdx.x * kReduceParams + pl;
  const int n_groups = (n_blocks + kFold - 1) / kFold;
  if (p <= kParams) {
    for (int sg = q; sg < n_groups; sg += 256 / kReduceParams) {
      const int r0 = sg * kFold;
      const int r1 = r0 + kFold < n_blocks ? r0 + kFold : n_blocks;
      float v[kFold];
#pragma unroll
      for (int k = 0; k < kFold; ++k) v[k] = r0 + k < r1 ? partial[(int64_t)(r0 + k) * (kParams + 1) + p] : 0.0f;
      float acc = 0.0f;
#pragma unroll
      for (int k = 0; k < kFold; ++k)
        if (r0 + k < r1) acc += v[k];
      gs[sg][pl] = acc;
    }
  }
  __syncthreads();
  if (q != 0 || p > kParams) return;
  float gsum = 0.0f;
  for (int sg = 0; sg < n_groups; ++sg) gsum += gs[sg][pl];
  if (p == kParams) {
    if (loss_out) *loss_out = gsum;
    return;
  }
  if (grad_out) grad_out[p] = gsum;
  float wv = params[p];
  if (do_adam) {
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    const float mm = b1 * m1[p] + (1.0f - b1) * gsum;
    const float vv = b2 * m2[p] + (1.0f - b2) * gsum * gsum;
    m1[p] = mm;
    m2[p] = vv;
    wv = wv - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    params[p] = wv;
  }
  pack_param(p, wv, pack, small);
}

// ----------------------------------------------------------------- training data

// Squared distance (voxel units) from each voxel to the zero set, seeded at
// voxels whose sign differs from a 6-neighbour with the linear-interpolation
// estimate of the sub-voxel distance; INF elsewhere.
__global__ void edt_seed_kernel(VolView V, float* __restrict__ g) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)V.dim[0] * V.dim[1] * V.dim[2];
  if (i >= n) return;
  const int x = (int)(i % V.dim[0]), y = (int)((i / V.dim[0]) % V.dim[1]), z = (int)(i / ((int64_t)V.dim[0] * V.dim[1]));
  const float f = V.data[vox_index(V.bdim, x, y, z)];
  float best = 1e20f;
  const int dx[6] = {1, -1, 0, 0, 0, 0}, dy[6] = {0, 0, 1, -1, 0, 0}, dz[6] = {0, 0, 0, 0, 1, -1};
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int xn = x + dx[k], yn = y + dy[k], zn = z + dz[k];
    if (xn < 0 || yn < 0 || zn < 0 || xn >= V.dim[0] || yn >= V.dim[1] || zn >= V.dim[2]) continue;
    const float fn = V.data[vox_index(V.bdim, xn, yn, zn)];
    if ((f > 0.0f) != (fn > 0.0f)) {
      const float fr = fabsf(f) / fabsf(f - fn);
      best = fminf(best, fr * fr);
    }
  }
  g[i] = best;
}

// One separable pass of the exact Euclidean distance transform: for every
// line along `axis`, D(x) = min_x' (x - x')^2 + g(x').  One work-group per
// line; the line is staged in LDS (at most 4096 voxels).
__global__ void __launch_bounds__(256) edt_pass_kernel(const float* __restrict__ gin, float* __restrict__ gout,
                                                       int32_t dx, int32_t dy, int32_t dz, int32_t axis) {
  __shared__ float line[4096];
  const int64_t lid = blockIdx.x;
  int len, a, b;
  int64_t base, stride;
  if (axis == 0) { len = dx; a = (int)(lid % dy); b = (int)(lid / dy); base = ((int64_t)b * dy + a) * dx; stride = 1; }
  else if (axis == 1) { len = dy; a = (int)(lid % dx); b = (int)(lid / dx); base = (int64_t)b * dy * dx + a; stride = dx; }
  else { len = dz; a = (int)(lid % dx); b = (int)(lid / dx); base = (int64_t)b * dx + a; stride = (int64_t)dx * dy; }
  for (int x = threadIdx.x; x < len; x += blockDim.x) line[x] = gin[base + x * stride];
  __syncthreads();
  for (int x = threadIdx.x; x < len; x += blockDim.x) {
    float best = 1e20f;
    for (int q = 0; q < len; ++q) {
      const float d = (float)(x - q);
      best = fminf(best, d * d + line[q]);
    }
    gout[base + x * stride] = best;
  }
}

// Signed distance in world units, written into a bricked volume.
__global__ void edt_finish_kernel(VolView V, const float* __restrict__ g, float* __restrict__ out, float spacing) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)V.dim[0] * V.dim[1] * V.dim[2];
  if (i >= n) return;
  const int x = (int)(i % V.dim[0]), y = (int)((i / V.dim[0]) % V.dim[1]), z = (int)(i / ((int64_t)V.dim[0] * V.dim[1]));
  const int64_t vi = vox_index(V.bdim, x, y, z);
  const float f = V.data[vi];
  const float d = sqrtf(fminf(g[i], 1e12f)) * spacing;
  out[vi] = f > 0.0f ? d : -d;
}

// 2^k training points: even indices uniform in the volume, odd indices near
// the surface (|signed distance| < near_band, by bounded rejection).  Each
// sample is (x, y, z, signed distance).
__global__ void sample_kernel(VolView edt, float lo, float hi, int64_t n, uint64_t seed, float near_band,
                              int32_t max_tries, float4* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  DevPcg rng(seed, (uint64_t)i);
  f3 p{0, 0, 0};
  float d = 0.0f;
  const bool near = (i & 1) != 0;
  for (int t = 0; t < (near ? max_tries : 1); ++t) {
    p = f3{lo + (hi - lo) * rng.uniform(), lo + (hi - lo) * rng.uniform(), lo + (hi - lo) * rng.uniform()};
    d = trilinear(edt, p);
    if (!near || fabsf(d) < near_band) break;
  }
  out[i] = float4{p.x, p.y, p.z, d};
}

// ----------------------------------------------------------------- hybrid tracer

// Surrogate distance at the points of the lanes that need it.  The wave packs
// those points densely into its LDS slots (lane order), evaluates
// ceil(count / 16) tiles of 16 points, and each lane reads its result back,
// so a partly idle wave evaluates no empty tile rows.  LDS is in order within
// a wave; the fences keep the compiler from moving the slot reads and writes
// across each other.
__device__ __forceinline__ float wave_surrogate(const bf16x8* spack, const float* ssmall, f3 p, bool need,
                                                float* __restrict__ wpts, float* __restrict__ wout) {
  const int lane = threadIdx.x & 63;
  const unsigned long long mask = __ballot(need);
  const int count = __popcll(mask);
  const int slot = __popcll(mask & ((1ull << lane) - 1ull));
  if (need) {
    wpts[slot] = p.x;
    wpts[64 + slot] = p.y;
    wpts[128 + slot] = p.z;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll 1
  for (int q = 0; q < count; q += 16) {
    // keeps the weight reads inside the loop (hoisted, they would hold ~160
    // VGPRs for the whole kernel and cut occupancy to one wave per SIMD)
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int idx = q + (lane & 15);
    const int src = idx < count ? idx : count - 1;
    const f3 pq{wpts[src], wpts[64 + src], wpts[128 + src]};
    bf16x8 xb[4];
    features(pq, lane >> 4, xb);
    TileAct T;
    const float o = forward_tile<false, true>(spack, ssmall, xb, lane, T);
    if (lane < 16 && idx < count) wout[idx] = o;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  return need ? wout[slot] : 0.0f;
}

// Exact-phase march: h = max(f * k_exact, h_min) from a point with f > 0;
// on the first sample with f <= 0 the crossing is bisected (n_bisect steps)
// between the last positive and the first non-positive sample.
// Surrogate phase: step by (s - margin) while the surrogate distance s
// exceeds tau, then hand over to the exact phase; after rejoin of exact
// travel without a crossing, return to the surrogate.
constexpr int kDeferTile = 16;

template <bool USE_MLP>
__global__ void __launch_bounds__(512) hybrid_trace_kernel(
    VolView V, float bx1, float by1, float bz1, const bf16x8* __restrict__ gpack, const float* __restrict__ gsmall,
    const float* __restrict__ ray_o, const float* __restrict__ ray_d, int64_t n, HybridParams hp,
    float* __restrict__ t_out, uint8_t* __restrict__ hit_out, int32_t* __restrict__ evals_out,
    int32_t* __restrict__ samples_out) {
  __shared__ bf16x8 spack[USE_MLP ? (oPackB1 >> 3) : 1];
  __shared__ __align__(16) float ssmall[kSmall];
  __shared__ float slots[USE_MLP ? kSurrBlock / 64 : 1][4 * 64];  // per wave: x, y, z, result
  if (USE_MLP) load_weights(gpack, gsmall, spack, ssmall, oPackB1 >> 3);
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t ic = i < n ? i : n - 1;
  const f3 o{ray_o[3 * ic], ray_o[3 * ic + 1], ray_o[3 * ic + 2]};
  const f3 d{ray_d[3 * ic], ray_d[3 * ic + 1], ray_d[3 * ic + 2]};
  const float bmin[3] = {0.0f, 0.0f, 0.0f}, bmax[3] = {bx1, by1, bz1};
  if (!USE_MLP) {
    // Without the network: the exact-field march from the volume entry.
    if (i < n) {
      float t;
      int32_t samples;
      const bool h = exact_march(V, bmin, bmax, o, d, hp.k_exact, hp.h_min, hp.n_bisect, hp.max_iters, t, samples);
      t_out[i] = t;
      hit_out[i] = h ? 1 : 0;
      if (evals_out) evals_out[i] = 0;
      if (samples_out) samples_out[i] = samples;
    }
    return;
  }
  enum { SURR = 0, EXACT = 1, DONE = 2 };
  int phase = DONE;
  bool hit = false;
  float t = 0.0f, t_far = 0.0f, f_prev = 0.0f, travel = 0.0f;
  int32_t evals = 0, samples = 0;
  float tn, tf;
  if (i < n && ray_box(bmin, bmax, o, d, tn, tf)) {
    t = tn;
    t_far = tf;
    phase = USE_MLP ? SURR : EXACT;
    if (!USE_MLP) {
      f_prev = trilinear(V, add(o, smul(t, d)));
      ++samples;
      if (f_prev <= 0.0f) { hit = true; phase = DONE; }
    }
  }
  // A lane is advanced at most max_iters times (one network evaluation and/or
  // one exact step per advance).  While some lane can take an exact step, the
  // network only runs full 16-point tiles of the waiting lanes (which ones is
  // rotated from iteration to iteration); when none can, it runs for all of
  // them.  A waiting lane keeps its state, so every ray takes the same steps
  // as when the network runs for every waiting lane at once, in fewer
  // part-filled tiles.  Each wave iteration advances at least one lane, which
  // bounds the loop.
  int32_t advances = 0;
  const int64_t wave_iters = 64ll * hp.max_iters;
#pragma unroll 1
  for (int64_t it = 0; it < wave_iters; ++it) {
    if (phase != DONE && advances >= hp.max_iters) phase = DONE;
    if (!__any(phase != DONE)) break;
    bool advanced = phase == EXACT;
    if (USE_MLP) {
      const bool wait = phase == SURR;
      const unsigned long long wmask = __ballot(wait);
      const int n_wait = __popcll(wmask);
      const int n_run = __any(phase == EXACT) ? (n_wait & ~(kDeferTile - 1)) : n_wait;
      if (n_run > 0) {
        const int rank = __popcll(wmask & ((1ull << (threadIdx.x & 63)) - 1ull));
        const bool need = wait && (rank + (int)((uint32_t)it % (uint32_t)n_wait)) % n_wait < n_run;
        float* ws = slots[threadIdx.x >> 6];
        const float s = wave_surrogate(spack, ssmall, add(o, smul(t, d)), need, ws, ws + 192);
        if (need) {
          advanced = true;
          ++evals;
          if (s > hp.tau) {
            t += s - hp.margin;
            if (t >= t_far) phase = DONE;
          } else {
            phase = EXACT;
            travel = 0.0f;
            f_prev = trilinear(V, add(o, smul(t, d)));
            ++samples;
            if (f_prev <= 0.0f) {
              hit = true;  // entered the surface region already inside: accept the current point
              phase = DONE;
            }
          }
        }
      }
    }
    if (phase == EXACT) {
      const float h = fmaxf(f_prev * hp.k_exact, hp.h_min);
      const float t_next = t + h;
      const float f_next = trilinear(V, add(o, smul(t_next, d)));
      ++samples;
      if (f_next <= 0.0f) {
        float a = t, b = t_next;
        for (int k = 0; k < hp.n_bisect; ++k) {
          const float mid = 0.5f * (a + b);
          const float fm = trilinear(V, add(o, smul(mid, d)));
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
        else if (USE_MLP && travel > hp.rejoin) phase = SURR;
      }
    }
    if (advanced) ++advances;
  }
  if (i < n) {
    t_out[i] = hit ? t : kInf;
    hit_out[i] = hit ? 1 : 0;
    if (evals_out) evals_out[i] = evals;
    if (samples_out) samples_out[i] = samples;
  }
}

}  // namespace

// ---------------------------------------------------------------- launchers

hipError_t launch_hybrid_trace(const DevVolume& field, const void* pack, const float* small, const float* ray_o,
                               const float* ray_d, int64_t n, const HybridParams& hp, bool use_mlp, float* t,
                               uint8_t* hit, int32_t* evals, int32_t* samples, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const VolView V = field.view();
  if (use_mlp)
    hipLaunchKernelGGL(hybrid_trace_kernel<true>, dim3((unsigned)((n + kSurrBlock - 1) / kSurrBlock)),
                       dim3(kSurrBlock), 0, s, V, field.box_max[0],
                       field.box_max[1], field.box_max[2], (const bf16x8*)pack, small, ray_o, ray_d, n, hp, t, hit,
                       evals, samples);
  else
    hipLaunchKernelGGL(hybrid_trace_kernel<false>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, V,
                       field.box_max[0],
                       field.box_max[1], field.box_max[2], (const bf16x8*)pack, small, ray_o, ray_d, n, hp, t, hit,
                       evals, samples);
  return hipGetLastError();
}

size_t mlp_train_lds_bytes() { return sizeof(TrainLds); }

hipError_t launch_mlp_infer(const void* pack, const float* small, const float* pts, int64_t n, float* out,
                            hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int64_t tiles = (n + 15) / 16;
  int64_t blocks = (tiles + kSurrBlock / 64 - 1) / (kSurrBlock / 64);
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(mlp_infer_kernel, dim3((unsigned)blocks), dim3(kSurrBlock), 0, s, (const bf16x8*)pack, small,
                     pts, n, out);
  return hipGetLastError();
}

hipError_t launch_mlp_train(const void* pack, const float* small, const float4* samples, int64_t n_samples,
                            int64_t batch_start, int32_t batch, float over_weight, float* partial, int32_t blocks,
                            hipStream_t s) {
  hipLaunchKernelGGL(mlp_train_kernel, dim3(blocks), dim3(kTrainBlock), sizeof(TrainLds), s, (const bf16x8*)pack, small,
                     samples, n_samples, batch_start, batch, over_weight, partial);
  return hipGetLastError();
}

hipError_t launch_mlp_update(float* params, float* m1, float* m2, float* partial, int32_t n_blocks,
                             int32_t step, float lr, float b1, float b2, float eps, int32_t do_adam, float* grad_out,
                             void* pack, float* small, float* loss_out, hipStream_t s) {
  const int threads = 256;
  const int blocks = (kParams + 1 + threads - 1) / threads;
  if (partial && n_blocks > 0 && (n_blocks + kFold - 1) / kFold <= kMaxFoldGroups) {
    hipLaunchKernelGGL(grad_reduce_update_kernel, dim3((kParams + kReduceParams) / kReduceParams), dim3(256), 0, s,
                       params, m1, m2, partial, n_blocks, step, lr, b1, b2, eps, do_adam, grad_out, (__bf16*)pack,
                       small, loss_out);
    return hipGetLastError();
  }
  if (partial && n_blocks > 0)
    hipLaunchKernelGGL(partial_fold_kernel, dim3(blocks, (n_blocks + kFold - 1) / kFold), dim3(threads), 0, s,
                       partial, n_blocks);
  hipLaunchKernelGGL(mlp_update_kernel, dim3(blocks), dim3(threads), 0, s, params, m1, m2, partial, n_blocks, step,
                     lr, b1, b2, eps, do_adam, grad_out, (__bf16*)pack, small, loss_out);
  return hipGetLastError();
}

hipError_t launch_edt(const DevVolume& field, float* scratch_a, float* scratch_b, float* out_bricked,
                      hipStream_t s) {
  const VolView V = field.view();
  const int64_t n = (int64_t)V.dim[0] * V.dim[1] * V.dim[2];
  const int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(edt_seed_kernel, dim3((unsigned)blocks), dim3(256), 0, s, V, scratch_a);
  hipLaunchKernelGGL(edt_pass_kernel, dim3((unsigned)((int64_t)V.dim[1] * V.dim[2])), dim3(256), 0, s, scratch_a,
                     scratch_b, V.dim[0], V.dim[1], V.dim[2], 0);
  hipLaunchKernelGGL(edt_pass_kernel, dim3((unsigned)((int64_t)V.dim[0] * V.dim[2])), dim3(256), 0, s, scratch_b,
                     scratch_a, V.dim[0], V.dim[1], V.dim[2], 1);
  hipLaunchKernelGGL(edt_pass_kernel, dim3((unsigned)((int64_t)V.dim[0] * V.dim[1])), dim3(256), 0, s, scratch_a,
                     scratch_b, V.dim[0], V.dim[1], V.dim[2], 2);
  hipLaunchKernelGGL(edt_finish_kernel, dim3((unsigned)blocks), dim3(256), 0, s, V, scratch_b, out_bricked,
                     V.spacing[0]);
  return hipGetLastError();
}

hipError_t launch_sample(const DevVolume& edt, int64_t n, uint64_t seed, float near_band, int32_t max_tries,
                         float4* out, hipStream_t s) {
  const VolView V = edt.view();
  const float hi = (float)(V.dim[0] - 1) * V.spacing[0];
  hipLaunchKernelGGL(sample_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, V, 0.0f, hi, n, seed,
                     near_band, max_tries, out);
  return hipGetLastError();
}

}  // namespace re
