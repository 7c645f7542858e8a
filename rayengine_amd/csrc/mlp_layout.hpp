// SDF surrogate network: parameter layout and the packed BF16 operand layout
// of its MFMA kernels.  The surrogate has no counterpart in the reference.
//
// Network: 99 inputs (16-frequency Fourier features of the 3 coordinates, then
// the raw coordinates), 4 hidden layers of 64 SiLU units, one linear output.
//   feature 32 d + 2 k     = sin(2 pi f_k p_d),  f_k = 2^(k/2), k < 16, d < 3
//   feature 32 d + 2 k + 1 = cos(2 pi f_k p_d)
//   features 96..98        = p.x, p.y, p.z
// Flat FP32 parameters (also the layout of gradients and Adam moments):
//   W0[64][99] b0[64] W1[64][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64] b4
//
// MFMA mapping (v_mfma_f32_16x16x32_bf16, one wave, 16 points per tile):
// activations are kept transposed, Z^T = W X^T, so the points are the N
// dimension.  Lane l works on point j = l % 16 and group g = l / 16:
//   A operand (16 x 32): lane l holds A[l % 16][8 g + s], s < 8
//   B operand (32 x 16): lane l holds B[8 g + s][l % 16]
//   C result  (16 x 16): lane l holds C[4 g + r][l % 16], r < 4
// A layer's result tiles m = 0..3 (neurons 16 m ..) leave lane l holding
// neurons 16 m + 4 g + r.  The next layer consumes them directly as its B
// operand: K-step c covers neurons 32 c .. 32 c + 31 and hardware slot
// 8 g + s holds neuron 32 c + perm(8 g + s) with
//   perm(8 g + s) = s < 4 ? 4 g + s : 16 + 4 g + (s - 4),
// so the weights are packed with the same column permutation and the
// activations never leave registers between layers.
#pragma once

#include <stdint.h>

#include "re_math.hpp"

namespace re {
namespace mlp {

constexpr int kIn = 99;
constexpr int kInPad = 128;  // layer-0 K padded to 4 MFMA K-steps
constexpr int kH = 64;
constexpr int kFreq = 16;

constexpr int oW0 = 0;
constexpr int ob0 = oW0 + kH * kIn;
constexpr int oW1 = ob0 + kH;
constexpr int ob1 = oW1 + kH * kH;
constexpr int oW2 = ob1 + kH;
constexpr int ob2 = oW2 + kH * kH;
constexpr int oW3 = ob2 + kH;
constexpr int ob3 = oW3 + kH * kH;
constexpr int oW4 = ob3 + kH;
constexpr int ob4 = oW4 + kH;
constexpr int kParams = ob4 + 1;  // 18945

// Packed BF16 operands, in elements: forward A operands of layers 0..3, then
// the transposed (backward) A operands of layers 1..3.
constexpr int kPackL0 = 4 * 4 * 64 * 8;  // 4 M-tiles x 4 K-steps x 64 lanes x 8
constexpr int kPackL = 4 * 2 * 64 * 8;   // 4 M-tiles x 2 K-steps x 64 lanes x 8
constexpr int oPackF0 = 0;
constexpr int oPackF1 = oPackF0 + kPackL0;
constexpr int oPackF2 = oPackF1 + kPackL;
constexpr int oPackF3 = oPackF2 + kPackL;
constexpr int oPackB1 = oPackF3 + kPackL;
constexpr int oPackB2 = oPackB1 + kPackL;
constexpr int oPackB3 = oPackB2 + kPackL;
constexpr int kPack = oPackB3 + kPackL;  // 32768 elements, 64 KB
// FP32 small parameters kept beside the pack: b0..b3, W4, b4.
constexpr int kSmall = 4 * kH + kH + 1;  // 321

// Inference copy of the forward operands, after the training pack (and of
// the small parameters, after the training ones), with the pre-activations
// carried as u = -log2(e) z: layer-0 weights and every bias are stored times
// -log2(e), the weights of layers 1..3 as they are, and W4 times -ln 2.  A
// hidden unit is then u / (1 + 2^u) = -log2(e) SiLU(z), one multiply fewer
// than SiLU(z) itself, and the next layer's pre-activation comes out as u
// again because log2(e) ln 2 = 1.
constexpr int oPackInfer = kPack;                // elements
constexpr int kPackAlloc = kPack + (kPackL0 + 3 * kPackL);
constexpr int oSmallInfer = kSmall;
constexpr int kSmallAlloc = 2 * kSmall;
constexpr float kNegLog2e = -1.44269504088896341f;
constexpr float kNegLn2 = -0.693147180559945309f;

// Neuron (within a 32-wide K-step) held in hardware K slot 8 g + s.
RE_HD int kperm(int g, int s) { return s < 4 ? 4 * g + s : 16 + 4 * g + (s - 4); }
// Inverse: neuron q (0..31) of a K-step -> (g, s).
RE_HD void kperm_inv(int q, int& g, int& s) {
  if (q < 16) { g = q >> 2; s = q & 3; }
  else { g = (q - 16) >> 2; s = 4 + ((q - 16) & 3); }
}

// Index into a packed layer of `csteps` K-steps.
RE_HD int pack_index(int m, int c, int csteps, int lane, int s) {
  return (((m * csteps + c) * 64 + lane) * 8) + s;
}

}  // namespace mlp
}  // namespace re
