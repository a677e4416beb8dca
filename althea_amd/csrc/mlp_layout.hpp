// Parameter layout of the prefiltered-radiance MLP (28 -> 64 -> 64 -> 64 -> 3)
// and the BF16 operand layouts of its MFMA kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace althea {

constexpr int kMlpIn = 28, kMlpHidden = 64, kMlpOut = 3;

// FP32 master copy, in order: W1[64][28] b1[64] W2[64][64] b2[64] W3[64][64]
// b3[64] W4[3][64] b4[3].
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpHidden * kMlpIn;
constexpr int kOffW2 = kOffB1 + kMlpHidden;
constexpr int kOffB2 = kOffW2 + kMlpHidden * kMlpHidden;
constexpr int kOffW3 = kOffB2 + kMlpHidden;
constexpr int kOffB3 = kOffW3 + kMlpHidden * kMlpHidden;
constexpr int kOffW4 = kOffB3 + kMlpHidden;
constexpr int kOffB4 = kOffW4 + kMlpOut * kMlpHidden;
constexpr int kMlpParams = kOffB4 + kMlpOut;  // 10371

// Weight-gradient partial rows (weights only, layer by layer).
constexpr int kWgW1 = 0;
constexpr int kWgW2 = kWgW1 + kMlpHidden * kMlpIn;
constexpr int kWgW3 = kWgW2 + kMlpHidden * kMlpHidden;
constexpr int kWgW4 = kWgW3 + kMlpHidden * kMlpHidden;
constexpr int kMlpWeights = kWgW4 + kMlpOut * kMlpHidden;  // 10176

// Bias + loss partial rows.
constexpr int kBiasB1 = 0, kBiasB2 = 64, kBiasB3 = 128, kBiasB4 = 192, kBiasLoss = 195;
constexpr int kMlpBiasPartial = 196;

static_assert(kMlpParams == 10371, "parameter count");
static_assert(kMlpWeights + 3 * kMlpHidden + kMlpOut == kMlpParams, "weights + biases");

// K permutation inside a 32-wide block: MFMA B slot (g, j) = 8 g + j holds
// feature perm(slot) so a layer's D registers feed the next layer directly.
__host__ __device__ constexpr int mlp_perm(int slot) {
  return (slot & 7) < 4 ? 4 * (slot >> 3) + (slot & 7) : 16 + 4 * (slot >> 3) + (slot & 7) - 4;
}

// The slot of feature k (0..31) in a 32-wide block: mlp_perm's inverse.
__host__ __device__ constexpr int mlp_perm_inv(int k) {
  return k < 16 ? 8 * (k >> 2) + (k & 3) : 8 * ((k - 16) >> 2) + 4 + ((k - 16) & 3);
}
static_assert(mlp_perm_inv(mlp_perm(0)) == 0 && mlp_perm_inv(mlp_perm(5)) == 5 && mlp_perm_inv(mlp_perm(13)) == 13 &&
                  mlp_perm_inv(mlp_perm(22)) == 22 && mlp_perm_inv(mlp_perm(31)) == 31,
              "mlp_perm_inv inverts mlp_perm");

// Master-copy index -> weight-gradient index (or -1 for a bias).
__host__ __device__ inline int mlp_param_to_wgrad(int t) {
  if (t < kOffB1) return kWgW1 + (t - kOffW1);
  if (t < kOffW2) return -1;
  if (t < kOffB2) return kWgW2 + (t - kOffW2);
  if (t < kOffW3) return -1;
  if (t < kOffB3) return kWgW3 + (t - kOffW3);
  if (t < kOffW4) return -1;
  if (t < kOffB4) return kWgW4 + (t - kOffW4);
  return -1;
}
// Master-copy index of a bias -> bias-partial index.
__host__ __device__ inline int mlp_param_to_bias(int t) {
  if (t >= kOffB1 && t < kOffW2) return kBiasB1 + (t - kOffB1);
  if (t >= kOffB2 && t < kOffW3) return kBiasB2 + (t - kOffB2);
  if (t >= kOffB3 && t < kOffW4) return kBiasB3 + (t - kOffB3);
  return kBiasB4 + (t - kOffB4);
}

// BF16 operand layouts (see mlp_kernels.hip):
//   w1  [64 o][32 k]           forward A, natural k (k >= 28 zero)
//   w2, w3 [64 o][64 k]        forward A, k permuted per 32-block
//   w4  [16 o][64 k]           forward A (o >= 3 zero), k permuted
//   w2t, w3t [64 k][64 o]      backward A: W^T with o permuted per 32-block
//   w4t [64 k][32 o]           backward A (o >= 3 zero), o permuted
//   b1..b3 [64], b4 [16]       FP32 biases (b4 zero-padded)
struct MlpDeviceWeights {
  __bf16 *w1, *w2, *w3, *w4, *w2t, *w3t, *w4t;
  float *b1, *b2, *b3, *b4;
};

// Feature-major BF16 copies of one training batch (rows x stride samples).
struct MlpActivations {
  __bf16 *x, *h1, *h2, *h3, *dy, *dz1, *dz2, *dz3;  // 32, 64, 64, 64, 16, 64, 64, 64 rows
  size_t stride;
};

}  // namespace althea
