// Learned stand-in for the prefiltered cube: a 28 -> 64 -> 64 -> 64 -> 3 MLP
// (SiLU hidden layers, log-radiance output) in BF16 with FP32 accumulation on
// gfx950 MFMA (v_mfma_f32_16x16x32_bf16).  The reference has no learned part.
//
// Layout.  Every layer is computed transposed, D[out][sample] = W . X^T, with
// 16 samples on the MFMA N axis.  With lane l = 16 g + n:
//   A (16 x 32): lane holds W[row = n][k = 8 g + j], j = 0..7
//   B (32 x 16): lane holds X^T[k = 8 g + j][sample = n]
//   D (16 x 16): lane holds D[row = 4 g + i][sample = n], i = 0..3
// A layer's D registers for output blocks 2kb and 2kb+1 are exactly the next
// layer's B operand for K block kb if the K index is permuted within each
// 32-block as slot (g, j) -> feature perm(g, j) = j < 4 ? 4g + j : 16 + 4g + j - 4.
// The packed weights carry that permutation, so activations never move
// between layers.  The backward pass (dH = W^T dZ) uses the same trick with
// transposed packed weights.
//
// Training per batch: mlp_fwd_bwd_kernel (forward, loss, backward-data,
// feature-major bf16 copies of X, H1..H3, dY, dZ1..dZ3, per-block bias and
// loss partials) -> mlp_wgrad_kernel (dW = dZ . H^T as split-K MFMA over
// samples, per-block partials) -> mlp_adam_kernel (fixed-order reduction of
// the partials, Adam on the FP32 master copy) -> mlp_pack_kernel (BF16
// operand layouts).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <vector>

#include "internal.hpp"
#include "mlp_eval.hpp"
#include "mlp_layout.hpp"

namespace althea {

// ------------------------------------------------------------ pack weights
// Builds the BF16 operand layouts from the FP32 master copy (mlp_layout.hpp).
__global__ void mlp_pack_kernel(const float* __restrict__ p, MlpDeviceWeights w) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  // Forward A operands.
  if (t < 64 * 32) {  // W1p[o][k], natural k, zero for k >= 28
    int o = t / 32, k = t % 32;
    w.w1[t] = static_cast<__bf16>(k < kMlpIn ? p[kOffW1 + o * kMlpIn + k] : 0.0f);
  }
  if (t < 64 * 64) {  // W2p / W3p [o][kb*32 + slot], slot -> perm
    int o = t / 64, kk = t % 64;
    int src_k = (kk & ~31) + mlp_perm(kk & 31);
    w.w2[t] = static_cast<__bf16>(p[kOffW2 + o * 64 + src_k]);
    w.w3[t] = static_cast<__bf16>(p[kOffW3 + o * 64 + src_k]);
    // Backward A operands: WT[k][ob*32 + slot] = W[ob*32 + perm(slot)][k].
    int k = t / 64, oo = t % 64;
    int src_o = (oo & ~31) + mlp_perm(oo & 31);
    w.w2t[t] = static_cast<__bf16>(p[kOffW2 + src_o * 64 + k]);
    w.w3t[t] = static_cast<__bf16>(p[kOffW3 + src_o * 64 + k]);
  }
  if (t < 16 * 64) {  // W4p[o][k] (o >= 3 zero), k permuted
    int o = t / 64, kk = t % 64;
    int src_k = (kk & ~31) + mlp_perm(kk & 31);
    w.w4[t] = static_cast<__bf16>(o < kMlpOut ? p[kOffW4 + o * 64 + src_k] : 0.0f);
  }
  if (t < 64 * 32) {  // W4t[k][slot] = W4[perm(slot)][k] (o >= 3 zero)
    int k = t / 32, slot = t % 32;
    int o = mlp_perm(slot);
    w.w4t[t] = static_cast<__bf16>(o < kMlpOut ? p[kOffW4 + o * 64 + k] : 0.0f);
  }
  if (t < 64) {
    w.b1[t] = p[kOffB1 + t];
    w.b2[t] = p[kOffB2 + t];
    w.b3[t] = p[kOffB3 + t];
  }
  if (t < 16) w.b4[t] = t < kMlpOut ? p[kOffB4 + t] : 0.0f;
}

// ------------------------------------------------- forward (+ backward-data)
struct FwdParams {
  MlpDeviceWeights w;
  const float4* dir_rough;
  const float4* targets;  // training only
  float4* out;            // inference only: rgb radiance = exp(y)
  int count;
  float grad_scale;       // 2 / (3 * batch)
  MlpActivations act;     // feature-major copies (training)
  float* bias_partial;    // [gridDim.x][kMlpBiasPartial] (training: the loss column)
};

// One 16-sample tile per iteration per wave.  kTrain adds the loss, the
// backward-data chain and the feature-major stores (the weight-gradient
// kernel forms every weight and bias gradient from those).
template <bool kTrain>
__global__ __launch_bounds__(256) void mlp_fwd_kernel(const FwdParams p) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, n = lane & 15;
  const int wave = threadIdx.x >> 6;
  const int waves_total = gridDim.x * 4;
  const int tiles = kTrain ? 2 * ((p.count + 31) / 32) : (p.count + 15) / 16;

  // Inference keeps the forward A operands in registers; training reloads
  // them (L1/L2 resident) at each use to leave registers for the backward.
  bf16x8 a1[4], a2[4][2], a3[4][2], a4[2];
  if (!kTrain) {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      a1[b] = load8(p.w.w1 + (16 * b + n) * 32 + 8 * g);
#pragma unroll
      for (int kb = 0; kb < 2; ++kb) {
        a2[b][kb] = load8(p.w.w2 + (16 * b + n) * 64 + 32 * kb + 8 * g);
        a3[b][kb] = load8(p.w.w3 + (16 * b + n) * 64 + 32 * kb + 8 * g);
      }
    }
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) a4[kb] = load8(p.w.w4 + n * 64 + 32 * kb + 8 * g);
  }
  auto A1 = [&](int b) { return kTrain ? load8(p.w.w1 + (16 * b + n) * 32 + 8 * g) : a1[b]; };
  auto A2 = [&](int b, int kb) { return kTrain ? load8(p.w.w2 + (16 * b + n) * 64 + 32 * kb + 8 * g) : a2[b][kb]; };
  auto A3 = [&](int b, int kb) { return kTrain ? load8(p.w.w3 + (16 * b + n) * 64 + 32 * kb + 8 * g) : a3[b][kb]; };
  auto A4 = [&](int kb) { return kTrain ? load8(p.w.w4 + n * 64 + 32 * kb + 8 * g) : a4[kb]; };

  float loss = 0.0f;  // training: this lane's squared log errors

  for (int tile = blockIdx.x * 4 + wave; tile < tiles; tile += waves_total) {
    const int s = tile * 16 + n;
    const bool valid = s < p.count;
    const float4 q = valid ? p.dir_rough[s] : make_float4(0.f, 0.f, 1.f, 0.f);
    // Layer 1 input (B operand): this lane's 8 features.
    float f[32];
    encode_all(q.x, q.y, q.z, q.w, f);
    bf16x8 x;
#pragma unroll
    for (int j = 0; j < 8; ++j)
      x[j] = static_cast<__bf16>(g == 0 ? f[j] : (g == 1 ? f[8 + j] : (g == 2 ? f[16 + j] : f[24 + j])));

    // Hidden layer: D = W . H^T (+ bias), SiLU; `sp` keeps SiLU'(z) for training.
    auto hidden = [&](auto&& wtile, const bf16x8* in, int nk, const float* bias, f32x4* h, f32x4* sp) {
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < nk; ++kb) acc = mfma(wtile(b, kb), in[kb], acc);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float z = acc[i] + bias[16 * b + 4 * g + i];
          const float sg = sigmoidf(z);
          h[b][i] = z * sg;
          if (kTrain) sp[b][i] = sg * (1.0f + z * (1.0f - sg));
        }
      }
    };
    // Feature-major copies for the weight gradients (training): element
    // (row, sample) at row * B + col, 32-bit offsets (batch <= 2^24).
    const MlpActivations& A = p.act;
    const uint32_t B = static_cast<uint32_t>(A.stride);
    const uint32_t col = static_cast<uint32_t>(tile) * 16u + n;
    auto store_tile = [&](__bf16* dst, const bf16x8* hb) {
      if (!kTrain) return;
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int i = 0; i < 4; ++i) dst[(16u * b + 4u * g + i) * B + col] = hb[b >> 1][4 * (b & 1) + i];
    };
    if (kTrain) {
#pragma unroll
      for (int j = 0; j < 8; ++j) A.x[(8u * g + j) * B + col] = x[j];
    }
    f32x4 h[4], sp1[4], sp2[4], sp3[4];
    hidden([&](int b, int) { return A1(b); }, &x, 1, p.w.b1, h, sp1);
    const bf16x8 hb1[2] = {pack8(h[0], h[1]), pack8(h[2], h[3])};
    store_tile(A.h1, hb1);
    hidden([&](int b, int kb) { return A2(b, kb); }, hb1, 2, p.w.b2, h, sp2);
    const bf16x8 hb2[2] = {pack8(h[0], h[1]), pack8(h[2], h[3])};
    store_tile(A.h2, hb2);
    hidden([&](int b, int kb) { return A3(b, kb); }, hb2, 2, p.w.b3, h, sp3);
    const bf16x8 hb3[2] = {pack8(h[0], h[1]), pack8(h[2], h[3])};
    store_tile(A.h3, hb3);
    f32x4 y = {0.f, 0.f, 0.f, 0.f};
    y = mfma(A4(0), hb3[0], y);
    y = mfma(A4(1), hb3[1], y);
#pragma unroll
    for (int i = 0; i < 4; ++i) y[i] += p.w.b4[4 * g + i];

    if (!kTrain) {
      if (g == 0 && valid) p.out[s] = make_float4(expf(y[0]), expf(y[1]), expf(y[2]), 1.0f);
      continue;
    }

    // ---- loss: lanes g == 0 own outputs 0..2 of sample n.
    f32x4 dy = {0.f, 0.f, 0.f, 0.f};
    if (g == 0 && valid) {
      const float4 t = p.targets[s];
      const float e0 = y[0] - logf(t.x), e1 = y[1] - logf(t.y), e2 = y[2] - logf(t.z);
      loss += e0 * e0 + e1 * e1 + e2 * e2;
      dy[0] = p.grad_scale * e0;
      dy[1] = p.grad_scale * e1;
      dy[2] = p.grad_scale * e2;
    }

    bf16x8 dyb;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      dyb[i] = static_cast<__bf16>(dy[i]);
      dyb[4 + i] = static_cast<__bf16>(0.0f);
      A.dy[(4u * g + i) * B + col] = dyb[i];
    }

    // ---- backward-data: dH = W^T dZ, dZ = dH * SiLU'(z), layer by layer.
    auto backward = [&](auto&& wtile, const bf16x8* in, int nk, const f32x4* sp, __bf16* store, f32x4* dz) {
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < nk; ++kb) acc = mfma(wtile(b, kb), in[kb], acc);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float d = acc[i] * sp[b][i];
          dz[b][i] = d;
          store[(16u * b + 4u * g + i) * B + col] = static_cast<__bf16>(d);
        }
      }
    };
    f32x4 dz[4];
    backward([&](int b, int) { return load8(p.w.w4t + (16 * b + n) * 32 + 8 * g); }, &dyb, 1, sp3, A.dz3, dz);
    bf16x8 d_in[2] = {pack8(dz[0], dz[1]), pack8(dz[2], dz[3])};
    backward([&](int b, int kb) { return load8(p.w.w3t + (16 * b + n) * 64 + 32 * kb + 8 * g); }, d_in, 2, sp2,
             A.dz2, dz);
    d_in[0] = pack8(dz[0], dz[1]);
    d_in[1] = pack8(dz[2], dz[3]);
    backward([&](int b, int kb) { return load8(p.w.w2t + (16 * b + n) * 64 + 32 * kb + 8 * g); }, d_in, 2, sp1,
             A.dz1, dz);
  }

  if (!kTrain) return;
  // The block's loss in the loss column of its partial row, fixed order.
  __shared__ float s_loss[4];
  float l = loss;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) l += __shfl_xor(l, o, 64);
  if (lane == 0) s_loss[wave] = l;
  __syncthreads();
  if (threadIdx.x == 0)
    p.bias_partial[static_cast<size_t>(blockIdx.x) * kMlpBiasPartial + kBiasLoss] =
        s_loss[0] + s_loss[1] + s_loss[2] + s_loss[3];
}

// ------------------------------------------------------ weight gradients
// dW_l[o][k] = sum_s dZ_l[o][s] H_{l-1}[k][s] with samples on the MFMA K axis:
// A = dZ tile (lane: 8 consecutive samples of one output row), B = H tile.
// The bias gradients db_l[o] = sum_s dZ_l[o][s] are the same product with a
// B tile of ones.  Each block owns a contiguous sample range; wave w owns
// output block w of layers 1-3 and input block w of layer 4 (wave 0 also the
// layer-4 bias).  Partials: one row of weights and one row of biases per
// block (the forward kernel writes the loss column of the bias row).
struct WgradParams {
  MlpActivations act;
  int count;         // samples (multiple of 32 after padding)
  int chunks_per_block;
  float* partial;       // [gridDim.x][kMlpWeights]
  float* bias_partial;  // [gridDim.x][kMlpBiasPartial], columns before kBiasLoss
};

__global__ __launch_bounds__(256) void mlp_wgrad_kernel(const WgradParams p) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, n = lane & 15;
  const int w = threadIdx.x >> 6;
  const MlpActivations& A = p.act;
  const size_t B = A.stride;
  f32x4 d1[2] = {}, d2[4] = {}, d3[4] = {}, d4 = {};
  f32x4 e1 = {}, e2 = {}, e3 = {}, e4 = {};  // bias sums (every column of D the same)
  bf16x8 ones;
#pragma unroll
  for (int j = 0; j < 8; ++j) ones[j] = static_cast<__bf16>(1.0f);
  const int chunk0 = blockIdx.x * p.chunks_per_block;
  const int nchunks = (p.count + 31) / 32;
  for (int c = chunk0; c < min(chunk0 + p.chunks_per_block, nchunks); ++c) {
    const size_t s0 = static_cast<size_t>(c) * 32 + 8 * g;
    const bf16x8 z1 = load8(A.dz1 + (16 * w + n) * B + s0);
    const bf16x8 z2 = load8(A.dz2 + (16 * w + n) * B + s0);
    const bf16x8 z3 = load8(A.dz3 + (16 * w + n) * B + s0);
    const bf16x8 y = load8(A.dy + n * B + s0);
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) d1[kb] = mfma(z1, load8(A.x + (16 * kb + n) * B + s0), d1[kb]);
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
      d2[kb] = mfma(z2, load8(A.h1 + (16 * kb + n) * B + s0), d2[kb]);
      d3[kb] = mfma(z3, load8(A.h2 + (16 * kb + n) * B + s0), d3[kb]);
    }
    d4 = mfma(y, load8(A.h3 + (16 * w + n) * B + s0), d4);
    e1 = mfma(z1, ones, e1);
    e2 = mfma(z2, ones, e2);
    e3 = mfma(z3, ones, e3);
    if (w == 0) e4 = mfma(y, ones, e4);
  }
  float* bout = p.bias_partial + static_cast<size_t>(blockIdx.x) * kMlpBiasPartial;
  if (n == 0) {  // column 0 of D: row o = 16 w + 4 g + i
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int o = 16 * w + 4 * g + i;
      bout[kBiasB1 + o] = e1[i];
      bout[kBiasB2 + o] = e2[i];
      bout[kBiasB3 + o] = e3[i];
      if (w == 0 && o < kMlpOut) bout[kBiasB4 + o] = e4[i];
    }
  }
  float* out = p.partial + static_cast<size_t>(blockIdx.x) * kMlpWeights;
  // D: lane holds dW[o = 16 w + 4 g + i][k = 16 kb + n].
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int o = 16 * w + 4 * g + i;
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
      const int k = 16 * kb + n;
      if (k < kMlpIn) out[kWgW1 + o * kMlpIn + k] = d1[kb][i];
    }
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
      out[kWgW2 + o * 64 + 16 * kb + n] = d2[kb][i];
      out[kWgW3 + o * 64 + 16 * kb + n] = d3[kb][i];
    }
    const int o4 = 4 * g + i;  // layer 4: output rows 0..15 (3 used), inputs 16 w + n
    if (o4 < kMlpOut) out[kWgW4 + o4 * 64 + 16 * w + n] = d4[i];
  }
}

// --------------------------------------------------------------- Adam step
struct AdamParams {
  const float* wpartial;  // [wblocks][kMlpWeights]
  int wblocks;
  const float* bpartial;  // [lblocks][kMlpBiasPartial]: bias columns in rows < bblocks, loss in rows < lblocks
  int bblocks;
  int lblocks;
  float* params;          // FP32 master copy, oracle order
  float* m;
  float* v;
  float* grad_out;        // optional: the reduced gradient (oracle order)
  float* loss_out;        // optional: mean loss
  float loss_scale;       // 1 / (3 * batch)
  int update;
  float lr, b1, b2, eps, c1, c2;  // c1 = 1 - b1^t, c2 = 1 - b2^t
  MlpDeviceWeights w;  // the BF16 operand copies, rewritten for each updated parameter
};

// Writes the operand copies of master parameter t (value v): the scatter form
// of mlp_pack_kernel for one parameter (the zero padding is never touched).
__device__ __forceinline__ void mlp_pack_one(const MlpDeviceWeights& w, int t, float v) {
  const __bf16 b = static_cast<__bf16>(v);
  auto slot64 = [](int k) { return (k & ~31) + mlp_perm_inv(k & 31); };
  if (t < kOffB1) {
    const int i = t - kOffW1, o = i / kMlpIn, k = i - o * kMlpIn;
    w.w1[o * 32 + k] = b;
  } else if (t < kOffW2) {
    w.b1[t - kOffB1] = v;
  } else if (t < kOffB2) {
    const int i = t - kOffW2, o = i >> 6, k = i & 63;
    w.w2[o * 64 + slot64(k)] = b;
    w.w2t[k * 64 + slot64(o)] = b;
  } else if (t < kOffW3) {
    w.b2[t - kOffB2] = v;
  } else if (t < kOffB3) {
    const int i = t - kOffW3, o = i >> 6, k = i & 63;
    w.w3[o * 64 + slot64(k)] = b;
    w.w3t[k * 64 + slot64(o)] = b;
  } else if (t < kOffW4) {
    w.b3[t - kOffB3] = v;
  } else if (t < kOffB4) {
    const int i = t - kOffW4, o = i >> 6, k = i & 63;
    w.w4[o * 64 + slot64(k)] = b;
    w.w4t[k * 32 + mlp_perm_inv(o)] = b;
  } else {
    w.b4[t - kOffB4] = v;
  }
}

// kAdamLanes lanes per parameter each sum every kAdamLanes-th partial row,
// then a fixed xor tree over those lanes (every lane ends with the same
// total); the last block reduces the loss partials.  Deterministic.
constexpr int kAdamLanes = 16;
constexpr int kAdamParamsPerBlock = 256 / kAdamLanes;

__global__ __launch_bounds__(256) void mlp_adam_kernel(const AdamParams p) {
  if (blockIdx.x == gridDim.x - 1) {
    if (!p.loss_out) return;
    __shared__ float s_loss[256];
    float l = 0.0f;
    for (int b = threadIdx.x; b < p.lblocks; b += 256)
      l += p.bpartial[static_cast<size_t>(b) * kMlpBiasPartial + kBiasLoss];
    s_loss[threadIdx.x] = l;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if (static_cast<int>(threadIdx.x) < o) s_loss[threadIdx.x] += s_loss[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) *p.loss_out = s_loss[0] * p.loss_scale;
    return;
  }
  const int lane = threadIdx.x & (kAdamLanes - 1);
  const int t = blockIdx.x * kAdamParamsPerBlock + static_cast<int>(threadIdx.x) / kAdamLanes;
  float gsum = 0.0f;
  if (t < kMlpParams) {
    const int wi = mlp_param_to_wgrad(t);
    if (wi >= 0) {
      for (int b = lane; b < p.wblocks; b += kAdamLanes) gsum += p.wpartial[static_cast<size_t>(b) * kMlpWeights + wi];
    } else {
      const int bi = mlp_param_to_bias(t);
      for (int b = lane; b < p.bblocks; b += kAdamLanes)
        gsum += p.bpartial[static_cast<size_t>(b) * kMlpBiasPartial + bi];
    }
  }
#pragma unroll
  for (int o = kAdamLanes / 2; o > 0; o >>= 1) gsum += __shfl_xor(gsum, o, kAdamLanes);
  if (t >= kMlpParams || lane != 0) return;
  if (p.grad_out) p.grad_out[t] = gsum;
  if (!p.update) return;
  float m = p.b1 * p.m[t] + (1.0f - p.b1) * gsum;
  float v = p.b2 * p.v[t] + (1.0f - p.b2) * gsum * gsum;
  p.m[t] = m;
  p.v[t] = v;
  float mh = m / p.c1, vh = v / p.c2;
  const float np = p.params[t] - p.lr * mh / (sqrtf(vh) + p.eps);
  p.params[t] = np;
  mlp_pack_one(p.w, t, np);
}

// --------------------------------------------------------- layout probe
// One v_mfma_f32_16x16x32_bf16 with A, B loaded and D stored through the lane
// layout every MLP kernel assumes (header comment): D must equal A . B.
__global__ void mfma_probe_kernel(const float* __restrict__ a, const float* __restrict__ b, float* __restrict__ d) {
  const int lane = threadIdx.x;
  const int g = lane >> 4, n = lane & 15;
  bf16x8 av, bv;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    av[j] = static_cast<__bf16>(a[n * 32 + 8 * g + j]);  // A[row n][k 8g+j]
    bv[j] = static_cast<__bf16>(b[(8 * g + j) * 16 + n]);  // B[k 8g+j][col n]
  }
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = mfma(av, bv, acc);
#pragma unroll
  for (int i = 0; i < 4; ++i) d[(4 * g + i) * 16 + n] = acc[i];  // D[row 4g+i][col n]
}

// ------------------------------------------------------------ error monitor
// One block per cell (mip m, face, 8 x 8 grid over the face).  Each texel
// centre of the cell is queried through the MLP (unit direction of the texel
// centre, roughness m / (mips - 1)) and compared with the cube texel, the
// reference it must match: e = mean over rgb of |mlp - cube| / max(|cube|,
// floor).  The
// cell error is the mean of e over its texels; the cell is accepted when
// that is at most the threshold.  Texel x of a face of size S lies in cell
// floor((x + 0.5) * 8 / S), the cell a direction through the texel centre
// selects; a cell with no texel centre (S < 8) is measured at the texel that
// contains the cell centre.
struct MonitorParams {
  MlpDeviceWeights w;
  const float4* cube;
  CubeMips cm;
  float threshold, floor;
  float* cell_error;  // [mip][face][cy][cx]
  uint8_t* cell_ok;
};

// Texels [lo, hi) of one axis whose centres fall in cell c.
__device__ __forceinline__ void monitor_range(int c, int S, int& lo, int& hi) {
  lo = (c * S - 4 + 7) >> 3;  // ceil((c S - 4) / 8)
  hi = ((c + 1) * S - 4 + 7) >> 3;
  if (hi <= lo) {
    lo = ((2 * c + 1) * S) >> 4;
    hi = lo + 1;
  }
}

__global__ __launch_bounds__(256) void mlp_monitor_kernel(const MonitorParams p) {
  __shared__ MlpSmem s;
  __shared__ MlpWaveStage st[4];
  __shared__ float s_sum[4];
  mlp_stage_weights(s, p.w, threadIdx.x, 256);
  __syncthreads();
  const int cx = blockIdx.x & 7, cy = blockIdx.x >> 3, face = blockIdx.y, m = blockIdx.z;
  const int S = max(1, p.cm.size >> m);
  int x0, x1, y0, y1;
  monitor_range(cx, S, x0, x1);
  monitor_range(cy, S, y0, y1);
  const int w = x1 - x0, area = w * (y1 - y0);
  const float rough = p.cm.mips > 1 ? static_cast<float>(m) / static_cast<float>(p.cm.mips - 1) : 0.0f;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const float4* mip = p.cube + p.cm.off[m] + static_cast<size_t>(face) * S * S;
  float sum = 0.0f;
  for (int base = 0; base < area; base += 256) {
    const int i = base + static_cast<int>(threadIdx.x);
    const bool valid = i < area;
    const int x = x0 + (valid ? i % w : 0), y = y0 + (valid ? i / w : 0);
    const f3 d = normalize_fast(cube_texel_raw_dir(face, x, y, S));
    const f3 ly = mlp_wave_eval(s, st[wave], d.x, d.y, d.z, rough, valid);
    if (valid) {
      const float4 c = mip[y * S + x];
      const float e = fabsf(expf(ly.x) - c.x) / fmaxf(fabsf(c.x), p.floor) +
                      fabsf(expf(ly.y) - c.y) / fmaxf(fabsf(c.y), p.floor) +
                      fabsf(expf(ly.z) - c.z) / fmaxf(fabsf(c.z), p.floor);
      sum += e * (1.0f / 3.0f);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
  if (lane == 0) s_sum[wave] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float err = (s_sum[0] + s_sum[1] + s_sum[2] + s_sum[3]) / static_cast<float>(area);
    const int cell = (m * 6 + face) * 64 + blockIdx.x;
    p.cell_error[cell] = err;
    p.cell_ok[cell] = err <= p.threshold ? 1 : 0;
  }
}

// ------------------------------------------------------------ host side
struct MlpState {
  althea_mlp_desc desc{};
  float* params = nullptr;  // device FP32 master
  float* m = nullptr;
  float* v = nullptr;
  MlpDeviceWeights w{};
  void* weights_mem = nullptr;
  MlpActivations act{};
  void* act_mem = nullptr;
  int act_capacity = 0;
  float* wpartial = nullptr;
  float* bpartial = nullptr;
  int step = 0;
  // Error monitor of the last althea_mlp_build_monitor call.
  float* cell_error = nullptr;
  uint8_t* cell_ok = nullptr;
  int mon_size = 0, mon_mips = 0, mon_accepted = 0;
  // Queue of the MLP-IBL resolve's deferred pixels: a counter, then records.
  void* queue_mem = nullptr;
  size_t queue_regions = 0;
};

// Training grids.  The forward (about 240 registers per lane: two waves per
// SIMD) runs two waves per SIMD; the weight gradients keep one partial row
// per block, so their grid stays at one wave per SIMD (10 MB of partials).
// Row b of the bias partials holds the weight-gradient block b's bias sums
// and the forward block b's loss.
static constexpr int kFwdBlocks = 512;
static constexpr int kWgradBlocks = 256;
static_assert(kFwdBlocks >= kWgradBlocks, "bias partial rows");

static size_t align256(size_t v) { return (v + 255) & ~static_cast<size_t>(255); }

static althea_status alloc_weights(MlpState* s) {
  size_t sizes[] = {64 * 32, 64 * 64, 64 * 64, 16 * 64, 64 * 64, 64 * 64, 64 * 32};
  size_t total = 0;
  for (size_t e : sizes) total += align256(e * 2);
  total += 4 * align256(64 * 4);
  ALTHEA_TRY_HIP(hipMalloc(&s->weights_mem, total));
  char* c = static_cast<char*>(s->weights_mem);
  __bf16** dst[] = {&s->w.w1, &s->w.w2, &s->w.w3, &s->w.w4, &s->w.w2t, &s->w.w3t, &s->w.w4t};
  for (int i = 0; i < 7; ++i) {
    *dst[i] = reinterpret_cast<__bf16*>(c);
    c += align256(sizes[i] * 2);
  }
  float** fb[] = {&s->w.b1, &s->w.b2, &s->w.b3, &s->w.b4};
  for (int i = 0; i < 4; ++i) {
    *fb[i] = reinterpret_cast<float*>(c);
    c += align256(64 * 4);
  }
  return ALTHEA_OK;
}

static althea_status ensure_activations(MlpState* s, int count) {
  int cap = (count + 31) / 32 * 32;
  if (cap <= s->act_capacity) return ALTHEA_OK;
  ALTHEA_TRY_HIP(hipDeviceSynchronize());
  (void)hipFree(s->act_mem);
  s->act_mem = nullptr;
  s->act_capacity = 0;
  size_t B = static_cast<size_t>(cap);
  size_t rows[] = {32, 64, 64, 64, 16, 64, 64, 64};
  size_t total = 0;
  for (size_t r : rows) total += align256(r * B * 2);
  ALTHEA_TRY_HIP(hipMalloc(&s->act_mem, total));
  // Padding columns must read as zero in the weight-gradient pass.
  ALTHEA_TRY_HIP(hipMemset(s->act_mem, 0, total));
  char* c = static_cast<char*>(s->act_mem);
  __bf16** dst[] = {&s->act.x, &s->act.h1, &s->act.h2, &s->act.h3, &s->act.dy, &s->act.dz1, &s->act.dz2,
                    &s->act.dz3};
  for (int i = 0; i < 8; ++i) {
    *dst[i] = reinterpret_cast<__bf16*>(c);
    c += align256(rows[i] * B * 2);
  }
  s->act.stride = B;
  s->act_capacity = cap;
  return ALTHEA_OK;
}

static althea_status launch_pack(MlpState* s, hipStream_t stream) {
  hipLaunchKernelGGL(mlp_pack_kernel, dim3((64 * 64 + 255) / 256), dim3(256), 0, stream, s->params, s->w);
  return check_launch("mlp_pack_kernel");
}

// Forward + backward + reduction (+ Adam when update).
static althea_status run_train(MlpState* s, const float* dir_rough, const float* targets, int count, float* grad_out,
                               float* loss_out, bool update, hipStream_t stream) {
  althea_status st = ensure_activations(s, count);
  if (st != ALTHEA_OK) return st;
  FwdParams fp{};
  fp.w = s->w;
  fp.dir_rough = reinterpret_cast<const float4*>(dir_rough);
  fp.targets = reinterpret_cast<const float4*>(targets);
  fp.count = count;
  fp.grad_scale = 2.0f / (3.0f * static_cast<float>(count));
  fp.act = s->act;
  fp.bias_partial = s->bpartial;
  hipLaunchKernelGGL(mlp_fwd_kernel<true>, dim3(kFwdBlocks), dim3(256), 0, stream, fp);
  st = check_launch("mlp_fwd_kernel<train>");
  if (st != ALTHEA_OK) return st;
  WgradParams wp{};
  wp.act = s->act;
  wp.count = count;
  int nchunks = (count + 31) / 32;
  wp.chunks_per_block = (nchunks + kWgradBlocks - 1) / kWgradBlocks;
  wp.partial = s->wpartial;
  wp.bias_partial = s->bpartial;
  hipLaunchKernelGGL(mlp_wgrad_kernel, dim3(kWgradBlocks), dim3(256), 0, stream, wp);
  st = check_launch("mlp_wgrad_kernel");
  if (st != ALTHEA_OK) return st;
  AdamParams ap{};
  ap.wpartial = s->wpartial;
  ap.wblocks = kWgradBlocks;
  ap.bpartial = s->bpartial;
  ap.bblocks = kWgradBlocks;
  ap.lblocks = kFwdBlocks;
  ap.params = s->params;
  ap.m = s->m;
  ap.v = s->v;
  ap.grad_out = grad_out;
  ap.loss_out = loss_out;
  ap.loss_scale = 1.0f / (3.0f * static_cast<float>(count));
  ap.update = update ? 1 : 0;
  ap.w = s->w;
  if (update) {
    s->step += 1;
    ap.lr = s->desc.learning_rate;
    ap.b1 = s->desc.beta1;
    ap.b2 = s->desc.beta2;
    ap.eps = s->desc.epsilon;
    ap.c1 = static_cast<float>(1.0 - std::pow(static_cast<double>(s->desc.beta1), s->step));
    ap.c2 = static_cast<float>(1.0 - std::pow(static_cast<double>(s->desc.beta2), s->step));
  }
  hipLaunchKernelGGL(mlp_adam_kernel, dim3((kMlpParams + kAdamParamsPerBlock - 1) / kAdamParamsPerBlock + 1),
                     dim3(256), 0, stream, ap);
  st = check_launch("mlp_adam_kernel");
  if (st != ALTHEA_OK) return st;
  return ALTHEA_OK;
}

}  // namespace althea

using namespace althea;

struct althea_mlp {
  MlpState s;
};

namespace althea {
althea_status mlp_inference_view(const althea_mlp* m, MlpInferenceView* v) {
  if (!m || !v) {
    set_error("mlp: null handle");
    return ALTHEA_ERROR_INVALID_ARGUMENT;
  }
  v->w = m->s.w;
  v->cell_ok = m->s.cell_ok;
  v->mon_size = m->s.mon_size;
  v->mon_mips = m->s.mon_mips;
  v->mon_accepted = m->s.mon_accepted;
  v->queue = nullptr;
  v->region_count = nullptr;
  v->queue_regions = 0;
  return ALTHEA_OK;
}

// Queue storage: the region counts (padded to 256 bytes), then the records.
static size_t queue_count_bytes(size_t regions) { return (regions * sizeof(uint32_t) + 255) / 256 * 256; }

althea_status mlp_resolve_queue(const althea_mlp* m, size_t regions, MlpInferenceView* v) {
  althea_status st = mlp_inference_view(m, v);
  if (st != ALTHEA_OK) return st;
  // The queue is the handle's scratch (like its activations), so a const
  // handle may grow it.
  MlpState* s = const_cast<MlpState*>(&m->s);
  if (regions > s->queue_regions) {
    ALTHEA_TRY_HIP(hipDeviceSynchronize());  // an earlier frame may still read the old queue
    (void)hipFree(s->queue_mem);
    s->queue_mem = nullptr;
    s->queue_regions = 0;
    ALTHEA_TRY_HIP(hipMalloc(&s->queue_mem,
                             queue_count_bytes(regions) + regions * kMlpQueueRegion * 3 * sizeof(float4)));
    s->queue_regions = regions;
  }
  v->region_count = static_cast<uint32_t*>(s->queue_mem);
  v->queue = reinterpret_cast<float4*>(static_cast<char*>(s->queue_mem) + queue_count_bytes(s->queue_regions));
  v->queue_regions = s->queue_regions;
  return ALTHEA_OK;
}
}  // namespace althea

extern "C" {

void althea_mlp_desc_default(althea_mlp_desc* d) {
  if (!d) return;
  d->seed = 0x4D4C50ull;
  d->learning_rate = 1e-3f;
  d->beta1 = 0.9f;
  d->beta2 = 0.999f;
  d->epsilon = 1e-8f;
  d->init_output_bias = 0.0f;
}

int althea_mlp_param_count(void) { return kMlpParams; }

althea_status althea_mlp_create(const althea_mlp_desc* desc, althea_mlp** out) {
  ALTHEA_REQUIRE(desc && out, "mlp_create: null argument");
  *out = nullptr;
  ALTHEA_REQUIRE(desc->learning_rate > 0.0f && desc->beta1 >= 0.0f && desc->beta1 < 1.0f && desc->beta2 >= 0.0f &&
                     desc->beta2 < 1.0f && desc->epsilon > 0.0f,
                 "mlp_create: invalid Adam hyper-parameters");
  althea_mlp* m = new (std::nothrow) althea_mlp;
  if (!m) return ALTHEA_ERROR_OUT_OF_MEMORY;
  MlpState* s = &m->s;
  s->desc = *desc;
  auto fail = [&](althea_status st) {
    althea_mlp_destroy(m);
    return st;
  };
  if (hipMalloc(&s->params, kMlpParams * 4) != hipSuccess || hipMalloc(&s->m, kMlpParams * 4) != hipSuccess ||
      hipMalloc(&s->v, kMlpParams * 4) != hipSuccess ||
      hipMalloc(&s->wpartial, static_cast<size_t>(kWgradBlocks) * kMlpWeights * 4) != hipSuccess ||
      hipMalloc(&s->bpartial, static_cast<size_t>(kFwdBlocks) * kMlpBiasPartial * 4) != hipSuccess) {
    set_error("mlp_create: device allocation failed");
    return fail(ALTHEA_ERROR_OUT_OF_MEMORY);
  }
  if (alloc_weights(s) != ALTHEA_OK) return fail(ALTHEA_ERROR_OUT_OF_MEMORY);
  // Initialisation: weights U(-1/sqrt(fan_in), 1/sqrt(fan_in)) from a seeded
  // host generator, hidden biases 0, output bias init_output_bias.
  std::vector<float> h(kMlpParams, 0.0f);
  std::mt19937_64 rng(desc->seed);
  auto fill = [&](int off, int rows, int cols) {
    std::uniform_real_distribution<float> u(-1.0f / std::sqrt(static_cast<float>(cols)),
                                            1.0f / std::sqrt(static_cast<float>(cols)));
    for (int i = 0; i < rows * cols; ++i) h[off + i] = u(rng);
  };
  fill(kOffW1, 64, kMlpIn);
  fill(kOffW2, 64, 64);
  fill(kOffW3, 64, 64);
  fill(kOffW4, kMlpOut, 64);
  for (int i = 0; i < kMlpOut; ++i) h[kOffB4 + i] = desc->init_output_bias;
  if (hipMemcpy(s->params, h.data(), kMlpParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(s->m, 0, kMlpParams * 4) != hipSuccess || hipMemset(s->v, 0, kMlpParams * 4) != hipSuccess) {
    set_error("mlp_create: initial upload failed");
    return fail(ALTHEA_ERROR_HIP);
  }
  if (launch_pack(s, nullptr) != ALTHEA_OK || hipDeviceSynchronize() != hipSuccess) return fail(ALTHEA_ERROR_HIP);
  *out = m;
  return ALTHEA_OK;
}

althea_status althea_mlp_destroy(althea_mlp* m) {
  if (!m) return ALTHEA_OK;
  MlpState* s = &m->s;
  (void)hipFree(s->params);
  (void)hipFree(s->m);
  (void)hipFree(s->v);
  (void)hipFree(s->weights_mem);
  (void)hipFree(s->act_mem);
  (void)hipFree(s->wpartial);
  (void)hipFree(s->bpartial);
  (void)hipFree(s->cell_error);
  (void)hipFree(s->cell_ok);
  (void)hipFree(s->queue_mem);
  delete m;
  return ALTHEA_OK;
}

althea_status althea_mlp_set_params(althea_mlp* m, const float* host_params, hipStream_t stream) {
  ALTHEA_REQUIRE(m && host_params, "mlp_set_params: null argument");
  ALTHEA_TRY_HIP(hipMemcpyAsync(m->s.params, host_params, kMlpParams * 4, hipMemcpyHostToDevice, stream));
  ALTHEA_TRY_HIP(hipMemsetAsync(m->s.m, 0, kMlpParams * 4, stream));
  ALTHEA_TRY_HIP(hipMemsetAsync(m->s.v, 0, kMlpParams * 4, stream));
  m->s.step = 0;
  althea_status st = launch_pack(&m->s, stream);
  if (st != ALTHEA_OK) return st;
  ALTHEA_TRY_HIP(hipStreamSynchronize(stream));
  return ALTHEA_OK;
}

althea_status althea_mlp_get_params(const althea_mlp* m, float* host_params, hipStream_t stream) {
  ALTHEA_REQUIRE(m && host_params, "mlp_get_params: null argument");
  ALTHEA_TRY_HIP(hipMemcpyAsync(host_params, m->s.params, kMlpParams * 4, hipMemcpyDeviceToHost, stream));
  ALTHEA_TRY_HIP(hipStreamSynchronize(stream));
  return ALTHEA_OK;
}

althea_status althea_mlp_infer(const althea_mlp* m, const float* d_dir_rough, int count, float* d_out,
                               hipStream_t stream) {
  ALTHEA_REQUIRE(m && count >= 0, "mlp_infer: invalid argument");
  if (count == 0) return ALTHEA_OK;
  ALTHEA_REQUIRE(d_dir_rough && d_out, "mlp_infer: null buffer");
  FwdParams fp{};
  fp.w = m->s.w;
  fp.dir_rough = reinterpret_cast<const float4*>(d_dir_rough);
  fp.out = reinterpret_cast<float4*>(d_out);
  fp.count = count;
  int tiles = (count + 15) / 16;
  int blocks = std::min((tiles + 3) / 4, 4096);
  hipLaunchKernelGGL(mlp_fwd_kernel<false>, dim3(blocks), dim3(256), 0, stream, fp);
  return check_launch("mlp_fwd_kernel<infer>");
}

althea_status althea_mlp_gradients(althea_mlp* m, const float* d_dir_rough, const float* d_targets, int count,
                                   float* d_grad, float* d_loss, hipStream_t stream) {
  ALTHEA_REQUIRE(m && d_dir_rough && d_targets && d_grad && count >= 1, "mlp_gradients: invalid argument");
  ALTHEA_REQUIRE(count <= (1 << 24), "mlp_gradients: batch %d too large", count);
  return run_train(&m->s, d_dir_rough, d_targets, count, d_grad, d_loss, false, stream);
}

althea_status althea_mlp_set_learning_rate(althea_mlp* m, float learning_rate) {
  ALTHEA_REQUIRE(m && learning_rate > 0.0f, "mlp_set_learning_rate: invalid argument");
  m->s.desc.learning_rate = learning_rate;
  return ALTHEA_OK;
}

althea_status althea_mlp_train_step(althea_mlp* m, const float* d_dir_rough, const float* d_targets, int count,
                                    float* d_loss, hipStream_t stream) {
  ALTHEA_REQUIRE(m && d_dir_rough && d_targets && count >= 1, "mlp_train_step: invalid argument");
  ALTHEA_REQUIRE(count <= (1 << 24), "mlp_train_step: batch %d too large", count);
  return run_train(&m->s, d_dir_rough, d_targets, count, nullptr, d_loss, true, stream);
}

althea_status althea_mfma_bf16_probe(const float* d_a, const float* d_b, float* d_d, hipStream_t stream) {
  ALTHEA_REQUIRE(d_a && d_b && d_d, "mfma_bf16_probe: null buffer");
  hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, stream, d_a, d_b, d_d);
  return check_launch("mfma_probe_kernel");
}

althea_status althea_mlp_build_monitor(althea_mlp* m, const float* d_prefiltered, int size, int mips,
                                      float threshold, float* d_cell_error, int* accepted, hipStream_t stream) {
  ALTHEA_REQUIRE(m && d_prefiltered, "mlp_build_monitor: null argument");
  ALTHEA_REQUIRE(is_pow2(size) && mips >= 1 && mips <= kMaxLevels && (size >> (mips - 1)) >= 1,
                 "mlp_build_monitor: cube %d / %d mips", size, mips);
  ALTHEA_REQUIRE(!(threshold != threshold), "mlp_build_monitor: threshold is NaN");
  MlpState* s = &m->s;
  const int cells = mips * 6 * ALTHEA_MLP_MONITOR_CELLS;
  if (s->mon_mips != mips || !s->cell_error) {
    ALTHEA_TRY_HIP(hipStreamSynchronize(stream));
    (void)hipFree(s->cell_error);
    (void)hipFree(s->cell_ok);
    s->cell_error = nullptr;
    s->cell_ok = nullptr;
    s->mon_size = s->mon_mips = 0;
    ALTHEA_TRY_HIP(hipMalloc(&s->cell_error, cells * sizeof(float)));
    ALTHEA_TRY_HIP(hipMalloc(&s->cell_ok, cells));
  }
  MonitorParams p{};
  p.w = s->w;
  p.cube = reinterpret_cast<const float4*>(d_prefiltered);
  p.cm.size = size;
  p.cm.mips = mips;
  for (int l = 0; l < mips; ++l) p.cm.off[l] = static_cast<uint32_t>(cube_mip_offset(size, l));
  p.threshold = threshold;
  p.floor = ALTHEA_MLP_MONITOR_FLOOR;
  p.cell_error = s->cell_error;
  p.cell_ok = s->cell_ok;
  hipLaunchKernelGGL(mlp_monitor_kernel, dim3(ALTHEA_MLP_MONITOR_CELLS, 6, mips), dim3(256), 0, stream, p);
  althea_status st = check_launch("mlp_monitor_kernel");
  if (st != ALTHEA_OK) return st;
  s->mon_size = size;
  s->mon_mips = mips;
  if (d_cell_error)
    ALTHEA_TRY_HIP(hipMemcpyAsync(d_cell_error, s->cell_error, cells * sizeof(float), hipMemcpyDeviceToDevice,
                                  stream));
  // The accepted count is kept so the resolve can skip the MLP variant when
  // no cell is accepted (built once per environment; the copy synchronises).
  std::vector<uint8_t> ok(cells);
  ALTHEA_TRY_HIP(hipMemcpyAsync(ok.data(), s->cell_ok, cells, hipMemcpyDeviceToHost, stream));
  ALTHEA_TRY_HIP(hipStreamSynchronize(stream));
  int n = 0;
  for (uint8_t v : ok) n += v;
  s->mon_accepted = n;
  if (accepted) *accepted = n;
  return ALTHEA_OK;
}

size_t althea_mlp_state_bytes(const althea_mlp* m) {
  if (!m) return 0;
  // Inference state: BF16 weights of the four layers + FP32 biases.
  return 2 * static_cast<size_t>(kMlpWeights) + 4 * static_cast<size_t>(kMlpParams - kMlpWeights);
}

}  // extern "C"
