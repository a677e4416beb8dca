// Device building blocks of the prefiltered-radiance MLP shared by the
// training kernels, the error monitor and the MLP-IBL resolve.
//
// MFMA lane layout (v_mfma_f32_16x16x32_bf16, lane l = 16 g + n):
//   A (16 x 32): lane holds W[row = n][k = 8 g + j], j = 0..7
//   B (32 x 16): lane holds X^T[k = 8 g + j][sample = n]
//   D (16 x 16): lane holds D[row = 4 g + i][sample = n], i = 0..3
// Hidden layers feed the next layer in registers through the K permutation
// baked into the packed weights (mlp_layout.hpp).
#pragma once

#include <hip/hip_runtime.h>

#include "althea/althea.h"
#include "device_math.hpp"
#include "mlp_layout.hpp"
#include "mlp_view.hpp"

namespace althea {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 load8(const __bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }

__device__ __forceinline__ bf16x8 pack8(f32x4 lo, f32x4 hi) {
  bf16x8 r;
  r[0] = static_cast<__bf16>(lo[0]);
  r[1] = static_cast<__bf16>(lo[1]);
  r[2] = static_cast<__bf16>(lo[2]);
  r[3] = static_cast<__bf16>(lo[3]);
  r[4] = static_cast<__bf16>(hi[0]);
  r[5] = static_cast<__bf16>(hi[1]);
  r[6] = static_cast<__bf16>(hi[2]);
  r[7] = static_cast<__bf16>(hi[3]);
  return r;
}

// Logistic sigmoid with the hardware exp2 and reciprocal (about 2 ulp); the
// only activation path, shared by training and inference.
__device__ __forceinline__ float sigmoidf(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.44269504088896341f * z));
}

// sin(pi x) and cos(pi x) for the components of a unit direction (|x| <= 1
// up to rounding): x is reflected into r in [-1/2, 1/2] (for |x| > 1/2,
// r = sign(x) - x: sin(pi r) = sin(pi x), cos(pi r) = -cos(pi x)), then
// sin(pi r) = r P(r^2) and cos(pi r) = Q(r^2) with P, Q of degree 4 fitted
// to the functions on [0, 1/2] by iteratively reweighted least squares on
// Chebyshev nodes (3.3e-9 and 4.7e-8 absolute in exact arithmetic, about
// 2e-7 evaluated in FP32: far below the BF16 rounding of the features).
__device__ __forceinline__ void sincospi_unit(float x, float& s, float& c) {
  const bool far = fabsf(x) > 0.5f;
  const float r = far ? copysignf(1.0f, x) - x : x;
  const float u = r * r;
  float ps = fmaf(u, 0.0772201270f, -0.5980451703f);
  ps = fmaf(ps, u, 2.5500314236f);
  ps = fmaf(ps, u, -5.1677069664f);
  ps = fmaf(ps, u, 3.1415925026f);
  float pc = fmaf(u, 0.2196981013f, -1.3318809271f);
  pc = fmaf(pc, u, 4.0584120750f);
  pc = fmaf(pc, u, -4.9347929955f);
  pc = fmaf(pc, u, 0.9999999404f);
  s = r * ps;
  c = far ? -pc : pc;
}

// The input features of one query (28, zero-padded to 32): dir (0..2), then
// for k = 0..3 sin(2^k pi d_c) at 3 + 6k + c and cos(2^k pi d_c) at
// 3 + 6k + 3 + c (c = x, y, z), then roughness (27).  sin/cos(pi d) come
// from sincospi_unit, the higher octaves from the double-angle identities
// sin 2a = 2 sin a cos a, cos 2a = (cos a - sin a)(cos a + sin a) (about
// 1e-6 absolute after three doublings, far below the BF16 input rounding).
__device__ __forceinline__ void encode_all(float dx, float dy, float dz, float rough, float f[32]) {
  f[0] = dx;
  f[1] = dy;
  f[2] = dz;
  const float d[3] = {dx, dy, dz};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float s, co;
    sincospi_unit(d[c], s, co);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      f[3 + 6 * k + c] = s;
      f[3 + 6 * k + 3 + c] = co;
      const float s2 = 2.0f * s * co;
      co = (co - s) * (co + s);
      s = s2;
    }
  }
  f[27] = rough;
  f[28] = f[29] = f[30] = f[31] = 0.0f;
}

// -------------------------------------------- register-resident inference
// The forward A operands of this lane (feature group g = lane / 16, row
// n = lane % 16 of each 16-row block), 88 VGPRs, loaded once per wave; the
// streaming inference layout of mlp_fwd_kernel.
struct MlpRegWeights {
  bf16x8 a1[4], a2[4][2], a3[4][2], a4[2];
};
__device__ __forceinline__ void mlp_load_reg_weights(const MlpDeviceWeights& w, int g, int n, MlpRegWeights& r) {
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    r.a1[b] = load8(w.w1 + (16 * b + n) * 32 + 8 * g);
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
      r.a2[b][kb] = load8(w.w2 + (16 * b + n) * 64 + 32 * kb + 8 * g);
      r.a3[b][kb] = load8(w.w3 + (16 * b + n) * 64 + 32 * kb + 8 * g);
    }
  }
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) r.a4[kb] = load8(w.w4 + n * 64 + 32 * kb + 8 * g);
}

// Log-radiance of one 16-query tile: lane (g, n) passes query n (every group
// of the tile passes the same query n); the result for query n sits in lane
// n of group 0 (rows 0..2 of y).  Arithmetic identical to mlp_fwd_kernel's
// inference: same features, MFMA order, bias and SiLU.
__device__ __forceinline__ f32x4 mlp_tile_eval(const MlpRegWeights& r, const MlpDeviceWeights& w, int g, float dx,
                                               float dy, float dz, float rough) {
  float f[32];
  encode_all(dx, dy, dz, rough, f);
  bf16x8 x;
#pragma unroll
  for (int j = 0; j < 8; ++j)
    x[j] = static_cast<__bf16>(g == 0 ? f[j] : (g == 1 ? f[8 + j] : (g == 2 ? f[16 + j] : f[24 + j])));
  f32x4 h[4];
  auto hidden = [&](auto&& wtile, const bf16x8* in, int nk, const float* bias) {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < nk; ++kb) acc = mfma(wtile(b, kb), in[kb], acc);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float z = acc[i] + bias[16 * b + 4 * g + i];
        h[b][i] = z * sigmoidf(z);
      }
    }
  };
  hidden([&](int b, int) { return r.a1[b]; }, &x, 1, w.b1);
  const bf16x8 hb1[2] = {pack8(h[0], h[1]), pack8(h[2], h[3])};
  hidden([&](int b, int kb) { return r.a2[b][kb]; }, hb1, 2, w.b2);
  const bf16x8 hb2[2] = {pack8(h[0], h[1]), pack8(h[2], h[3])};
  hidden([&](int b, int kb) { return r.a3[b][kb]; }, hb2, 2, w.b3);
  const bf16x8 hb3[2] = {pack8(h[0], h[1]), pack8(h[2], h[3])};
  f32x4 y = {0.f, 0.f, 0.f, 0.f};
  y = mfma(r.a4[0], hb3[0], y);
  y = mfma(r.a4[1], hb3[1], y);
#pragma unroll
  for (int i = 0; i < 4; ++i) y[i] += w.b4[4 * g + i];
  return y;
}

// ------------------------------------------------- LDS-resident inference
// The packed forward operands in LDS, 16-byte groups XOR-swizzled so that
// each aligned group of 8 lanes of an A-operand read hits distinct banks:
// 64-wide rows (8 groups) swap group G with G ^ (row & 7), 32-wide rows
// (4 groups) with G ^ ((row >> 1) & 3).
struct MlpSmem {
  bf16x8 w1[64 * 4];
  bf16x8 w2[64 * 8];
  bf16x8 w3[64 * 8];
  bf16x8 w4[16 * 8];
  float b[3 * 64 + 16];  // b1, b2, b3, b4 (zero-padded to 16)
};
// Per-wave staging of the 64 encoded queries (32 BF16 features each).
struct MlpWaveStage {
  bf16x8 feat[64 * 4];
};

__device__ __forceinline__ int mlp_swz64(int row, int grp) { return row * 8 + (grp ^ (row & 7)); }
__device__ __forceinline__ int mlp_swz32(int row, int grp) { return row * 4 + (grp ^ ((row >> 1) & 3)); }

// Block-cooperative copy of the packed weights into LDS; the caller
// synchronises the block before the first evaluation.
__device__ inline void mlp_stage_weights(MlpSmem& s, const MlpDeviceWeights& w, int tid, int nthreads) {
  for (int i = tid; i < 64 * 4; i += nthreads) s.w1[mlp_swz32(i >> 2, i & 3)] = load8(w.w1 + 8 * i);
  for (int i = tid; i < 64 * 8; i += nthreads) {
    s.w2[mlp_swz64(i >> 3, i & 7)] = load8(w.w2 + 8 * i);
    s.w3[mlp_swz64(i >> 3, i & 7)] = load8(w.w3 + 8 * i);
  }
  for (int i = tid; i < 16 * 8; i += nthreads) s.w4[mlp_swz64(i >> 3, i & 7)] = load8(w.w4 + 8 * i);
  for (int i = tid; i < 64; i += nthreads) {
    s.b[i] = w.b1[i];
    s.b[64 + i] = w.b2[i];
    s.b[128 + i] = w.b3[i];
    if (i < 16) s.b[192 + i] = w.b4[i];
  }
}

// Orders this wave's LDS writes before its later LDS reads from other lanes.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Log-radiance of this lane's query (d, rough), evaluated for the wave as
// four 16-sample MFMA tiles (lanes 16t .. 16t+15); a tile none of whose lanes
// sets `want` is skipped and its lanes get 0.  Every lane of the wave must be
// active.  Arithmetic identical to the streaming inference kernel: same
// features, same MFMA order, same bias and SiLU.
__device__ inline f3 mlp_wave_eval(const MlpSmem& s, MlpWaveStage& st, float dx, float dy, float dz, float rough,
                                   bool want = true) {
  const uint64_t wanted = __builtin_amdgcn_ballot_w64(want);
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, n = lane & 15;
  wave_lds_sync();  // the previous evaluation's reads are done
  float f[32];
  encode_all(dx, dy, dz, rough, f);
#pragma unroll
  for (int G = 0; G < 4; ++G) {
    bf16x8 v;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = static_cast<__bf16>(f[8 * G + j]);
    st.feat[mlp_swz32(lane, G)] = v;
  }
  wave_lds_sync();
  f3 out = mk(0.0f, 0.0f, 0.0f);
#pragma unroll 1
  for (int t = 0; t < 4; ++t) {
    if (((wanted >> (16 * t)) & 0xFFFFull) == 0) continue;  // wave-uniform
    // Reload the weights per tile: hoisting the 88 loop-invariant operand
    // reads out of this loop would hold 352 VGPRs.
    asm volatile("" ::: "memory");
    const bf16x8 x = st.feat[mlp_swz32(16 * t + n, g)];
    f32x4 h[4];
    auto act = [&](f32x4 acc, int bias0, int b) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float z = acc[i] + s.b[bias0 + 16 * b + 4 * g + i];
        h[b][i] = z * sigmoidf(z);
      }
    };
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = mfma(s.w1[mlp_swz32(16 * b + n, g)], x, acc);
      act(acc, 0, b);
    }
    bf16x8 hb[2] = {pack8(h[0], h[1]), pack8(h[2], h[3])};
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < 2; ++kb) acc = mfma(s.w2[mlp_swz64(16 * b + n, 4 * kb + g)], hb[kb], acc);
      act(acc, 64, b);
    }
    hb[0] = pack8(h[0], h[1]);
    hb[1] = pack8(h[2], h[3]);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < 2; ++kb) acc = mfma(s.w3[mlp_swz64(16 * b + n, 4 * kb + g)], hb[kb], acc);
      act(acc, 128, b);
    }
    hb[0] = pack8(h[0], h[1]);
    hb[1] = pack8(h[2], h[3]);
    f32x4 y = {0.f, 0.f, 0.f, 0.f};
    y = mfma(s.w4[mlp_swz64(n, g)], hb[0], y);
    y = mfma(s.w4[mlp_swz64(n, 4 + g)], hb[1], y);
#pragma unroll
    for (int i = 0; i < 4; ++i) y[i] += s.b[192 + 4 * g + i];
    // Sample 16 t + n's outputs sit in lane n (g == 0), rows 0..2.
    const float y0 = __shfl(y[0], n, 64), y1 = __shfl(y[1], n, 64), y2 = __shfl(y[2], n, 64);
    if (g == t) out = mk(y0, y1, y2);
  }
  return out;
}

}  // namespace althea
