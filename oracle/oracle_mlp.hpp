// FP64 oracle for the learned prefiltered-radiance surrogate (TEST
// INFRASTRUCTURE ONLY; see oracle.hpp).  The reference has no learned part;
// this states the network the task specifies:
//   input  = reflection dir (3) + sin/cos(2^k pi dir), k = 0..3 (24) + roughness (1) = 28
//   hidden = 3 x 64, SiLU;  output = 3 (log radiance)
//   loss   = mean over batch and channels of (y - log(target))^2
//   Adam   (beta1 0.9, beta2 0.999, eps 1e-8)
// Parameter order (FP32 master copy): W1[64][28] b1[64] W2[64][64] b2[64]
// W3[64][64] b3[64] W4[3][64] b4[3]  -> 10371 values.
//
// Two precisions:
//   exact  - every value in FP64 from the given FP32 parameters.  Pinned by
//            central finite differences of the loss (tests/test_oracle_mlp.py).
//   bf16   - the rounding points of the BF16 MFMA kernels: weights, the 28
//            encoded inputs, every hidden activation and every back-propagated
//            pre-activation gradient are rounded to bfloat16 (nearest even)
//            before they enter a product (the bias gradients are the sums of
//            the rounded pre-activation gradients); sums stay FP64.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace oracle {
namespace mlp {

constexpr int kIn = 28, kHidden = 64, kOut = 3;
constexpr int kW1 = 0, kB1 = kW1 + kHidden * kIn;
constexpr int kW2 = kB1 + kHidden, kB2 = kW2 + kHidden * kHidden;
constexpr int kW3 = kB2 + kHidden, kB3 = kW3 + kHidden * kHidden;
constexpr int kW4 = kB3 + kHidden, kB4 = kW4 + kOut * kHidden;
constexpr int kParams = kB4 + kOut;  // 10371

inline float bf16_round(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return x;  // inf / nan unchanged
  uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  u &= 0xffff0000u;
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}
inline double q(double x, bool bf16) {
  return bf16 ? static_cast<double>(bf16_round(static_cast<float>(x))) : x;
}

inline void encode(const double dir[3], double roughness, double out[kIn]) {
  for (int c = 0; c < 3; ++c) out[c] = dir[c];
  for (int k = 0; k < 4; ++k) {
    double f = std::ldexp(M_PI, k);
    for (int c = 0; c < 3; ++c) {
      out[3 + 6 * k + c] = std::sin(f * dir[c]);
      out[3 + 6 * k + 3 + c] = std::cos(f * dir[c]);
    }
  }
  out[27] = roughness;
}

inline double sigmoid(double z) { return 1.0 / (1.0 + std::exp(-z)); }
inline double silu(double z) { return z * sigmoid(z); }
inline double silu_grad(double z) {
  double s = sigmoid(z);
  return s * (1.0 + z * (1.0 - s));
}

struct Activations {
  double x[kIn];
  double z1[kHidden], h1[kHidden], z2[kHidden], h2[kHidden], z3[kHidden], h3[kHidden];
  double y[kOut];
};

// Forward pass of one sample.  Inputs: unit-ish direction and roughness.
// P is float (the library's FP32 master copy) or double (finite differences).
template <class P>
inline void forward(const P* params, const double dir[3], double roughness, bool bf16,
                    Activations& a) {
  double enc[kIn];
  encode(dir, roughness, enc);
  for (int i = 0; i < kIn; ++i) a.x[i] = q(enc[i], bf16);
  auto w = [&](int base, int idx) { return q(params[base + idx], bf16); };
  for (int o = 0; o < kHidden; ++o) {
    double s = params[kB1 + o];
    for (int i = 0; i < kIn; ++i) s += w(kW1, o * kIn + i) * a.x[i];
    a.z1[o] = s;
    a.h1[o] = q(silu(s), bf16);
  }
  for (int o = 0; o < kHidden; ++o) {
    double s = params[kB2 + o];
    for (int i = 0; i < kHidden; ++i) s += w(kW2, o * kHidden + i) * a.h1[i];
    a.z2[o] = s;
    a.h2[o] = q(silu(s), bf16);
  }
  for (int o = 0; o < kHidden; ++o) {
    double s = params[kB3 + o];
    for (int i = 0; i < kHidden; ++i) s += w(kW3, o * kHidden + i) * a.h2[i];
    a.z3[o] = s;
    a.h3[o] = q(silu(s), bf16);
  }
  for (int o = 0; o < kOut; ++o) {
    double s = params[kB4 + o];
    for (int i = 0; i < kHidden; ++i) s += w(kW4, o * kHidden + i) * a.h3[i];
    a.y[o] = s;
  }
}

// Loss and gradient summed over `count` samples.  Targets are linear radiance;
// the network predicts log radiance.  grad has kParams entries (overwritten).
// `norm_count` is the batch size the mean is taken over (count for a whole
// batch, larger when this call covers a slice of it).
template <class P>
inline double loss_grad(const P* params, const float* dir_rough, const float* targets, int count,
                        int norm_count, bool bf16, double* grad) {
  std::memset(grad, 0, sizeof(double) * kParams);
  double loss = 0.0;
  const double scale = 2.0 / (3.0 * norm_count);
  auto w = [&](int base, int idx) { return q(params[base + idx], bf16); };
  for (int s = 0; s < count; ++s) {
    const float* in = dir_rough + 4 * s;
    double dir[3] = {in[0], in[1], in[2]};
    Activations a;
    forward(params, dir, in[3], bf16, a);
    double dy[kOut], dz3[kHidden], dz2[kHidden], dz1[kHidden];
    for (int o = 0; o < kOut; ++o) {
      double t = std::log(static_cast<double>(targets[4 * s + o]));
      double e = a.y[o] - t;
      loss += e * e;
      dy[o] = scale * e;
    }
    for (int o = 0; o < kOut; ++o) {
      double g = q(dy[o], bf16);
      grad[kB4 + o] += g;
      for (int i = 0; i < kHidden; ++i) grad[kW4 + o * kHidden + i] += g * a.h3[i];
    }
    for (int i = 0; i < kHidden; ++i) {
      double s3 = 0.0;
      for (int o = 0; o < kOut; ++o) s3 += w(kW4, o * kHidden + i) * q(dy[o], bf16);
      dz3[i] = s3 * silu_grad(a.z3[i]);
    }
    for (int o = 0; o < kHidden; ++o) {
      double g = q(dz3[o], bf16);
      grad[kB3 + o] += g;
      for (int i = 0; i < kHidden; ++i) grad[kW3 + o * kHidden + i] += g * a.h2[i];
    }
    for (int i = 0; i < kHidden; ++i) {
      double s2 = 0.0;
      for (int o = 0; o < kHidden; ++o) s2 += w(kW3, o * kHidden + i) * q(dz3[o], bf16);
      dz2[i] = s2 * silu_grad(a.z2[i]);
    }
    for (int o = 0; o < kHidden; ++o) {
      double g = q(dz2[o], bf16);
      grad[kB2 + o] += g;
      for (int i = 0; i < kHidden; ++i) grad[kW2 + o * kHidden + i] += g * a.h1[i];
    }
    for (int i = 0; i < kHidden; ++i) {
      double s1 = 0.0;
      for (int o = 0; o < kHidden; ++o) s1 += w(kW2, o * kHidden + i) * q(dz2[o], bf16);
      dz1[i] = s1 * silu_grad(a.z1[i]);
    }
    for (int o = 0; o < kHidden; ++o) {
      double g = q(dz1[o], bf16);
      grad[kB1 + o] += g;
      for (int i = 0; i < kIn; ++i) grad[kW1 + o * kIn + i] += g * a.x[i];
    }
  }
  return loss / (3.0 * norm_count);
}

// One Adam step (Kingma & Ba 2015, bias-corrected), in FP64.
inline void adam(double* p, double* m, double* v, const double* g, int n, int step, double lr,
                 double b1, double b2, double eps) {
  double c1 = 1.0 - std::pow(b1, step), c2 = 1.0 - std::pow(b2, step);
  for (int i = 0; i < n; ++i) {
    m[i] = b1 * m[i] + (1.0 - b1) * g[i];
    v[i] = b2 * v[i] + (1.0 - b2) * g[i] * g[i];
    double mh = m[i] / c1, vh = v[i] / c2;
    p[i] -= lr * mh / (std::sqrt(vh) + eps);
  }
}

}  // namespace mlp
}  // namespace oracle
