// Double-float ("df32") arithmetic: a value is the unevaluated sum hi + lo of
// two FP32 numbers, about 44 significant bits.  Used by the few address
// computations whose FP32 rounding the IBL sources amplify: a mirror-like
// lookup of a texture whose neighbouring texels differ by orders of magnitude
// turns a texel-coordinate error e into a relative error of about e times that
// contrast, and FP32 carries e ~ 1e-4 texel at 4096-texel extents.  Nothing
// here uses FP64.
//
// Building blocks (error-free transformations with FMA): two_sum (Knuth),
// fast_two_sum (Dekker, |a| >= |b|), two_prod (p = ab, e = fma(a, b, -p)).
// atan on [0, 1] reduces t to the nearest k/64 with a 65-entry table of
// atan(k/64) split into FP32 pairs: atan(t) = atan(k/64) + atan(u),
// u = (t - k/64) / (1 + t k/64), |u| <= 1/128, and atan(u) = u - u^3/3 + u^5/5
// with the cubic and quintic terms in FP32 (their size is below 2e-7, so FP32
// rounding leaves them correct to ~1e-14).  Absolute error of df32_atan2 is
// about 1e-13 rad.
//
// Host and device: the functions are __host__ __device__ and use only IEEE
// FP32 operations (+, -, *, /, sqrtf, fmaf, floorf) and a reciprocal (df_rcp:
// v_rcp_f32 in device code, the division on the host), so a host build gives
// the kernels' results up to the last bits of quotients, which df_div's
// remainder correction holds to the same bound on both.
#pragma once

#include <math.h>

namespace althea {

// The error-free transformations hold only if every product and sum rounds
// as written: a compiler that fuses a * b with a later + c (HIP's default
// contraction does, across statements) turns p = a b; e = fma(a, b, -p);
// c - p - e into c - ab - e and counts the low part twice.  Every function
// here therefore forbids contraction in its body; the fused multiply-adds it
// wants are explicit fmaf.
#define ALTHEA_DF32_NO_CONTRACT _Pragma("clang fp contract(off)")

struct df32 {
  float hi, lo;
};

constexpr float kAtanK64Hi[65] = {
    0.000000000e+00f, 1.562372874e-02f, 3.123983368e-02f, 4.684071243e-02f, 6.241881102e-02f,
    7.796663046e-02f, 9.347677976e-02f, 1.089419574e-01f, 1.243549958e-01f, 1.397088766e-01f,
    1.549967378e-01f, 1.702119261e-01f, 1.853479445e-01f, 2.003985494e-01f, 2.153577060e-01f,
    2.302195877e-01f, 2.449786663e-01f, 2.596296370e-01f, 2.741674483e-01f, 2.885873616e-01f,
    3.028848767e-01f, 3.170557618e-01f, 3.310960829e-01f, 3.450021744e-01f, 3.587706685e-01f,
    3.723984361e-01f, 3.858826756e-01f, 3.992207646e-01f, 4.124104381e-01f, 4.254496396e-01f,
    4.383365512e-01f, 4.510696530e-01f, 4.636476040e-01f, 4.760693312e-01f, 4.883339405e-01f,
    5.004408360e-01f, 5.123894811e-01f, 5.241796374e-01f, 5.358112454e-01f, 5.472843647e-01f,
    5.585992932e-01f, 5.697564483e-01f, 5.807563663e-01f, 5.915997028e-01f, 6.022873521e-01f,
    6.128202081e-01f, 6.231993437e-01f, 6.334258914e-01f, 6.435011029e-01f, 6.534263492e-01f,
    6.632030010e-01f, 6.728325486e-01f, 6.823165417e-01f, 6.916566491e-01f, 7.008544207e-01f,
    7.099116445e-01f, 7.188299894e-01f, 7.276113033e-01f, 7.362574339e-01f, 7.447701097e-01f,
    7.531512976e-01f, 7.614027858e-01f, 7.695264816e-01f, 7.775242925e-01f, 7.853981853e-01f,
};
constexpr float kAtanK64Lo[65] = {
    0.000000000e+00f, -1.242088238e-10f, -2.525072096e-10f, 4.878766835e-10f, -1.027277929e-09f,
    3.372710555e-09f, 1.399659277e-09f, -3.646797864e-10f, -1.240382241e-09f, -2.320638570e-09f,
    4.086149641e-09f, -8.171638366e-10f, 5.497632571e-09f, 4.388355457e-09f, -6.252999007e-09f,
    -4.068339554e-10f, -3.178677765e-09f, -7.594687546e-09f, 2.837416968e-09f, 3.199044363e-10f,
    -8.353086223e-09f, -8.604970070e-09f, -6.221664517e-09f, 2.829663615e-09f, 1.763949875e-09f,
    1.060726529e-08f, -6.249661677e-09f, 4.938259135e-09f, 3.536626769e-09f, -2.188749848e-09f,
    8.668535223e-09f, 2.954313238e-09f, 5.012158688e-09f, -8.463671985e-10f, 1.055042453e-08f,
    -2.280546418e-08f, -2.075691974e-08f, -8.649185368e-09f, -7.480973174e-09f, 1.628711921e-08f,
    2.211159789e-08f, 5.214087828e-09f, -1.268522887e-08f, 7.500028154e-09f, -5.950149262e-09f,
    -5.907421219e-09f, -1.374726999e-08f, -8.430238729e-09f, 5.868937336e-09f, -7.982293937e-09f,
    -8.316245470e-09f, -1.024553442e-09f, 1.320299514e-08f, -2.725950132e-08f, -1.277747597e-08f,
    -2.599524507e-08f, 1.018833551e-08f, 2.929704301e-08f, -4.909868068e-09f, 1.606260192e-08f,
    -1.660708016e-08f, -1.597246779e-08f, -1.222759671e-09f, 1.790432336e-08f, -2.185569414e-08f,
};

__host__ __device__ __forceinline__ df32 df_two_sum(float a, float b) {
  ALTHEA_DF32_NO_CONTRACT
  const float s = a + b;
  const float bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}
// |a| >= |b| (or a == 0).
__host__ __device__ __forceinline__ df32 df_fast_two_sum(float a, float b) {
  ALTHEA_DF32_NO_CONTRACT
  const float s = a + b;
  return {s, b - (s - a)};
}
__host__ __device__ __forceinline__ df32 df_two_prod(float a, float b) {
  ALTHEA_DF32_NO_CONTRACT
  const float p = a * b;
  return {p, fmaf(a, b, -p)};
}
__host__ __device__ __forceinline__ df32 df_of(float a) { return {a, 0.0f}; }
__host__ __device__ __forceinline__ df32 df_neg(df32 a) { return {-a.hi, -a.lo}; }
__host__ __device__ __forceinline__ df32 df_abs(df32 a) { return a.hi < 0.0f ? df_neg(a) : a; }

// a + b with both parts summed exactly (accurate also under cancellation).
__host__ __device__ __forceinline__ df32 df_add(df32 a, df32 b) {
  ALTHEA_DF32_NO_CONTRACT
  df32 s = df_two_sum(a.hi, b.hi);
  const df32 t = df_two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = df_fast_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return df_fast_two_sum(s.hi, s.lo);
}
__host__ __device__ __forceinline__ df32 df_sub(df32 a, df32 b) { return df_add(a, df_neg(b)); }
// a + b with the low parts added in FP32 (Dekker's add): absolute error about
// 2^-44 (|a| + |b|), so relative only to the operands, not to a + b.  For sums
// whose result must be accurate on the scale of the operands (the components
// of a direction, against its length), not under cancellation.
__host__ __device__ __forceinline__ df32 df_add_fast(df32 a, df32 b) {
  ALTHEA_DF32_NO_CONTRACT
  df32 s = df_two_sum(a.hi, b.hi);
  s.lo += a.lo + b.lo;
  return df_fast_two_sum(s.hi, s.lo);
}
__host__ __device__ __forceinline__ df32 df_sub_fast(df32 a, df32 b) { return df_add_fast(a, df_neg(b)); }
// 1 / x to within an ulp or two: v_rcp_f32 in device code, the division on
// the host.  Quotients built on it are corrected by an exact remainder, so
// host and device results meet the same bound and may differ in the last bits
// of the low part.
__host__ __device__ __forceinline__ float df_rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}
__host__ __device__ __forceinline__ df32 df_add_f(df32 a, float b) {
  ALTHEA_DF32_NO_CONTRACT
  df32 s = df_two_sum(a.hi, b);
  s.lo += a.lo;
  return df_fast_two_sum(s.hi, s.lo);
}
__host__ __device__ __forceinline__ df32 df_mul(df32 a, df32 b) {
  ALTHEA_DF32_NO_CONTRACT
  df32 p = df_two_prod(a.hi, b.hi);
  p.lo = fmaf(a.hi, b.lo, fmaf(a.lo, b.hi, p.lo));
  return df_fast_two_sum(p.hi, p.lo);
}
__host__ __device__ __forceinline__ df32 df_mul_f(df32 a, float b) {
  ALTHEA_DF32_NO_CONTRACT
  df32 p = df_two_prod(a.hi, b);
  p.lo = fmaf(a.lo, b, p.lo);
  return df_fast_two_sum(p.hi, p.lo);
}
// a / b: an FP32 quotient q1 from the reciprocal, corrected by the
// remainder r = a - q1 b.  q1 b.hi is taken exactly (two_prod); q1 is within
// a few ulps of a.hi / b.hi, so a.hi - (q1 b.hi).hi is exact (Sterbenz) and
// the rest of r is a few FP32 roundings relative to r itself.  The correction
// q2 = r / b.hi is about 2^-23 of q1, so its own error is about 2^-46 of the
// result.
__host__ __device__ __forceinline__ df32 df_div(df32 a, df32 b) {
  ALTHEA_DF32_NO_CONTRACT
  const float rb = df_rcp(b.hi);
  const float q1 = a.hi * rb;
  const df32 p = df_two_prod(q1, b.hi);
  const float r = (((a.hi - p.hi) - p.lo) + a.lo) - q1 * b.lo;
  const float q2 = r * rb;
  return df_fast_two_sum(q1, q2);
}
// sqrt(a), a >= 0: one Newton correction of the FP32 root.
__host__ __device__ __forceinline__ df32 df_sqrt(df32 a) {
  ALTHEA_DF32_NO_CONTRACT
  if (!(a.hi > 0.0f)) return {0.0f, 0.0f};
  const float s = sqrtf(a.hi);
  const df32 p = df_two_prod(s, s);
  const float r = ((a.hi - p.hi) - p.lo) + a.lo;
  return df_fast_two_sum(s, r * df_rcp(2.0f * s));
}
// x^2 + z^2 (two non-negative terms: no cancellation, so the fast add).
__host__ __device__ __forceinline__ df32 df_sum_squares(df32 x, df32 z) {
  ALTHEA_DF32_NO_CONTRACT
  return df_add_fast(df_mul(x, x), df_mul(z, z));
}

constexpr df32 kDfPi = {3.141592741e+00f, -8.742277657e-08f};
constexpr df32 kDfHalfPi = {1.570796371e+00f, -4.371138829e-08f};
constexpr df32 kDfInv2Pi = {1.591549367e-01f, 6.420638243e-09f};
constexpr df32 kDfInvPi = {3.183098733e-01f, 1.284127649e-08f};

// atan(t) for t in [0, 1].
__host__ __device__ __forceinline__ df32 df_atan_unit(df32 t) {
  ALTHEA_DF32_NO_CONTRACT
  const int k = static_cast<int>(fmaf(t.hi, 64.0f, 0.5f));
  const float c = static_cast<float>(k) * 0.015625f;
  const df32 num = df_two_sum(t.hi - c, t.lo);  // t.hi - c is exact (Sterbenz; c = 0 for k = 0)
  const df32 den = df_add_f(df_mul_f(t, c), 1.0f);
  const df32 u = df_div(num, den);
  const float p = u.hi * u.hi;
  const float tail = u.hi * p * fmaf(p, fmaf(p, -0.142857149f, 0.2f), -0.333333343f);
  // atan(k/64) >= |u| for k >= 1 and 0 for k = 0: no cancellation.
  return df_add_f(df_add_fast(df32{kAtanK64Hi[k], kAtanK64Lo[k]}, u), tail);
}

// atan2(y, x) in (-pi, pi].
__host__ __device__ __forceinline__ df32 df_atan2(df32 y, df32 x) {
  ALTHEA_DF32_NO_CONTRACT
  const df32 ax = df_abs(x), ay = df_abs(y);
  const bool swap = ay.hi > ax.hi;
  const df32 num = swap ? ax : ay, den = swap ? ay : ax;
  df32 r = den.hi > 0.0f ? df_atan_unit(df_div(num, den)) : df32{0.0f, 0.0f};
  // r <= pi/4 before the first reflection and <= pi/2 before the second.
  if (swap) r = df_sub_fast(kDfHalfPi, r);
  if (x.hi < 0.0f) r = df_sub_fast(kDfPi, r);
  return y.hi < 0.0f ? df_neg(r) : r;
}

// A texel coordinate x split into floor(x) and the bilinear weights
// a = x - floor(x) and b = 1 - a, each rounded once to FP32 from the df
// value (a may round up to 1 while b keeps the small remainder) (the filter forms t0 b + t1 a, whose relative error stays at a few
// FP32 ulps whatever the contrast of t0 and t1).
struct TexelSplit {
  int i;
  float a, b;
};
__host__ __device__ __forceinline__ TexelSplit df_texel(df32 x) {
  ALTHEA_DF32_NO_CONTRACT
  float f = floorf(x.hi);
  // x.hi - floor(x.hi) is not always exact (x.hi in (-1, 0) has finer bits
  // than x.hi + 1 can hold), so it is a two_sum.
  df32 a = df_add_f(df_two_sum(x.hi, -f), x.lo);  // x - floor(x.hi), in (-1, 2)
  // The fraction is the df value: 1 - 1e-9 has hi = 1 and lo < 0.
  if (a.hi < 0.0f || (a.hi == 0.0f && a.lo < 0.0f)) {
    f -= 1.0f;
    a = df_add_f(a, 1.0f);
  } else if (a.hi > 1.0f || (a.hi == 1.0f && a.lo >= 0.0f)) {
    f += 1.0f;
    a = df_add_f(a, -1.0f);
  }
  const df32 b = df_add_f(df_neg(a), 1.0f);
  return {static_cast<int>(f), a.hi, b.hi};
}

}  // namespace althea
