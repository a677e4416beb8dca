// Mirror-like texture lookups with double-float (df32) texel addressing.
//
// A lookup that reads a texture at one direction (a roughness-0 prefiltered
// texel, the background, a reflection that reads mip 0 of the prefiltered
// cube) inherits the full texel-to-texel contrast of what it reads.  On a
// source with per-texel noise, neighbours differ by up to 1000x, so the
// bilinear value moves by (texel-coordinate error) x (contrast).  These
// functions carry the direction, the angles or face coordinates and the
// texel coordinate in df32 (about 1e-13 relative), round each bilinear weight
// once to FP32, and filter in the form t0 b + t1 a with both weights
// non-negative, whose relative error is a few FP32 ulps whatever the
// contrast.  The texel fetches and the Vulkan filtering rules (u W - 0.5,
// floor, REPEAT / CLAMP_TO_EDGE, seamless cube edges) are those of
// device_math.hpp.
#pragma once

#include "device_math.hpp"
#include "df32.hpp"

namespace althea {

struct df3 {
  df32 x, y, z;
};

__device__ __forceinline__ f3 weighted4(f3 t00, f3 t10, f3 t01, f3 t11, const TexelSplit& sx,
                                        const TexelSplit& sy) {
  const f3 top = mk(fmaf(t10.x, sx.a, t00.x * sx.b), fmaf(t10.y, sx.a, t00.y * sx.b), fmaf(t10.z, sx.a, t00.z * sx.b));
  const f3 bot = mk(fmaf(t11.x, sx.a, t01.x * sx.b), fmaf(t11.y, sx.a, t01.y * sx.b), fmaf(t11.z, sx.a, t01.z * sx.b));
  return mk(fmaf(bot.x, sy.a, top.x * sy.b), fmaf(bot.y, sy.a, top.y * sy.b), fmaf(bot.z, sy.a, top.z * sy.b));
}

// Equirect angles of direction d: yaw = atan2(z, x), pitch = atan2(y, |xz|)
// (the shading orientation of GlobalResources.glsl:48-54 the whole build
// uses; neither depends on |d|).  `pole`: the precompute's pole branch
// (PreFilterEnvMap.comp:104-112), decided by the caller: yaw 0, pitch +-pi/2.
__device__ __forceinline__ void equirect_angles_df(const df3& d, int pole, df32& yaw, df32& pitch) {
  if (pole != 0) {
    yaw = df_of(0.0f);
    pitch = pole > 0 ? kDfHalfPi : df_neg(kDfHalfPi);
  } else {
    yaw = df_atan2(d.z, d.x);
    pitch = df_atan2(d.y, df_sqrt(df_sum_squares(d.x, d.z)));
  }
}

// Texel coordinates of (yaw, pitch) on a w x h level (w, h powers of two, so
// the scale factors are exact): u = yaw / 2pi + 0.5, v = 0.5 - pitch / pi,
// x = u w - 0.5, y = v h - 0.5.
__device__ __forceinline__ void equirect_texel_df(df32 yaw, df32 pitch, int w, int h, df32& x, df32& y) {
  const float fw = static_cast<float>(w), fh = static_cast<float>(h);
  x = df_add_f(df_mul(yaw, df32{kDfInv2Pi.hi * fw, kDfInv2Pi.lo * fw}), 0.5f * fw - 0.5f);
  y = df_add_f(df_mul(pitch, df32{-kDfInvPi.hi * fh, -kDfInvPi.lo * fh}), 0.5f * fh - 0.5f);
}

__device__ __forceinline__ void equirect_coords_df(const df3& d, int w, int h, int pole, df32& x, df32& y) {
  df32 yaw, pitch;
  equirect_angles_df(d, pole, yaw, pitch);
  equirect_texel_df(yaw, pitch, w, h, x, y);
}

// Bilinear REPEAT fetch of a pyramid level at df32 texel coordinates.
__device__ __forceinline__ f3 level_bilinear_df(const char* __restrict__ pyr_bytes, const LevelC& c, df32 x,
                                                df32 y) {
  const TexelSplit sx = df_texel(x), sy = df_texel(y);
  const uint32_t x0 = static_cast<uint32_t>(sx.i) & c.wmask, x1 = static_cast<uint32_t>(sx.i + 1) & c.wmask;
  const uint32_t r0 = (static_cast<uint32_t>(sy.i) & c.hmask) << c.wshift;
  const uint32_t r1 = (static_cast<uint32_t>(sy.i + 1) & c.hmask) << c.wshift;
  const char* lvl = pyr_bytes + (static_cast<size_t>(c.off) << 4);
  auto ld = [&](uint32_t t) { return xyz(*reinterpret_cast<const float4*>(lvl + (t << 4))); };
  return weighted4(ld(r0 + x0), ld(r0 + x1), ld(r1 + x0), ld(r1 + x1), sx, sy);
}

// Bilinear CLAMP_TO_EDGE fetch of a w x h image at df32 texel coordinates.
__device__ __forceinline__ f3 clamp_bilinear_df(const float4* __restrict__ level, int w, int h, df32 x, df32 y) {
  const TexelSplit sx = df_texel(x), sy = df_texel(y);
  const int x0 = min(max(sx.i, 0), w - 1), x1 = min(max(sx.i + 1, 0), w - 1);
  const int y0 = min(max(sy.i, 0), h - 1), y1 = min(max(sy.i + 1, 0), h - 1);
  return weighted4(xyz(level[y0 * w + x0]), xyz(level[y0 * w + x1]), xyz(level[y1 * w + x0]),
                   xyz(level[y1 * w + x1]), sx, sy);
}

// Seamless bilinear fetch of one bordered cube mip (faces of size S, a power
// of two, stored with their seamless border as BorderedCube) in direction R:
// face by the major axis of R's leading parts (the Vulkan table of
// cube_select), (sc / ma, tc / ma) and the texel coordinates in df32.
__device__ __forceinline__ f3 cube_bilinear_df(const float4* __restrict__ mip, int S, const df3& R) {
  const float ax = fabsf(R.x.hi), ay = fabsf(R.y.hi), az = fabsf(R.z.hi);
  int face;
  df32 sc, tc, ma;
  if (ax >= ay && ax >= az) {
    ma = df_abs(R.x);
    if (R.x.hi >= 0.0f) { face = 0; sc = df_neg(R.z); tc = df_neg(R.y); }
    else { face = 1; sc = R.z; tc = df_neg(R.y); }
  } else if (ay >= az) {
    ma = df_abs(R.y);
    if (R.y.hi >= 0.0f) { face = 2; sc = R.x; tc = R.z; }
    else { face = 3; sc = R.x; tc = df_neg(R.z); }
  } else {
    ma = df_abs(R.z);
    if (R.z.hi >= 0.0f) { face = 4; sc = R.x; tc = df_neg(R.y); }
    else { face = 5; sc = df_neg(R.x); tc = df_neg(R.y); }
  }
  // S is a power of two (the prefiltered cube's sizes are), so scaling by S / 2
  // is exact on both parts.
  const float half = 0.5f * static_cast<float>(S);
  const df32 u = df_div(sc, ma), v = df_div(tc, ma);
  const df32 x = df_add_f(df32{u.hi * half, u.lo * half}, half - 0.5f);
  const df32 y = df_add_f(df32{v.hi * half, v.lo * half}, half - 0.5f);
  // |sc|, |tc| <= ma, so floor(x), floor(y) lie in [-1, S - 1]: the taps are
  // inside the bordered face (BorderedCube), four plain loads.
  const TexelSplit sx = df_texel(x), sy = df_texel(y);
  const int W = S + 2;
  const float4* r0 = mip + (face * W + sy.i + 1) * W + (sx.i + 1);
  return weighted4(xyz(r0[0]), xyz(r0[1]), xyz(r0[W]), xyz(r0[W + 1]), sx, sy);
}

// Reflection of the view ray v (camera -> surface, any positive scale) about
// the normal n (any positive scale), up to a positive scale:
// reflect(v/|v|, n/|n|) |v| |n|^2 = v (n.n) - 2 (n.v) n.  Cube lookups depend
// only on the direction, so no normalisation is needed.
// The sums are accurate on the scale of |v| |n|^2 (df_add_fast), which is
// what a direction needs.
__device__ __forceinline__ df3 reflect_df(const df3& v, f3 n) {
  const df32 nn =
      df_add_fast(df_add_fast(df_two_prod(n.x, n.x), df_two_prod(n.y, n.y)), df_two_prod(n.z, n.z));
  const df32 nv = df_add_fast(df_add_fast(df_mul_f(v.x, n.x), df_mul_f(v.y, n.y)), df_mul_f(v.z, n.z));
  const df32 two_nv = {2.0f * nv.hi, 2.0f * nv.lo};
  return {df_sub_fast(df_mul(v.x, nn), df_mul_f(two_nv, n.x)), df_sub_fast(df_mul(v.y, nn), df_mul_f(two_nv, n.y)),
          df_sub_fast(df_mul(v.z, nn), df_mul_f(two_nv, n.z))};
}

// The view ray of pixel centre (cx, cy) from df32 affine coefficients:
// ray0 + rayx cx + rayy cy (DeferredPass.vert:10-22 is linear in NDC), each
// component accurate on the scale of the terms (df_add_fast).
struct ViewRayDf {
  float ray0[3][2], rayx[3][2], rayy[3][2];  // (hi, lo) per component
};
__device__ __forceinline__ df3 view_ray_df(const ViewRayDf& r, float cx, float cy) {
  auto comp = [&](int k) {
    return df_add_fast(
        df_add_fast(df32{r.ray0[k][0], r.ray0[k][1]}, df_mul_f(df32{r.rayx[k][0], r.rayx[k][1]}, cx)),
        df_mul_f(df32{r.rayy[k][0], r.rayy[k][1]}, cy));
  };
  return {comp(0), comp(1), comp(2)};
}

// Unit vector of a df32 vector.
__device__ __forceinline__ df3 normalize_df(const df3& v) {
  const df32 n2 = df_add_fast(df_add_fast(df_mul(v.x, v.x), df_mul(v.y, v.y)), df_mul(v.z, v.z));
  const df32 inv = df_div(df_of(1.0f), df_sqrt(n2));
  return {df_mul(v.x, inv), df_mul(v.y, inv), df_mul(v.z, inv)};
}

}  // namespace althea
