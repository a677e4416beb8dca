// Reflection buffer (Src/ReflectionBuffer.cpp:20-94) as the kernels see it:
// frame-sized levels stored back to back, RGBA per texel in FP16 or FP32.
//
// Filtered reads follow the Vulkan LINEAR / CLAMP_TO_EDGE rules: texel-space
// coordinate u * extent - 0.5, floor, linear weights, clamped indices.  Where
// the coordinate is a rational of integers (texel centres of another level,
// the convolution's texel corners) it is split into an integer quotient and a
// fraction so that FP32 rounding never moves a tap across a texel: only the
// small fractional part carries rounding.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "althea/althea.h"
#include "device_math.hpp"

namespace althea {

struct ReflChain {
  void* data;
  int f16;
  int mips;
  int w[ALTHEA_REFLECTION_MAX_MIPS], h[ALTHEA_REFLECTION_MAX_MIPS];
  uint32_t off[ALTHEA_REFLECTION_MAX_MIPS];  // texel offset of each level
};

__device__ __forceinline__ float4 refl_load(const ReflChain& c, uint32_t i) {
  if (c.f16) {
    const ushort4 h = static_cast<const ushort4*>(c.data)[i];
    return make_float4(__half2float(__ushort_as_half(h.x)), __half2float(__ushort_as_half(h.y)),
                       __half2float(__ushort_as_half(h.z)), __half2float(__ushort_as_half(h.w)));
  }
  return static_cast<const float4*>(c.data)[i];
}

__device__ __forceinline__ void refl_store(const ReflChain& c, uint32_t i, float4 v) {
  if (c.f16) {
    ushort4 h;
    h.x = __half_as_ushort(__float2half_rn(v.x));
    h.y = __half_as_ushort(__float2half_rn(v.y));
    h.z = __half_as_ushort(__float2half_rn(v.z));
    h.w = __half_as_ushort(__float2half_rn(v.w));
    static_cast<ushort4*>(c.data)[i] = h;
  } else {
    static_cast<float4*>(c.data)[i] = v;
  }
}

// num / den = q + frac exactly in the integers, frac in [0, 1) rounded once.
__device__ __forceinline__ void rational_split(uint32_t num, uint32_t den, int& q, float& frac) {
  const uint32_t qq = num / den;
  q = static_cast<int>(qq);
  frac = static_cast<float>(num - qq * den) / static_cast<float>(den);
}

// Left tap index and weight of the coordinate q + c (c small).
__device__ __forceinline__ void tap_split(int q, float c, int& i0, float& w) {
  const float fl = floorf(c);
  i0 = q + static_cast<int>(fl);
  w = c - fl;
}

// Bilinear CLAMP_TO_EDGE read of level l at texel-space coordinate
// (qx + cx, qy + cy), where the texel-space coordinate already includes the
// -0.5 of the filter.
__device__ __forceinline__ float4 refl_bilinear(const ReflChain& c, int l, int qx, float cx, int qy, float cy) {
  const int W = c.w[l], H = c.h[l];
  int ix, iy;
  float ax, ay;
  tap_split(qx, cx, ix, ax);
  tap_split(qy, cy, iy, ay);
  const int x0 = min(max(ix, 0), W - 1), x1 = min(max(ix + 1, 0), W - 1);
  const int y0 = min(max(iy, 0), H - 1), y1 = min(max(iy + 1, 0), H - 1);
  const uint32_t base = c.off[l];
  const float4 t00 = refl_load(c, base + static_cast<uint32_t>(y0 * W + x0));
  const float4 t10 = refl_load(c, base + static_cast<uint32_t>(y0 * W + x1));
  const float4 t01 = refl_load(c, base + static_cast<uint32_t>(y1 * W + x0));
  const float4 t11 = refl_load(c, base + static_cast<uint32_t>(y1 * W + x1));
  const float4 top = lerp4(t00, t10, ax), bot = lerp4(t01, t11, ax);
  return lerp4(top, bot, ay);
}

// The buffer at the centre of frame pixel (px, py), LOD `lod` with LINEAR
// mipmapping clamped to [0, mips - 1] (DeferredPass.frag:29-31,
// textureLod(reflectionBuffer, uv, 4 roughness) on a 5-level buffer).
// uv = (p + 0.5) / frame is (2p + 1) w_l / (2 W) - 0.5 texels on level l.
__device__ __forceinline__ float4 refl_trilinear_pixel(const ReflChain& c, int px, int py, float lod) {
  const float d = fminf(fmaxf(lod, 0.0f), static_cast<float>(c.mips - 1));
  const float fl = floorf(d);
  const int l0 = static_cast<int>(fl), l1 = min(l0 + 1, c.mips - 1);
  const float t = d - fl;
  auto level = [&](int l) {
    int qx, qy;
    float fx, fy;
    rational_split(static_cast<uint32_t>(2 * px + 1) * static_cast<uint32_t>(c.w[l]),
                   2u * static_cast<uint32_t>(c.w[0]), qx, fx);
    rational_split(static_cast<uint32_t>(2 * py + 1) * static_cast<uint32_t>(c.h[l]),
                   2u * static_cast<uint32_t>(c.h[0]), qy, fy);
    return refl_bilinear(c, l, qx, fx - 0.5f, qy, fy - 0.5f);
  };
  const float4 a = level(l0);
  if (t == 0.0f) return a;
  return lerp4(a, level(l1), t);
}

}  // namespace althea
