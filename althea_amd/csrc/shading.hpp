// pbrMaterial (Shaders/PBR/PBRMaterial.glsl:72-162) as device functions,
// shared by the deferred resolve and the SSR hit shading
// (SSR.frag:55-78 calls the same pbrMaterial).
//
// Conventions (the live call of SSR.frag:65-75): V = unit view ray,
// camera -> surface, so NdotV = max(dot(N, -V), 0) (PBRMaterial.glsl:82);
// H = normalize(V + L) (PBRMaterial.glsl:145); alpha = roughness^2,
// kDirect = (alpha + 1)^2 / 8 (PBRMaterial.glsl:93-96); Schlick-roughness
// Fresnel F0 + (max(1 - r, F0) - F0) (1 - x)^5 (PBRMaterial.glsl:51-53).
#pragma once

#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "precise_lookup.hpp"

namespace althea {

// Smith-Schlick view term gv = NdotV / (NdotV (1 - k) + k) (PBRMaterial.glsl:55-57).
__device__ __forceinline__ float gv_of(float n_dot_v, float k) {
  return n_dot_v * rcp_fast(fmaf(n_dot_v, 1.0f - k, k));
}

__device__ __forceinline__ float pow5(float x) {
  float x2 = x * x;
  return x2 * x2 * x;
}

// Split-sum LUT at (u, v) = (NdotV, roughness): bilinear, CLAMP_TO_EDGE,
// RG per texel, row = roughness.
__device__ __forceinline__ void lut_sample(const float2* __restrict__ lut, int n, float u, float v, float& A,
                                           float& B) {
  const float x = fmaf(u, static_cast<float>(n), -0.5f), y = fmaf(v, static_cast<float>(n), -0.5f);
  const float fx = floorf(x), fy = floorf(y);
  const float ax = x - fx, ay = y - fy;
  const int ix = static_cast<int>(fx), iy = static_cast<int>(fy);
  const int x0 = min(max(ix, 0), n - 1), x1 = min(max(ix + 1, 0), n - 1);
  const int y0 = min(max(iy, 0), n - 1), y1 = min(max(iy + 1, 0), n - 1);
  const float2 t00 = lut[y0 * n + x0], t10 = lut[y0 * n + x1];
  const float2 t01 = lut[y1 * n + x0], t11 = lut[y1 * n + x1];
  const float ta = fmaf(t10.x - t00.x, ax, t00.x), ba = fmaf(t11.x - t01.x, ax, t01.x);
  const float tb = fmaf(t10.y - t00.y, ax, t00.y), bb = fmaf(t11.y - t01.y, ax, t01.y);
  A = fmaf(ba - ta, ay, ta);
  B = fmaf(bb - tb, ay, tb);
}

// Per shading point: the material terms the environment and light parts share.
struct PbrPoint {
  f3 F0, dF, kd;
  float n_dot_v, a2, kdir;
};

__device__ __forceinline__ PbrPoint pbr_point(f3 base, float metallic, float roughness, float NI) {
  PbrPoint m;
  m.n_dot_v = fmaxf(-NI, 0.0f);
  m.F0 = mk(fmaf(metallic, base.x - 0.04f, 0.04f), fmaf(metallic, base.y - 0.04f, 0.04f),
            fmaf(metallic, base.z - 0.04f, 0.04f));
  const float a = roughness * roughness;
  m.a2 = a * a;
  m.kdir = (a + 1.0f) * (a + 1.0f) * 0.125f;
  const float om_r = 1.0f - roughness;
  m.dF = mk(fmaxf(om_r, m.F0.x) - m.F0.x, fmaxf(om_r, m.F0.y) - m.F0.y, fmaxf(om_r, m.F0.z) - m.F0.z);
  m.kd = base * (1.0f - metallic);
  return m;
}

// Environment term of pbrMaterial (PBRMaterial.glsl:101-114):
// (irr (1 - F) kd + refl (F A + B)) ao with F at NdotV.
__device__ __forceinline__ f3 pbr_environment(const PbrPoint& m, f3 irr, f3 refl, float A, float B, float ao) {
  const float fx = pow5(1.0f - m.n_dot_v);
  const f3 F = mk(fmaf(m.dF.x, fx, m.F0.x), fmaf(m.dF.y, fx, m.F0.y), fmaf(m.dF.z, fx, m.F0.z));
  return mk((irr.x * ((1.0f - F.x) * m.kd.x) + refl.x * fmaf(F.x, A, B)) * ao,
            (irr.y * ((1.0f - F.y) * m.kd.y) + refl.y * fmaf(F.y, A, B)) * ao,
            (irr.z * ((1.0f - F.z) * m.kd.z) + refl.z * fmaf(F.z, A, B)) * ao);
}

// Point lights (PBRMaterial.glsl:117-159), added into `color`.  Per light:
//   d = |Lp - P|, L = (Lp - P)/d, NdotL = max(N.L, 0),
//   |V + L|^2 = 2 + 2 V.L (unit V, L), NdotH = max((N.V + N.L)/|V + L|, 0),
//   F = F0 + dF (1 - NdotH)^5, D = a2 / (pi t^2), t = NdotH^2 (a2 - 1) + 1,
//   spec = D F G / (4 NdotL NdotV + 1e-4), G = gv gl, gl = NdotL / (NdotL (1-k) + k),
//   diffuse = (1 - F) kd / pi, color += (diffuse + spec) E NdotL / d^2.
// The three denominators share one reciprocal.  Grey lights (E.x = E.y =
// E.z) accumulate four per-pixel scalars instead of three channels:
//   sum (F_c s + (1 - F_c) kdp_c) e w
//     = F0_c (A2 - kdp_c A1) + dF_c (A4 - kdp_c A3) + kdp_c A1
//   with ew = e w, S = s ew, A1 = sum ew, A2 = sum S, A3 = sum x5 ew, A4 = sum x5 S;
//   the per-pixel factor c1 = a2 gv / pi of s is applied once to A2 and A4.
// V.L enters only as 2 V.L, formed as (2V).L (exact scaling).
// A light below the horizon of every lane of the wave adds exactly zero and
// is skipped with a wave-uniform branch; lanes with N = 0 pass no horizon.
// Lights are wave-uniform: scalar loads into SGPRs, and the grey test is a
// scalar compare of the emission bits.  `any_lane`: some lane of the wave
// shades (wave-uniform).  `shadow` (ShadowCubes or NoShadow): a light it
// hides at a lane adds nothing there (PBRMaterial.glsl:138-140); it is asked
// only where the light is above the horizon.
struct NoShadow {
  static constexpr bool kActive = false;
  __device__ __forceinline__ bool hidden(int, f3, float) const { return false; }
};

// The reference's half vector H = normalize(V + L) (PBRMaterial.glsl:145, Q2)
// is ill-conditioned where L nearly equals -V, a light near the line of
// sight behind the viewer: |V + L|^2 = 2 + 2 V.L cancels, so the FP32 form
// carries an error of about 1e-7 / |V + L|^2 into NdotH.  Where some lane of
// the wave has |V + L|^2 < kHalfVecRefine (about one light per wave in the 4K
// frame), NdotH is recomputed from the explicit vector h = V + L, whose error
// is about 1e-7 / |V + L|; below kHalfVecDf (one light in 60 waves) from V and
// L in df32, V from the caller's df32 view ray (`view_df`, any positive scale).
constexpr float kHalfVecRefine = 1e-2f;
constexpr float kHalfVecDf = 1e-4f;

// The view direction as given (FP32), for callers without a df32 view ray.
struct ViewOfF3 {
  f3 v;
  __device__ __forceinline__ df3 operator()() const { return {df_of(v.x), df_of(v.y), df_of(v.z)}; }
};

template <class ViewDf>
__device__ __forceinline__ float n_dot_h_refined(f3 N, f3 I, f3 P, float4 lp, f3 L, float inv_d, float h2,
                                              const ViewDf& view_df) {
  if (h2 >= kHalfVecDf) {
    const f3 h = mk(fmaf(L.x, inv_d, I.x), fmaf(L.y, inv_d, I.y), fmaf(L.z, inv_d, I.z));
    return fmaxf(dot(N, h) * rsq_fast(dot(h, h)), 0.0f);
  }
  const df3 Lu = normalize_df({df_two_sum(lp.x, -P.x), df_two_sum(lp.y, -P.y), df_two_sum(lp.z, -P.z)});
  const df3 Vu = normalize_df(view_df());
  const df3 h = {df_add(Vu.x, Lu.x), df_add(Vu.y, Lu.y), df_add(Vu.z, Lu.z)};
  const df32 hh = df_add(df_add(df_mul(h.x, h.x), df_mul(h.y, h.y)), df_mul(h.z, h.z));
  const df32 nh = df_add(df_add(df_mul_f(h.x, N.x), df_mul_f(h.y, N.y)), df_mul_f(h.z, N.z));
  return fmaxf(df_div(nh, df_sqrt(hh)).hi, 0.0f);
}

template <class Shadow = NoShadow, class ViewDf = ViewOfF3>
__device__ __forceinline__ void pbr_lights(const PbrPoint& m, f3 P, f3 N, f3 I, float NI, bool any_lane,
                                           const float4* __restrict__ lights, int light_count, f3& color,
                                           const Shadow& shadow = Shadow(), const ViewDf& view_df = ViewDf{}) {
  const float c1 = m.a2 * gv_of(m.n_dot_v, m.kdir) * kInvPi;
  const float four_nv = 4.0f * m.n_dot_v;
  const float om_k = 1.0f - m.kdir;
  const float a2m1 = m.a2 - 1.0f;
  // gl * (4 NdotV NdotL + 1e-4) = (NdotL (1 - k) + k)(4 NdotV NdotL + 1e-4),
  // a quadratic in NdotL: (q2 NdotL + q1) NdotL + q0.
  const float q2 = om_k * four_nv, q1 = fmaf(om_k, 0.0001f, m.kdir * four_nv), q0 = m.kdir * 0.0001f;
  const f3 kdp = m.kd * kInvPi;
  const f3 I2 = I * 2.0f;
  float A1 = 0.0f, A2 = 0.0f, A3 = 0.0f, A4 = 0.0f;
  if (any_lane) {
    for (int i = 0; i < light_count; ++i) {
      const float4 lp = lights[2 * i];
      const float4 em = lights[2 * i + 1];
      const f3 L = mk(lp.x - P.x, lp.y - P.y, lp.z - P.z);
      const float nl_raw = dot(N, L);
      const uint64_t lit = __builtin_amdgcn_ballot_w64(nl_raw > 0.0f);
      if (lit == 0) continue;
      const float d2 = dot(L, L);
      const float inv_d = rsq_fast(d2);
      const float h2 = fmaf(dot(I2, L), inv_d, 2.0f);
      // max(., 0) of the reference, saturated: the cosines of unit vectors
      // exceed 1 only by rounding, and the clamp costs no instruction.
      const float n_dot_l = saturate(nl_raw * inv_d);
      float n_dot_h = saturate(fmaf(nl_raw, inv_d, NI) * rsq_fast(h2));
      // Lanes above the horizon with |V + L|^2 below the refinement bound.
      if ((__builtin_amdgcn_ballot_w64(h2 < kHalfVecRefine) & lit) != 0) {
        if (h2 < kHalfVecRefine && nl_raw > 0.0f) n_dot_h = n_dot_h_refined(N, I, P, lp, L, inv_d, h2, view_df);
      }
      const float x = 1.0f - n_dot_h;
      const float x5 = pow5(x);
      const float t = fmaf(n_dot_h * n_dot_h, a2m1, 1.0f);
      const float den = (t * t) * fmaf(fmaf(q2, n_dot_l, q1), n_dot_l, q0);
      const float s1 = n_dot_l * rcp_fast(den);  // s / c1
      float w = (inv_d * inv_d) * n_dot_l;
      if (Shadow::kActive && n_dot_l > 0.0f && shadow.hidden(i, L, d2 * inv_d)) w = 0.0f;
      const bool grey =
          __float_as_uint(em.x) == __float_as_uint(em.y) && __float_as_uint(em.y) == __float_as_uint(em.z);
      if (grey) {
        const float ew = em.x * w;
        const float S = s1 * ew;
        A1 += ew;
        A2 += S;
        A3 = fmaf(x5, ew, A3);
        A4 = fmaf(x5, S, A4);
      } else {
        const float s = c1 * s1;
        const f3 F = mk(fmaf(m.dF.x, x5, m.F0.x), fmaf(m.dF.y, x5, m.F0.y), fmaf(m.dF.z, x5, m.F0.z));
        color.x = fmaf(fmaf(F.x, s, fmaf(-F.x, kdp.x, kdp.x)) * em.x, w, color.x);
        color.y = fmaf(fmaf(F.y, s, fmaf(-F.y, kdp.y, kdp.y)) * em.y, w, color.y);
        color.z = fmaf(fmaf(F.z, s, fmaf(-F.z, kdp.z, kdp.z)) * em.z, w, color.z);
      }
    }
  }
  A2 *= c1;
  A4 *= c1;
  color.x += fmaf(m.F0.x, fmaf(-kdp.x, A1, A2), fmaf(m.dF.x, fmaf(-kdp.x, A3, A4), kdp.x * A1));
  color.y += fmaf(m.F0.y, fmaf(-kdp.y, A1, A2), fmaf(m.dF.y, fmaf(-kdp.y, A3, A4), kdp.y * A1));
  color.z += fmaf(m.F0.z, fmaf(-kdp.z, A1, A2), fmaf(m.dF.z, fmaf(-kdp.z, A3, A4), kdp.z * A1));
}

}  // namespace althea
