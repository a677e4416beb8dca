// Deferred resolve for gfx950: one lane per pixel, a 256-lane block covers a
// 64 x 4 pixel tile so each wave reads one 64-pixel run of every G-buffer
// plane (1 KiB contiguous per plane in FP32).
//
// Per pixel (DeferredPass.vert:10-22, DeferredPass.frag:41-94):
//   view ray, affine in the pixel centre and precomputed on the host from
//     mat3(inverseView) * (inverseProjection * (ndc, 0, 1)).xyz, in FP32 and
//     (for the mirror-like lookups) with df32 coefficients;
//   background equirect lookup (LOD 0, CLAMP) when position.w == 0 (Q13),
//     addressed in df32;
//   reflection = seamless trilinear prefiltered-cube lookup at
//     LOD roughness*(mips-1) (PBRMaterial.glsl:4-11 with cube maps); where the
//     LOD is below 1 the mip-0 tap (a mirror of the source's texel contrast)
//     is addressed in df32 from the df32 reflection of the view ray;
//   irradiance = seamless bilinear irradiance-cube lookup (PBRMaterial.glsl:13-19);
//   SSAO hemisphere kernel into the occlusion channel (Q9);
//   pbrMaterial (PBRMaterial.glsl:72-162, live convention of SSR.frag:65-75);
//   tonemap 1 - exp(-c * exposure) (DeferredPass.frag:86-88).
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>

#include "device_math.hpp"
#include "internal.hpp"
#include "mlp_eval.hpp"
#include "precise_lookup.hpp"
#include "reflection.hpp"
#include "shading.hpp"
#include "shadow.hpp"

namespace althea {

struct ResolveParams {
  const void* position;
  const void* normal;
  const void* albedo;
  const void* mro;
  int width, height;
  const float4* env;  // equirect level 0
  int env_w, env_h;
  const float4* prefiltered;  // seamless-bordered copy (BorderedCube)
  BorderedCube pref;
  const float4* irradiance;  // seamless-bordered copy, one mip
  int irr_size;
  const float2* lut;
  int lut_size;
  int light_count;
  float ray0[3], rayx[3], rayy[3];  // view ray = ray0 + rayx*(px+0.5) + rayy*(py+0.5)
  ViewRayDf ray_df;                 // the same coefficients as df32 pairs
  float view_z[4];                  // row 2 of the view matrix
  float vp[16];                     // projection * view, column-major
  float exposure;
  int tonemap;
  int ssao_on;
  int ssao_w_depth;  // clip w is the view depth bit for bit (w row of vp = -view_z)
  int ssao_samples;
  float ssao_inv_samples;
  float ssao_radius, ssao_bias;
  float kernel[ALTHEA_SSAO_MAX_SAMPLES][4];
  float noise[16][4];
  const float* ssao_depth;  // per pixel view_depth(position), +inf where uncovered
  void* out;
  float* ao_out;
  MlpInferenceView mlp;  // MLP variant only
  uint8_t* mlp_used;
  ReflChain refl;      // reflection-buffer variant only
  ShadowCubes shadow;  // shadowed variant only
};

template <int kFmt>
__device__ __forceinline__ float4 load_gbuf(const void* plane, size_t idx) {
  if (kFmt == ALTHEA_GBUFFER_FP32) {
    return static_cast<const float4*>(plane)[idx];
  } else {
    const ushort4 h = static_cast<const ushort4*>(plane)[idx];
    return make_float4(__half2float(__ushort_as_half(h.x)), __half2float(__ushort_as_half(h.y)),
                       __half2float(__ushort_as_half(h.z)), __half2float(__ushort_as_half(h.w)));
  }
}

template <bool kOutF16>
__device__ __forceinline__ void store_out(void* out, size_t idx, f3 c) {
  if (kOutF16) {
    ushort4 h;
    h.x = __half_as_ushort(__float2half_rn(c.x));
    h.y = __half_as_ushort(__float2half_rn(c.y));
    h.z = __half_as_ushort(__float2half_rn(c.z));
    h.w = __half_as_ushort(__float2half_rn(1.0f));
    static_cast<ushort4*>(out)[idx] = h;
  } else {
    static_cast<float4*>(out)[idx] = make_float4(c.x, c.y, c.z, 1.0f);
  }
}

__device__ __forceinline__ f3 tonemap(f3 c, float exposure) {
  return mk(1.0f - expf(-c.x * exposure), 1.0f - expf(-c.y * exposure), 1.0f - expf(-c.z * exposure));
}

__device__ __forceinline__ float smoothstep01(float x) {
  const float t = saturate(x);
  return t * t * fmaf(-2.0f, t, 3.0f);
}

__device__ __forceinline__ float view_depth(const ResolveParams& p, f3 q) {
  return -fmaf(p.view_z[0], q.x, fmaf(p.view_z[1], q.y, fmaf(p.view_z[2], q.z, p.view_z[3])));
}

constexpr int kSsaoBatch = 4;  // position gathers issued together

// kWDepth: the projection's w row is (0, 0, -1, 0), so the w row of
// projection * view is -view_z entry for entry (the host checks the bits)
// and the clip w of every tap, formed by the same fused operations with
// every operand negated, equals its view depth: the depth stands in for w.
template <int kFmt, bool kWDepth>
__device__ float ssao(const ResolveParams& p, const float4* s_noise, int px, int py, f3 P, f3 N) {
  const float4 nz = s_noise[(px & 3) + 4 * (py & 3)];
  const f3 r = mk(nz.x, nz.y, nz.z);
  const f3 T = normalize_fast(r - N * dot(r, N));
  const f3 B = cross(N, T);
  const float frag_depth = view_depth(p, P);
  const float fw = static_cast<float>(p.width), fh = static_cast<float>(p.height);
  // Clip coordinates of S = P + radius * TBN k are affine in k: precompute
  // the clip and depth of P and of the three scaled basis vectors.
  const f3 Tr = T * p.ssao_radius, Br = B * p.ssao_radius, Nr = N * p.ssao_radius;
  auto clip3 = [&](f3 v, float w1, float& cx, float& cy, float& cw) {
    cx = fmaf(p.vp[0], v.x, fmaf(p.vp[4], v.y, fmaf(p.vp[8], v.z, p.vp[12] * w1)));
    cy = fmaf(p.vp[1], v.x, fmaf(p.vp[5], v.y, fmaf(p.vp[9], v.z, p.vp[13] * w1)));
    cw = fmaf(p.vp[3], v.x, fmaf(p.vp[7], v.y, fmaf(p.vp[11], v.z, p.vp[15] * w1)));
  };
  float px0, py0, pw0, tx, ty, tw, bx, by, bw, nx, ny, nw;
  clip3(P, 1.0f, px0, py0, pw0);
  clip3(Tr, 0.0f, tx, ty, tw);
  clip3(Br, 0.0f, bx, by, bw);
  clip3(Nr, 0.0f, nx, ny, nw);
  // Screen position in pixels, (0.5 x / w + 0.5) * W: fold W / 2 into the
  // clip x and y coefficients so each tap needs one FMA per axis.
  const float hw = 0.5f * fw, hh = 0.5f * fh;
  px0 *= hw, tx *= hw, bx *= hw, nx *= hw;
  py0 *= hh, ty *= hh, by *= hh, ny *= hh;
  auto vz = [&](f3 v) { return -fmaf(p.view_z[0], v.x, fmaf(p.view_z[1], v.y, p.view_z[2] * v.z)); };
  const float dT = vz(Tr), dB = vz(Br), dN = vz(Nr);
  // Taps in groups of kSsaoBatch: their depth loads are issued together
  // (an invalid tap loads texel 0 and is ignored), then evaluated in tap
  // order, so the per-tap arithmetic and the sum order are those of a
  // one-tap-at-a-time loop.  The depths come from the frame's view-depth
  // plane (ssao_depth_kernel): 4 bytes per tap instead of the 16-byte
  // position, the same value view_depth gives the position, and +inf where
  // the G-buffer is uncovered, which fails the occlusion compare exactly as
  // the uncovered-tap skip does.
  float occ = 0.0f;
  const int n = p.ssao_samples;
  for (int k0 = 0; k0 < n; k0 += kSsaoBatch) {
    uint32_t idx[kSsaoBatch];
    bool ok[kSsaoBatch];
    float sd[kSsaoBatch];
#pragma unroll
    for (int j = 0; j < kSsaoBatch; ++j) {
      const int k = min(k0 + j, n - 1);
      const float kx = p.kernel[k][0], ky = p.kernel[k][1], kz = p.kernel[k][2];
      sd[j] = fmaf(dT, kx, fmaf(dB, ky, fmaf(dN, kz, frag_depth)));
      const float cx = fmaf(tx, kx, fmaf(bx, ky, fmaf(nx, kz, px0)));
      const float cy = fmaf(ty, kx, fmaf(by, ky, fmaf(ny, kz, py0)));
      const float cw = kWDepth ? sd[j] : fmaf(tw, kx, fmaf(bw, ky, fmaf(nw, kz, pw0)));
      const float inv_w = rcp_fast(cw);
      const float sx = fmaf(cx, inv_w, hw), sy = fmaf(cy, inv_w, hh);  // pixels
      // floor(s) lies in [0, extent) exactly where s does, and equals the
      // truncation there: one unsigned compare per axis is the bounds test.
      const uint32_t qx = static_cast<uint32_t>(cvt_floor(sx)), qy = static_cast<uint32_t>(cvt_floor(sy));
      ok[j] = (k0 + j < n) & (cw > 0.0f) & (qx < static_cast<uint32_t>(p.width)) &
              (qy < static_cast<uint32_t>(p.height));
      const uint32_t lin = qy * static_cast<uint32_t>(p.width) + qx;
      idx[j] = ok[j] ? lin : 0u;
    }
    float qd[kSsaoBatch];
#pragma unroll
    for (int j = 0; j < kSsaoBatch; ++j) qd[j] = p.ssao_depth[idx[j]];
    // The first load is consumed here, so it is issued with the others
    // rather than moved under that tap's validity test.
    asm volatile("" : "+v"(qd[0]));
#pragma unroll
    for (int j = 0; j < kSsaoBatch; ++j) {
      if (ok[j] && qd[j] <= sd[j] - p.ssao_bias)
        occ += smoothstep01(p.ssao_radius * rcp_fast(fabsf(frag_depth - qd[j])));
    }
  }
  return fmaf(-occ, p.ssao_inv_samples, 1.0f);
}


// One 64 x 4 tile, row-major tile id.  Tiles are taken in launch order: a
// mapping that gives each XCD (with its own L2) contiguous runs of tiles was
// measured slower at 4K (runs of 2, 8 or 68 tile rows: 0.87, 0.89, 0.84 ms
// against 0.83 ms).
//
// kMlp (the MLP-IBL resolve): a pixel whose reflection cell the MLP's error
// monitor accepts at both mips the cube lookup would blend skips the cube
// lookup and is not stored here; it goes to the handle's queue as three
// float4 (R, roughness | (F A + B) ao per channel, pixel index | the colour
// without the prefiltered term), which resolve_mlp_finish_kernel completes
// with the MLP's radiance.  Every other pixel is shaded exactly as by the
// plain resolve.  Wave w of the grid (block * 4 + wave) owns queue records
// [64 w, 64 w + 64): its deferred lanes take consecutive records in lane
// order and lane 0 writes the wave's count, zero included, so the queue needs
// neither atomics nor clearing.
template <int kFmt, bool kOutF16, bool kMlp, bool kRefl = false, bool kShadow = false, bool kWDepth = false>
__device__ __forceinline__ void resolve_tile(const ResolveParams& p, int tile, int tiles_x,
                                             const float4* __restrict__ lights, const float4* s_noise) {
  const int px = (tile % tiles_x) * 64 + (threadIdx.x & 63);
  const int py = (tile / tiles_x) * 4 + (threadIdx.x >> 6);
  const bool inside = px < p.width && py < p.height;
  const size_t idx = static_cast<size_t>(min(py, p.height - 1)) * p.width + min(px, p.width - 1);
  const float cxp = static_cast<float>(px) + 0.5f, cyp = static_cast<float>(py) + 0.5f;
  const f3 dir = mk(fmaf(p.rayy[0], cyp, fmaf(p.rayx[0], cxp, p.ray0[0])),
                    fmaf(p.rayy[1], cyp, fmaf(p.rayx[1], cxp, p.ray0[1])),
                    fmaf(p.rayy[2], cyp, fmaf(p.rayx[2], cxp, p.ray0[2])));
  const float4 pos = load_gbuf<kFmt>(p.position, idx);
  const bool background = pos.w == 0.0f;
  if (background) {
    // Background (DeferredPass.frag:44-53): env LOD 0, CLAMP_TO_EDGE, the atan2
    // form without a pole branch; a mirror of the source, addressed in df32.
    df32 tx, ty;
    equirect_coords_df(view_ray_df(p.ray_df, cxp, cyp), p.env_w, p.env_h, 0, tx, ty);
    f3 c = clamp_bilinear_df(p.env, p.env_w, p.env_h, tx, ty);
    if (p.tonemap) c = tonemap(c, p.exposure);
    if (inside) {
      if (p.ao_out) p.ao_out[idx] = 1.0f;
      store_out<kOutF16>(p.out, idx, c);
    }
  }
  // Every lane stays for the block barriers; background lanes and lanes
  // outside the image shade nothing, store nothing and skip the other
  // G-buffer planes.
  const bool shade = inside && !background;
  const float4 zero4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float4 nrm = shade ? load_gbuf<kFmt>(p.normal, idx) : zero4;
  const float4 mro = shade ? load_gbuf<kFmt>(p.mro, idx) : zero4;
  const f3 P = mk(pos.x, pos.y, pos.z);
  // Lanes that shade nothing get N = 0, so no light passes their horizon test.
  const f3 N = shade ? normalize_fast(mk(nrm.x, nrm.y, nrm.z)) : mk(0.0f, 0.0f, 0.0f);
  const float metallic = mro.x, roughness = shade ? mro.y : 1.0f;
  float ao = mro.z;
  const f3 I = normalize_fast(dir);
  const float NI = dot(N, I);
  const float nd = 2.0f * NI;
  const f3 R = mk(fmaf(-nd, N.x, I.x), fmaf(-nd, N.y, I.y), fmaf(-nd, N.z, I.z));
  const float lod = roughness * static_cast<float>(p.pref.mips - 1);
  f3 refl = mk(0.0f, 0.0f, 0.0f);
  bool from_mlp = false;
  uint32_t mlp_slot = 0;
  if (kMlp) {
    // Monitor cell of R at the two mips the cube lookup would blend.
    int rface;
    float rpc, rpt;
    cube_select(R, rface, rpc, rpt);
    const float d = fminf(fmaxf(lod, 0.0f), static_cast<float>(p.pref.mips - 1));
    const float fl = floorf(d);
    const int m0 = static_cast<int>(fl), m1 = min(m0 + 1, p.pref.mips - 1);
    const int cx = min(max(static_cast<int>((rpc + 1.0f) * 4.0f), 0), 7);
    const int cy = min(max(static_cast<int>((rpt + 1.0f) * 4.0f), 0), 7);
    const int cell = rface * ALTHEA_MLP_MONITOR_CELLS + cy * 8 + cx;
    const int per_mip = 6 * ALTHEA_MLP_MONITOR_CELLS;
    from_mlp = shade && p.mlp.cell_ok[m0 * per_mip + cell] != 0 &&
               (d == fl || p.mlp.cell_ok[m1 * per_mip + cell] != 0);
    if (inside && p.mlp_used) p.mlp_used[idx] = from_mlp ? 1 : 0;
    // The query goes out now, the colour after the lights.
    const uint64_t mask = __builtin_amdgcn_ballot_w64(from_mlp);
    const int lane = threadIdx.x & 63;
    const uint32_t region = static_cast<uint32_t>(tile) * 4u + (threadIdx.x >> 6);
    if (lane == 0) p.mlp.region_count[region] = static_cast<uint32_t>(__popcll(mask));
    mlp_slot = region * kMlpQueueRegion + static_cast<uint32_t>(__popcll(mask & ((1ull << lane) - 1ull)));
    if (from_mlp) p.mlp.queue[3 * static_cast<size_t>(mlp_slot)] = make_float4(R.x, R.y, R.z, roughness);
  }
  const float4 alb = shade ? load_gbuf<kFmt>(p.albedo, idx) : zero4;
  const PbrPoint m = pbr_point(mk(alb.x, alb.y, alb.z), metallic, roughness, NI);
  f3 color = mk(0.0f, 0.0f, 0.0f);
  if (shade) {
    if (!from_mlp) {
      const float d = fminf(fmaxf(lod, 0.0f), static_cast<float>(p.pref.mips - 1));
      if (d < 1.0f) {
        // The mip-0 tap mirrors the source's texel contrast: reflection and
        // face coordinates in df32, the blend with mip 1 in weight form.
        const df3 Rd = reflect_df(view_ray_df(p.ray_df, cxp, cyp), mk(nrm.x, nrm.y, nrm.z));
        refl = cube_bilinear_df(p.prefiltered + p.pref.off[0], p.pref.size, Rd);
        if (d > 0.0f) {
          int face;
          float pc, pt;
          cube_select(R, face, pc, pt);
          const f3 b = bcube_bilinear(p.prefiltered + p.pref.off[1], max(1, p.pref.size >> 1), face, pc, pt);
          const float w0 = 1.0f - d;
          refl = mk(fmaf(b.x, d, refl.x * w0), fmaf(b.y, d, refl.y * w0), fmaf(b.z, d, refl.z * w0));
        }
      } else {
        refl = bcube_sample(p.prefiltered, p.pref, R, lod);
      }
    }
    if (kRefl) {
      // Reflection buffer at LOD roughness * (mips - 1); where its alpha a
      // reaches 0.01, mix(prefiltered, rgb / a, a) (DeferredPass.frag:59-68).
      const float4 rb = refl_trilinear_pixel(p.refl, px, py, roughness * static_cast<float>(p.refl.mips - 1));
      if (rb.w >= 0.01f) {
        const float ia = 1.0f / rb.w;
        refl = mk(fmaf(fmaf(rb.x, ia, -refl.x), rb.w, refl.x), fmaf(fmaf(rb.y, ia, -refl.y), rb.w, refl.y),
                  fmaf(fmaf(rb.z, ia, -refl.z), rb.w, refl.z));
      }
    }
    int face;
    float pc, pt;
    cube_select(N, face, pc, pt);
    const f3 irr = bcube_bilinear(p.irradiance, p.irr_size, face, pc, pt);
    if (p.ssao_on) ao = ssao<kFmt, kWDepth>(p, s_noise, px, py, P, N);
    if (p.ao_out) p.ao_out[idx] = ao;
    float A, B;
    lut_sample(p.lut, p.lut_size, m.n_dot_v, roughness, A, B);
    color = pbr_environment(m, irr, refl, A, B, ao);
    if (kMlp && from_mlp) {
      // The prefiltered term's factor (F A + B) ao per channel, F at NdotV
      // (pbr_environment), and the pixel index.
      const float fx = pow5(1.0f - m.n_dot_v);
      p.mlp.queue[3 * static_cast<size_t>(mlp_slot) + 1] =
          make_float4(fmaf(fmaf(m.dF.x, fx, m.F0.x), A, B) * ao, fmaf(fmaf(m.dF.y, fx, m.F0.y), A, B) * ao,
                      fmaf(fmaf(m.dF.z, fx, m.F0.z), A, B) * ao, __uint_as_float(static_cast<uint32_t>(idx)));
    }
  }
  const auto view_df = [&]() { return view_ray_df(p.ray_df, cxp, cyp); };
  if (kShadow)
    pbr_lights(m, P, N, I, NI, __any(shade), lights, p.light_count, color, p.shadow, view_df);
  else
    pbr_lights(m, P, N, I, NI, __any(shade), lights, p.light_count, color, NoShadow(), view_df);
  if (kMlp && from_mlp)  // the colour without the prefiltered term
    p.mlp.queue[3 * static_cast<size_t>(mlp_slot) + 2] = make_float4(color.x, color.y, color.z, 0.0f);
  if (shade && !(kMlp && from_mlp)) {
    if (p.tonemap) color = tonemap(color, p.exposure);
    store_out<kOutF16>(p.out, idx, color);
  }
}

// The frame's SSAO view-depth plane: view_depth(position) per pixel, +inf
// where the G-buffer is uncovered (position.w == 0); the expression and the
// FP16 conversion are the ones the resolve applies to a tap's position.
template <int kFmt>
__global__ __launch_bounds__(256) void ssao_depth_kernel(const ResolveParams p, float* __restrict__ plane) {
  const size_t n = static_cast<size_t>(p.width) * p.height;
  const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const float4 q = load_gbuf<kFmt>(p.position, i);
  plane[i] = q.w == 0.0f ? __builtin_inff() : view_depth(p, mk(q.x, q.y, q.z));
}

// The resolve: one block per tile, the SSAO noise tile in LDS; kRefl blends
// the reflection buffer, kShadow applies the point-light shadow test, kMlp
// defers the pixels the MLP serves; kWDepth (the plain resolve under a
// perspective projection) takes each SSAO tap's view depth as its clip w.
template <int kFmt, bool kOutF16, bool kRefl = false, bool kShadow = false, bool kMlp = false, bool kWDepth = false>
__global__ __launch_bounds__(256) void resolve_kernel(const ResolveParams p, const float4* __restrict__ lights) {
  __shared__ float4 s_noise[16];
  if (threadIdx.x < 16)
    s_noise[threadIdx.x] =
        make_float4(p.noise[threadIdx.x][0], p.noise[threadIdx.x][1], p.noise[threadIdx.x][2], 0.0f);
  __syncthreads();
  const int tiles_x = (p.width + 63) / 64;
  resolve_tile<kFmt, kOutF16, kMlp, kRefl, kShadow, kWDepth>(p, blockIdx.x, tiles_x, lights, s_noise);
}

// The MLP-IBL resolve held to the plain resolve's waves per SIMD (six, or
// seven where SSAO takes the depth for w: 80 or 72 registers per lane), so
// that a frame whose MLP serves few pixels costs what the cube-IBL frame
// costs.
template <int kFmt, bool kOutF16, bool kWDepth>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kWDepth ? 7 : 6))) void resolve_mlp_kernel(
    const ResolveParams p, const float4* __restrict__ lights) {
  __shared__ float4 s_noise[16];
  if (threadIdx.x < 16)
    s_noise[threadIdx.x] =
        make_float4(p.noise[threadIdx.x][0], p.noise[threadIdx.x][1], p.noise[threadIdx.x][2], 0.0f);
  __syncthreads();
  const int tiles_x = (p.width + 63) / 64;
  resolve_tile<kFmt, kOutF16, true, false, false, kWDepth>(p, blockIdx.x, tiles_x, lights, s_noise);
}

// Completes the pixels the MLP-IBL resolve deferred.  Persistent waves take
// queue regions in turn (wave-uniform counts); a wave loads the MLP's forward
// operands into registers at its first non-empty region and evaluates each
// region in 16-record tiles: lane (g, n) evaluates feature group g of record
// n, and lane n of group 0 forms the pixel colour + exp(y) (F A + B) ao,
// tonemapped as the resolve would, and stores it.  The MLP arithmetic is the
// streaming inference kernel's (althea_mlp_infer).  No barriers, no atomics.
template <bool kOutF16>
__global__ __launch_bounds__(256) void resolve_mlp_finish_kernel(const ResolveParams p, uint32_t regions) {
  const int lane = threadIdx.x & 63, g = lane >> 4, n = lane & 15;
  const uint32_t waves = gridDim.x * 4u;
  auto count_of = [&](uint32_t region) {
    return min(__builtin_amdgcn_readfirstlane(p.mlp.region_count[region]), static_cast<uint32_t>(kMlpQueueRegion));
  };
  uint32_t region = blockIdx.x * 4u + static_cast<uint32_t>(threadIdx.x >> 6);
  while (region < regions && count_of(region) == 0) region += waves;
  if (region >= regions) return;
  MlpRegWeights r;
  mlp_load_reg_weights(p.mlp.w, g, n, r);
  for (; region < regions; region += waves) {
    const uint32_t count = count_of(region);
    for (uint32_t t = 0; t < count; t += 16u) {
      const uint32_t k = t + static_cast<uint32_t>(n);
      const bool valid = k < count;
      const float4* q = p.mlp.queue + 3 * (static_cast<size_t>(region) * kMlpQueueRegion + (valid ? k : 0u));
      const float4 q0 = valid ? q[0] : make_float4(0.0f, 0.0f, 1.0f, 0.0f);
      const f32x4 y = mlp_tile_eval(r, p.mlp.w, g, q0.x, q0.y, q0.z, q0.w);
      if (g == 0 && valid) {
        const float4 q1 = q[1], q2 = q[2];
        f3 c = mk(fmaf(expf(y[0]), q1.x, q2.x), fmaf(expf(y[1]), q1.y, q2.y), fmaf(expf(y[2]), q1.z, q2.z));
        if (p.tonemap) c = tonemap(c, p.exposure);
        store_out<kOutF16>(p.out, __float_as_uint(q1.w), c);
      }
    }
  }
}

// Blocks of the MLP finishing kernel per CU: about 160 registers per lane
// (88 of them the weights) allow 3 waves per SIMD.
constexpr int kMlpFinishBlocksPerCu = 3;

__global__ void fill_u8_kernel(uint8_t* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Host: the view ray of DeferredPass.vert is linear in NDC, so it is affine in
// the pixel centre; fold mat3(inverseView) * inverseProjection into it.
void view_ray_affine(const althea_globals& g, int W, int H, float ray0[3], float rayx[3], float rayy[3],
                     ViewRayDf* df) {
  const float* ip = g.inverse_projection;
  const float* iv = g.inverse_view;
  auto e_of = [&](double nx, double ny, double out[3]) {
    double e[3];
    for (int r = 0; r < 3; ++r) e[r] = ip[0 * 4 + r] * nx + ip[1 * 4 + r] * ny + ip[3 * 4 + r];
    for (int r = 0; r < 3; ++r) out[r] = iv[0 * 4 + r] * e[0] + iv[1 * 4 + r] * e[1] + iv[2 * 4 + r] * e[2];
  };
  double o[3], ax[3], ay[3];
  e_of(-1.0, -1.0, o);
  e_of(1.0, -1.0, ax);
  e_of(-1.0, 1.0, ay);
  auto split = [](double v, float out[2]) {
    out[0] = static_cast<float>(v);
    out[1] = static_cast<float>(v - static_cast<double>(out[0]));
  };
  for (int r = 0; r < 3; ++r) {
    const double c0 = o[r], cx = (ax[r] - o[r]) / W, cy = (ay[r] - o[r]) / H;
    ray0[r] = static_cast<float>(c0);
    rayx[r] = static_cast<float>(cx);
    rayy[r] = static_cast<float>(cy);
    if (df) {
      split(c0, df->ray0[r]);
      split(cx, df->rayx[r]);
      split(cy, df->rayy[r]);
    }
  }
}

void view_projection(const althea_globals& g, float vp[16]) {
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) {
      double acc = 0.0;
      for (int k = 0; k < 4; ++k) acc += static_cast<double>(g.projection[k * 4 + r]) * g.view[c * 4 + k];
      vp[c * 4 + r] = static_cast<float>(acc);
    }
}

static althea_status launch_resolve_with(ResolveParams& p, const althea_gbuffer_view* g,
                                        const althea_point_light* lights, unsigned flags,
                                        const MlpInferenceView* mlp, uint8_t* mlp_used, hipStream_t s,
                                        const althea_reflection_view* reflection, const float* shadow_cubes,
                                        int shadow_size);

// The SSAO view-depth plane lives in a stream-ordered allocation for the
// duration of the call (W x H floats).
althea_status launch_resolve(const althea_gbuffer_view* g, const althea_ibl_views* ibl,
                             const althea_point_light* lights, int light_count, const althea_globals* globals,
                             const althea_ssao_desc* ssao, unsigned flags, void* out, float* ao_out,
                             const MlpInferenceView* mlp, uint8_t* mlp_used, hipStream_t s,
                             const althea_reflection_view* reflection, const float* shadow_cubes,
                             int shadow_size) {
  ResolveParams p{};
  p.position = g->position;
  p.normal = g->normal;
  p.albedo = g->albedo;
  p.mro = g->mro;
  p.width = g->width;
  p.height = g->height;
  p.env = reinterpret_cast<const float4*>(ibl->environment);
  p.env_w = ibl->env_width;
  p.env_h = ibl->env_height;
  p.prefiltered = reinterpret_cast<const float4*>(ibl->prefiltered_bordered);
  p.pref = bordered_cube_of(ibl->prefiltered_size, ibl->prefiltered_mips);
  p.irradiance = reinterpret_cast<const float4*>(ibl->irradiance_bordered);
  p.irr_size = ibl->irradiance_size;
  p.lut = reinterpret_cast<const float2*>(ibl->brdf_lut);
  p.lut_size = ibl->brdf_lut_size;
  p.light_count = light_count;
  view_ray_affine(*globals, g->width, g->height, p.ray0, p.rayx, p.rayy, &p.ray_df);
  for (int c = 0; c < 4; ++c) p.view_z[c] = globals->view[c * 4 + 2];
  view_projection(*globals, p.vp);
  p.ssao_w_depth = p.vp[3] == -p.view_z[0] && p.vp[7] == -p.view_z[1] && p.vp[11] == -p.view_z[2] &&
                   p.vp[15] == -p.view_z[3];
  p.exposure = globals->exposure;
  p.tonemap = (flags & ALTHEA_RESOLVE_TONEMAP) ? 1 : 0;
  p.ssao_on = (ssao && ssao->enabled) ? 1 : 0;
  if (p.ssao_on) {
    p.ssao_samples = ssao->samples;
    p.ssao_inv_samples = 1.0f / static_cast<float>(ssao->samples);
    p.ssao_radius = ssao->radius;
    p.ssao_bias = ssao->bias;
    for (int k = 0; k < ssao->samples; ++k)
      for (int c = 0; c < 4; ++c) p.kernel[k][c] = ssao->kernel[k][c];
    for (int k = 0; k < 16; ++k)
      for (int c = 0; c < 4; ++c) p.noise[k][c] = ssao->noise[k][c];
  }
  p.out = out;
  p.ao_out = ao_out;
  // One stream-ordered allocation for what the call must build: the SSAO
  // depth plane (SSAO on) and the bordered cubes (views without them).
  const size_t pixels = static_cast<size_t>(g->width) * g->height;
  const bool border = !p.prefiltered || !p.irradiance;
  const size_t plane_bytes = p.ssao_on ? (pixels * sizeof(float) + 255) & ~static_cast<size_t>(255) : 0;
  const size_t pref_bytes =
      border ? 16 * cube_bordered_texels(ibl->prefiltered_size, ibl->prefiltered_mips) : 0;
  const size_t irr_bytes = border ? 16 * cube_bordered_texels(ibl->irradiance_size, 1) : 0;
  const size_t bytes = plane_bytes + ((pref_bytes + 255) & ~static_cast<size_t>(255)) + irr_bytes;
  if (bytes == 0)
    return launch_resolve_with(p, g, lights, flags, mlp, mlp_used, s, reflection, shadow_cubes, shadow_size);
  char* mem = nullptr;
  const hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&mem), bytes, s);
  if (e != hipSuccess) {
    set_error("resolve: hipMallocAsync of %zu bytes (SSAO depth plane, bordered cubes) failed: %s", bytes,
              hipGetErrorString(e));
    return ALTHEA_ERROR_HIP;
  }
  althea_status st = ALTHEA_OK;
  if (border) {
    float* pref_b = reinterpret_cast<float*>(mem + plane_bytes);
    float* irr_b = reinterpret_cast<float*>(mem + plane_bytes + ((pref_bytes + 255) & ~static_cast<size_t>(255)));
    st = launch_cube_border(ibl->prefiltered, ibl->prefiltered_size, ibl->prefiltered_mips, pref_b, s);
    if (st == ALTHEA_OK) st = launch_cube_border(ibl->irradiance, ibl->irradiance_size, 1, irr_b, s);
    p.prefiltered = reinterpret_cast<const float4*>(pref_b);
    p.irradiance = reinterpret_cast<const float4*>(irr_b);
  }
  if (st == ALTHEA_OK && p.ssao_on) {
    float* plane = reinterpret_cast<float*>(mem);
    p.ssao_depth = plane;
    const dim3 dgrid(static_cast<unsigned>((pixels + 255) / 256));
    if (g->format == ALTHEA_GBUFFER_FP32)
      hipLaunchKernelGGL(ssao_depth_kernel<ALTHEA_GBUFFER_FP32>, dgrid, dim3(256), 0, s, p, plane);
    else
      hipLaunchKernelGGL(ssao_depth_kernel<ALTHEA_GBUFFER_FP16>, dgrid, dim3(256), 0, s, p, plane);
    st = check_launch("ssao_depth_kernel");
  }
  if (st == ALTHEA_OK)
    st = launch_resolve_with(p, g, lights, flags, mlp, mlp_used, s, reflection, shadow_cubes, shadow_size);
  (void)hipFreeAsync(mem, s);
  return st;
}

static althea_status launch_resolve_with(ResolveParams& p, const althea_gbuffer_view* g,
                                        const althea_point_light* lights, unsigned flags,
                                        const MlpInferenceView* mlp, uint8_t* mlp_used, hipStream_t s,
                                        const althea_reflection_view* reflection, const float* shadow_cubes,
                                        int shadow_size) {
  const int tiles = ((g->width + 63) / 64) * ((g->height + 3) / 4);
  const float4* lights4 = reinterpret_cast<const float4*>(lights);
  const bool f16_out = (flags & ALTHEA_RESOLVE_OUT_FP16) != 0;
  const bool fp32 = g->format == ALTHEA_GBUFFER_FP32;
  if (mlp && mlp->mon_accepted == 0) {
    // No cell accepted: the MLP variant would shade exactly like the plain
    // kernel at lower occupancy.  Run the plain kernel and report no MLP use.
    if (mlp_used) {
      hipLaunchKernelGGL(fill_u8_kernel, dim3((g->width * g->height + 255) / 256), dim3(256), 0, s, mlp_used,
                         g->width * g->height);
      althea_status st = check_launch("fill_u8_kernel");
      if (st != ALTHEA_OK) return st;
    }
    mlp = nullptr;
  }
  if (mlp) {
    // The resolve at full occupancy with the MLP's pixels queued, then the
    // finishing kernel over the queue: the MLP's cost follows its pixels.
    p.mlp = *mlp;
    p.mlp_used = mlp_used;
    const uint32_t regions = static_cast<uint32_t>(tiles) * 4u;
    if (p.mlp.queue_regions < regions) {
      set_error("deferred_resolve_mlp: queue of %zu regions, the frame needs %u", p.mlp.queue_regions, regions);
      return ALTHEA_ERROR_INVALID_ARGUMENT;
    }
    const dim3 grid(tiles);
    const int mv = (f16_out ? 1 : 0) | (fp32 ? 0 : 2) | (p.ssao_w_depth ? 4 : 0);
    switch (mv) {
#define ALTHEA_RESOLVE_MLP_CASE(v, fmt, f16, wdepth)                                                    \
  case v:                                                                                             \
    hipLaunchKernelGGL((resolve_mlp_kernel<fmt, f16, wdepth>), grid, dim3(256), 0, s, p, lights4); \
    break;
      ALTHEA_RESOLVE_MLP_CASE(0, ALTHEA_GBUFFER_FP32, false, false)
      ALTHEA_RESOLVE_MLP_CASE(1, ALTHEA_GBUFFER_FP32, true, false)
      ALTHEA_RESOLVE_MLP_CASE(2, ALTHEA_GBUFFER_FP16, false, false)
      ALTHEA_RESOLVE_MLP_CASE(3, ALTHEA_GBUFFER_FP16, true, false)
      ALTHEA_RESOLVE_MLP_CASE(4, ALTHEA_GBUFFER_FP32, false, true)
      ALTHEA_RESOLVE_MLP_CASE(5, ALTHEA_GBUFFER_FP32, true, true)
      ALTHEA_RESOLVE_MLP_CASE(6, ALTHEA_GBUFFER_FP16, false, true)
      ALTHEA_RESOLVE_MLP_CASE(7, ALTHEA_GBUFFER_FP16, true, true)
#undef ALTHEA_RESOLVE_MLP_CASE
      default: break;
    }
    althea_status st = check_launch("resolve_mlp_kernel");
    if (st != ALTHEA_OK) return st;
    int dev = 0, cus = 256;
    ALTHEA_TRY_HIP(hipGetDevice(&dev));
    ALTHEA_TRY_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const dim3 fgrid(kMlpFinishBlocksPerCu * cus);
    if (f16_out)
      hipLaunchKernelGGL(resolve_mlp_finish_kernel<true>, fgrid, dim3(256), 0, s, p, regions);
    else
      hipLaunchKernelGGL(resolve_mlp_finish_kernel<false>, fgrid, dim3(256), 0, s, p, regions);
    return check_launch("resolve_mlp_finish_kernel");
  }
  const dim3 grid(tiles);
  if (reflection || shadow_cubes) {
    if (reflection) refl_chain(*reflection, p.refl);
    p.shadow.cubes = shadow_cubes;
    p.shadow.size = shadow_size;
    const int variant = (reflection ? 1 : 0) | (shadow_cubes ? 2 : 0) | (f16_out ? 4 : 0) | (fp32 ? 0 : 8) |
                        (p.ssao_w_depth ? 16 : 0);
    switch (variant) {
#define ALTHEA_RESOLVE_CASE2(v, fmt, f16, refl, shadow, wdepth)                                                  \
  case v:                                                                                                        \
    hipLaunchKernelGGL((resolve_kernel<fmt, f16, refl, shadow, false, wdepth>), grid, dim3(256), 0, s, p, lights4); \
    break;
#define ALTHEA_RESOLVE_CASE(v, fmt, f16, refl, shadow)    \
  ALTHEA_RESOLVE_CASE2(v, fmt, f16, refl, shadow, false) \
  ALTHEA_RESOLVE_CASE2(v + 16, fmt, f16, refl, shadow, true)
      ALTHEA_RESOLVE_CASE(1, ALTHEA_GBUFFER_FP32, false, true, false)
      ALTHEA_RESOLVE_CASE(2, ALTHEA_GBUFFER_FP32, false, false, true)
      ALTHEA_RESOLVE_CASE(3, ALTHEA_GBUFFER_FP32, false, true, true)
      ALTHEA_RESOLVE_CASE(5, ALTHEA_GBUFFER_FP32, true, true, false)
      ALTHEA_RESOLVE_CASE(6, ALTHEA_GBUFFER_FP32, true, false, true)
      ALTHEA_RESOLVE_CASE(7, ALTHEA_GBUFFER_FP32, true, true, true)
      ALTHEA_RESOLVE_CASE(9, ALTHEA_GBUFFER_FP16, false, true, false)
      ALTHEA_RESOLVE_CASE(10, ALTHEA_GBUFFER_FP16, false, false, true)
      ALTHEA_RESOLVE_CASE(11, ALTHEA_GBUFFER_FP16, false, true, true)
      ALTHEA_RESOLVE_CASE(13, ALTHEA_GBUFFER_FP16, true, true, false)
      ALTHEA_RESOLVE_CASE(14, ALTHEA_GBUFFER_FP16, true, false, true)
      ALTHEA_RESOLVE_CASE(15, ALTHEA_GBUFFER_FP16, true, true, true)
#undef ALTHEA_RESOLVE_CASE
#undef ALTHEA_RESOLVE_CASE2
      default: break;
    }
    return check_launch("resolve_kernel");
  }
  const int variant = (f16_out ? 1 : 0) | (fp32 ? 0 : 2) | (p.ssao_w_depth ? 4 : 0);
  switch (variant) {
#define ALTHEA_RESOLVE_CASE(v, fmt, f16, wdepth)                                                                  \
  case v:                                                                                                         \
    hipLaunchKernelGGL((resolve_kernel<fmt, f16, false, false, false, wdepth>), grid, dim3(256), 0, s, p, lights4); \
    break;
    ALTHEA_RESOLVE_CASE(0, ALTHEA_GBUFFER_FP32, false, false)
    ALTHEA_RESOLVE_CASE(1, ALTHEA_GBUFFER_FP32, true, false)
    ALTHEA_RESOLVE_CASE(2, ALTHEA_GBUFFER_FP16, false, false)
    ALTHEA_RESOLVE_CASE(3, ALTHEA_GBUFFER_FP16, true, false)
    ALTHEA_RESOLVE_CASE(4, ALTHEA_GBUFFER_FP32, false, true)
    ALTHEA_RESOLVE_CASE(5, ALTHEA_GBUFFER_FP32, true, true)
    ALTHEA_RESOLVE_CASE(6, ALTHEA_GBUFFER_FP16, false, true)
    ALTHEA_RESOLVE_CASE(7, ALTHEA_GBUFFER_FP16, true, true)
#undef ALTHEA_RESOLVE_CASE
    default: break;
  }
  return check_launch("resolve_kernel");
}

}  // namespace althea
