// Screen-space reflection and the glossy reflection buffer for gfx950.
//
// ssr_capture_kernel (Shaders/SSR.frag:80-149 with SSR.vert:15-22): one lane
// per pixel (tile order below).  A lane leaves the march at its hit, its exit
// from the screen or step 128; the wave leaves when no lane is marching.  The hit shading
// (SSR.frag:55-78, pbrMaterial with the IBL state and the point lights) runs
// once after the march for all hit lanes of the wave together.
//
// reflection_convolve_kernel (Shaders/SSRGlossyConvolve.comp:22-55, scheduled
// by Src/ReflectionBuffer.cpp:163-277): one lane per texel of level m, 7
// bilinear taps into level m-1 along one axis.
//
// Positions along the march are rebuilt from the depth buffer as the
// reference does (SSR.frag:42-45, Misc/ReconstructPosition.glsl): bilinear
// depth (the LINEAR / CLAMP_TO_EDGE depth sampler, DeferredRendering.cpp:67-69),
// then the point at that view depth on the pixel's view ray.  This G-buffer
// carries world positions instead of a depth attachment, so ssr_depth_kernel
// first writes the depth plane: 1 / view distance per pixel, which is affine
// in the [0, 1] depth the rasterizer stores (filtering either gives the same
// result in exact arithmetic, and 1/z keeps full FP32 relative precision where
// [0, 1] depth crowds towards 1); uncovered pixels hold the far plane (depth
// cleared to 1, DeferredRendering.cpp:105-106).  4 bytes per step instead of a
// 64-byte filtered position read.  The near and far planes come from the
// projection; ReconstructPosition.glsl hard-codes the default camera's 0.01
// and 1000.
//
// G-buffer coverage is position.w != 0 (the flag the resolve uses,
// DeferredPass.frag:44); the reference's SSR tests normal.a (SSR.frag:137).
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "device_math.hpp"
#include "internal.hpp"
#include "reflection.hpp"
#include "shading.hpp"
#include "shadow.hpp"

namespace althea {

constexpr int kSsrSteps = 128;        // RAYMARCH_STEPS, SSR.frag:80
constexpr float kSsrStep = 0.005f;    // uv per step, SSR.frag:89
constexpr float kSsrCone = 0.999f;    // cos of the accepted angle, SSR.frag:116-117
constexpr float kSsrFar = 10000.0f;   // march target distance, SSR.frag:83

struct SsrParams {
  const void* position;
  const void* normal;
  const void* albedo;
  const void* mro;
  int width, height;
  const float4* prefiltered;  // seamless-bordered copy (BorderedCube)
  BorderedCube pref;
  const float4* irradiance;  // seamless-bordered copy, one mip
  int irr_size;
  const float2* lut;
  int lut_size;
  int light_count;
  float ray0[3], rayu[3], rayv[3];  // view ray at uv: ray0 + rayu u + rayv v
  float vp[16];                     // projection * view, column-major
  float view_z[4];                  // row 2 of the view matrix
  float eye[3];
  float inv_far;                    // depth-plane value of an uncovered pixel
  float* depth;                     // depth plane in 8 x 4 texel blocks (taddr)
  int depth_tw;                     // blocks per block row, ceil(W / 8)
  ReflChain out;                    // level 0 is written
  ShadowCubes shadow;               // shadowed variant only
};

template <int kFmt>
__device__ __forceinline__ float4 gb_load(const void* plane, int i) {
  if (kFmt == ALTHEA_GBUFFER_FP32) return static_cast<const float4*>(plane)[i];
  const ushort4 h = static_cast<const ushort4*>(plane)[i];
  return make_float4(__half2float(__ushort_as_half(h.x)), __half2float(__ushort_as_half(h.y)),
                     __half2float(__ushort_as_half(h.z)), __half2float(__ushort_as_half(h.w)));
}

// Bilinear CLAMP_TO_EDGE read of a G-buffer plane at uv in [0, 1]^2.
template <int kFmt>
__device__ __forceinline__ float4 gb_bilinear(const void* plane, int W, int H, float u, float v) {
  const float x = fmaf(u, static_cast<float>(W), -0.5f), y = fmaf(v, static_cast<float>(H), -0.5f);
  const float fx = floorf(x), fy = floorf(y);
  const float ax = x - fx, ay = y - fy;
  const int ix = static_cast<int>(fx), iy = static_cast<int>(fy);
  const int x0 = min(max(ix, 0), W - 1), x1 = min(max(ix + 1, 0), W - 1);
  const int y0 = min(max(iy, 0), H - 1), y1 = min(max(iy + 1, 0), H - 1);
  const float4 t00 = gb_load<kFmt>(plane, y0 * W + x0), t10 = gb_load<kFmt>(plane, y0 * W + x1);
  const float4 t01 = gb_load<kFmt>(plane, y1 * W + x0), t11 = gb_load<kFmt>(plane, y1 * W + x1);
  return lerp4(lerp4(t00, t10, ax), lerp4(t01, t11, ax), ay);
}

// Depth-plane layout: blocks of 8 x 4 texels, one 128-byte cache line each,
// block rows of ceil(W / 8) blocks.  A bilinear footprint then falls in one
// line 2 times in 3 (row-major rows always take two), and a wave's nearby
// march points share lines in both directions.  Measured at 4K: 2.07 ms
// against 2.43 ms row-major (4 x 8 blocks 2.24, 16 x 2 blocks 2.23).
__device__ __forceinline__ int taddr(int tw, int x, int y) {
  return (((y >> 2) * tw + (x >> 3)) << 5) + ((y & 3) << 3) + (x & 7);
}

// Depth plane: 1 / view distance of each covered pixel, 1 / far elsewhere.
template <int kFmt>
__global__ __launch_bounds__(256) void ssr_depth_kernel(const SsrParams p) {
  const int n = p.width * p.height;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float4 q = gb_load<kFmt>(p.position, i);
  const float vz = fmaf(p.view_z[0], q.x, fmaf(p.view_z[1], q.y, fmaf(p.view_z[2], q.z, p.view_z[3])));
  const int x = i % p.width, y = i / p.width;
  p.depth[taddr(p.depth_tw, x, y)] = q.w != 0.0f ? -1.0f / vz : p.inv_far;
}

// Bilinear CLAMP_TO_EDGE depth at uv: the four texel reads (depth_tap), then
// the blend (depth_value).
struct DepthTap {
  float a0, b0, a1, b1, ax, ay;
};

__device__ __forceinline__ DepthTap depth_tap(const float* __restrict__ d, int tw, int W, int H, float u, float v) {
  const float x = fmaf(u, static_cast<float>(W), -0.5f), y = fmaf(v, static_cast<float>(H), -0.5f);
  const float fx = floorf(x), fy = floorf(y);
  DepthTap t;
  t.ax = x - fx;
  t.ay = y - fy;
  const int ix = static_cast<int>(fx), iy = static_cast<int>(fy);
  const int x0 = min(max(ix, 0), W - 1), x1 = min(max(ix + 1, 0), W - 1);
  const int y0 = min(max(iy, 0), H - 1), y1 = min(max(iy + 1, 0), H - 1);
  // The other three texels from the first: +1 in a block, +25 into the next
  // block; +8 in a block, one block row on (32 tw) less 3 rows of 8.
  const int i00 = taddr(tw, x0, y0);
  const int dx = x1 == x0 ? 0 : ((x0 & 7) == 7 ? 25 : 1);
  const int dy = y1 == y0 ? 0 : ((y0 & 3) == 3 ? (tw << 5) - 24 : 8);
  t.a0 = d[i00];
  t.b0 = d[i00 + dx];
  t.a1 = d[i00 + dy];
  t.b1 = d[i00 + dy + dx];
  return t;
}

__device__ __forceinline__ float depth_value(const DepthTap& t) {
  const float top = fmaf(t.b0 - t.a0, t.ax, t.a0), bot = fmaf(t.b1 - t.a1, t.ax, t.a1);
  return fmaf(bot - top, t.ay, top);
}

// View ray through uv (SSR.vert:20, interpolated across the screen).
__device__ __forceinline__ f3 view_ray_at(const SsrParams& p, float u, float v) {
  return mk(fmaf(p.rayv[0], v, fmaf(p.rayu[0], u, p.ray0[0])), fmaf(p.rayv[1], v, fmaf(p.rayu[1], u, p.ray0[1])),
            fmaf(p.rayv[2], v, fmaf(p.rayu[2], u, p.ray0[2])));
}

__device__ __forceinline__ float view_z_of(const SsrParams& p, f3 r) {
  return fmaf(p.view_z[0], r.x, fmaf(p.view_z[1], r.y, p.view_z[2] * r.z));
}

// reconstructPosition (Misc/ReconstructPosition.glsl:4-22): the point at
// view distance 1 / s on the view ray through uv.
__device__ __forceinline__ f3 reconstruct(const SsrParams& p, float u, float v, float s) {
  const f3 ray = view_ray_at(p, u, v);
  const float t = -rcp_fast(s * view_z_of(p, ray));
  return mk(fmaf(ray.x, t, p.eye[0]), fmaf(ray.y, t, p.eye[1]), fmaf(ray.z, t, p.eye[2]));
}

// Tiles: 16 x 16 pixels per block, each wave an 8 x 8 quadrant, so a wave's
// 64 rays start close together and their filtered depth reads share cache
// lines.  The march reads far along the frame (mostly up it), so L2 reuse
// decides the speed: blocks are dealt to the 8 XCDs in turn (block b runs on
// XCD b % 8), so XCD x takes the tiles of column band x, walking it column by
// column.  Each XCD's L2 then holds one band's depth (about 4 MB at 4K)
// instead of the whole frame's.  Measured at 4K: 64 x 4 tiles in launch order
// 4.6 ms, column bands 3.5 ms, bands of 16 x 16 tiles 2.5 ms.
constexpr int kSsrTile = 16;
constexpr int kXcds = 8;

__host__ __device__ inline int ssr_band_tiles(int tiles_x, int tiles_y) {
  return ((tiles_x + kXcds - 1) / kXcds) * tiles_y;
}

template <int kFmt, bool kShadow = false>
__global__ __launch_bounds__(256) void ssr_capture_kernel(const SsrParams p, const float4* __restrict__ lights) {
  const int tiles_x = (p.width + kSsrTile - 1) / kSsrTile, tiles_y = (p.height + kSsrTile - 1) / kSsrTile;
  const int band = blockIdx.x % kXcds, j = blockIdx.x / kXcds;
  const int c0 = band * tiles_x / kXcds, bw = (band + 1) * tiles_x / kXcds - c0;
  if (j >= bw * tiles_y) return;  // the band has fewer tiles than the widest one
  const int tx = c0 + j / tiles_y, ty = j % tiles_y;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int px = tx * kSsrTile + (lane & 7) + 8 * (wave & 1);
  const int py = ty * kSsrTile + (lane >> 3) + 8 * (wave >> 1);
  const bool inside = px < p.width && py < p.height;
  const int idx = min(py, p.height - 1) * p.width + min(px, p.width - 1);
  const float4 pos = gb_load<kFmt>(p.position, idx);
  const bool covered = inside && pos.w != 0.0f;
  const float4 nrm = covered ? gb_load<kFmt>(p.normal, idx) : make_float4(0.0f, 0.0f, 1.0f, 0.0f);
  const f3 N = normalize_fast(mk(nrm.x, nrm.y, nrm.z));
  const float u0 = (static_cast<float>(px) + 0.5f) / static_cast<float>(p.width);
  const float v0 = (static_cast<float>(py) + 0.5f) / static_cast<float>(p.height);
  // The pixel's own position from its depth texel (SSR.frag:143).
  const float s0 = p.depth[taddr(p.depth_tw, min(px, p.width - 1), min(py, p.height - 1))];
  const f3 P = reconstruct(p, u0, v0, s0);
  // reflect(normalize(view ray), N) (SSR.frag:145).
  const f3 ray0 = view_ray_at(p, u0, v0);
  const f3 I = normalize_fast(ray0);
  const float ni = 2.0f * dot(N, I);
  const f3 R = mk(fmaf(-ni, N.x, I.x), fmaf(-ni, N.y, I.y), fmaf(-ni, N.z, I.z));
  // March direction in uv towards the projection of P + 10000 R
  // (SSR.frag:82-90), and the reference plane normal (SSR.frag:91-92).
  const f3 E = mk(fmaf(R.x, kSsrFar, P.x), fmaf(R.y, kSsrFar, P.y), fmaf(R.z, kSsrFar, P.z));
  const float ex = fmaf(p.vp[0], E.x, fmaf(p.vp[4], E.y, fmaf(p.vp[8], E.z, p.vp[12])));
  const float ey = fmaf(p.vp[1], E.x, fmaf(p.vp[5], E.y, fmaf(p.vp[9], E.z, p.vp[13])));
  const float ew = fmaf(p.vp[3], E.x, fmaf(p.vp[7], E.y, fmaf(p.vp[11], E.z, p.vp[15])));
  const float du = fmaf(0.5f, ex / ew, 0.5f) - u0, dv = fmaf(0.5f, ey / ew, 0.5f) - v0;
  const float sl = kSsrStep * rsq_fast(fmaf(du, du, dv * dv));
  // A non-finite march direction (the target projects onto the pixel, or to
  // w = 0) gives no reflection.
  bool active = covered && isfinite(du * sl) && isfinite(dv * sl);
  const float su = active ? du * sl : 0.0f, sv = active ? dv * sl : 0.0f;
  const f3 perp = normalize_fast(cross(cross(R, N), R));
  // Along the march uv, the view ray and its view-space z are affine in the
  // step number k: uv0 + k step (currentUV += uvStep * stepSize in the
  // reference, equal in exact arithmetic, one rounding instead of k).
  const f3 rk = mk(fmaf(p.rayv[0], sv, p.rayu[0] * su), fmaf(p.rayv[1], sv, p.rayu[1] * su),
                   fmaf(p.rayv[2], sv, p.rayu[2] * su));
  const float vz0 = view_z_of(p, ray0), vzk = view_z_of(p, rk);
  const f3 EP = mk(p.eye[0] - P.x, p.eye[1] - P.y, p.eye[2] - P.z);
  bool hit = false;
  float hk = 0.0f;
  f3 Q = mk(0.0f, 0.0f, 0.0f), Nh = mk(0.0f, 0.0f, 0.0f);
  float prev = 0.0f;
  // Lanes that left the march issue no loads (exec-masked): measured faster
  // than marching every lane to the wave's end, with or without issuing the
  // next step's loads ahead.
  for (int i = 0; i < kSsrSteps; ++i) {
    if (!__any(active)) break;
    if (active) {
      const float k = static_cast<float>(i + 1);
      const float u = fmaf(k, su, u0), v = fmaf(k, sv, v0);
      if (!(u >= 0.0f && u <= 1.0f && v >= 0.0f && v <= 1.0f)) {
        active = false;  // off screen: SSR.frag:102-104
      } else {
        const DepthTap tap = depth_tap(p.depth, p.depth_tw, p.width, p.height, u, v);
        const f3 ray = mk(fmaf(k, rk.x, ray0.x), fmaf(k, rk.y, ray0.y), fmaf(k, rk.z, ray0.z));
        const float t = -rcp_fast(depth_value(tap) * fmaf(k, vzk, vz0));
        const f3 dq = mk(fmaf(ray.x, t, EP.x), fmaf(ray.y, t, EP.y), fmaf(ray.z, t, EP.z));
        const float rl = rsq_fast(dot(dq, dq));
        const float proj = dot(dq, perp) * rl;
        const float f = dot(dq, R) * rl;
        if (i > 0 && proj * prev <= 0.0f && f > kSsrCone) {
          const float4 n4 = gb_bilinear<kFmt>(p.normal, p.width, p.height, u, v);
          const f3 nh = normalize_fast(mk(n4.x, n4.y, n4.z));
          if (dot(nh, R) < 0.0f) {
            hit = true;
            active = false;
            hk = k;
            Nh = nh;
            Q = mk(fmaf(ray.x, t, p.eye[0]), fmaf(ray.y, t, p.eye[1]), fmaf(ray.z, t, p.eye[2]));
          }
        }
        prev = proj;
      }
    }
  }
  const float hu = fmaf(hk, su, u0), hv = fmaf(hk, sv, v0);
  // environmentLitSample (SSR.frag:55-78): V = the reflected ray, occlusion 1.
  const float4 zero4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float4 alb = hit ? gb_bilinear<kFmt>(p.albedo, p.width, p.height, hu, hv) : zero4;
  const float4 mro = hit ? gb_bilinear<kFmt>(p.mro, p.width, p.height, hu, hv) : zero4;
  const float roughness = hit ? mro.y : 1.0f;
  const float NI = dot(Nh, R);
  const PbrPoint m = pbr_point(mk(alb.x, alb.y, alb.z), mro.x, roughness, NI);
  f3 color = mk(0.0f, 0.0f, 0.0f);
  if (hit) {
    const float nd = 2.0f * NI;
    const f3 Rr = mk(fmaf(-nd, Nh.x, R.x), fmaf(-nd, Nh.y, R.y), fmaf(-nd, Nh.z, R.z));
    const f3 refl = bcube_sample(p.prefiltered, p.pref, Rr, roughness * static_cast<float>(p.pref.mips - 1));
    int face;
    float pc, pt;
    cube_select(Nh, face, pc, pt);
    const f3 irr = bcube_bilinear(p.irradiance, p.irr_size, face, pc, pt);
    float A, B;
    lut_sample(p.lut, p.lut_size, m.n_dot_v, roughness, A, B);
    color = pbr_environment(m, irr, refl, A, B, 1.0f);
  }
  // pbrMaterial's light loop, with the shadow test when cubes are given
  // (PBRMaterial.glsl:119-140 runs it for every caller).
  if (kShadow)
    pbr_lights(m, Q, Nh, R, NI, __any(hit), lights, p.light_count, color, p.shadow);
  else
    pbr_lights(m, Q, Nh, R, NI, __any(hit), lights, p.light_count, color);
  if (inside) refl_store(p.out, static_cast<uint32_t>(idx), hit ? make_float4(color.x, color.y, color.z, 1.0f) : zero4);
}

// Level m from level m-1.  The texel's uv is texel / extent(m) with no
// half-texel offset (SSRGlossyConvolve.comp:35-36); a tap at uv + o d / W_m
// lands at (q + c) source texels, q integer, c = frac + o s - 0.5 with
// s = W_{m-1} / W_m (along x) or H_{m-1} / W_m (along y).
struct ConvolveArgs {
  ReflChain c;
  int level;
  int vertical;  // blur along y (odd levels, ReflectionBuffer.cpp:254-255)
  float scale;   // s above
};

__global__ __launch_bounds__(256) void reflection_convolve_kernel(const ConvolveArgs a) {
  const int l = a.level;
  const int dw = a.c.w[l], dh = a.c.h[l];
  const int tiles_x = (dw + 63) / 64;
  const int x = (blockIdx.x % tiles_x) * 64 + (threadIdx.x & 63);
  const int y = (blockIdx.x / tiles_x) * 4 + (threadIdx.x >> 6);
  if (x >= dw || y >= dh) return;
  int qx, qy;
  float fx, fy;
  rational_split(static_cast<uint32_t>(x) * static_cast<uint32_t>(a.c.w[l - 1]), static_cast<uint32_t>(dw), qx, fx);
  rational_split(static_cast<uint32_t>(y) * static_cast<uint32_t>(a.c.h[l - 1]), static_cast<uint32_t>(dh), qy, fy);
  fx -= 0.5f;
  fy -= 0.5f;
  // 13-tap binomial kernel (row 16 of Pascal's triangle without its two
  // outer taps per side) folded into 7 linear taps (SSRGlossyConvolve.comp:42-52).
  constexpr float kOff[3] = {1.411764705882353f, 3.2941176470588234f, 5.176470588235294f};
  constexpr float kW[3] = {0.2969069646728344f, 0.09447039785044732f, 0.010381362401148057f};
  auto tap = [&](float o) {
    const float d = o * a.scale;
    return a.vertical ? refl_bilinear(a.c, l - 1, qx, fx, qy, fy + d) : refl_bilinear(a.c, l - 1, qx, fx + d, qy, fy);
  };
  const float4 c0 = tap(0.0f);
  float4 acc = make_float4(c0.x * 0.1964825501511404f, c0.y * 0.1964825501511404f, c0.z * 0.1964825501511404f,
                           c0.w * 0.1964825501511404f);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float4 t0 = tap(kOff[k]);
    acc = make_float4(fmaf(t0.x, kW[k], acc.x), fmaf(t0.y, kW[k], acc.y), fmaf(t0.z, kW[k], acc.z),
                      fmaf(t0.w, kW[k], acc.w));
    const float4 t1 = tap(-kOff[k]);
    acc = make_float4(fmaf(t1.x, kW[k], acc.x), fmaf(t1.y, kW[k], acc.y), fmaf(t1.z, kW[k], acc.z),
                      fmaf(t1.w, kW[k], acc.w));
  }
  refl_store(a.c, a.c.off[l] + static_cast<uint32_t>(y * dw + x), acc);
}

void refl_chain(const althea_reflection_view& v, ReflChain& c) {
  c.data = v.data;
  c.f16 = v.format == ALTHEA_REFLECTION_FP16 ? 1 : 0;
  c.mips = v.mips;
  uint32_t off = 0;
  for (int m = 0; m < ALTHEA_REFLECTION_MAX_MIPS; ++m) {
    c.w[m] = std::max(1, v.width >> m);
    c.h[m] = std::max(1, v.height >> m);
    c.off[m] = off;
    if (m < v.mips) off += static_cast<uint32_t>(c.w[m]) * static_cast<uint32_t>(c.h[m]);
  }
}

size_t ssr_workspace_bytes(int width, int height) {
  return static_cast<size_t>((width + 7) / 8) * static_cast<size_t>((height + 3) / 4) * 32 * sizeof(float);
}

althea_status launch_ssr_capture(const althea_gbuffer_view* g, const althea_ibl_views* ibl,
                                 const althea_point_light* lights, int light_count, const althea_globals* globals,
                                 const althea_reflection_view* reflection, void* workspace, hipStream_t s,
                                 const float* shadow_cubes, int shadow_size) {
  SsrParams p{};
  p.position = g->position;
  p.normal = g->normal;
  p.albedo = g->albedo;
  p.mro = g->mro;
  p.width = g->width;
  p.height = g->height;
  p.prefiltered = reinterpret_cast<const float4*>(ibl->prefiltered_bordered);
  p.pref = bordered_cube_of(ibl->prefiltered_size, ibl->prefiltered_mips);
  p.irradiance = reinterpret_cast<const float4*>(ibl->irradiance_bordered);
  p.irr_size = ibl->irradiance_size;
  p.lut = reinterpret_cast<const float2*>(ibl->brdf_lut);
  p.lut_size = ibl->brdf_lut_size;
  p.light_count = light_count;
  // View ray per unit uv: the pixel-affine form scaled by the extent.
  float rx[3], ry[3];
  view_ray_affine(*globals, g->width, g->height, p.ray0, rx, ry);
  for (int c = 0; c < 3; ++c) {
    p.rayu[c] = static_cast<float>(static_cast<double>(rx[c]) * g->width);
    p.rayv[c] = static_cast<float>(static_cast<double>(ry[c]) * g->height);
  }
  view_projection(*globals, p.vp);
  for (int c = 0; c < 4; ++c) p.view_z[c] = globals->view[c * 4 + 2];
  for (int c = 0; c < 3; ++c) p.eye[c] = globals->inverse_view[12 + c];
  // Far plane of a [0, 1]-depth perspective projection: P[3][2] / (P[2][2] + 1).
  const double far =
      static_cast<double>(globals->projection[14]) / (static_cast<double>(globals->projection[10]) + 1.0);
  p.inv_far = static_cast<float>(1.0 / far);
  p.depth = static_cast<float*>(workspace);
  p.depth_tw = (g->width + 7) / 8;
  refl_chain(*reflection, p.out);
  const int n = g->width * g->height;
  const dim3 dgrid((n + 255) / 256);
  const int tiles_x = (g->width + kSsrTile - 1) / kSsrTile, tiles_y = (g->height + kSsrTile - 1) / kSsrTile;
  const dim3 grid(kXcds * ssr_band_tiles(tiles_x, tiles_y));
  const float4* lights4 = reinterpret_cast<const float4*>(lights);
  p.shadow.cubes = shadow_cubes;
  p.shadow.size = shadow_size;
  const bool sh = shadow_cubes != nullptr;
  // Views without the seamless-bordered cubes: build them in a
  // stream-ordered allocation for this call.
  char* border_mem = nullptr;
  if (!p.prefiltered || !p.irradiance) {
    const size_t pref_bytes = (16 * cube_bordered_texels(ibl->prefiltered_size, ibl->prefiltered_mips) + 255) &
                              ~static_cast<size_t>(255);
    const size_t bytes = pref_bytes + 16 * cube_bordered_texels(ibl->irradiance_size, 1);
    const hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&border_mem), bytes, s);
    if (e != hipSuccess) {
      set_error("ssr_capture: hipMallocAsync of %zu bytes (bordered cubes) failed: %s", bytes, hipGetErrorString(e));
      return ALTHEA_ERROR_HIP;
    }
    float* pref_b = reinterpret_cast<float*>(border_mem);
    float* irr_b = reinterpret_cast<float*>(border_mem + pref_bytes);
    althea_status st = launch_cube_border(ibl->prefiltered, ibl->prefiltered_size, ibl->prefiltered_mips, pref_b, s);
    if (st == ALTHEA_OK) st = launch_cube_border(ibl->irradiance, ibl->irradiance_size, 1, irr_b, s);
    if (st != ALTHEA_OK) {
      (void)hipFreeAsync(border_mem, s);
      return st;
    }
    p.prefiltered = reinterpret_cast<const float4*>(pref_b);
    p.irradiance = reinterpret_cast<const float4*>(irr_b);
  }
  if (g->format == ALTHEA_GBUFFER_FP32) {
    hipLaunchKernelGGL((ssr_depth_kernel<ALTHEA_GBUFFER_FP32>), dgrid, dim3(256), 0, s, p);
    if (sh)
      hipLaunchKernelGGL((ssr_capture_kernel<ALTHEA_GBUFFER_FP32, true>), grid, dim3(256), 0, s, p, lights4);
    else
      hipLaunchKernelGGL((ssr_capture_kernel<ALTHEA_GBUFFER_FP32>), grid, dim3(256), 0, s, p, lights4);
  } else {
    hipLaunchKernelGGL((ssr_depth_kernel<ALTHEA_GBUFFER_FP16>), dgrid, dim3(256), 0, s, p);
    if (sh)
      hipLaunchKernelGGL((ssr_capture_kernel<ALTHEA_GBUFFER_FP16, true>), grid, dim3(256), 0, s, p, lights4);
    else
      hipLaunchKernelGGL((ssr_capture_kernel<ALTHEA_GBUFFER_FP16>), grid, dim3(256), 0, s, p, lights4);
  }
  const althea_status st = check_launch("ssr_capture_kernel");
  if (border_mem) (void)hipFreeAsync(border_mem, s);
  return st;
}

althea_status launch_reflection_convolve(const althea_reflection_view* reflection, hipStream_t s) {
  ConvolveArgs a{};
  refl_chain(*reflection, a.c);
  for (int l = 1; l < reflection->mips; ++l) {
    a.level = l;
    a.vertical = (l & 1) ? 1 : 0;
    a.scale = static_cast<float>(static_cast<double>(a.vertical ? a.c.h[l - 1] : a.c.w[l - 1]) / a.c.w[l]);
    const dim3 grid(((a.c.w[l] + 63) / 64) * ((a.c.h[l] + 3) / 4));
    hipLaunchKernelGGL(reflection_convolve_kernel, grid, dim3(256), 0, s, a);
    althea_status st = check_launch("reflection_convolve_kernel");
    if (st != ALTHEA_OK) return st;
  }
  return ALTHEA_OK;
}

}  // namespace althea
