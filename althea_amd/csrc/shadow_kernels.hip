// Point-light shadow cubes for gfx950 (the reference's shadow pass,
// Src/PointLight.cpp:235-282 with Shaders/ShadowMapBindless.{vert,frag}).
//
// What the reference rasterizes per face texel is the distance to the
// nearest front-facing triangle hit by the ray through the texel centre,
// kept where the hit's view depth lies in [0.01, 1000], over 1000, clamped to
// 1 (ShadowMap.frag: gl_FragDepth = length(worldPosCS) / zFar; LESS depth
// test; depth cleared to 1).  That is computed here directly: the ray
// through the centre of texel direction d from the light meets triangle
// (a, b, c) (light-relative) where every edge function d . (b x c),
// d . (c x a), d . (a x b) has the sign of det = a . (b x c); front-facing
// (counter-clockwise as the light sees it, the reference's default) means
// det < 0, and the hit lies at view depth t = det / (d . N), distance t |d|,
// where N = (b - a) x (c - a) is the sum of the three edge vectors.  N and
// det = a . N are formed from the triangle's edges, not as a . (b x c) of the
// light-relative vertices, which cancels in FP32 for small, distant
// triangles.
//
// One lane per (light, triangle): it clips the triangle to the part of each
// face's view volume where a hit can be written (view depth ma >= near / 2:
// the view depth of a hit on a face is its coordinate ma along the face axis,
// and hits nearer than the near plane are not written), takes the texel box
// that part projects to (the projection of a convex polygon in front of the
// camera is spanned by its projected vertices: the triangle's vertices in
// front and the points where its edges cross ma = near / 2), and writes each
// hit with an integer atomicMin on
// the depth's bits (non-negative floats order as their bit patterns), so the
// result does not depend on the order of the writes.  Two schedules run the
// same texel tests: with a workspace, every item's boxes are counted and
// scanned and their texels dealt out evenly over a fixed grid of waves (the
// shadow_slot_* kernels); without one, small boxes are walked by their own
// lane and a wave walks larger boxes together, one lane's box at a time
// (shadow_raster_kernel).  A last pass fills each face's border from the
// adjacent faces (the seamless filter's reads beyond the edge), so the lookup
// needs no folding.
//
// glTF primitives (althea_shadow_cubes_gltf) add the shadow shaders' alpha
// test (ShadowMapBindless.frag, ShadowMap.frag): a setup pass flattens the
// primitives into a triangle soup with one alpha-test record per triangle,
// and a texel whose base colour alpha (texture x factor) is below the
// material's cut-off is discarded before its depth is written.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "device_math.hpp"
#include "fill.hpp"
#include "gltf_texture.hpp"
#include "internal.hpp"
#include "shadow.hpp"
#include "slot_scan.hpp"

namespace althea {

constexpr int kShadowLaneTexels = 16;  // boxes up to this many texels stay on their lane
constexpr float kShadowClip = 0.5f * ALTHEA_SHADOW_NEAR;  // clip plane ma = near / 2
// Box margin in texels.  The FP32 projection of a vertex is off by a few ulp
// of a coordinate up to S + 1 (under 2e-3 texel at S = 4096), a texel-centre
// ray by about 2e-4 texel and a non-grazing edge test by about 1e-7 rad (2e-4
// texel): a texel the edge tests accept lies well within 1/64 texel of the
// exact span.
constexpr float kShadowBoxMargin = 1.0f / 64.0f;

struct ShadowSlots {
  unsigned long long* offsets;  // items + 1 slot offsets
  uint2* boxes;                 // per item and face: x0 | y0 << 16, w | h << 16 (items with texels only)
  unsigned long long* blocks;   // per setup block: slot total, then first slot
};

struct ShadowParams {
  const float* positions;    // float3 vertices, or with no indices a triangle soup (9 floats a triangle)
  const uint32_t* indices;   // 3 a triangle, or null
  const GbAlpha* alpha;      // per triangle (glTF draws), or null: every triangle opaque
  int vertex_count;
  int triangle_count;
  const float4* lights;  // [light][2]: position, emission
  int light_count;
  int size;
  float* cubes;
  double inv_triangles;  // 1 / triangle_count
  ShadowSlots slots;     // the scheduled path only
};

// Texel (i, j) of face f: the ray direction R_f v (major component +-1) and
// its length.
__device__ __forceinline__ f3 shadow_texel_dir(int face, int i, int j, float inv_s, float& len) {
  const float pc = fmaf(static_cast<float>(2 * i + 1), inv_s, -1.0f);
  const float pt = fmaf(static_cast<float>(2 * j + 1), inv_s, -1.0f);
  len = sqrt_fast(fmaf(pc, pc, fmaf(pt, pt, 1.0f)));
  return shadow_face_turn(face, cube_face_dir(face, pc, pt));
}

struct ShadowTri {
  f3 n0, n1, n2;  // edge functions: b x c, c x a, a x b
  f3 n;           // plane normal (b - a) x (c - a)
  float det;      // a . n
};

// The shadow shaders' alpha test at texel (i, j) of `face`, e the texel-centre
// ray's edge functions: the base colour alpha (texture x factor) at the ray's
// point on the triangle, the texture coordinate interpolated with the
// barycentrics e / (e0 + e1 + e2) and its derivatives taken as forward
// differences to the rays of texels (i + 1, j) and (i, j + 1) on the
// triangle's plane (the face image's x and y), below the cut-off.
__device__ bool shadow_alpha_discards(const GbAlpha& a, const ShadowTri& tri, float e0, float e1, float e2, int face,
                                      int i, int j, float inv_s) {
  auto bary = [&](f3 d, float l[3]) {
    const float f0 = dot(d, tri.n0), f1 = dot(d, tri.n1), f2 = dot(d, tri.n2);
    const float inv = 1.0f / (f0 + f1 + f2);
    l[0] = f0 * inv, l[1] = f1 * inv, l[2] = f2 * inv;
  };
  float len;
  const float inv = 1.0f / (e0 + e1 + e2);
  const float l[3] = {e0 * inv, e1 * inv, e2 * inv};
  float lx[3], ly[3];
  bary(shadow_texel_dir(face, i + 1, j, inv_s, len), lx);
  bary(shadow_texel_dir(face, i, j + 1, inv_s, len), ly);
  const GbUv c = gb_uv_of(a.uv[0], a.uv[1], a.uv[2], l, lx, ly);
  return gb_sample<true>(a.tex, c).w * a.factor < a.cutoff;
}

// kAlpha: `a` is the triangle's alpha-test record (no test where its texels
// are null).
template <bool kAlpha = false>
__device__ __forceinline__ void shadow_texel(const ShadowTri& tri, int face, int i, int j, int S, float inv_s,
                                             float* __restrict__ cube, const GbAlpha* a = nullptr) {
  float len;
  const f3 d = shadow_texel_dir(face, i, j, inv_s, len);
  const float e0 = dot(d, tri.n0), e1 = dot(d, tri.n1), e2 = dot(d, tri.n2);
  if (e0 > 0.0f || e1 > 0.0f || e2 > 0.0f) return;  // outside (det < 0)
  const float rn = dot(d, tri.n);
  if (!(rn < 0.0f)) return;
  const float t = tri.det * rcp_fast(rn);  // view depth of the hit
  if (t < ALTHEA_SHADOW_NEAR || t > ALTHEA_SHADOW_FAR) return;
  if (kAlpha && a->tex.texels && shadow_alpha_discards(*a, tri, e0, e1, e2, face, i, j, inv_s)) return;
  const float depth = fminf(t * len * (1.0f / ALTHEA_SHADOW_FAR), 1.0f);
  atomicMin(reinterpret_cast<unsigned int*>(cube) + shadow_texel_index(S, face, i, j), __float_as_uint(depth));
}

// Light-relative vertices and edge functions of item = light * T + t (t into
// `tri_index`); live is false for items past the end and for back-facing or
// edge-on triangles.
__device__ __forceinline__ void shadow_item(const ShadowParams& p, int item, int& light, f3 v[3], ShadowTri& tri,
                                            bool& live, int& tri_index) {
  // item / T through the FP64 reciprocal: for item < 2^31 the truncated
  // product is off by at most one, which the remainder corrects.
  light = static_cast<int>(static_cast<double>(item) * p.inv_triangles);
  int t = item - light * p.triangle_count;
  if (t < 0) {
    --light, t += p.triangle_count;
  } else if (t >= p.triangle_count) {
    ++light, t -= p.triangle_count;
  }
  tri_index = t;
  live = item < p.light_count * p.triangle_count;
  tri = ShadowTri{};
  v[0] = v[1] = v[2] = mk(0.0f, 0.0f, 0.0f);
  if (!live) return;
  const float4 lp = p.lights[2 * light];
  for (int k = 0; k < 3; ++k) {
    const float* q;
    if (p.indices) {
      const uint32_t vi = p.indices[3 * t + k];
      q = p.positions + 3 * static_cast<size_t>(min(vi, static_cast<uint32_t>(p.vertex_count - 1)));
    } else {
      q = p.positions + 9 * static_cast<size_t>(t) + 3 * k;
    }
    v[k] = mk(q[0] - lp.x, q[1] - lp.y, q[2] - lp.z);
  }
  // Edge functions from the edges e1 = b - a, e2 = c - a (c x a = e2 x a,
  // a x b = a x e1, b x c = n - both): rounding scales with |a| |e|, not
  // |a| |b|.
  const f3 e1 = v[1] - v[0], e2 = v[2] - v[0];
  tri.n = cross(e1, e2);
  tri.n1 = cross(e2, v[0]);
  tri.n2 = cross(v[0], e1);
  tri.n0 = tri.n - tri.n1 - tri.n2;
  tri.det = dot(v[0], tri.n);
  live = tri.det < 0.0f;  // back faces and edge-on triangles are culled
}

// Texel box [x0, x1] x [y0, y1] of the item on `face` (empty: x1 < x0).
__device__ __forceinline__ void shadow_face_box(int face, const f3 v[3], bool live, int S, int& x0, int& x1, int& y0,
                                                int& y1) {
  const float half = 0.5f * static_cast<float>(S);
  x0 = 0, x1 = -1, y0 = 0, y1 = -1;
  if (live) {
    float sc[3], tc[3], ma[3];
    bool in[3];
    for (int k = 0; k < 3; ++k) {
      shadow_face_coords(face, v[k], sc[k], tc[k], ma[k]);
      in[k] = ma[k] >= kShadowClip;
    }
    if (in[0] || in[1] || in[2]) {
      float lo_x = 1e30f, hi_x = -1e30f, lo_y = 1e30f, hi_y = -1e30f;
      auto add = [&](float s_, float t_, float m_) {
        const float r = rcp_fast(m_);  // 1 ulp: far inside the box margin
        const float x = fmaf(s_ * r, half, half - 0.5f), y = fmaf(t_ * r, half, half - 0.5f);
        lo_x = fminf(lo_x, x), hi_x = fmaxf(hi_x, x), lo_y = fminf(lo_y, y), hi_y = fmaxf(hi_y, y);
      };
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int k2 = k == 2 ? 0 : k + 1;
        if (in[k]) add(sc[k], tc[k], ma[k]);
        if (in[k] != in[k2]) {  // the edge crosses the clip plane
          const float u = (kShadowClip - ma[k]) * rcp_fast(ma[k2] - ma[k]);
          add(fmaf(u, sc[k2] - sc[k], sc[k]), fmaf(u, tc[k2] - tc[k], tc[k]), kShadowClip);
        }
      }
      // Texel centres inside the span widened by kShadowBoxMargin
      // (coordinates clamped to [-2, S + 1] before the integer conversion).
      const float lim = static_cast<float>(S + 1);
      x0 = max(static_cast<int>(ceilf(fminf(fmaxf(lo_x - kShadowBoxMargin, -2.0f), lim))), 0);
      x1 = min(static_cast<int>(floorf(fminf(fmaxf(hi_x + kShadowBoxMargin, -2.0f), lim))), S - 1);
      y0 = max(static_cast<int>(ceilf(fminf(fmaxf(lo_y - kShadowBoxMargin, -2.0f), lim))), 0);
      y1 = min(static_cast<int>(floorf(fminf(fmaxf(hi_y + kShadowBoxMargin, -2.0f), lim))), S - 1);
    }
  }
}

template <bool kAlpha>
__global__ __launch_bounds__(256) void shadow_raster_kernel(const ShadowParams p) {
  const int item = blockIdx.x * blockDim.x + threadIdx.x;
  int light, t;
  f3 v[3];
  ShadowTri tri;
  bool live;
  shadow_item(p, item, light, v, tri, live, t);
  const int S = p.size;
  const float inv_s = 1.0f / static_cast<float>(S);
  const size_t cube_floats = static_cast<size_t>(6) * (S + 2) * (S + 2);
  float* cube = p.cubes + light * cube_floats;
  for (int face = 0; face < 6; ++face) {
    int x0, x1, y0, y1;
    shadow_face_box(face, v, live, S, x0, x1, y0, y1);
    const int bw = x1 - x0 + 1, bh = y1 - y0 + 1;
    const int area = (bw > 0 && bh > 0) ? bw * bh : 0;
    if (area > 0 && area <= kShadowLaneTexels) {
      for (int j = y0; j <= y1; ++j)
        for (int i = x0; i <= x1; ++i) shadow_texel<kAlpha>(tri, face, i, j, S, inv_s, cube, kAlpha ? p.alpha + t : nullptr);
    }
    // Larger boxes: the wave walks them one lane's box at a time.
    uint64_t big = __builtin_amdgcn_ballot_w64(area > kShadowLaneTexels);
    while (big) {
      const int src = __builtin_ctzll(big);
      big &= big - 1;
      ShadowTri w;
      w.n0 = mk(__shfl(tri.n0.x, src), __shfl(tri.n0.y, src), __shfl(tri.n0.z, src));
      w.n1 = mk(__shfl(tri.n1.x, src), __shfl(tri.n1.y, src), __shfl(tri.n1.z, src));
      w.n2 = mk(__shfl(tri.n2.x, src), __shfl(tri.n2.y, src), __shfl(tri.n2.z, src));
      w.n = mk(__shfl(tri.n.x, src), __shfl(tri.n.y, src), __shfl(tri.n.z, src));
      w.det = __shfl(tri.det, src);
      const int wx0 = __shfl(x0, src), wy0 = __shfl(y0, src), ww = __shfl(bw, src), wh = __shfl(bh, src);
      float* wcube = p.cubes + __shfl(light, src) * cube_floats;
      const GbAlpha* wa = kAlpha ? p.alpha + __shfl(t, src) : nullptr;
      const int n = ww * wh;
      for (int k = threadIdx.x & 63; k < n; k += 64) {
        const int j = wy0 + k / ww, i = wx0 + k % ww;
        shadow_texel<kAlpha>(w, face, i, j, S, inv_s, wcube, wa);
      }
    }
  }
}

// The scheduled path: every item's six boxes, their total texels as its run
// of slots, then a grid of waves over equal contiguous slot ranges (the
// G-buffer rasterizer's scheme), so boxes of every size cost the same per
// texel.  A wave finds the items of its 64 slots from a window of 64
// consecutive prefix sums (one coalesced load) by a binary search across the
// lanes, instead of dependent loads per slot; the item's edge functions are
// rebuilt per lane when its item changes.
__global__ __launch_bounds__(kSlotBlock) void shadow_slot_setup_kernel(const ShadowParams p) {
  __shared__ unsigned long long wave_total[kSlotBlock / 64];
  const int item = blockIdx.x * kSlotBlock + threadIdx.x;
  const int items = p.light_count * p.triangle_count;
  int light, t;
  f3 v[3];
  ShadowTri tri;
  bool live;
  shadow_item(p, item, light, v, tri, live, t);
  unsigned long long area = 0;
  unsigned bx[12];
#pragma unroll
  for (int face = 0; face < 6; ++face) {
    int x0, x1, y0, y1;
    shadow_face_box(face, v, live, p.size, x0, x1, y0, y1);
    const int bw = x1 >= x0 ? x1 - x0 + 1 : 0, bh = y1 >= y0 ? y1 - y0 + 1 : 0;
    const int w = bh > 0 ? bw : 0, h = bw > 0 ? bh : 0;
    area += static_cast<unsigned long long>(w) * static_cast<unsigned long long>(h);
    bx[2 * face] = static_cast<unsigned>(max(x0, 0)) | (static_cast<unsigned>(max(y0, 0)) << 16);
    bx[2 * face + 1] = static_cast<unsigned>(w) | (static_cast<unsigned>(h) << 16);
  }
  // Only items with texels are ever looked up: the rest (back faces, boxes
  // off every face) write no boxes.
  if (item < items && area > 0) {
    uint4* q = reinterpret_cast<uint4*>(p.slots.boxes + static_cast<size_t>(item) * 6);
    q[0] = make_uint4(bx[0], bx[1], bx[2], bx[3]);
    q[1] = make_uint4(bx[4], bx[5], bx[6], bx[7]);
    q[2] = make_uint4(bx[8], bx[9], bx[10], bx[11]);
  }
  unsigned long long block_total;
  const unsigned long long excl = slot_block_scan(area, wave_total, block_total);
  if (item < items) p.slots.offsets[item] = excl;
  if (threadIdx.x == 0) p.slots.blocks[blockIdx.x] = block_total;
}

__global__ __launch_bounds__(kSlotScan) void shadow_scan_kernel(unsigned long long* __restrict__ blocks, int n,
                                                               unsigned long long* __restrict__ total) {
  __shared__ unsigned long long sums[kSlotScan / 64];
  slot_scan_blocks(blocks, n, total, sums);
}

__global__ __launch_bounds__(256) void shadow_offset_kernel(unsigned long long* __restrict__ offsets,
                                                            const unsigned long long* __restrict__ blocks, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) offsets[t] += blocks[t / kSlotBlock];
}

template <bool kAlpha>
__global__ __launch_bounds__(256) void shadow_slot_raster_kernel(const ShadowParams p) {
  const unsigned long long* __restrict__ offsets = p.slots.offsets;
  const int n = p.light_count * p.triangle_count;
  const unsigned long long total = offsets[n];
  unsigned long long begin, end;
  slot_range(total, begin, end);
  if (begin >= end) return;
  const int S = p.size;
  const float inv_s = 1.0f / static_cast<float>(S);
  const size_t cube_floats = static_cast<size_t>(6) * (S + 2) * (S + 2);
  int w0 = slot_find(offsets, 0, n, begin);  // the item holding `begin`
  int cur = -1, light = 0, t = 0;
  ShadowTri tri{};
  for (unsigned long long base = begin; base < end; base += 64) {
    unsigned item_start;
    const int item = slot_step_items(offsets, n, base, end, w0, item_start);
    if (item < 0) continue;
    unsigned local = static_cast<unsigned>(base + (threadIdx.x & 63)) - item_start;
    const uint4* bx = reinterpret_cast<const uint4*>(p.slots.boxes + static_cast<size_t>(item) * 6);
    const uint4 q0 = bx[0], q1 = bx[1], q2 = bx[2];
    const uint2 boxes[6] = {make_uint2(q0.x, q0.y), make_uint2(q0.z, q0.w), make_uint2(q1.x, q1.y),
                            make_uint2(q1.z, q1.w), make_uint2(q2.x, q2.y), make_uint2(q2.z, q2.w)};
    int face = 0;
    uint2 b = boxes[0];
    bool found = false;
#pragma unroll
    for (int f = 0; f < 6; ++f) {  // the face whose texels hold `local`
      const unsigned a = (boxes[f].y & 0xffffu) * (boxes[f].y >> 16);
      if (!found && local < a) {
        face = f, b = boxes[f], found = true;
      } else if (!found) {
        local -= a;
      }
    }
    if (!found) continue;  // not reached: the runs are the boxes' areas
    // Row and column of `local` in the box: local < w h <= 2^24 is exact in
    // FP32, so the reciprocal's row is off by at most one, corrected here.
    const unsigned w = b.y & 0xffffu;
    int row = static_cast<int>(static_cast<float>(local) * rcp_fast(static_cast<float>(w)));
    int col = static_cast<int>(local) - row * static_cast<int>(w);
    if (col < 0) {
      --row, col += static_cast<int>(w);
    } else if (col >= static_cast<int>(w)) {
      ++row, col -= static_cast<int>(w);
    }
    const int i = static_cast<int>(b.x & 0xffffu) + col, j = static_cast<int>(b.x >> 16) + row;
    if (item != cur) {
      f3 v[3];
      bool live;
      shadow_item(p, item, light, v, tri, live, t);
      cur = item;
    }
    shadow_texel<kAlpha>(tri, face, i, j, S, inv_s, p.cubes + light * cube_floats, kAlpha ? p.alpha + t : nullptr);
  }
}

// Border texel k of face f of light l: the seamless filter's tap one texel
// beyond the face (cube_fold onto the adjacent face; at a corner the mean of
// the clamped texel and its two folds, summed in that order).
__device__ __forceinline__ float shadow_interior(const float* c, int S, int face, int i, int j) {
  return c[shadow_texel_index(S, face, i, j)];
}

__global__ __launch_bounds__(256) void shadow_border_kernel(float* __restrict__ cubes, int S, int light_count) {
  const int ring = 4 * S + 4;
  const int item = blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= light_count * 6 * ring) return;
  const int k = item % ring, lf = item / ring;
  const int face = lf % 6, light = lf / 6;
  int i, j;
  if (k < S + 2) {
    i = k - 1, j = -1;
  } else if (k < 2 * (S + 2)) {
    i = k - (S + 2) - 1, j = S;
  } else if (k < 2 * (S + 2) + S) {
    i = -1, j = k - 2 * (S + 2);
  } else {
    i = S, j = k - 2 * (S + 2) - S;
  }
  float* c = cubes + static_cast<size_t>(light) * 6 * (S + 2) * (S + 2);
  const bool in_i = i >= 0 && i < S, in_j = j >= 0 && j < S;
  int f, a, b;
  float v;
  if (in_i || in_j) {
    cube_fold(face, i, j, S, f, a, b);
    v = shadow_interior(c, S, f, a, b);
  } else {
    const int ic = min(max(i, 0), S - 1), jc = min(max(j, 0), S - 1);
    float acc = shadow_interior(c, S, face, ic, jc);
    cube_fold(face, i, jc, S, f, a, b);
    acc += shadow_interior(c, S, f, a, b);
    cube_fold(face, ic, j, S, f, a, b);
    acc += shadow_interior(c, S, f, a, b);
    v = acc * (1.0f / 3.0f);
  }
  c[shadow_texel_index(S, face, i, j)] = v;
}

namespace {
size_t align256(size_t v) { return (v + 255) & ~static_cast<size_t>(255); }
size_t slot_blocks(size_t items) { return (items + kSlotBlock - 1) / kSlotBlock; }
}  // namespace

size_t shadow_workspace_bytes(int light_count, int triangle_count) {
  if (light_count < 1 || triangle_count < 0) return 0;
  const size_t items = static_cast<size_t>(light_count) * triangle_count;
  return align256((items + 1) * 8) + align256(items * 6 * sizeof(uint2)) + align256(slot_blocks(items) * 8);
}

// The cubes' clear, the raster (scheduled with a workspace, else the
// lane-walk) and the borders, for the triangles p describes.
static althea_status shadow_passes(ShadowParams& p, void* workspace, hipStream_t s) {
  const int light_count = p.light_count, size = p.size;
  float* cubes = p.cubes;
  const size_t texels = static_cast<size_t>(light_count) * 6 * (size + 2) * (size + 2);
  hipLaunchKernelGGL(fill_kernel<float>, dim3(fill_blocks(texels, sizeof(float))), dim3(256), 0, s, cubes, texels,
                     1.0f);
  althea_status st = check_launch("fill_kernel<float> (shadow clear)");
  if (st != ALTHEA_OK || p.triangle_count == 0) return st;  // borders of a cleared cube are 1 too
  p.inv_triangles = 1.0 / static_cast<double>(p.triangle_count);
  const size_t items = static_cast<size_t>(light_count) * p.triangle_count;
  const bool alpha = p.alpha != nullptr;
  if (workspace) {
    char* ws = static_cast<char*>(workspace);
    p.slots.offsets = reinterpret_cast<unsigned long long*>(ws);
    ws += align256((items + 1) * 8);
    p.slots.boxes = reinterpret_cast<uint2*>(ws);
    ws += align256(items * 6 * sizeof(uint2));
    p.slots.blocks = reinterpret_cast<unsigned long long*>(ws);
    const int nb = static_cast<int>(slot_blocks(items));
    const int ni = static_cast<int>(items);
    hipLaunchKernelGGL(shadow_slot_setup_kernel, dim3(nb), dim3(kSlotBlock), 0, s, p);
    if ((st = check_launch("shadow_slot_setup_kernel")) != ALTHEA_OK) return st;
    hipLaunchKernelGGL(shadow_scan_kernel, dim3(1), dim3(kSlotScan), 0, s, p.slots.blocks, nb, p.slots.offsets + ni);
    if ((st = check_launch("shadow_scan_kernel")) != ALTHEA_OK) return st;
    hipLaunchKernelGGL(shadow_offset_kernel, dim3((ni + 255) / 256), dim3(256), 0, s, p.slots.offsets,
                       p.slots.blocks, ni);
    if ((st = check_launch("shadow_offset_kernel")) != ALTHEA_OK) return st;
    if (alpha)
      hipLaunchKernelGGL(shadow_slot_raster_kernel<true>, dim3(4096), dim3(256), 0, s, p);
    else
      hipLaunchKernelGGL(shadow_slot_raster_kernel<false>, dim3(4096), dim3(256), 0, s, p);
    st = check_launch("shadow_slot_raster_kernel");
  } else {
    const dim3 grid(static_cast<unsigned>((items + 255) / 256));
    if (alpha)
      hipLaunchKernelGGL(shadow_raster_kernel<true>, grid, dim3(256), 0, s, p);
    else
      hipLaunchKernelGGL(shadow_raster_kernel<false>, grid, dim3(256), 0, s, p);
    st = check_launch("shadow_raster_kernel");
  }
  if (st != ALTHEA_OK) return st;
  const size_t border = static_cast<size_t>(light_count) * 6 * (4 * size + 4);
  hipLaunchKernelGGL(shadow_border_kernel, dim3(static_cast<unsigned>((border + 255) / 256)), dim3(256), 0, s, cubes,
                     size, light_count);
  return check_launch("shadow_border_kernel");
}

althea_status launch_shadow_cubes(const althea_mesh_view* mesh, const althea_point_light* lights, int light_count,
                                  int size, float* cubes, hipStream_t s, void* workspace) {
  ShadowParams p{};
  p.positions = mesh->positions;
  p.indices = mesh->indices;
  p.vertex_count = mesh->vertex_count;
  p.triangle_count = mesh->triangle_count;
  p.lights = reinterpret_cast<const float4*>(lights);
  p.light_count = light_count;
  p.size = size;
  p.cubes = cubes;
  return shadow_passes(p, workspace, s);
}

// ------------------------------------------------------- glTF primitives
// A primitive of the shadow pass as the setup reads it: world-space geometry,
// the base colour coordinate set its material reads, and the alpha test's
// factor, cut-off and texture.
struct ShadowPrim {
  const float* positions;
  const float* uv;  // float2 per vertex
  const uint32_t* indices;
  int vertex_count, tri_begin, triangle_count;
  float factor, cutoff;
  althea_texture_view tex;
};

// One lane per triangle of the draw (primitives numbered one after the
// other): its three world-space vertices into the triangle soup the raster
// reads, and its alpha-test record.
__global__ __launch_bounds__(256) void shadow_gltf_setup_kernel(const ShadowPrim* __restrict__ prims, int prim_count,
                                                                int triangles, float* __restrict__ soup,
                                                                GbAlpha* __restrict__ alpha) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= triangles) return;
  int lo = 0, hi = prim_count;  // the last primitive whose first triangle is <= t
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prims[mid].tri_begin <= t)
      lo = mid;
    else
      hi = mid;
  }
  const ShadowPrim& pr = prims[lo];
  const int lt = t - pr.tri_begin;
  const uint32_t last = static_cast<uint32_t>(pr.vertex_count - 1);
  GbAlpha a{};
  const float2* uv = reinterpret_cast<const float2*>(pr.uv);
  for (int k = 0; k < 3; ++k) {
    const uint32_t vi = min(pr.indices[3 * lt + k], last);
    for (int c = 0; c < 3; ++c) soup[9 * static_cast<size_t>(t) + 3 * k + c] = pr.positions[3 * static_cast<size_t>(vi) + c];
    a.uv[k] = uv[vi];
  }
  a.factor = pr.factor;
  a.cutoff = pr.cutoff;
  a.tex = pr.tex;
  alpha[t] = a;
}

static size_t shadow_gltf_tables_bytes(int triangles, int prim_count) {
  return align256(static_cast<size_t>(triangles) * 9 * sizeof(float)) +
         align256(static_cast<size_t>(triangles) * sizeof(GbAlpha)) +
         align256(static_cast<size_t>(prim_count) * sizeof(ShadowPrim));
}

size_t shadow_gltf_workspace_bytes(int light_count, int triangles, int prim_count) {
  return shadow_workspace_bytes(light_count, triangles) + shadow_gltf_tables_bytes(triangles, prim_count);
}

althea_status launch_shadow_cubes_gltf(const althea_gltf_primitive* prims, int prim_count,
                                       const althea_gltf_material* materials, const althea_texture_view* textures,
                                       const althea_point_light* lights, int light_count, int size, float* cubes,
                                       void* workspace, hipStream_t s) {
  std::vector<ShadowPrim> table;
  int total = 0;
  for (int k = 0; k < prim_count; ++k) {
    if (prims[k].triangle_count == 0) continue;
    const althea_gltf_material& m = materials[prims[k].material];
    ShadowPrim q{};
    q.positions = prims[k].positions;
    q.uv = prims[k].uvs[m.base_uv];
    q.indices = prims[k].indices;
    q.vertex_count = prims[k].vertex_count;
    q.tri_begin = total;
    q.triangle_count = prims[k].triangle_count;
    q.factor = m.base_color_factor[3];
    q.cutoff = m.alpha_cutoff;
    q.tex = textures[m.base_texture];
    table.push_back(q);
    total += prims[k].triangle_count;
  }
  char* ws = static_cast<char*>(workspace);
  float* soup = reinterpret_cast<float*>(ws + shadow_workspace_bytes(light_count, total));
  GbAlpha* alpha = reinterpret_cast<GbAlpha*>(reinterpret_cast<char*>(soup) +
                                              align256(static_cast<size_t>(total) * 9 * sizeof(float)));
  ShadowPrim* d_prims = reinterpret_cast<ShadowPrim*>(reinterpret_cast<char*>(alpha) +
                                                      align256(static_cast<size_t>(total) * sizeof(GbAlpha)));
  if (total > 0) {
    // Pageable host table: the copy orders after the stream's earlier work
    // (the draw reads it from host memory, so it is not captured in graphs).
    ALTHEA_TRY_HIP(hipMemcpyAsync(d_prims, table.data(), table.size() * sizeof(ShadowPrim), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(shadow_gltf_setup_kernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, s,
                       d_prims, static_cast<int>(table.size()), total, soup, alpha);
    const althea_status st = check_launch("shadow_gltf_setup_kernel");
    if (st != ALTHEA_OK) return st;
  }
  ShadowParams p{};
  p.positions = soup;
  p.indices = nullptr;
  p.alpha = alpha;
  p.vertex_count = 3 * total;
  p.triangle_count = total;
  p.lights = reinterpret_cast<const float4*>(lights);
  p.light_count = light_count;
  p.size = size;
  p.cubes = cubes;
  return shadow_passes(p, workspace, s);
}

}  // namespace althea
