// G-buffer producer for gfx950: the reference's SceneToGBufferPass
// (Src/DeferredRendering.cpp:268-332) with the pass-through G-buffer shader
// (Shaders/GBuffer/GBufferPassThrough.frag, drawn by Src/Scene/Floor.cpp:14-38),
// as a compute rasterizer.
//
// What the raster pipeline stores per pixel is the nearest front-facing
// triangle whose interior contains the pixel centre, with its view depth in
// [near, far] (Vulkan clips to 0 <= z <= w), earlier triangles winning ties
// (LESS depth test, GraphicsPipeline.h:171-172 back-face culling with
// counter-clockwise front faces), and the vertex attributes interpolated
// perspective-correctly at the pixel centre.  The ray from the eye through the
// pixel centre meets the same triangle at the same point, and the triangle's
// barycentric coordinates of that point are the perspective-correct weights;
// that is what these kernels compute.
//
// Passes (one stream):
//   fill_kernel            the visibility keys set to empty;
//   gbuffer_setup_kernel   one lane per triangle: edge normals, plane, the
//                          pixel box of its part in front of half the near
//                          plane, the box area and its prefix sum within the
//                          block (64-bit pixel slots);
//   gbuffer_scan_kernel    one block: prefix sum of the block totals;
//   gbuffer_offset_kernel  block offset added to every triangle's slot;
//   gbuffer_raster_kernel  a fixed grid of waves, each taking an equal
//                          contiguous run of slots, 64 at a time: the wave
//                          finds the triangles of its 64 slots from a window
//                          of 64 consecutive offsets by a binary search
//                          across the lanes, each lane casts its
//                          pixel's ray and keeps the
//                          nearest hit as (depth bits << 32 | triangle) with a
//                          64-bit atomic min, so big and small triangles cost
//                          the same per pixel and the result does not depend
//                          on the order of the writes;
//   gbuffer_resolve_kernel one lane per pixel (an instance with and one
//                          without the glTF material path): the winning triangle's
//                          barycentrics at the pixel centre and the four
//                          G-buffer planes, from the pass-through inputs or
//                          the glTF material (textures sampled at the mip
//                          level of the texture coordinate's screen
//                          derivatives).
// A draw holds one or more primitives, their triangles numbered one after
// the other; with the glTF subpass a candidate fragment whose base colour
// alpha is below the material's cut-off is discarded before the depth test.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "device_math.hpp"
#include "fill.hpp"
#include "internal.hpp"
#include "gltf_texture.hpp"
#include "slot_scan.hpp"

namespace althea {

constexpr unsigned long long kGbEmpty = ~0ull;
constexpr int kGbScanThreads = kSlotScan;
constexpr int kGbSetupThreads = kSlotBlock;
constexpr int kGbRasterBlocks = 4096;

// Per-triangle setup, 64 bytes.
struct GbTri {
  float n0[3], n1[3], n2[3];  // edge functions: b x c, c x a, a x b (eye-relative a, b, c)
  float n[3];                 // plane normal (b - a) x (c - a)
  float det;                  // a . n (< 0 for a front face)
  int x0, y0, w;              // pixel box
};
static_assert(sizeof(GbTri) == 64, "GbTri is one 64-byte record");

// One drawn primitive.  material < 0: the pass-through inputs (normals may
// be null: all zero); otherwise the glTF inputs and that material.
struct GbPrim {
  const float* positions;
  const float* normals;
  const float* colors;
  const float* mro;
  const float* tangents;
  const float* bitangents;
  const float* uvs[4];
  const uint32_t* indices;
  int vertex_count, tri_begin, triangle_count, material;
};


struct GbParams {
  GbPrim one;                 // the primitive when prims is null
  const GbPrim* prims;        // device table of prim_count primitives, or null
  int prim_count;
  const althea_gltf_material* materials;
  const althea_texture_view* textures;
  int alpha_test;             // some primitive is a glTF one
  GbAlpha* alpha;             // triangle_count records when alpha_test
  int triangle_count;         // over all primitives
  int width, height;
  float eye[3];
  float ray0[3], rayx[3], rayy[3];  // pixel ray = ray0 + rayx (x + 0.5) + rayy (y + 0.5)
  float view_z[3];                  // row 2 of the view rotation: view depth of d = -view_z . d
  float vp[16];                     // projection * view, column-major
  float near_clip;                  // half the near plane: no hit nearer than the near plane is kept
  float near, far;
  GbTri* tris;
  unsigned long long* offsets;  // triangle_count + 1 slot offsets
  unsigned long long* blocks;   // per setup block: its slot total, then its first slot
  unsigned long long* vis;      // width * height keys
  void* position_out;
  void* normal_out;
  void* albedo_out;
  void* mro_out;
  int fp16;
};

__device__ __forceinline__ f3 gb_vertex(const float* __restrict__ a, uint32_t i) {
  return mk(a[3 * i], a[3 * i + 1], a[3 * i + 2]);
}

// Primitive of triangle t (numbered over the draw).  kTable: the draw passes
// a device table (glTF draws always do), so the kernel argument's single
// primitive is never referenced (a reference to it would copy the argument
// to scratch memory).
template <bool kTable = false>
__device__ __forceinline__ const GbPrim& gb_prim(const GbParams& p, int t) {
  if (!kTable && !p.prims) return p.one;
  int lo = 0, hi = p.prim_count;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (p.prims[mid].tri_begin <= t)
      lo = mid;
    else
      hi = mid;
  }
  return p.prims[lo];
}

__device__ __forceinline__ void gb_indices(const GbPrim& pr, int lt, uint32_t vi[3]) {
  const uint32_t last = static_cast<uint32_t>(pr.vertex_count - 1);
  for (int k = 0; k < 3; ++k) vi[k] = min(pr.indices[3 * lt + k], last);
}

// Texture coordinate set k of a primitive, selected without indexing the
// array (a runtime index into it would place it in scratch memory).
__device__ __forceinline__ const float* gb_uv_set(const GbPrim& pr, int k) {
  return k == 0 ? pr.uvs[0] : (k == 1 ? pr.uvs[1] : (k == 2 ? pr.uvs[2] : pr.uvs[3]));
}

__device__ GbAlpha gb_alpha_record(const GbParams& p, int t) {
  GbAlpha a{};
  const GbPrim& pr = gb_prim<true>(p, t);
  if (pr.material < 0) return a;  // pass-through: texels stay null
  const althea_gltf_material& m = p.materials[pr.material];
  const float2* uv = reinterpret_cast<const float2*>(gb_uv_set(pr, m.base_uv));
  uint32_t vi[3];
  gb_indices(pr, t - pr.tri_begin, vi);
  for (int k = 0; k < 3; ++k) a.uv[k] = uv ? uv[vi[k]] : make_float2(0.0f, 0.0f);
  a.factor = m.base_color_factor[3];
  a.cutoff = m.alpha_cutoff;
  a.tex = p.textures[m.base_texture];
  return a;
}

__device__ __forceinline__ f3 gb_ray(const GbParams& p, int x, int y) {
  const float cx = static_cast<float>(x) + 0.5f, cy = static_cast<float>(y) + 0.5f;
  return mk(fmaf(p.rayy[0], cy, fmaf(p.rayx[0], cx, p.ray0[0])), fmaf(p.rayy[1], cy, fmaf(p.rayx[1], cx, p.ray0[1])),
            fmaf(p.rayy[2], cy, fmaf(p.rayx[2], cx, p.ray0[2])));
}

__device__ void gbuffer_setup_triangle(const GbParams& p, int t, GbTri& tri, unsigned long long& area) {
  const GbPrim& pr = gb_prim(p, t);
  const int lt = t - pr.tri_begin;
  const uint32_t last = static_cast<uint32_t>(pr.vertex_count - 1);
  const f3 eye = mk(p.eye[0], p.eye[1], p.eye[2]);
  f3 v[3];
  for (int k = 0; k < 3; ++k) v[k] = gb_vertex(pr.positions, min(pr.indices[3 * lt + k], last)) - eye;
  // Edge functions b x c, c x a, a x b formed from the edges e1 = b - a,
  // e2 = c - a (c x a = e2 x a, a x b = a x e1, b x c = n - both), so their
  // rounding scales with |a| |e| rather than |a| |b|: far, small triangles
  // keep their barycentrics.
  const f3 e1 = v[1] - v[0], e2 = v[2] - v[0];
  const f3 n = cross(e1, e2);
  const f3 n1 = cross(e2, v[0]), n2 = cross(v[0], e1);
  const f3 n0 = n - n1 - n2;
  tri.n0[0] = n0.x, tri.n0[1] = n0.y, tri.n0[2] = n0.z;
  tri.n1[0] = n1.x, tri.n1[1] = n1.y, tri.n1[2] = n1.z;
  tri.n2[0] = n2.x, tri.n2[1] = n2.y, tri.n2[2] = n2.z;
  tri.n[0] = n.x, tri.n[1] = n.y, tri.n[2] = n.z;
  tri.det = dot(v[0], n);
  int x0 = 0, x1 = -1, y0 = 0, y1 = -1;
  if (tri.det < 0.0f) {  // back faces and edge-on triangles are culled
    // The part of the triangle at view depth >= near / 2 (a convex polygon in
    // front of the camera) projects into the box of its projected vertices:
    // the triangle's vertices there and its edges' crossings of that plane.
    const f3 vz = mk(p.view_z[0], p.view_z[1], p.view_z[2]);
    float dep[3];
    bool in[3];
    for (int k = 0; k < 3; ++k) {
      dep[k] = -dot(vz, v[k]);
      in[k] = dep[k] >= p.near_clip;
    }
    if (in[0] || in[1] || in[2]) {
      float lo_x = 1e30f, hi_x = -1e30f, lo_y = 1e30f, hi_y = -1e30f;
      const float W = static_cast<float>(p.width), H = static_cast<float>(p.height);
      auto add = [&](f3 q) {  // eye-relative point -> pixel-centre coordinates
        const f3 w = q + eye;
        const float cx = fmaf(p.vp[8], w.z, fmaf(p.vp[4], w.y, fmaf(p.vp[0], w.x, p.vp[12])));
        const float cy = fmaf(p.vp[9], w.z, fmaf(p.vp[5], w.y, fmaf(p.vp[1], w.x, p.vp[13])));
        const float cw = fmaf(p.vp[11], w.z, fmaf(p.vp[7], w.y, fmaf(p.vp[3], w.x, p.vp[15])));
        const float r = 1.0f / cw;
        const float x = fmaf(fmaf(cx, r, 1.0f), 0.5f * W, -0.5f), y = fmaf(fmaf(cy, r, 1.0f), 0.5f * H, -0.5f);
        lo_x = fminf(lo_x, x), hi_x = fmaxf(hi_x, x), lo_y = fminf(lo_y, y), hi_y = fmaxf(hi_y, y);
      };
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int k2 = k == 2 ? 0 : k + 1;
        if (in[k]) add(v[k]);
        if (in[k] != in[k2]) {
          const float u = (p.near_clip - dep[k]) / (dep[k2] - dep[k]);
          add(mk(fmaf(u, v[k2].x - v[k].x, v[k].x), fmaf(u, v[k2].y - v[k].y, v[k].y),
                 fmaf(u, v[k2].z - v[k].z, v[k].z)));
        }
      }
      // Pixel centres inside the span, one pixel of margin for rounding
      // (coordinates clamped before the integer conversion).
      const float limx = W + 1.0f, limy = H + 1.0f;
      x0 = max(static_cast<int>(ceilf(fminf(fmaxf(lo_x, -2.0f), limx))) - 1, 0);
      x1 = min(static_cast<int>(floorf(fminf(fmaxf(hi_x, -2.0f), limx))) + 1, p.width - 1);
      y0 = max(static_cast<int>(ceilf(fminf(fmaxf(lo_y, -2.0f), limy))) - 1, 0);
      y1 = min(static_cast<int>(floorf(fminf(fmaxf(hi_y, -2.0f), limy))) + 1, p.height - 1);
    }
  }
  const int bw = x1 - x0 + 1, bh = y1 - y0 + 1;
  tri.x0 = x0;
  tri.y0 = y0;
  tri.w = bw > 0 ? bw : 1;
  area = (bw > 0 && bh > 0) ? static_cast<unsigned long long>(bw) * static_cast<unsigned long long>(bh) : 0ull;
}

// Setup of one triangle per lane, and the exclusive prefix sum of the box
// areas within the block (wave shuffles, then the waves' totals).
__global__ __launch_bounds__(kGbSetupThreads) void gbuffer_setup_kernel(const GbParams p) {
  __shared__ unsigned long long wave_total[kGbSetupThreads / 64];
  const int t = blockIdx.x * kGbSetupThreads + threadIdx.x;
  unsigned long long area = 0;
  if (t < p.triangle_count) {
    GbTri tri;
    gbuffer_setup_triangle(p, t, tri, area);
    p.tris[t] = tri;
    if (p.alpha) p.alpha[t] = gb_alpha_record(p, t);
  }
  unsigned long long block_total;
  const unsigned long long excl = slot_block_scan(area, wave_total, block_total);
  if (t < p.triangle_count) p.offsets[t] = excl;
  if (threadIdx.x == 0) p.blocks[blockIdx.x] = block_total;
}

// Exclusive prefix sum of the n block totals in place, in one block (each
// thread a contiguous run), and the total slot count into *total.
__global__ __launch_bounds__(kGbScanThreads) void gbuffer_scan_kernel(unsigned long long* __restrict__ blocks, int n,
                                                                      unsigned long long* __restrict__ total) {
  __shared__ unsigned long long sums[kGbScanThreads / 64];
  slot_scan_blocks(blocks, n, total, sums);
}

__global__ __launch_bounds__(256) void gbuffer_offset_kernel(unsigned long long* __restrict__ offsets,
                                                             const unsigned long long* __restrict__ blocks, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) offsets[t] += blocks[t / kGbSetupThreads];
}

// Ray of pixel (x, y) against triangle `tri`: the hit's ray parameter, or a
// negative value when the pixel centre is outside, the triangle is behind the
// eye or the hit lies outside [near, far].  e receives the edge functions.
__device__ __forceinline__ float gb_hit(const GbParams& p, const GbTri& tri, f3 d, float e[3]) {
  e[0] = fmaf(d.z, tri.n0[2], fmaf(d.y, tri.n0[1], d.x * tri.n0[0]));
  e[1] = fmaf(d.z, tri.n1[2], fmaf(d.y, tri.n1[1], d.x * tri.n1[0]));
  e[2] = fmaf(d.z, tri.n2[2], fmaf(d.y, tri.n2[1], d.x * tri.n2[0]));
  if (e[0] > 0.0f || e[1] > 0.0f || e[2] > 0.0f) return -1.0f;
  const float rn = fmaf(d.z, tri.n[2], fmaf(d.y, tri.n[1], d.x * tri.n[0]));
  if (!(rn < 0.0f)) return -1.0f;
  const float t = tri.det / rn;
  const float depth = -t * fmaf(p.view_z[2], d.z, fmaf(p.view_z[1], d.y, p.view_z[0] * d.x));
  if (!(depth >= p.near && depth <= p.far)) return -1.0f;
  return t;
}

// sRGB decode of a filtered value (the mip blit): the transfer function.
__device__ __forceinline__ float gb_srgb_to_linear(float c) {
  return c <= 0.04045f ? c * (1.0f / 12.92f) : powf((c + 0.055f) * (1.0f / 1.055f), 2.4f);
}

// Barycentrics of ray d on the triangle's plane (outside the triangle too).
__device__ __forceinline__ void gb_bary(const GbTri& tri, f3 d, float l[3]) {
  const float e0 = fmaf(d.z, tri.n0[2], fmaf(d.y, tri.n0[1], d.x * tri.n0[0]));
  const float e1 = fmaf(d.z, tri.n1[2], fmaf(d.y, tri.n1[1], d.x * tri.n1[0]));
  const float e2 = fmaf(d.z, tri.n2[2], fmaf(d.y, tri.n2[1], d.x * tri.n2[0]));
  const float inv = 1.0f / (e0 + e1 + e2);
  l[0] = e0 * inv, l[1] = e1 * inv, l[2] = e2 * inv;
}


__device__ __forceinline__ GbUv gb_uv(const float* __restrict__ uv, const uint32_t vi[3], const float l[3],
                                      const float lx[3], const float ly[3]) {
  if (!uv) return GbUv{};
  const float2* q = reinterpret_cast<const float2*>(uv);
  return gb_uv_of(q[vi[0]], q[vi[1]], q[vi[2]], l, lx, ly);
}

// Base colour alpha of a glTF fragment (the alpha test of Gltf.frag:40-42).
__device__ float gb_base_alpha(const GbParams& p, const GbAlpha& a, const GbTri& tri, int x, int y, f3 d) {
  float l[3], lx[3], ly[3];
  gb_bary(tri, d, l);
  gb_bary(tri, gb_ray(p, x + 1, y), lx);
  gb_bary(tri, gb_ray(p, x, y + 1), ly);
  const GbUv c = gb_uv_of(a.uv[0], a.uv[1], a.uv[2], l, lx, ly);
  return gb_sample<true>(a.tex, c).w * a.factor;
}

__global__ __launch_bounds__(256) void gbuffer_raster_kernel(const GbParams p) {
  const unsigned long long* __restrict__ offsets = p.offsets;
  const int T = p.triangle_count;
  const unsigned long long total = offsets[T];
  const int lane = threadIdx.x & 63;
  // Each wave takes an equal contiguous run of slots, 64 per iteration.
  unsigned long long begin, end;
  slot_range(total, begin, end);
  if (begin >= end) return;
  int w0 = slot_find(offsets, 0, T, begin);  // the triangle holding `begin`
  for (unsigned long long base = begin; base < end; base += 64) {
    unsigned start;
    const int lo = slot_step_items(offsets, T, base, end, w0, start);
    if (lo < 0) continue;
    const GbTri tri = p.tris[lo];
    const unsigned local = static_cast<unsigned>(base + lane) - start;
    const int x = tri.x0 + static_cast<int>(local % static_cast<unsigned>(tri.w));
    const int y = tri.y0 + static_cast<int>(local / static_cast<unsigned>(tri.w));
    float e[3];
    const f3 d = gb_ray(p, x, y);
    const float t = gb_hit(p, tri, d, e);
    if (t < 0.0f) continue;
    if (p.alpha_test) {
      const GbAlpha& a = p.alpha[lo];
      if (a.tex.texels && gb_base_alpha(p, a, tri, x, y, d) < a.cutoff) continue;  // discarded: no depth written
    }
    const unsigned long long key =
        (static_cast<unsigned long long>(__float_as_uint(t)) << 32) | static_cast<unsigned>(lo);
    atomicMin(p.vis + static_cast<size_t>(y) * p.width + x, key);
  }
}

// Planes are written once and read by a later pass: streaming (non-temporal)
// vector stores.
typedef float gb_f4 __attribute__((ext_vector_type(4)));
typedef unsigned gb_u2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void gb_store(void* plane, int fp16, size_t i, float x, float y, float z, float w) {
  if (fp16) {
    const unsigned lo = __half_as_ushort(__float2half_rn(x)) |
                        (static_cast<unsigned>(__half_as_ushort(__float2half_rn(y))) << 16);
    const unsigned hi = __half_as_ushort(__float2half_rn(z)) |
                        (static_cast<unsigned>(__half_as_ushort(__float2half_rn(w))) << 16);
    const gb_u2 v = {lo, hi};
    __builtin_nontemporal_store(v, static_cast<gb_u2*>(plane) + i);
  } else {
    const gb_f4 v = {x, y, z, w};
    __builtin_nontemporal_store(v, static_cast<gb_f4*>(plane) + i);
  }
}

__device__ __forceinline__ f3 gb_interp(const float* __restrict__ a, const uint32_t vi[3], const float l[3]) {
  const f3 p0 = gb_vertex(a, vi[0]), p1 = gb_vertex(a, vi[1]), p2 = gb_vertex(a, vi[2]);
  return mk(fmaf(l[2], p2.x, fmaf(l[1], p1.x, l[0] * p0.x)), fmaf(l[2], p2.y, fmaf(l[1], p1.y, l[0] * p0.y)),
            fmaf(l[2], p2.z, fmaf(l[1], p1.z, l[0] * p0.z)));
}

// GBufferPassThrough.frag: normal (the interpolated normal, or where it is
// exactly zero normalize(cross(dFdx(pos), dFdy(pos))), 1), albedo (color, 1),
// metallic-roughness-occlusion (mro, 1); the position plane holds (pos, 1).
// Pixels no triangle covers keep the clear colour (0, 0, 0, 1) of the
// reference's attachments (DeferredRendering.cpp:274-275), with position
// (0, 0, 0, 0): position.w is this G-buffer's coverage flag.
// kGltf: the draw has glTF primitives (their material path is compiled in);
// the pass-through instance keeps its registers for the plain attributes.
template <bool kGltf>
__global__ __launch_bounds__(256) void gbuffer_resolve_kernel(const GbParams p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.width * p.height) return;
  const unsigned long long key = p.vis[i];
  if (key == kGbEmpty) {
    gb_store(p.position_out, p.fp16, i, 0.0f, 0.0f, 0.0f, 0.0f);
    gb_store(p.normal_out, p.fp16, i, 0.0f, 0.0f, 0.0f, 1.0f);
    gb_store(p.albedo_out, p.fp16, i, 0.0f, 0.0f, 0.0f, 1.0f);
    gb_store(p.mro_out, p.fp16, i, 0.0f, 0.0f, 0.0f, 1.0f);
    return;
  }
  const int t = static_cast<int>(key & 0xffffffffu);
  const int x = i % p.width, y = i / p.width;
  const GbTri tri = p.tris[t];
  const GbPrim& pr = gb_prim<kGltf>(p, t);
  const int lt = t - pr.tri_begin;
  const f3 d = gb_ray(p, x, y);
  float e[3];
  gb_hit(p, tri, d, e);
  const float inv = 1.0f / (e[0] + e[1] + e[2]);
  const float l[3] = {e[0] * inv, e[1] * inv, e[2] * inv};
  uint32_t vi[3];
  gb_indices(pr, lt, vi);
  const f3 pos = gb_interp(pr.positions, vi, l);
  gb_store(p.position_out, p.fp16, i, pos.x, pos.y, pos.z, 1.0f);
  if (kGltf && pr.material >= 0) {
    // Gltf.frag with fetchMaterial (InstanceData.glsl:27-63).
    const althea_gltf_material& m = p.materials[pr.material];
    float lx[3], ly[3];
    gb_bary(tri, gb_ray(p, x + 1, y), lx);
    gb_bary(tri, gb_ray(p, x, y + 1), ly);
    const GbUv cb = gb_uv(gb_uv_set(pr, m.base_uv), vi, l, lx, ly);
    const GbUv cm = gb_uv(gb_uv_set(pr, m.metallic_roughness_uv), vi, l, lx, ly);
    const float4 bt = gb_sample(p.textures[m.base_texture], cb);
    const float4 base = make_float4(bt.x * m.base_color_factor[0], bt.y * m.base_color_factor[1],
                                    bt.z * m.base_color_factor[2], bt.w * m.base_color_factor[3]);
    const float4 nt = gb_sample(p.textures[m.normal_texture], cb);  // the base colour's coordinates
    const float tx = fmaf(2.0f, nt.x, -1.0f) * m.normal_scale, ty = fmaf(2.0f, nt.y, -1.0f) * m.normal_scale;
    const float tz = fmaf(2.0f, nt.z, -1.0f);
    const f3 T = gb_interp(pr.tangents, vi, l), B = gb_interp(pr.bitangents, vi, l), N = gb_interp(pr.normals, vi, l);
    f3 n = mk(fmaf(tz, N.x, fmaf(ty, B.x, tx * T.x)), fmaf(tz, N.y, fmaf(ty, B.y, tx * T.y)),
              fmaf(tz, N.z, fmaf(ty, B.z, tx * T.z)));
    n = n * rsqrtf(dot(n, n));
    const float4 mr = gb_sample(p.textures[m.metallic_roughness_texture], cm);
    gb_store(p.normal_out, p.fp16, i, n.x, n.y, n.z, base.w);
    gb_store(p.albedo_out, p.fp16, i, base.x, base.y, base.z, base.w);
    gb_store(p.mro_out, p.fp16, i, mr.z * m.metallic_factor, mr.y * m.roughness_factor, 0.0f, base.w);
    return;
  }
  f3 nrm = pr.normals ? gb_interp(pr.normals, vi, l) : mk(0.0f, 0.0f, 0.0f);
  if (nrm.x == 0.0f && nrm.y == 0.0f && nrm.z == 0.0f) {
    // Screen derivatives of the position on the triangle's plane: the hits
    // of the next pixel's ray to the right (x + 1) and below (y + 1, the
    // framebuffer's y grows downwards) minus this one's.
    const f3 n = mk(tri.n[0], tri.n[1], tri.n[2]);
    const f3 dx = gb_ray(p, x + 1, y), dy = gb_ray(p, x, y + 1);
    const float t0 = tri.det / dot(d, n), tx = tri.det / dot(dx, n), ty = tri.det / dot(dy, n);
    const f3 px = dx * tx - d * t0, py = dy * ty - d * t0;
    const f3 c = cross(px, py);
    nrm = c * rsqrtf(dot(c, c));
  }
  const f3 alb = gb_interp(pr.colors, vi, l);
  const f3 m = gb_interp(pr.mro, vi, l);
  gb_store(p.normal_out, p.fp16, i, nrm.x, nrm.y, nrm.z, 1.0f);
  gb_store(p.albedo_out, p.fp16, i, alb.x, alb.y, alb.z, 1.0f);
  gb_store(p.mro_out, p.fp16, i, m.x, m.y, m.z, 1.0f);
}

namespace {
size_t align256(size_t v) { return (v + 255) & ~static_cast<size_t>(255); }
}  // namespace

size_t gb_setup_blocks(int triangle_count) { return (static_cast<size_t>(triangle_count) + kGbSetupThreads - 1) / kGbSetupThreads; }

size_t gbuffer_workspace_bytes(int width, int height, int triangle_count) {
  if (width < 1 || height < 1 || triangle_count < 0) return 0;
  const size_t t = static_cast<size_t>(triangle_count);
  return align256(static_cast<size_t>(width) * height * 8) + align256(t * sizeof(GbTri)) + align256((t + 1) * 8) +
         align256(gb_setup_blocks(triangle_count) * 8);
}

namespace {

// Camera, planes and the workspace's fixed parts (keys, setup, offsets,
// block totals); returns the first byte after them.
char* gb_common(GbParams& p, const althea_globals* g, const althea_gbuffer_view* out, void* workspace,
                int triangle_count) {
  p.triangle_count = triangle_count;
  p.width = out->width;
  p.height = out->height;
  for (int c = 0; c < 3; ++c) {
    p.eye[c] = g->inverse_view[12 + c];
    p.view_z[c] = g->view[c * 4 + 2];
  }
  view_ray_affine(*g, out->width, out->height, p.ray0, p.rayx, p.rayy);
  view_projection(*g, p.vp);
  // Near and far of a [0, 1]-depth perspective projection: P[3][2] / P[2][2]
  // and P[3][2] / (P[2][2] + 1).
  const double p10 = g->projection[10], p14 = g->projection[14];
  p.near = static_cast<float>(p14 / p10);
  p.far = static_cast<float>(p14 / (p10 + 1.0));
  p.near_clip = 0.5f * p.near;
  char* ws = static_cast<char*>(workspace);
  const size_t npx = static_cast<size_t>(out->width) * out->height;
  const size_t t = static_cast<size_t>(triangle_count);
  p.vis = reinterpret_cast<unsigned long long*>(ws);
  ws += align256(npx * 8);
  p.tris = reinterpret_cast<GbTri*>(ws);
  ws += align256(t * sizeof(GbTri));
  p.offsets = reinterpret_cast<unsigned long long*>(ws);
  ws += align256((t + 1) * 8);
  p.blocks = reinterpret_cast<unsigned long long*>(ws);
  ws += align256(gb_setup_blocks(triangle_count) * 8);
  p.position_out = const_cast<void*>(out->position);
  p.normal_out = const_cast<void*>(out->normal);
  p.albedo_out = const_cast<void*>(out->albedo);
  p.mro_out = const_cast<void*>(out->mro);
  p.fp16 = out->format == ALTHEA_GBUFFER_FP16 ? 1 : 0;
  return ws;
}

althea_status gb_run(const GbParams& p, hipStream_t s) {
  const int n = p.width * p.height;
  const int T = p.triangle_count;
  hipLaunchKernelGGL(fill_kernel<unsigned long long>, dim3(fill_blocks(static_cast<size_t>(n), 8)), dim3(256), 0, s,
                     p.vis, static_cast<size_t>(n), kGbEmpty);
  althea_status st = check_launch("fill_kernel<unsigned long long> (visibility clear)");
  if (st != ALTHEA_OK) return st;
  if (T > 0) {
    const int nb = static_cast<int>(gb_setup_blocks(T));
    hipLaunchKernelGGL(gbuffer_setup_kernel, dim3(nb), dim3(kGbSetupThreads), 0, s, p);
    if ((st = check_launch("gbuffer_setup_kernel")) != ALTHEA_OK) return st;
    hipLaunchKernelGGL(gbuffer_scan_kernel, dim3(1), dim3(kGbScanThreads), 0, s, p.blocks, nb, p.offsets + T);
    if ((st = check_launch("gbuffer_scan_kernel")) != ALTHEA_OK) return st;
    hipLaunchKernelGGL(gbuffer_offset_kernel, dim3((T + 255) / 256), dim3(256), 0, s, p.offsets, p.blocks, T);
    if ((st = check_launch("gbuffer_offset_kernel")) != ALTHEA_OK) return st;
    hipLaunchKernelGGL(gbuffer_raster_kernel, dim3(kGbRasterBlocks), dim3(256), 0, s, p);
    if ((st = check_launch("gbuffer_raster_kernel")) != ALTHEA_OK) return st;
  }
  if (p.alpha_test)
    hipLaunchKernelGGL(gbuffer_resolve_kernel<true>, dim3((n + 255) / 256), dim3(256), 0, s, p);
  else
    hipLaunchKernelGGL(gbuffer_resolve_kernel<false>, dim3((n + 255) / 256), dim3(256), 0, s, p);
  return check_launch("gbuffer_resolve_kernel");
}

}  // namespace

althea_status launch_gbuffer_draw(const althea_gbuffer_mesh* mesh, const althea_globals* g,
                                  const althea_gbuffer_view* out, void* workspace, hipStream_t s) {
  GbParams p{};
  p.one.positions = mesh->positions;
  p.one.normals = mesh->normals;
  p.one.colors = mesh->colors;
  p.one.mro = mesh->mro;
  p.one.indices = mesh->indices;
  p.one.vertex_count = mesh->vertex_count;
  p.one.triangle_count = mesh->triangle_count;
  p.one.material = -1;
  gb_common(p, g, out, workspace, mesh->triangle_count);
  return gb_run(p, s);
}

size_t gbuffer_gltf_workspace_bytes(int width, int height, int triangle_count, int primitive_count,
                                    int material_count, int texture_count) {
  const size_t base = gbuffer_workspace_bytes(width, height, triangle_count);
  if (base == 0) return 0;
  return base + align256(static_cast<size_t>(primitive_count) * sizeof(GbPrim)) +
         align256(static_cast<size_t>(material_count) * sizeof(althea_gltf_material)) +
         align256(static_cast<size_t>(texture_count) * sizeof(althea_texture_view)) +
         align256(static_cast<size_t>(triangle_count) * sizeof(GbAlpha));
}

althea_status launch_gbuffer_draw_gltf(const althea_gltf_primitive* prims, int prim_count,
                                       const althea_gltf_material* materials, int material_count,
                                       const althea_texture_view* textures, int texture_count,
                                       const althea_globals* g, const althea_gbuffer_view* out, void* workspace,
                                       hipStream_t s) {
  std::vector<GbPrim> table(prim_count);
  int total = 0;
  for (int k = 0; k < prim_count; ++k) {
    GbPrim& q = table[k];
    q = GbPrim{};
    q.positions = prims[k].positions;
    q.normals = prims[k].normals;
    q.tangents = prims[k].tangents;
    q.bitangents = prims[k].bitangents;
    for (int u = 0; u < 4; ++u) q.uvs[u] = prims[k].uvs[u];
    q.indices = prims[k].indices;
    q.vertex_count = prims[k].vertex_count;
    q.tri_begin = total;
    q.triangle_count = prims[k].triangle_count;
    q.material = prims[k].material;
    total += prims[k].triangle_count;
  }
  GbParams p{};
  char* ws = gb_common(p, g, out, workspace, total);
  GbPrim* d_prims = reinterpret_cast<GbPrim*>(ws);
  ws += align256(static_cast<size_t>(prim_count) * sizeof(GbPrim));
  althea_gltf_material* d_mat = reinterpret_cast<althea_gltf_material*>(ws);
  ws += align256(static_cast<size_t>(material_count) * sizeof(althea_gltf_material));
  althea_texture_view* d_tex = reinterpret_cast<althea_texture_view*>(ws);
  ws += align256(static_cast<size_t>(texture_count) * sizeof(althea_texture_view));
  p.alpha = reinterpret_cast<GbAlpha*>(ws);
  // Pageable host tables: the copies order after the stream's earlier work
  // (this draw reads its tables from host memory, so it is not captured in
  // graphs).
  ALTHEA_TRY_HIP(hipMemcpyAsync(d_prims, table.data(), table.size() * sizeof(GbPrim), hipMemcpyHostToDevice, s));
  ALTHEA_TRY_HIP(hipMemcpyAsync(d_mat, materials, static_cast<size_t>(material_count) * sizeof(althea_gltf_material),
                                hipMemcpyHostToDevice, s));
  ALTHEA_TRY_HIP(hipMemcpyAsync(d_tex, textures, static_cast<size_t>(texture_count) * sizeof(althea_texture_view),
                                hipMemcpyHostToDevice, s));
  p.prims = d_prims;
  p.prim_count = prim_count;
  p.materials = d_mat;
  p.textures = d_tex;
  p.alpha_test = 1;
  return gb_run(p, s);
}

// Image::generateMipMaps (Src/Image.cpp:135-230): level m from level m - 1 by
// a linear blit, i.e. the bilinear value (clamped to the edge) at each
// destination texel centre mapped into the source, in linear space for sRGB
// (alpha stays linear), rounded to 8 bits.
__device__ __forceinline__ float gb_linear_to_srgb(float l) {
  return l <= 0.0031308f ? 12.92f * l : fmaf(1.055f, powf(l, 1.0f / 2.4f), -0.055f);
}

__device__ __forceinline__ unsigned gb_unorm8(float x) {
  return static_cast<unsigned>(fminf(fmaxf(floorf(fmaf(x, 255.0f, 0.5f)), 0.0f), 255.0f));
}

__global__ __launch_bounds__(256) void texture_mip_kernel(uint8_t* __restrict__ texels, size_t src_off, int sw,
                                                          int sh, size_t dst_off, int dw, int dh, int srgb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dw * dh) return;
  const int x = i % dw, y = i / dw;
  const float u = (static_cast<float>(x) + 0.5f) * (static_cast<float>(sw) / static_cast<float>(dw)) - 0.5f;
  const float v = (static_cast<float>(y) + 0.5f) * (static_cast<float>(sh) / static_cast<float>(dh)) - 0.5f;
  const float fx = floorf(u), fy = floorf(v);
  const float ax = u - fx, ay = v - fy;
  const int i0 = min(max(static_cast<int>(fx), 0), sw - 1), i1 = min(max(static_cast<int>(fx) + 1, 0), sw - 1);
  const int j0 = min(max(static_cast<int>(fy), 0), sh - 1), j1 = min(max(static_cast<int>(fy) + 1, 0), sh - 1);
  const uchar4* src = reinterpret_cast<const uchar4*>(texels) + src_off;
  const uchar4 c[4] = {src[j0 * sw + i0], src[j0 * sw + i1], src[j1 * sw + i0], src[j1 * sw + i1]};
  float r[4][4];
  for (int k = 0; k < 4; ++k) {
    r[k][0] = c[k].x * (1.0f / 255.0f);
    r[k][1] = c[k].y * (1.0f / 255.0f);
    r[k][2] = c[k].z * (1.0f / 255.0f);
    r[k][3] = c[k].w * (1.0f / 255.0f);
    if (srgb)
      for (int ch = 0; ch < 3; ++ch) r[k][ch] = gb_srgb_to_linear(r[k][ch]);
  }
  unsigned q[4];
  for (int ch = 0; ch < 4; ++ch) {
    const float top = fmaf(r[1][ch] - r[0][ch], ax, r[0][ch]), bot = fmaf(r[3][ch] - r[2][ch], ax, r[2][ch]);
    float val = fmaf(bot - top, ay, top);
    if (srgb && ch < 3) val = gb_linear_to_srgb(val);
    q[ch] = gb_unorm8(val);
  }
  reinterpret_cast<uchar4*>(texels)[dst_off + i] =
      make_uchar4(static_cast<unsigned char>(q[0]), static_cast<unsigned char>(q[1]), static_cast<unsigned char>(q[2]),
                  static_cast<unsigned char>(q[3]));
}

althea_status launch_texture_mips(uint8_t* texels, int w, int h, int mips, int srgb, hipStream_t s) {
  size_t off = 0;
  for (int m = 1; m < mips; ++m) {
    const int sw = std::max(1, w >> (m - 1)), sh = std::max(1, h >> (m - 1));
    const int dw = std::max(1, w >> m), dh = std::max(1, h >> m);
    const size_t dst = off + static_cast<size_t>(sw) * sh;
    hipLaunchKernelGGL(texture_mip_kernel, dim3((dw * dh + 255) / 256), dim3(256), 0, s, texels, off, sw, sh, dst, dw,
                       dh, srgb);
    const althea_status st = check_launch("texture_mip_kernel");
    if (st != ALTHEA_OK) return st;
    off = dst;
  }
  return ALTHEA_OK;
}

}  // namespace althea
