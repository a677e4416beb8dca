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
//   gbuffer_setup_kernel   one lane per triangle: edge normals, plane, the
//                          pixel box of its part in front of half the near
//                          plane, the box area and its prefix sum within the
//                          block (64-bit pixel slots);
//   gbuffer_scan_kernel    one block: prefix sum of the block totals;
//   gbuffer_offset_kernel  block offset added to every triangle's slot;
//   gbuffer_raster_kernel  a fixed grid of waves, each taking an equal
//                          contiguous run of slots, 64 at a time: the wave
//                          follows the triangles of its first and last slot
//                          forward with wave-uniform galloping searches, each
//                          lane finds its own within that range, casts its
//                          pixel's ray and keeps the
//                          nearest hit as (depth bits << 32 | triangle) with a
//                          64-bit atomic min, so big and small triangles cost
//                          the same per pixel and the result does not depend
//                          on the order of the writes;
//   gbuffer_resolve_kernel one lane per pixel: the winning triangle's
//                          barycentrics at the pixel centre and the four
//                          G-buffer planes.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cmath>

#include "device_math.hpp"
#include "internal.hpp"

namespace althea {

constexpr unsigned long long kGbEmpty = ~0ull;
constexpr int kGbScanThreads = 1024;
constexpr int kGbSetupThreads = 256;
constexpr int kGbRasterBlocks = 4096;

// Per-triangle setup, 64 bytes.
struct GbTri {
  float n0[3], n1[3], n2[3];  // edge functions: b x c, c x a, a x b (eye-relative a, b, c)
  float n[3];                 // plane normal (b - a) x (c - a)
  float det;                  // a . n (< 0 for a front face)
  int x0, y0, w;              // pixel box
};
static_assert(sizeof(GbTri) == 64, "GbTri is one 64-byte record");

struct GbParams {
  const float* positions;
  const float* normals;  // may be null (zero normals)
  const float* colors;
  const float* mro;
  const uint32_t* indices;
  int vertex_count, triangle_count;
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

__device__ __forceinline__ f3 gb_ray(const GbParams& p, int x, int y) {
  const float cx = static_cast<float>(x) + 0.5f, cy = static_cast<float>(y) + 0.5f;
  return mk(fmaf(p.rayy[0], cy, fmaf(p.rayx[0], cx, p.ray0[0])), fmaf(p.rayy[1], cy, fmaf(p.rayx[1], cx, p.ray0[1])),
            fmaf(p.rayy[2], cy, fmaf(p.rayx[2], cx, p.ray0[2])));
}

__device__ void gbuffer_setup_triangle(const GbParams& p, int t, GbTri& tri, unsigned long long& area) {
  const uint32_t last = static_cast<uint32_t>(p.vertex_count - 1);
  const f3 eye = mk(p.eye[0], p.eye[1], p.eye[2]);
  f3 v[3];
  for (int k = 0; k < 3; ++k) v[k] = gb_vertex(p.positions, min(p.indices[3 * t + k], last)) - eye;
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
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned long long x = area;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned long long y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) wave_total[wave] = x;
  __syncthreads();
  unsigned long long before = 0;
  for (int k = 0; k < wave; ++k) before += wave_total[k];
  if (t < p.triangle_count) p.offsets[t] = before + x - area;
  if (threadIdx.x == kGbSetupThreads - 1) p.blocks[blockIdx.x] = before + x;
}

// Exclusive prefix sum of the n block totals in place, in one block (each
// thread a contiguous run), and the total slot count into *total.
__global__ __launch_bounds__(kGbScanThreads) void gbuffer_scan_kernel(unsigned long long* __restrict__ blocks, int n,
                                                                      unsigned long long* __restrict__ total) {
  __shared__ unsigned long long sums[kGbScanThreads];
  const int tid = threadIdx.x;
  const int run = (n + kGbScanThreads - 1) / kGbScanThreads;
  const int b = min(tid * run, n), e = min(b + run, n);
  unsigned long long s = 0;
  for (int i = b; i < e; ++i) s += blocks[i];
  sums[tid] = s;
  __syncthreads();
  for (int d = 1; d < kGbScanThreads; d <<= 1) {  // inclusive Hillis-Steele scan
    const unsigned long long add = tid >= d ? sums[tid - d] : 0ull;
    __syncthreads();
    sums[tid] += add;
    __syncthreads();
  }
  unsigned long long acc = sums[tid] - s;
  for (int i = b; i < e; ++i) {
    const unsigned long long v = blocks[i];
    blocks[i] = acc;
    acc += v;
  }
  if (tid == kGbScanThreads - 1) *total = sums[tid];
}

__global__ __launch_bounds__(256) void gbuffer_offset_kernel(unsigned long long* __restrict__ offsets,
                                                             const unsigned long long* __restrict__ blocks, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) offsets[t] += blocks[t / kGbSetupThreads];
}

__global__ void gbuffer_clear_kernel(unsigned long long* __restrict__ vis, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) vis[i] = kGbEmpty;
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

// Largest t in [lo, hi) with offsets[t] <= slot, given offsets[lo] <= slot <
// offsets[hi].
__device__ __forceinline__ int gb_find(const unsigned long long* __restrict__ offsets, int lo, int hi,
                                       unsigned long long slot) {
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (offsets[mid] <= slot)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

__device__ __forceinline__ unsigned long long gb_uniform(unsigned long long v) {
  const unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(v));
  const unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(v >> 32));
  return (static_cast<unsigned long long>(hi) << 32) | lo;
}

// Largest t >= lo with offsets[t] <= slot (offsets[lo] <= slot < offsets[T]):
// probes lo + 1, lo + 3, lo + 7, ... then a binary search in the last step.
__device__ __forceinline__ int gb_gallop(const unsigned long long* __restrict__ offsets, int lo, int T,
                                         unsigned long long slot) {
  int step = 1;
  while (lo + step < T && offsets[lo + step] <= slot) {
    lo += step;
    step <<= 1;
  }
  return gb_find(offsets, lo, min(lo + step, T), slot);
}

__global__ __launch_bounds__(256) void gbuffer_raster_kernel(const GbParams p) {
  const unsigned long long* __restrict__ offsets = p.offsets;
  const int T = p.triangle_count;
  const unsigned long long total = offsets[T];
  const int lane = threadIdx.x & 63;
  const unsigned waves = gridDim.x * (blockDim.x >> 6);
  const unsigned wave = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  // Each wave takes an equal contiguous run of slots, 64 per iteration.
  const unsigned long long per = ((total + waves - 1) / waves + 63) & ~63ull;
  const unsigned long long begin = gb_uniform(wave * per);
  const unsigned long long end = begin + per < total ? begin + per : total;
  if (begin >= end) return;
  int tf = gb_find(offsets, 0, T, begin);
  for (unsigned long long base = begin; base < end; base += 64) {
    const unsigned long long last = base + 63 < end ? base + 63 : end - 1;
    tf = gb_gallop(offsets, tf, T, base);
    const int tl = gb_gallop(offsets, tf, T, last);
    const unsigned long long slot = base + lane;
    if (slot > last) continue;
    const int lo = gb_find(offsets, tf, tl + 1, slot);
    const GbTri tri = p.tris[lo];
    const unsigned local = static_cast<unsigned>(slot - offsets[lo]);
    const int x = tri.x0 + static_cast<int>(local % static_cast<unsigned>(tri.w));
    const int y = tri.y0 + static_cast<int>(local / static_cast<unsigned>(tri.w));
    float e[3];
    const float t = gb_hit(p, tri, gb_ray(p, x, y), e);
    if (t < 0.0f) continue;
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
  const f3 d = gb_ray(p, x, y);
  float e[3];
  gb_hit(p, tri, d, e);
  const float inv = 1.0f / (e[0] + e[1] + e[2]);
  const float l[3] = {e[0] * inv, e[1] * inv, e[2] * inv};
  const uint32_t last = static_cast<uint32_t>(p.vertex_count - 1);
  const uint32_t vi[3] = {min(p.indices[3 * t], last), min(p.indices[3 * t + 1], last),
                          min(p.indices[3 * t + 2], last)};
  const f3 pos = gb_interp(p.positions, vi, l);
  f3 nrm = p.normals ? gb_interp(p.normals, vi, l) : mk(0.0f, 0.0f, 0.0f);
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
  const f3 alb = gb_interp(p.colors, vi, l);
  const f3 m = gb_interp(p.mro, vi, l);
  gb_store(p.position_out, p.fp16, i, pos.x, pos.y, pos.z, 1.0f);
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

void view_ray_affine(const althea_globals& g, int W, int H, float ray0[3], float rayx[3], float rayy[3]);
void view_projection(const althea_globals& g, float vp[16]);

althea_status launch_gbuffer_draw(const althea_gbuffer_mesh* mesh, const althea_globals* g,
                                  const althea_gbuffer_view* out, void* workspace, hipStream_t s) {
  GbParams p{};
  p.positions = mesh->positions;
  p.normals = mesh->normals;
  p.colors = mesh->colors;
  p.mro = mesh->mro;
  p.indices = mesh->indices;
  p.vertex_count = mesh->vertex_count;
  p.triangle_count = mesh->triangle_count;
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
  const size_t t = static_cast<size_t>(mesh->triangle_count);
  p.vis = reinterpret_cast<unsigned long long*>(ws);
  ws += align256(npx * 8);
  p.tris = reinterpret_cast<GbTri*>(ws);
  ws += align256(t * sizeof(GbTri));
  p.offsets = reinterpret_cast<unsigned long long*>(ws);
  ws += align256((t + 1) * 8);
  p.blocks = reinterpret_cast<unsigned long long*>(ws);
  p.position_out = const_cast<void*>(out->position);
  p.normal_out = const_cast<void*>(out->normal);
  p.albedo_out = const_cast<void*>(out->albedo);
  p.mro_out = const_cast<void*>(out->mro);
  p.fp16 = out->format == ALTHEA_GBUFFER_FP16 ? 1 : 0;
  const int n = static_cast<int>(npx);
  hipLaunchKernelGGL(gbuffer_clear_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p.vis, n);
  althea_status st = check_launch("gbuffer_clear_kernel");
  if (st != ALTHEA_OK) return st;
  if (mesh->triangle_count > 0) {
    const int nb = static_cast<int>(gb_setup_blocks(mesh->triangle_count));
    hipLaunchKernelGGL(gbuffer_setup_kernel, dim3(nb), dim3(kGbSetupThreads), 0, s, p);
    if ((st = check_launch("gbuffer_setup_kernel")) != ALTHEA_OK) return st;
    hipLaunchKernelGGL(gbuffer_scan_kernel, dim3(1), dim3(kGbScanThreads), 0, s, p.blocks, nb,
                       p.offsets + mesh->triangle_count);
    if ((st = check_launch("gbuffer_scan_kernel")) != ALTHEA_OK) return st;
    hipLaunchKernelGGL(gbuffer_offset_kernel, dim3((mesh->triangle_count + 255) / 256), dim3(256), 0, s, p.offsets,
                       p.blocks, mesh->triangle_count);
    if ((st = check_launch("gbuffer_offset_kernel")) != ALTHEA_OK) return st;
    hipLaunchKernelGGL(gbuffer_raster_kernel, dim3(kGbRasterBlocks), dim3(256), 0, s, p);
    if ((st = check_launch("gbuffer_raster_kernel")) != ALTHEA_OK) return st;
  }
  hipLaunchKernelGGL(gbuffer_resolve_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p);
  return check_launch("gbuffer_resolve_kernel");
}

}  // namespace althea
