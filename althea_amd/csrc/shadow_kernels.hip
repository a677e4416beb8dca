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
// One lane per (light, triangle): it projects the triangle onto each face,
// takes the texel box its vertices span (the whole face when a vertex lies
// behind the face's plane), and writes each hit with an integer atomicMin on
// the depth's bits (non-negative floats order as their bit patterns), so the
// result does not depend on the order of the writes.  Small boxes are walked
// by their own lane; a wave walks larger boxes together, one lane's box at a
// time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "device_math.hpp"
#include "internal.hpp"
#include "shadow.hpp"

namespace althea {

constexpr int kShadowLaneTexels = 16;  // boxes up to this many texels stay on their lane

struct ShadowParams {
  const float* positions;
  const uint32_t* indices;
  int vertex_count;
  int triangle_count;
  const float4* lights;  // [light][2]: position, emission
  int light_count;
  int size;
  float* cubes;
};

__global__ void shadow_clear_kernel(float* __restrict__ c, size_t n) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) c[i] = 1.0f;
}

// Texel (i, j) of face f: the ray direction R_f v (major component +-1) and
// its length.
__device__ __forceinline__ f3 shadow_texel_dir(int face, int i, int j, float inv_s, float& len) {
  const float pc = fmaf(static_cast<float>(2 * i + 1), inv_s, -1.0f);
  const float pt = fmaf(static_cast<float>(2 * j + 1), inv_s, -1.0f);
  len = sqrtf(fmaf(pc, pc, fmaf(pt, pt, 1.0f)));
  return shadow_face_turn(face, cube_face_dir(face, pc, pt));
}

struct ShadowTri {
  f3 n0, n1, n2;  // edge functions: b x c, c x a, a x b
  f3 n;           // plane normal (b - a) x (c - a)
  float det;      // a . n
};

__device__ __forceinline__ void shadow_texel(const ShadowTri& tri, int face, int i, int j, int S, float inv_s,
                                             float* __restrict__ cube) {
  float len;
  const f3 d = shadow_texel_dir(face, i, j, inv_s, len);
  const float e0 = dot(d, tri.n0), e1 = dot(d, tri.n1), e2 = dot(d, tri.n2);
  if (e0 > 0.0f || e1 > 0.0f || e2 > 0.0f) return;  // outside (det < 0)
  const float rn = dot(d, tri.n);
  if (!(rn < 0.0f)) return;
  const float t = tri.det / rn;  // view depth of the hit
  if (t < ALTHEA_SHADOW_NEAR || t > ALTHEA_SHADOW_FAR) return;
  const float depth = fminf(t * len * (1.0f / ALTHEA_SHADOW_FAR), 1.0f);
  atomicMin(reinterpret_cast<unsigned int*>(cube) + (face * S + j) * S + i, __float_as_uint(depth));
}

__global__ __launch_bounds__(256) void shadow_raster_kernel(const ShadowParams p) {
  const int item = blockIdx.x * blockDim.x + threadIdx.x;
  const int light = item / max(p.triangle_count, 1), t = item - light * p.triangle_count;
  bool live = item < p.light_count * p.triangle_count;
  ShadowTri tri{};
  f3 v[3] = {mk(0.0f, 0.0f, 0.0f), mk(0.0f, 0.0f, 0.0f), mk(0.0f, 0.0f, 0.0f)};
  if (live) {
    const float4 lp = p.lights[2 * light];
    for (int k = 0; k < 3; ++k) {
      const uint32_t vi = p.indices[3 * t + k];
      const float* q = p.positions + 3 * static_cast<size_t>(min(vi, static_cast<uint32_t>(p.vertex_count - 1)));
      v[k] = mk(q[0] - lp.x, q[1] - lp.y, q[2] - lp.z);
    }
    tri.n0 = cross(v[1], v[2]);
    tri.n1 = cross(v[2], v[0]);
    tri.n2 = cross(v[0], v[1]);
    tri.n = cross(v[1] - v[0], v[2] - v[0]);
    tri.det = dot(v[0], tri.n);
    live = tri.det < 0.0f;  // back faces and edge-on triangles are culled
  }
  const int S = p.size;
  const float inv_s = 1.0f / static_cast<float>(S);
  const float half = 0.5f * static_cast<float>(S);
  float* cube = p.cubes + static_cast<size_t>(light) * 6 * S * S;
  for (int face = 0; face < 6; ++face) {
    int x0 = 0, x1 = -1, y0 = 0, y1 = -1;
    if (live) {
      float sc[3], tc[3], ma[3];
      int front = 0;
      for (int k = 0; k < 3; ++k) {
        shadow_face_coords(face, v[k], sc[k], tc[k], ma[k]);
        front += ma[k] > 0.0f;
      }
      if (front == 3) {
        float lo_x = 1e30f, hi_x = -1e30f, lo_y = 1e30f, hi_y = -1e30f;
        for (int k = 0; k < 3; ++k) {
          const float x = fmaf(sc[k] / ma[k], half, half - 0.5f), y = fmaf(tc[k] / ma[k], half, half - 0.5f);
          lo_x = fminf(lo_x, x), hi_x = fmaxf(hi_x, x), lo_y = fminf(lo_y, y), hi_y = fmaxf(hi_y, y);
        }
        // Texel centres inside the span, one texel of margin for rounding
        // (coordinates clamped to [-2, S + 1] before the integer conversion).
        const float lim = static_cast<float>(S + 1);
        x0 = max(static_cast<int>(ceilf(fminf(fmaxf(lo_x, -2.0f), lim))) - 1, 0);
        x1 = min(static_cast<int>(floorf(fminf(fmaxf(hi_x, -2.0f), lim))) + 1, S - 1);
        y0 = max(static_cast<int>(ceilf(fminf(fmaxf(lo_y, -2.0f), lim))) - 1, 0);
        y1 = min(static_cast<int>(floorf(fminf(fmaxf(hi_y, -2.0f), lim))) + 1, S - 1);
      } else if (front > 0) {
        x0 = 0, x1 = S - 1, y0 = 0, y1 = S - 1;  // crosses the face's plane
      }
    }
    const int bw = x1 - x0 + 1, bh = y1 - y0 + 1;
    const int area = (bw > 0 && bh > 0) ? bw * bh : 0;
    if (area > 0 && area <= kShadowLaneTexels) {
      for (int j = y0; j <= y1; ++j)
        for (int i = x0; i <= x1; ++i) shadow_texel(tri, face, i, j, S, inv_s, cube);
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
      float* wcube = p.cubes + static_cast<size_t>(__shfl(light, src)) * 6 * S * S;
      const int n = ww * wh;
      for (int k = threadIdx.x & 63; k < n; k += 64) {
        const int j = wy0 + k / ww, i = wx0 + k % ww;
        shadow_texel(w, face, i, j, S, inv_s, wcube);
      }
    }
  }
}

althea_status launch_shadow_cubes(const althea_mesh_view* mesh, const althea_point_light* lights, int light_count,
                                  int size, float* cubes, hipStream_t s) {
  const size_t texels = static_cast<size_t>(light_count) * 6 * size * size;
  hipLaunchKernelGGL(shadow_clear_kernel, dim3(static_cast<unsigned>((texels + 255) / 256)), dim3(256), 0, s, cubes,
                     texels);
  althea_status st = check_launch("shadow_clear_kernel");
  if (st != ALTHEA_OK || mesh->triangle_count == 0) return st;
  ShadowParams p{};
  p.positions = mesh->positions;
  p.indices = mesh->indices;
  p.vertex_count = mesh->vertex_count;
  p.triangle_count = mesh->triangle_count;
  p.lights = reinterpret_cast<const float4*>(lights);
  p.light_count = light_count;
  p.size = size;
  p.cubes = cubes;
  const size_t items = static_cast<size_t>(light_count) * mesh->triangle_count;
  hipLaunchKernelGGL(shadow_raster_kernel, dim3(static_cast<unsigned>((items + 255) / 256)), dim3(256), 0, s, p);
  return check_launch("shadow_raster_kernel");
}

}  // namespace althea
