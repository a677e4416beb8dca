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
// result does not depend on the order of the writes.  Small boxes are walked
// by their own lane; a wave walks larger boxes together, one lane's box at a
// time.  A last pass fills each face's border from the adjacent faces (the
// seamless filter's reads beyond the edge), so the lookup needs no folding.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "device_math.hpp"
#include "internal.hpp"
#include "shadow.hpp"

namespace althea {

constexpr int kShadowLaneTexels = 16;  // boxes up to this many texels stay on their lane
constexpr float kShadowClip = 0.5f * ALTHEA_SHADOW_NEAR;  // clip plane ma = near / 2

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
  atomicMin(reinterpret_cast<unsigned int*>(cube) + shadow_texel_index(S, face, i, j), __float_as_uint(depth));
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
  const int S = p.size;
  const float inv_s = 1.0f / static_cast<float>(S);
  const float half = 0.5f * static_cast<float>(S);
  const size_t cube_floats = static_cast<size_t>(6) * (S + 2) * (S + 2);
  float* cube = p.cubes + light * cube_floats;
  for (int face = 0; face < 6; ++face) {
    int x0 = 0, x1 = -1, y0 = 0, y1 = -1;
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
          const float r = 1.0f / m_;
          const float x = fmaf(s_ * r, half, half - 0.5f), y = fmaf(t_ * r, half, half - 0.5f);
          lo_x = fminf(lo_x, x), hi_x = fmaxf(hi_x, x), lo_y = fminf(lo_y, y), hi_y = fmaxf(hi_y, y);
        };
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const int k2 = k == 2 ? 0 : k + 1;
          if (in[k]) add(sc[k], tc[k], ma[k]);
          if (in[k] != in[k2]) {  // the edge crosses the clip plane
            const float u = (kShadowClip - ma[k]) / (ma[k2] - ma[k]);
            add(fmaf(u, sc[k2] - sc[k], sc[k]), fmaf(u, tc[k2] - tc[k], tc[k]), kShadowClip);
          }
        }
        // Texel centres inside the span, one texel of margin for rounding
        // (coordinates clamped to [-2, S + 1] before the integer conversion).
        const float lim = static_cast<float>(S + 1);
        x0 = max(static_cast<int>(ceilf(fminf(fmaxf(lo_x, -2.0f), lim))) - 1, 0);
        x1 = min(static_cast<int>(floorf(fminf(fmaxf(hi_x, -2.0f), lim))) + 1, S - 1);
        y0 = max(static_cast<int>(ceilf(fminf(fmaxf(lo_y, -2.0f), lim))) - 1, 0);
        y1 = min(static_cast<int>(floorf(fminf(fmaxf(hi_y, -2.0f), lim))) + 1, S - 1);
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
      float* wcube = p.cubes + __shfl(light, src) * cube_floats;
      const int n = ww * wh;
      for (int k = threadIdx.x & 63; k < n; k += 64) {
        const int j = wy0 + k / ww, i = wx0 + k % ww;
        shadow_texel(w, face, i, j, S, inv_s, wcube);
      }
    }
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

althea_status launch_shadow_cubes(const althea_mesh_view* mesh, const althea_point_light* lights, int light_count,
                                  int size, float* cubes, hipStream_t s) {
  const size_t texels = static_cast<size_t>(light_count) * 6 * (size + 2) * (size + 2);
  hipLaunchKernelGGL(shadow_clear_kernel, dim3(static_cast<unsigned>((texels + 255) / 256)), dim3(256), 0, s, cubes,
                     texels);
  althea_status st = check_launch("shadow_clear_kernel");
  if (st != ALTHEA_OK || mesh->triangle_count == 0) return st;  // borders of a cleared cube are 1 too
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
  st = check_launch("shadow_raster_kernel");
  if (st != ALTHEA_OK) return st;
  const size_t border = static_cast<size_t>(light_count) * 6 * (4 * size + 4);
  hipLaunchKernelGGL(shadow_border_kernel, dim3(static_cast<unsigned>((border + 255) / 256)), dim3(256), 0, s, cubes,
                     size, light_count);
  return check_launch("shadow_border_kernel");
}

}  // namespace althea
