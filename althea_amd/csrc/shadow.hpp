// Point-light shadow cubes as the reference lays them out and reads them.
//
// The reference renders each light's cube with six yaw / pitch cameras
// (Src/PointLight.cpp:73-131, Src/Camera.cpp:49-66) into the layers of a cube
// array and reads it at (L.x, -L.y, -L.z) (PBRMaterial.glsl:130), L the unit
// vector from the surface to the light.  Worked through the Vulkan face table,
// layer f's texel with Vulkan direction v holds the depth along R_f v:
// R_f = diag(-1, 1, 1) for the +-X and +-Z layers and diag(1, 1, -1) for +-Y.
// The lookup then reads the depth along -L on the +-X and +-Z faces and along
// -L mirrored through the light's vertical axis, (L.x, -L.y, L.z), on the
// +-Y faces: the +-Y images are turned by 180 degrees against the lookup.
// This build keeps both halves as they are, so its shadows are the
// reference's.
#pragma once

#include <hip/hip_runtime.h>

#include "althea/althea.h"
#include "device_math.hpp"

namespace althea {

// R_f (its own inverse).
__device__ __forceinline__ f3 shadow_face_turn(int face, f3 v) {
  return (face == 2 || face == 3) ? mk(v.x, v.y, -v.z) : mk(-v.x, v.y, v.z);
}

// Vulkan (sc, tc, ma) of the light-relative point p on face f of the
// reference's images, whose texel direction is R_f v.
__device__ __forceinline__ void shadow_face_coords(int face, f3 p, float& sc, float& tc, float& ma) {
  const f3 q = shadow_face_turn(face, p);
  switch (face) {
    case 0: ma = q.x; sc = -q.z; tc = -q.y; break;
    case 1: ma = -q.x; sc = q.z; tc = -q.y; break;
    case 2: ma = q.y; sc = q.x; tc = q.z; break;
    case 3: ma = -q.y; sc = q.x; tc = -q.z; break;
    case 4: ma = q.z; sc = q.x; tc = -q.y; break;
    default: ma = -q.z; sc = -q.x; tc = -q.y; break;
  }
}

// Seamless bilinear read of one scalar cube (face-major [face][y][x]) at
// projected face coordinates, with the same edge and corner rules as the
// float4 cubes (cube_tap).
__device__ __forceinline__ float cube1_texel(const float* __restrict__ c, int S, int face, int i, int j) {
  return c[(face * S + j) * S + i];
}

__device__ inline float cube1_tap(const float* __restrict__ c, int S, int face, int i, int j) {
  const bool in_i = static_cast<unsigned>(i) < static_cast<unsigned>(S);
  const bool in_j = static_cast<unsigned>(j) < static_cast<unsigned>(S);
  if (in_i && in_j) return cube1_texel(c, S, face, i, j);
  int f, a, b;
  if (in_i || in_j) {
    cube_fold(face, i, j, S, f, a, b);
    return cube1_texel(c, S, f, a, b);
  }
  const int ic = min(max(i, 0), S - 1), jc = min(max(j, 0), S - 1);
  float acc = cube1_texel(c, S, face, ic, jc);
  cube_fold(face, i, jc, S, f, a, b);
  acc += cube1_texel(c, S, f, a, b);
  cube_fold(face, ic, j, S, f, a, b);
  acc += cube1_texel(c, S, f, a, b);
  return acc * (1.0f / 3.0f);
}

__device__ inline float cube1_sample(const float* __restrict__ c, int S, f3 dir) {
  int face;
  float pc, pt;
  cube_select(dir, face, pc, pt);
  const float half = 0.5f * static_cast<float>(S);
  // pc, pt lie in [-1, 1]; the clamp only keeps a degenerate direction (NaN
  // coordinates) on the texels.
  const float hi = static_cast<float>(S) - 0.5f;
  const float x = fminf(fmaxf(fmaf(pc, half, half - 0.5f), -0.5f), hi);
  const float y = fminf(fmaxf(fmaf(pt, half, half - 0.5f), -0.5f), hi);
  const float fx = floorf(x), fy = floorf(y);
  const float ax = x - fx, ay = y - fy;
  const int i0 = static_cast<int>(fx), j0 = static_cast<int>(fy);
  float t00, t10, t01, t11;
  if (i0 >= 0 && j0 >= 0 && i0 + 1 < S && j0 + 1 < S) {
    const float* r0 = c + (face * S + j0) * S + i0;
    t00 = r0[0];
    t10 = r0[1];
    t01 = r0[S];
    t11 = r0[S + 1];
  } else {
    t00 = cube1_tap(c, S, face, i0, j0);
    t10 = cube1_tap(c, S, face, i0 + 1, j0);
    t01 = cube1_tap(c, S, face, i0, j0 + 1);
    t11 = cube1_tap(c, S, face, i0 + 1, j0 + 1);
  }
  const float top = fmaf(t10 - t00, ax, t00), bot = fmaf(t11 - t01, ax, t01);
  return fmaf(bot - top, ay, top);
}

// The shadow test of PBRMaterial.glsl:119-140 for light i: true when the
// light is hidden (closestDepth < Ldist - 0.5).  L = Lp - P, unnormalised.
struct ShadowCubes {
  static constexpr bool kActive = true;
  const float* cubes;  // [light][face][y][x]
  int size;
  __device__ __forceinline__ bool hidden(int light, f3 L, float ldist) const {
    const float* c = cubes + static_cast<size_t>(light) * 6 * size * size;
    const float closest = cube1_sample(c, size, mk(L.x, -L.y, -L.z)) * ALTHEA_SHADOW_FAR;
    return closest < ldist - 0.5f;
  }
};

}  // namespace althea
