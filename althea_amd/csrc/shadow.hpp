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

// Storage: each face image with a one-texel border, (S + 2)^2 floats per
// face, face texel (i, j) at row j + 1, column i + 1.  The border holds what
// the seamless filter reads one texel beyond the face (the edge texels of the
// adjacent faces; at the corners the mean of the three texels that meet
// there), written once by the producer, so a lookup is four plain loads.
__device__ __forceinline__ int shadow_texel_index(int S, int face, int i, int j) {
  const int P = S + 2;
  return (face * P + j + 1) * P + i + 1;
}

// Face and projected coordinates of a direction, as cube_select, with
// selects instead of branches.
__device__ __forceinline__ void shadow_cube_select(f3 d, int& face, float& pc, float& pt) {
  const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  const bool xm = ax >= ay && ax >= az;
  const bool ym = !xm && ay >= az;
  const bool px = d.x >= 0.0f, py = d.y >= 0.0f, pz = d.z >= 0.0f;
  const float ma = xm ? ax : (ym ? ay : az);
  const float sc = xm ? (px ? -d.z : d.z) : (ym ? d.x : (pz ? d.x : -d.x));
  const float tc = xm ? -d.y : (ym ? (py ? d.z : -d.z) : -d.y);
  face = xm ? (px ? 0 : 1) : (ym ? (py ? 2 : 3) : (pz ? 4 : 5));
  const float inv = rcp_fast(ma);
  pc = sc * inv;
  pt = tc * inv;
}

// Seamless bilinear read of one bordered scalar cube.
__device__ __forceinline__ float shadow_cube_sample(const float* __restrict__ c, int S, f3 dir) {
  int face;
  float pc, pt;
  shadow_cube_select(dir, face, pc, pt);
  const float half = 0.5f * static_cast<float>(S);
  // pc, pt lie in [-1, 1]; the clamp only keeps a degenerate direction (NaN
  // coordinates) on the texels.
  const float hi = static_cast<float>(S) - 0.5f;
  const float x = fminf(fmaxf(fmaf(pc, half, half - 0.5f), -0.5f), hi);
  const float y = fminf(fmaxf(fmaf(pt, half, half - 0.5f), -0.5f), hi);
  const float fx = floorf(x), fy = floorf(y);
  const float ax = x - fx, ay = y - fy;
  const int i0 = static_cast<int>(fx), j0 = static_cast<int>(fy);  // in [-1, S - 1]
  const float* r0 = c + shadow_texel_index(S, face, i0, j0);
  const int P = S + 2;
  const float t00 = r0[0], t10 = r0[1], t01 = r0[P], t11 = r0[P + 1];
  const float top = fmaf(t10 - t00, ax, t00), bot = fmaf(t11 - t01, ax, t01);
  return fmaf(bot - top, ay, top);
}

// The shadow test of PBRMaterial.glsl:119-140 for light i: true when the
// light is hidden (closestDepth < Ldist - 0.5).  L = Lp - P, unnormalised.
struct ShadowCubes {
  static constexpr bool kActive = true;
  const float* cubes;  // [light][face][y + 1][x + 1], bordered faces
  int size;
  __device__ __forceinline__ bool hidden(int light, f3 L, float ldist) const {
    const float* c = cubes + static_cast<size_t>(light) * 6 * (size + 2) * (size + 2);
    const float closest = shadow_cube_sample(c, size, mk(L.x, -L.y, -L.z)) * ALTHEA_SHADOW_FAR;
    return closest < ldist - 0.5f;
  }
};

}  // namespace althea
