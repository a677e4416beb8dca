// Device-side math shared by the gfx950 kernels: Hammersley points, tangent
// frames, the software equirect and seamless-cube texture filters, and the
// GGX / Fresnel terms of the shading model.
//
// Texture filtering follows the public Vulkan texel-filtering rules the
// reference's samplers use (Src/Sampler.cpp:90-125): coordinates u*W - 0.5,
// floor, linear weights, REPEAT or CLAMP on integer texel indices, LINEAR
// mipmapping between floor(lod) and floor(lod)+1.  Texel coordinates are formed
// with one FMA from the angle (x = yaw * W/2pi + (W/2 - 0.5)) so FP32 rounding
// is applied once, in texel units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace althea {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.318309886183790671538f;
constexpr float kInv2Pi = 0.159154943091895335769f;
constexpr int kMaxLevels = 16;

// ------------------------------------------------------------------ basics
struct f3 {
  float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 mul(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ f3 normalize(f3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
__device__ __forceinline__ f3 lerp3(f3 a, f3 b, float t) {
  return {fmaf(b.x - a.x, t, a.x), fmaf(b.y - a.y, t, a.y), fmaf(b.z - a.z, t, a.z)};
}
__device__ __forceinline__ f3 xyz(float4 v) { return {v.x, v.y, v.z}; }

// Hardware reciprocal / reciprocal square root / square root (1 ulp), for the
// hot loops where the correctly rounded sequences (about ten instructions
// each) dominate the instruction count.
__device__ __forceinline__ float rcp_fast(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float rsq_fast(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float sqrt_fast(float x) { return __builtin_amdgcn_sqrtf(x); }
// min(max(x, 0), 1); folds into the producing instruction's clamp bit.
__device__ __forceinline__ float saturate(float x) { return __builtin_amdgcn_fmed3f(x, 0.0f, 1.0f); }
__device__ __forceinline__ f3 normalize_fast(f3 a) { return a * rsq_fast(fmaf(a.x, a.x, fmaf(a.y, a.y, a.z * a.z))); }
__device__ __forceinline__ float4 lerp4(float4 a, float4 b, float t) {
  return {fmaf(b.x - a.x, t, a.x), fmaf(b.y - a.y, t, a.y), fmaf(b.z - a.z, t, a.z), fmaf(b.w - a.w, t, a.w)};
}

// atan2 by octant reduction to t = min(|x|,|y|) / max(|x|,|y|) in [0, 1] and
// atan(t) = t P(t^2), P of degree 8 fitted to atan(t)/t on [0, 1] by
// iteratively reweighted least squares on Chebyshev nodes (absolute error
// 5.8e-9 in exact arithmetic, 1.1e-7 rad evaluated in FP32 by Horner).
__device__ __forceinline__ float atan_unit(float t) {
  const float s = t * t;
  float p = 2.456719056e-03f;
  p = fmaf(p, s, -1.440133527e-02f);
  p = fmaf(p, s, 3.978118673e-02f);
  p = fmaf(p, s, -7.234854251e-02f);
  p = fmaf(p, s, 1.049894467e-01f);
  p = fmaf(p, s, -1.416122913e-01f);
  p = fmaf(p, s, 1.998590678e-01f);
  p = fmaf(p, s, -3.333259821e-01f);
  p = fmaf(p, s, 9.999998808e-01f);
  return p * t;
}
// atan2 for (y, x) != (0, 0); at the origin the result is NaN (mn / mx =
// 0 * inf), which equirect_angles replaces by its pole branch.
__device__ __forceinline__ float atan2_fast(float y, float x) {
  const float ax = fabsf(x), ay = fabsf(y);
  const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
  float r = atan_unit(mn * rcp_fast(mx));
  r = ay > ax ? 1.570796327f - r : r;
  r = x < 0.0f ? 3.141592654f - r : r;
  return copysignf(r, y);
}
// atan2(y, x) for x >= 0 (the elevation angle atan(y / |xz|)), (y, x) != 0.
__device__ __forceinline__ float atan2_xpos(float y, float x) {
  const float ay = fabsf(y);
  const float mx = fmaxf(x, ay), mn = fminf(x, ay);
  float r = atan_unit(mn * rcp_fast(mx));
  r = ay > x ? 1.570796327f - r : r;
  return copysignf(r, y);
}

// -------------------------------------------------------------- Hammersley
// Point i of an n-point set: (i/n correctly rounded, float(bitreverse(i)) * 2^-32).
__device__ __forceinline__ uint32_t vdc_bits(uint32_t i) { return __builtin_bitreverse32(i); }
__device__ __forceinline__ void hammersley(uint32_t i, uint32_t n, float& x1, float& x2) {
  x1 = static_cast<float>(i) / static_cast<float>(n);
  x2 = static_cast<float>(vdc_bits(i)) * 2.3283064365386963e-10f;
}

// ------------------------------------------------------------ tangent frame
// coordinateSystem / LocalToWorld (Shaders/Misc/Sampling.glsl:58-70).  The
// |x| > |y| branch is taken on the raw vector so every precision agrees on it.
struct Frame {
  f3 t, b, n;
};
__device__ __forceinline__ Frame local_to_world(f3 raw) {
  Frame f;
  f.n = normalize(raw);
  f3 n = f.n;
  if (fabsf(raw.x) > fabsf(raw.y)) {
    float inv = 1.0f / sqrtf(fmaf(n.x, n.x, n.z * n.z));
    f.t = mk(-n.z * inv, 0.0f, n.x * inv);
  } else {
    float inv = 1.0f / sqrtf(fmaf(n.y, n.y, n.z * n.z));
    f.t = mk(0.0f, n.z * inv, -n.y * inv);
  }
  f.b = cross(n, f.t);
  return f;
}

// -------------------------------------------------------- environment maps
// Per-level sampler constants of a power-of-two equirect level, 32 bytes so a
// wave-uniform level loads with one scalar load:
//   x = yaw * kx + cx   (kx = w / 2pi, cx = w/2 - 0.5)
//   y = pitch * ky + cy (ky = -h / pi, cy = h/2 - 0.5)
//   texel (i, j) at byte offset ((off + (j << wshift) + i) << 4), REPEAT by masks.
struct LevelC {
  float kx, cx, ky, cy;
  uint32_t wmask, hmask, wshift, off;
};

__host__ inline LevelC make_level(int w, int h, uint32_t off) {
  LevelC c;
  c.kx = static_cast<float>(static_cast<double>(w) * 0.15915494309189533577);
  c.cx = 0.5f * static_cast<float>(w) - 0.5f;
  c.ky = static_cast<float>(-static_cast<double>(h) * 0.31830988618379067154);
  c.cy = 0.5f * static_cast<float>(h) - 0.5f;
  c.wmask = static_cast<uint32_t>(w - 1);
  c.hmask = static_cast<uint32_t>(h - 1);
  uint32_t s = 0;
  while ((1 << s) < w) ++s;
  c.wshift = s;
  c.off = off;
  return c;
}

// floor(x) as an integer in one instruction (v_cvt_flr_i32_f32).
__device__ __forceinline__ int cvt_floor(float x) {
  int r;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

// The four REPEAT texels around (yaw, pitch) on a level (the FP32 texel
// coordinate is one FMA from the angle) and the bilinear fractions.  Rows
// wrap through the level's texel mask: the row after the last is row 0.
struct LevelTaps {
  f3 t00, t10, t01, t11;
  float ax, ay;
};
__device__ __forceinline__ LevelTaps level_taps(const char* __restrict__ pyr_bytes, const LevelC& c, float yaw,
                                                float pitch) {
  const float x = fmaf(yaw, c.kx, c.cx), y = fmaf(pitch, c.ky, c.cy);
  LevelTaps t;
  t.ax = __builtin_amdgcn_fractf(x);
  t.ay = __builtin_amdgcn_fractf(y);
  const int ix = cvt_floor(x), iy = cvt_floor(y);
  const uint32_t x0 = static_cast<uint32_t>(ix) & c.wmask, x1 = static_cast<uint32_t>(ix + 1) & c.wmask;
  const uint32_t lmask = ((c.hmask + 1u) << c.wshift) - 1u;  // texels of the level - 1 (wave-uniform)
  const uint32_t r0 = (static_cast<uint32_t>(iy) << c.wshift) & lmask;
  const uint32_t r1 = (r0 + c.wmask + 1u) & lmask;
  // The level offset goes into the (wave-uniform, scalar) base address.
  const char* lvl = pyr_bytes + (static_cast<size_t>(c.off) << 4);
  auto ld = [&](uint32_t i) { return xyz(*reinterpret_cast<const float4*>(lvl + (i << 4))); };
  t.t00 = ld(r0 + x0);
  t.t10 = ld(r0 + x1);
  t.t01 = ld(r1 + x0);
  t.t11 = ld(r1 + x1);
  return t;
}

// Bilinear REPEAT fetch of (yaw, pitch) from a level.
__device__ __forceinline__ f3 level_bilinear(const char* __restrict__ pyr_bytes, const LevelC& c, float yaw,
                                             float pitch) {
  const LevelTaps t = level_taps(pyr_bytes, c, yaw, pitch);
  const f3 top = lerp3(t.t00, t.t10, t.ax), bot = lerp3(t.t01, t.t11, t.ax);
  return lerp3(top, bot, t.ay);
}

// Gather copy of a pyramid level, read by the precompute's sample loops:
// the level with its REPEAT neighbours around it, (w + 2) x (h + 2) texels,
// padded (r, c) = level((r - 1) mod h, (c - 1) mod w).  The texel coordinate
// is formed exactly as level_taps forms it; with the wrap stored, a tap pair
// sits at one byte offset and offset + 16, the next row at + rowbytes.
struct GatherLevel {
  float kx, cx, ky, cy;  // x = yaw kx + cx, y = pitch ky + cy (as LevelC)
  uint32_t rowbytes;     // (w + 2) * 16
  uint32_t off;          // first texel of the padded level, float4 units
  uint32_t pad0, pad1;
};

__host__ inline GatherLevel make_gather_level(int w, int h, uint32_t off) {
  const LevelC c = make_level(w, h, 0);
  GatherLevel g{};
  g.kx = c.kx;
  g.cx = c.cx;
  g.ky = c.ky;
  g.cy = c.cy;
  g.rowbytes = static_cast<uint32_t>(w + 2) * 16u;
  g.off = off;
  return g;
}

// level_taps on a gather level: the same four texels and fractions.  The
// angles are those of equirect_angles, |yaw| <= float(pi) and |pitch| <=
// float(pi) / 2, so x lies in [-0.5, w - 0.5] and y in [-0.5, h - 0.5] up to
// a rounding of w * 5e-8 texel: floor(x) + 1 is in [0, w] and floor(y) + 1
// in [0, h], inside the padded level (levels are at most 2^15 wide; a NaN
// angle converts to index 0).
__device__ __forceinline__ LevelTaps gather_taps(const char* __restrict__ gpyr_bytes, const GatherLevel& g, float yaw,
                                                 float pitch) {
  const float x = fmaf(yaw, g.kx, g.cx), y = fmaf(pitch, g.ky, g.cy);
  LevelTaps t;
  t.ax = __builtin_amdgcn_fractf(x);
  t.ay = __builtin_amdgcn_fractf(y);
  const int ix = cvt_floor(x), iy = cvt_floor(y);
  // (iy + 1) rowbytes + (ix + 1) 16 as iy * rowbytes + (16 ix + rowbytes + 16):
  // one shift-add and one signed 24-bit multiply-add (iy >= -1, rowbytes < 2^20).
  const int rb = static_cast<int>(g.rowbytes);
  int col, o;
  asm("v_lshl_add_u32 %0, %1, 4, %2" : "=v"(col) : "v"(ix), "s"(rb + 16));
  asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(o) : "v"(iy), "s"(rb), "v"(col));
  const uint32_t o0 = static_cast<uint32_t>(o);
  const uint32_t o1 = o0 + g.rowbytes;
  // Global address space: the loads take the scalar base plus a 32-bit lane
  // offset and the immediate +16, also where the base comes from memory.
  typedef const __attribute__((address_space(1))) char gchar;
  typedef const __attribute__((address_space(1))) float gfloat;
  gchar* lvl = (gchar*)gpyr_bytes + (static_cast<size_t>(g.off) << 4);
  gfloat* r0 = (gfloat*)(lvl + o0);
  gfloat* r1 = (gfloat*)(lvl + o1);
  t.t00 = mk(r0[0], r0[1], r0[2]);
  t.t10 = mk(r0[4], r0[5], r0[6]);
  t.t01 = mk(r1[0], r1[1], r1[2]);
  t.t11 = mk(r1[4], r1[5], r1[6]);
  return t;
}

// Split gather copy: the padded texels of a gather level as an RG plane
// (float2 per texel) and a B plane (float per texel).  A tap pair's red and
// green are one 16-byte load and its blue one 8-byte load, so a wave whose
// lanes read a patch of neighbouring texels touches about half the cache
// lines that the float4 copy's two 12-byte loads per pair touch.
struct SplitLevel {
  float kx, cx, ky, cy;  // as GatherLevel
  uint32_t rowbytes;     // (w + 2) * 8, one row of the RG plane
  uint32_t rg;           // byte offset of the level's RG plane
  uint32_t b;            // byte offset of the level's B plane
  uint32_t pad;
};

// The four taps of a split level around padded entry (ix + 1, iy + 1) with
// fractions (ax, ay).
__device__ __forceinline__ LevelTaps split_taps_at(const char* __restrict__ base, const SplitLevel& g, int ix, int iy,
                                                   float ax, float ay) {
  LevelTaps t;
  t.ax = ax;
  t.ay = ay;
  // RG offset (iy + 1) rowbytes + (ix + 1) 8 as iy * rowbytes + (8 ix +
  // rowbytes + 8); the B offset is half of it.
  const int rb = static_cast<int>(g.rowbytes);
  int col, o;
  asm("v_lshl_add_u32 %0, %1, 3, %2" : "=v"(col) : "v"(ix), "s"(rb + 8));
  asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(o) : "v"(iy), "s"(rb), "v"(col));
  const uint32_t o0 = static_cast<uint32_t>(o), ob = o0 >> 1;
  typedef const __attribute__((address_space(1))) char gchar;
  typedef const __attribute__((address_space(1))) float gfloat;
  gchar* rgp = (gchar*)base + g.rg;
  gchar* bp = (gchar*)base + g.b;
  gfloat* a0 = (gfloat*)(rgp + o0);
  gfloat* a1 = (gfloat*)(rgp + (o0 + g.rowbytes));
  gfloat* b0 = (gfloat*)(bp + ob);
  gfloat* b1 = (gfloat*)(bp + (ob + (g.rowbytes >> 1)));
  t.t00 = mk(a0[0], a0[1], b0[0]);
  t.t10 = mk(a0[2], a0[3], b0[1]);
  t.t01 = mk(a1[0], a1[1], b1[0]);
  t.t11 = mk(a1[2], a1[3], b1[1]);
  return t;
}

// gather_taps on a split level: the same four texels and fractions.
__device__ __forceinline__ LevelTaps split_taps(const char* __restrict__ base, const SplitLevel& g, float yaw,
                                                float pitch) {
  const float x = fmaf(yaw, g.kx, g.cx), y = fmaf(pitch, g.ky, g.cy);
  return split_taps_at(base, g, cvt_floor(x), cvt_floor(y), __builtin_amdgcn_fractf(x), __builtin_amdgcn_fractf(y));
}

// Blend grid of a two-level lookup at a fixed LOD (the irradiance's): the
// value (1 - delta) B0 + delta B1 of the fine level's bilinear B0 and the
// next level's B1 is bilinear between the lines where either piecewise
// bilinear function bends, the fine level's texel centres (integer fine
// coordinates x0) and the coarse level's (x0 = 2k + 1/2).  Its nodes sit at
// x0 = 2k, 2k + 1/2 and 2k + 1 (three per two fine texels, along each
// axis), so one bilinear lookup of the node values replaces the two levels'
// lookups.  Stored as a split level of (3 w0 / 2 + 2) x (3 h0 / 2 + 2)
// nodes (REPEAT border); kx, cx, ky, cy are the fine level's.
//
// Lower node of the cell holding x0 and the fraction across it: r = x0 - 2k
// in [0, 2) is exact (Sterbenz) for k >= 0 and rounded once for x0 in
// [-1/2, 0); the fractions 2r, 2r - 1 and r - 1 are exact.
__device__ __forceinline__ int blend_cell(float x0, float& t) {
  const float k = floorf(0.5f * x0);
  const float r = fmaf(-2.0f, k, x0);
  const bool c1 = r >= 0.5f, c2 = r >= 1.0f;
  t = c2 ? r - 1.0f : (c1 ? fmaf(2.0f, r, -1.0f) : 2.0f * r);
  return 3 * static_cast<int>(k) + static_cast<int>(c1) + static_cast<int>(c2);
}


// acc += w * bilinear in weight form: the four texel weights carry w, so a
// sample's two source levels (w = its weight times 1 - delta and times
// delta) add straight into the running sums.  w (1 - a) is one FMA.
__device__ __forceinline__ void taps_acc(const LevelTaps& t, float w, float& ar, float& ag, float& ab) {
  const float wy1 = t.ay * w, wy0 = fmaf(-t.ay, w, w);
  const float w10 = t.ax * wy0, w00 = fmaf(-t.ax, wy0, wy0);
  const float w11 = t.ax * wy1, w01 = fmaf(-t.ax, wy1, wy1);
  ar = fmaf(t.t00.x, w00, ar);
  ag = fmaf(t.t00.y, w00, ag);
  ab = fmaf(t.t00.z, w00, ab);
  ar = fmaf(t.t10.x, w10, ar);
  ag = fmaf(t.t10.y, w10, ag);
  ab = fmaf(t.t10.z, w10, ab);
  ar = fmaf(t.t01.x, w01, ar);
  ag = fmaf(t.t01.y, w01, ag);
  ab = fmaf(t.t01.z, w01, ab);
  ar = fmaf(t.t11.x, w11, ar);
  ag = fmaf(t.t11.y, w11, ag);
  ab = fmaf(t.t11.z, w11, ab);
}

__device__ __forceinline__ void level_bilinear_acc(const char* __restrict__ pyr_bytes, const LevelC& c, float yaw,
                                                   float pitch, float w, float& ar, float& ag, float& ab) {
  taps_acc(level_taps(pyr_bytes, c, yaw, pitch), w, ar, ag, ab);
}

__device__ __forceinline__ void gather_bilinear_acc(const char* __restrict__ gpyr_bytes, const GatherLevel& g,
                                                    float yaw, float pitch, float w, float& ar, float& ag,
                                                    float& ab) {
  taps_acc(gather_taps(gpyr_bytes, g, yaw, pitch), w, ar, ag, ab);
}

__device__ __forceinline__ void split_bilinear_acc(const char* __restrict__ base, const SplitLevel& g, float yaw,
                                                   float pitch, float w, float& ar, float& ag, float& ab) {
  taps_acc(split_taps(base, g, yaw, pitch), w, ar, ag, ab);
}

__device__ __forceinline__ void blend_bilinear_acc(const char* __restrict__ base, const SplitLevel& g, float yaw,
                                                   float pitch, float w, float& ar, float& ag, float& ab) {
  // Coordinates below -1/2 by rounding (and NaN) go to -1/2, whose cell's
  // lower node is the border node -1.
  const float x0 = fmaxf(fmaf(yaw, g.kx, g.cx), -0.5f), y0 = fmaxf(fmaf(pitch, g.ky, g.cy), -0.5f);
  float ax, ay;
  const int nx = blend_cell(x0, ax), ny = blend_cell(y0, ay);
  taps_acc(split_taps_at(base, g, nx, ny, ax, ay), w, ar, ag, ab);
}

// Bilinear CLAMP_TO_EDGE fetch (the shading-side samplers).
__device__ __forceinline__ f3 env_bilinear_clamp(const float4* __restrict__ level, int w, int h, float x,
                                                 float y) {
  float fx = floorf(x), fy = floorf(y);
  float ax = x - fx, ay = y - fy;
  int ix = static_cast<int>(fx), iy = static_cast<int>(fy);
  int x0 = min(max(ix, 0), w - 1), x1 = min(max(ix + 1, 0), w - 1);
  int y0 = min(max(iy, 0), h - 1), y1 = min(max(iy + 1, 0), h - 1);
  float4 t00 = level[y0 * w + x0];
  float4 t10 = level[y0 * w + x1];
  float4 t01 = level[y1 * w + x0];
  float4 t11 = level[y1 * w + x1];
  f3 top = lerp3(xyz(t00), xyz(t10), ax);
  f3 bot = lerp3(xyz(t01), xyz(t11), ax);
  return lerp3(top, bot, ay);
}

// Precompute-side equirect angles (PreFilterEnvMap.comp:98-122): pole branch
// at |xz| <= 0.001, yaw = atan2(z, x), pitch = atan(y / |xz|).  Returns the
// angles; the caller turns them into texel coordinates per level.
__device__ __forceinline__ void equirect_angles(f3 d, float& yaw, float& pitch) {
  const float len_xz = sqrt_fast(fmaf(d.x, d.x, d.z * d.z));
  const bool off_pole = len_xz > 0.001f;
  const float yaw_v = atan2_fast(d.z, d.x);
  const float pitch_v = atan2_xpos(d.y, len_xz);  // atan(y / |xz|) for |xz| > 0
  yaw = off_pole ? yaw_v : 0.0f;
  pitch = off_pole ? pitch_v : (d.y > 0.0f ? 0.5f * kPi : -0.5f * kPi);
}

// Texel coordinates of (yaw, pitch) on a w x h level: u = yaw/2pi + 0.5,
// v = 0.5 - pitch/pi, x = u*w - 0.5, y = v*h - 0.5.
__device__ __forceinline__ float equirect_x(float yaw, int w) {
  float fw = static_cast<float>(w);
  return fmaf(yaw, fw * kInv2Pi, 0.5f * fw - 0.5f);
}
__device__ __forceinline__ float equirect_y(float pitch, int h) {
  float fh = static_cast<float>(h);
  return fmaf(pitch, -fh * kInvPi, 0.5f * fh - 0.5f);
}

// -------------------------------------------------------------- cube maps
// Faces +X,-X,+Y,-Y,+Z,-Z; (sc, tc, ma) per the Vulkan cube-face-selection
// table.  Storage [mip][face][y][x] float4.
__device__ __forceinline__ f3 cube_face_dir(int face, float sc, float tc) {
  switch (face) {
    case 0: return mk(1.0f, -tc, -sc);
    case 1: return mk(-1.0f, -tc, sc);
    case 2: return mk(sc, 1.0f, tc);
    case 3: return mk(sc, -1.0f, -tc);
    case 4: return mk(sc, -tc, 1.0f);
    default: return mk(-sc, -tc, -1.0f);
  }
}
// Raw direction of texel (x, y) of an S x S face on the integer lattice of
// half-texels, (2x+1-S, 2y+1-S, +-S) through the face table: exact, so the
// frame's |x| > |y| tie-break matches the FP64 oracle for any S.
__device__ __forceinline__ f3 cube_texel_raw_dir(int face, int x, int y, int S) {
  const float sc = static_cast<float>(2 * x + 1 - S), tc = static_cast<float>(2 * y + 1 - S);
  const float m = static_cast<float>(S);
  switch (face) {
    case 0: return mk(m, -tc, -sc);
    case 1: return mk(-m, -tc, sc);
    case 2: return mk(sc, m, tc);
    case 3: return mk(sc, -m, -tc);
    case 4: return mk(sc, -tc, m);
    default: return mk(-sc, -tc, -m);
  }
}

// Major-axis selection: face and the projected (sc/ma, tc/ma) in [-1, 1].
__device__ __forceinline__ void cube_select(f3 d, int& face, float& pc, float& pt) {
  float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  float sc, tc, ma;
  if (ax >= ay && ax >= az) {
    ma = ax;
    if (d.x >= 0.0f) { face = 0; sc = -d.z; tc = -d.y; }
    else { face = 1; sc = d.z; tc = -d.y; }
  } else if (ay >= az) {
    ma = ay;
    if (d.y >= 0.0f) { face = 2; sc = d.x; tc = d.z; }
    else { face = 3; sc = d.x; tc = -d.z; }
  } else {
    ma = az;
    if (d.z >= 0.0f) { face = 4; sc = d.x; tc = -d.y; }
    else { face = 5; sc = -d.x; tc = -d.y; }
  }
  const float inv = rcp_fast(ma);
  pc = sc * inv;
  pt = tc * inv;
}

// Seamless edge handling on the integer lattice of half-texel units: a tap
// one texel beyond an edge is folded onto the neighbouring face.
__device__ inline void cube_fold(int face, int i, int j, int S, int& fo, int& io, int& jo) {
  int sc = 2 * i + 1 - S, tc = 2 * j + 1 - S;
  int p[3];
  switch (face) {
    case 0: p[0] = S; p[1] = -tc; p[2] = -sc; break;
    case 1: p[0] = -S; p[1] = -tc; p[2] = sc; break;
    case 2: p[0] = sc; p[1] = S; p[2] = tc; break;
    case 3: p[0] = sc; p[1] = -S; p[2] = -tc; break;
    case 4: p[0] = sc; p[1] = -tc; p[2] = S; break;
    default: p[0] = -sc; p[1] = -tc; p[2] = -S; break;
  }
  int major = (face >> 1);
  int over = -1;
#pragma unroll
  for (int a = 0; a < 3; ++a)
    if (abs(p[a]) == S + 1) over = a;
  p[over] = p[over] > 0 ? S : -S;
  p[major] = p[major] > 0 ? S - 1 : -(S - 1);
  int nsc, ntc;
  if (p[0] == S) { fo = 0; nsc = -p[2]; ntc = -p[1]; }
  else if (p[0] == -S) { fo = 1; nsc = p[2]; ntc = -p[1]; }
  else if (p[1] == S) { fo = 2; nsc = p[0]; ntc = p[2]; }
  else if (p[1] == -S) { fo = 3; nsc = p[0]; ntc = -p[2]; }
  else if (p[2] == S) { fo = 4; nsc = p[0]; ntc = -p[1]; }
  else { fo = 5; nsc = -p[0]; ntc = -p[1]; }
  io = (nsc + S - 1) >> 1;
  jo = (ntc + S - 1) >> 1;
}

__device__ __forceinline__ f3 cube_texel(const float4* __restrict__ mip, int S, int face, int i, int j) {
  return xyz(mip[(face * S + j) * S + i]);
}

__device__ inline f3 cube_tap(const float4* __restrict__ mip, int S, int face, int i, int j) {
  bool in_i = static_cast<unsigned>(i) < static_cast<unsigned>(S);
  bool in_j = static_cast<unsigned>(j) < static_cast<unsigned>(S);
  if (in_i && in_j) return cube_texel(mip, S, face, i, j);
  int f, a, b;
  if (in_i || in_j) {
    cube_fold(face, i, j, S, f, a, b);
    return cube_texel(mip, S, f, a, b);
  }
  int ic = min(max(i, 0), S - 1), jc = min(max(j, 0), S - 1);
  f3 acc = cube_texel(mip, S, face, ic, jc);
  cube_fold(face, i, jc, S, f, a, b);
  acc = acc + cube_texel(mip, S, f, a, b);
  cube_fold(face, ic, j, S, f, a, b);
  acc = acc + cube_texel(mip, S, f, a, b);
  return acc * (1.0f / 3.0f);
}

__device__ inline f3 cube_bilinear(const float4* __restrict__ mip, int S, int face, float pc, float pt) {
  float half = 0.5f * static_cast<float>(S);
  float x = fmaf(pc, half, half - 0.5f), y = fmaf(pt, half, half - 0.5f);
  float fx = floorf(x), fy = floorf(y);
  float ax = x - fx, ay = y - fy;
  int i0 = static_cast<int>(fx), j0 = static_cast<int>(fy);
  f3 t00, t10, t01, t11;
  if (i0 >= 0 && j0 >= 0 && i0 + 1 < S && j0 + 1 < S) {
    const float4* r0 = mip + (face * S + j0) * S + i0;
    t00 = xyz(r0[0]);
    t10 = xyz(r0[1]);
    t01 = xyz(r0[S]);
    t11 = xyz(r0[S + 1]);
  } else {
    t00 = cube_tap(mip, S, face, i0, j0);
    t10 = cube_tap(mip, S, face, i0 + 1, j0);
    t01 = cube_tap(mip, S, face, i0, j0 + 1);
    t11 = cube_tap(mip, S, face, i0 + 1, j0 + 1);
  }
  f3 top = lerp3(t00, t10, ax), bot = lerp3(t01, t11, ax);
  return lerp3(top, bot, ay);
}

struct CubeMips {
  int size, mips;
  uint32_t off[kMaxLevels];  // texel offset of each mip
};

// Seamless-bordered copy of a cube: mip m (faces of S = max(1, size >> m))
// stored as 6 faces of (S + 2) x (S + 2) texels, bordered (r, c) holding
// cube_tap(face, c - 1, r - 1): the face's own texel inside, beyond an edge
// the neighbouring face's edge texel, at a corner the mean of the three
// texels meeting there.  Those are exactly the taps cube_bilinear reads, so a
// lookup becomes four plain loads with the same result.  off[m] is in texels.
struct BorderedCube {
  int size, mips;
  uint32_t off[kMaxLevels];
};

// cube_bilinear on a bordered mip (`mip` at that mip's first texel).  The
// projected coordinates of cube_select satisfy |pc|, |pt| <= 1 (up to the
// reciprocal's rounding), so floor(x), floor(y) lie in [-1, S - 1] and the
// taps (floor + 1, floor + 2) stay inside the bordered face; a NaN direction
// converts to index 0.
__device__ __forceinline__ f3 bcube_bilinear(const float4* __restrict__ mip, int S, int face, float pc, float pt) {
  const float half = 0.5f * static_cast<float>(S);
  const float x = fmaf(pc, half, half - 0.5f), y = fmaf(pt, half, half - 0.5f);
  const float fx = floorf(x), fy = floorf(y);
  const float ax = x - fx, ay = y - fy;
  const int i0 = static_cast<int>(fx), j0 = static_cast<int>(fy);
  const int W = S + 2;
  const float4* r0 = mip + (face * W + j0 + 1) * W + (i0 + 1);
  const f3 t00 = xyz(r0[0]), t10 = xyz(r0[1]), t01 = xyz(r0[W]), t11 = xyz(r0[W + 1]);
  const f3 top = lerp3(t00, t10, ax), bot = lerp3(t01, t11, ax);
  return lerp3(top, bot, ay);
}

__device__ inline f3 cube_sample(const float4* __restrict__ cube, const CubeMips& cm, f3 dir, float lod) {
  int face;
  float pc, pt;
  cube_select(dir, face, pc, pt);
  float d = fminf(fmaxf(lod, 0.0f), static_cast<float>(cm.mips - 1));
  float fl = floorf(d);
  int l0 = static_cast<int>(fl);
  int l1 = min(l0 + 1, cm.mips - 1);
  float delta = d - fl;
  f3 a = cube_bilinear(cube + cm.off[l0], max(1, cm.size >> l0), face, pc, pt);
  if (delta == 0.0f) return a;
  f3 b = cube_bilinear(cube + cm.off[l1], max(1, cm.size >> l1), face, pc, pt);
  return lerp3(a, b, delta);
}

// cube_sample on a bordered copy: the same trilinear value.
__device__ inline f3 bcube_sample(const float4* __restrict__ cube, const BorderedCube& cm, f3 dir, float lod) {
  int face;
  float pc, pt;
  cube_select(dir, face, pc, pt);
  float d = fminf(fmaxf(lod, 0.0f), static_cast<float>(cm.mips - 1));
  float fl = floorf(d);
  int l0 = static_cast<int>(fl);
  int l1 = min(l0 + 1, cm.mips - 1);
  float delta = d - fl;
  f3 a = bcube_bilinear(cube + cm.off[l0], max(1, cm.size >> l0), face, pc, pt);
  if (delta == 0.0f) return a;
  f3 b = bcube_bilinear(cube + cm.off[l1], max(1, cm.size >> l1), face, pc, pt);
  return lerp3(a, b, delta);
}

// ------------------------------------------------------------ wave helpers
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

}  // namespace althea
