// FP32 CPU port of the deferred resolve (the MI355X kernel's arithmetic,
// restated for the host cores with OpenMP): the CPU baseline that bench.py
// times beside the GPU, and a CPU-side check that the kernel's FP32
// formulation with its df32 addressing meets the 1e-4 contract against the
// FP64 oracle.
//
// TEST INFRASTRUCTURE ONLY, like the rest of oracle/: loaded by tests/ and
// bench.py's cpu_baseline leg, never by the product library.
//
// Per pixel (DeferredPass.vert:10-22, DeferredPass.frag:41-94, as
// althea_amd/csrc/resolve_kernels.hip computes it): view ray affine in the
// pixel centre; background equirect lookup (LOD 0, CLAMP_TO_EDGE) addressed in
// df32; reflection = seamless trilinear prefiltered-cube lookup at LOD
// roughness (mips - 1), the mip-0 tap addressed in df32 below LOD 1;
// irradiance cube; 16-tap hemisphere SSAO; pbrMaterial (PBRMaterial.glsl:72-162)
// with the grey-light accumulation and the half-vector refinement near
// V = -L; tonemap.  FP32 G-buffer and FP32 output.
#define __host__
#define __device__
#define __forceinline__ inline
#include "../althea_amd/csrc/df32.hpp"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>

using althea::df32;

namespace port {

struct f3 {
  float x, y, z;
};
static inline f3 mk(float x, float y, float z) { return {x, y, z}; }
static inline f3 add(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline f3 sub(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline f3 scale(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
static inline float dot(f3 a, f3 b) { return std::fma(a.x, b.x, std::fma(a.y, b.y, a.z * b.z)); }
static inline f3 cross(f3 a, f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
static inline float rsq(float x) { return 1.0f / std::sqrt(x); }
static inline f3 normalize(f3 a) { return scale(a, rsq(dot(a, a))); }
static inline f3 lerp3(f3 a, f3 b, float t) {
  return {std::fma(b.x - a.x, t, a.x), std::fma(b.y - a.y, t, a.y), std::fma(b.z - a.z, t, a.z)};
}
static inline f3 texel3(const float* t) { return {t[0], t[1], t[2]}; }

struct df3 {
  df32 x, y, z;
};
using namespace althea;

// ------------------------------------------------------------------ cubes
static void cube_select(f3 d, int& face, float& pc, float& pt) {
  const float ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
  float sc, tc, ma;
  if (ax >= ay && ax >= az) {
    ma = ax;
    if (d.x >= 0.0f) { face = 0; sc = -d.z; tc = -d.y; } else { face = 1; sc = d.z; tc = -d.y; }
  } else if (ay >= az) {
    ma = ay;
    if (d.y >= 0.0f) { face = 2; sc = d.x; tc = d.z; } else { face = 3; sc = d.x; tc = -d.z; }
  } else {
    ma = az;
    if (d.z >= 0.0f) { face = 4; sc = d.x; tc = -d.y; } else { face = 5; sc = -d.x; tc = -d.y; }
  }
  const float inv = 1.0f / ma;
  pc = sc * inv;
  pt = tc * inv;
}

// A tap one texel beyond one edge folded onto the neighbouring face, on the
// integer lattice of half texels (the seamless filter).
static void cube_fold(int face, int i, int j, int S, int& fo, int& io, int& jo) {
  const int sc = 2 * i + 1 - S, tc = 2 * j + 1 - S;
  int p[3];
  switch (face) {
    case 0: p[0] = S; p[1] = -tc; p[2] = -sc; break;
    case 1: p[0] = -S; p[1] = -tc; p[2] = sc; break;
    case 2: p[0] = sc; p[1] = S; p[2] = tc; break;
    case 3: p[0] = sc; p[1] = -S; p[2] = -tc; break;
    case 4: p[0] = sc; p[1] = -tc; p[2] = S; break;
    default: p[0] = -sc; p[1] = -tc; p[2] = -S; break;
  }
  const int major = face >> 1;
  int over = 0;
  for (int a = 0; a < 3; ++a)
    if (std::abs(p[a]) == S + 1) over = a;
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

static f3 cube_tap(const float* mip, int S, int face, int i, int j) {
  auto tx = [&](int f, int a, int b) { return texel3(mip + 4 * ((static_cast<size_t>(f) * S + b) * S + a)); };
  const bool in_i = i >= 0 && i < S, in_j = j >= 0 && j < S;
  if (in_i && in_j) return tx(face, i, j);
  int f, a, b;
  if (in_i || in_j) {
    cube_fold(face, i, j, S, f, a, b);
    return tx(f, a, b);
  }
  const int ic = std::min(std::max(i, 0), S - 1), jc = std::min(std::max(j, 0), S - 1);
  f3 acc = tx(face, ic, jc);
  cube_fold(face, i, jc, S, f, a, b);
  acc = add(acc, tx(f, a, b));
  cube_fold(face, ic, j, S, f, a, b);
  acc = add(acc, tx(f, a, b));
  return scale(acc, 1.0f / 3.0f);
}

static f3 cube_bilinear(const float* mip, int S, int face, float pc, float pt) {
  const float half = 0.5f * static_cast<float>(S);
  const float x = std::fma(pc, half, half - 0.5f), y = std::fma(pt, half, half - 0.5f);
  const float fx = std::floor(x), fy = std::floor(y);
  const int i0 = static_cast<int>(fx), j0 = static_cast<int>(fy);
  const f3 top = lerp3(cube_tap(mip, S, face, i0, j0), cube_tap(mip, S, face, i0 + 1, j0), x - fx);
  const f3 bot = lerp3(cube_tap(mip, S, face, i0, j0 + 1), cube_tap(mip, S, face, i0 + 1, j0 + 1), x - fx);
  return lerp3(top, bot, y - fy);
}

static f3 weighted4(f3 t00, f3 t10, f3 t01, f3 t11, const TexelSplit& sx, const TexelSplit& sy) {
  const f3 top = mk(std::fma(t10.x, sx.a, t00.x * sx.b), std::fma(t10.y, sx.a, t00.y * sx.b),
                    std::fma(t10.z, sx.a, t00.z * sx.b));
  const f3 bot = mk(std::fma(t11.x, sx.a, t01.x * sx.b), std::fma(t11.y, sx.a, t01.y * sx.b),
                    std::fma(t11.z, sx.a, t01.z * sx.b));
  return mk(std::fma(bot.x, sy.a, top.x * sy.b), std::fma(bot.y, sy.a, top.y * sy.b),
            std::fma(bot.z, sy.a, top.z * sy.b));
}

static f3 cube_bilinear_df(const float* mip, int S, const df3& R) {
  const float ax = std::fabs(R.x.hi), ay = std::fabs(R.y.hi), az = std::fabs(R.z.hi);
  int face;
  df32 sc, tc, ma;
  if (ax >= ay && ax >= az) {
    ma = df_abs(R.x);
    if (R.x.hi >= 0.0f) { face = 0; sc = df_neg(R.z); tc = df_neg(R.y); } else { face = 1; sc = R.z; tc = df_neg(R.y); }
  } else if (ay >= az) {
    ma = df_abs(R.y);
    if (R.y.hi >= 0.0f) { face = 2; sc = R.x; tc = R.z; } else { face = 3; sc = R.x; tc = df_neg(R.z); }
  } else {
    ma = df_abs(R.z);
    if (R.z.hi >= 0.0f) { face = 4; sc = R.x; tc = df_neg(R.y); } else { face = 5; sc = df_neg(R.x); tc = df_neg(R.y); }
  }
  const float half = 0.5f * static_cast<float>(S);
  const TexelSplit sx = df_texel(df_add_f(df_mul_f(df_div(sc, ma), half), half - 0.5f));
  const TexelSplit sy = df_texel(df_add_f(df_mul_f(df_div(tc, ma), half), half - 0.5f));
  return weighted4(cube_tap(mip, S, face, sx.i, sy.i), cube_tap(mip, S, face, sx.i + 1, sy.i),
                   cube_tap(mip, S, face, sx.i, sy.i + 1), cube_tap(mip, S, face, sx.i + 1, sy.i + 1), sx, sy);
}

// ------------------------------------------------------------------ inputs
struct Inputs {
  const float *position, *normal, *albedo, *mro;  // RGBA32F planes
  int W, H;
  const float* env;  // equirect level 0, RGBA32F
  int env_w, env_h;
  const float* prefiltered;  // [mip][face][y][x] RGBA32F
  int pref_size, pref_mips;
  const float* irradiance;
  int irr_size;
  const float* lut;  // RG32F, row = roughness
  int lut_size;
  const float* lights;  // [n][8]
  int light_count;
  const float* view;      // column-major 4x4
  const float* proj_view; // projection * view, column-major
  float ray0[3], rayx[3], rayy[3];
  float ray_df[3][3][2];  // (ray0, rayx, rayy)[component] as (hi, lo)
  float exposure;
  int tonemap;
  int ssao_on, ssao_samples;
  float ssao_radius, ssao_bias;
  const float* ssao_kernel;  // [K][4]
  const float* ssao_noise;   // [16][4]
  unsigned fp32_steps;       // PORT_FP32_* : steps addressed in plain FP32 (error attribution)
};

// Error attribution: each flag puts one df32 step back to plain FP32.
constexpr unsigned kFp32Background = 1u;  // background equirect lookup
constexpr unsigned kFp32Reflection = 2u;  // mip-0 reflection tap
constexpr unsigned kFp32HalfVector = 4u;  // no half-vector refinement near V = -L

static float atan_poly(float t) {  // the kernels' FP32 atan on [0, 1]
  const float s = t * t;
  float p = 2.456719056e-03f;
  p = std::fma(p, s, -1.440133527e-02f);
  p = std::fma(p, s, 3.978118673e-02f);
  p = std::fma(p, s, -7.234854251e-02f);
  p = std::fma(p, s, 1.049894467e-01f);
  p = std::fma(p, s, -1.416122913e-01f);
  p = std::fma(p, s, 1.998590678e-01f);
  p = std::fma(p, s, -3.333259821e-01f);
  p = std::fma(p, s, 9.999998808e-01f);
  return p * t;
}
static float atan2_f32(float y, float x) {
  const float ax = std::fabs(x), ay = std::fabs(y);
  const float mx = std::max(ax, ay), mn = std::min(ax, ay);
  float r = atan_poly(mx > 0.0f ? mn / mx : 0.0f);
  r = ay > ax ? 1.570796327f - r : r;
  r = x < 0.0f ? 3.141592654f - r : r;
  return std::copysign(r, y);
}

static df3 view_ray_df(const Inputs& in, float cx, float cy) {
  auto comp = [&](int k) {
    const df32 r0{in.ray_df[0][k][0], in.ray_df[0][k][1]}, rx{in.ray_df[1][k][0], in.ray_df[1][k][1]};
    const df32 ry{in.ray_df[2][k][0], in.ray_df[2][k][1]};
    return df_add(df_add(r0, df_mul_f(rx, cx)), df_mul_f(ry, cy));
  };
  return {comp(0), comp(1), comp(2)};
}

static df3 normalize_df(const df3& v) {
  const df32 n2 = df_add(df_add(df_mul(v.x, v.x), df_mul(v.y, v.y)), df_mul(v.z, v.z));
  const df32 inv = df_div(df_of(1.0f), df_sqrt(n2));
  return {df_mul(v.x, inv), df_mul(v.y, inv), df_mul(v.z, inv)};
}

static float smoothstep01(float x) {
  const float t = std::min(std::max(x, 0.0f), 1.0f);
  return t * t * std::fma(-2.0f, t, 3.0f);
}

static float view_depth(const Inputs& in, f3 q) {
  const float* v = in.view;
  return -std::fma(v[2], q.x, std::fma(v[6], q.y, std::fma(v[10], q.z, v[14])));
}

static float ssao(const Inputs& in, int px, int py, f3 P, f3 N) {
  const float* nz = in.ssao_noise + 4 * ((px & 3) + 4 * (py & 3));
  const f3 r = mk(nz[0], nz[1], nz[2]);
  const f3 T = normalize(sub(r, scale(N, dot(r, N))));
  const f3 B = cross(N, T);
  const float frag_depth = view_depth(in, P);
  const float fw = static_cast<float>(in.W), fh = static_cast<float>(in.H);
  const float* vp = in.proj_view;
  float occ = 0.0f;
  for (int k = 0; k < in.ssao_samples; ++k) {
    const float* kk = in.ssao_kernel + 4 * k;
    const f3 S = add(P, scale(add(add(scale(T, kk[0]), scale(B, kk[1])), scale(N, kk[2])), in.ssao_radius));
    const float sd = view_depth(in, S);
    const float cx = std::fma(vp[0], S.x, std::fma(vp[4], S.y, std::fma(vp[8], S.z, vp[12])));
    const float cy = std::fma(vp[1], S.x, std::fma(vp[5], S.y, std::fma(vp[9], S.z, vp[13])));
    const float cw = std::fma(vp[3], S.x, std::fma(vp[7], S.y, std::fma(vp[11], S.z, vp[15])));
    if (!(cw > 0.0f)) continue;
    const float sx = (0.5f * cx / cw + 0.5f) * fw, sy = (0.5f * cy / cw + 0.5f) * fh;
    if (!(sx >= 0.0f && sx < fw && sy >= 0.0f && sy < fh)) continue;
    const int qx = std::min(static_cast<int>(sx), in.W - 1), qy = std::min(static_cast<int>(sy), in.H - 1);
    const float* q = in.position + 4 * (static_cast<size_t>(qy) * in.W + qx);
    if (q[3] == 0.0f) continue;
    const float qd = view_depth(in, mk(q[0], q[1], q[2]));
    if (qd <= sd - in.ssao_bias) occ += smoothstep01(in.ssao_radius / std::fabs(frag_depth - qd));
  }
  return std::fma(-occ, 1.0f / static_cast<float>(in.ssao_samples), 1.0f);
}

static void lut_sample(const Inputs& in, float u, float v, float& A, float& B) {
  const int n = in.lut_size;
  const float x = std::fma(u, static_cast<float>(n), -0.5f), y = std::fma(v, static_cast<float>(n), -0.5f);
  const float fx = std::floor(x), fy = std::floor(y);
  const float ax = x - fx, ay = y - fy;
  const int ix = static_cast<int>(fx), iy = static_cast<int>(fy);
  const int x0 = std::min(std::max(ix, 0), n - 1), x1 = std::min(std::max(ix + 1, 0), n - 1);
  const int y0 = std::min(std::max(iy, 0), n - 1), y1 = std::min(std::max(iy + 1, 0), n - 1);
  const float* t00 = in.lut + 2 * (y0 * n + x0);
  const float* t10 = in.lut + 2 * (y0 * n + x1);
  const float* t01 = in.lut + 2 * (y1 * n + x0);
  const float* t11 = in.lut + 2 * (y1 * n + x1);
  const float ta = std::fma(t10[0] - t00[0], ax, t00[0]), ba = std::fma(t11[0] - t01[0], ax, t01[0]);
  const float tb = std::fma(t10[1] - t00[1], ax, t00[1]), bb = std::fma(t11[1] - t01[1], ax, t01[1]);
  A = std::fma(ba - ta, ay, ta);
  B = std::fma(bb - tb, ay, tb);
}

static size_t cube_mip_offset(int size, int m) {
  size_t off = 0;
  for (int k = 0; k < m; ++k) {
    const size_t s = static_cast<size_t>(std::max(1, size >> k));
    off += 6 * s * s;
  }
  return off;
}

static f3 resolve_pixel(const Inputs& in, int px, int py) {
  const size_t idx = static_cast<size_t>(py) * in.W + px;
  const float cxp = static_cast<float>(px) + 0.5f, cyp = static_cast<float>(py) + 0.5f;
  const f3 dir = mk(std::fma(in.rayy[0], cyp, std::fma(in.rayx[0], cxp, in.ray0[0])),
                    std::fma(in.rayy[1], cyp, std::fma(in.rayx[1], cxp, in.ray0[1])),
                    std::fma(in.rayy[2], cyp, std::fma(in.rayx[2], cxp, in.ray0[2])));
  const float* pos = in.position + 4 * idx;
  auto tonemap = [&](f3 c) {
    return mk(1.0f - std::exp(-c.x * in.exposure), 1.0f - std::exp(-c.y * in.exposure),
              1.0f - std::exp(-c.z * in.exposure));
  };
  if (pos[3] == 0.0f && (in.fp32_steps & kFp32Background)) {
    const float yaw = atan2_f32(dir.z, dir.x), pitch = atan2_f32(dir.y, std::sqrt(dir.x * dir.x + dir.z * dir.z));
    const float fw = static_cast<float>(in.env_w), fh = static_cast<float>(in.env_h);
    const float x = std::fma(yaw, fw * 0.159154943f, 0.5f * fw - 0.5f);
    const float y = std::fma(pitch, -fh * 0.318309886f, 0.5f * fh - 0.5f);
    const float fx = std::floor(x), fy = std::floor(y);
    auto tx = [&](int i, int j) {
      i = std::min(std::max(i, 0), in.env_w - 1);
      j = std::min(std::max(j, 0), in.env_h - 1);
      return texel3(in.env + 4 * (static_cast<size_t>(j) * in.env_w + i));
    };
    const int i0 = static_cast<int>(fx), j0 = static_cast<int>(fy);
    const f3 c = lerp3(lerp3(tx(i0, j0), tx(i0 + 1, j0), x - fx), lerp3(tx(i0, j0 + 1), tx(i0 + 1, j0 + 1), x - fx),
                       y - fy);
    return in.tonemap ? tonemap(c) : c;
  }
  if (pos[3] == 0.0f) {
    const df3 d = view_ray_df(in, cxp, cyp);
    const df32 yaw = df_atan2(d.z, d.x), pitch = df_atan2(d.y, df_sqrt(df_sum_squares(d.x, d.z)));
    const float fw = static_cast<float>(in.env_w), fh = static_cast<float>(in.env_h);
    const TexelSplit sx = df_texel(df_add_f(df_mul(yaw, df32{kDfInv2Pi.hi * fw, kDfInv2Pi.lo * fw}), 0.5f * fw - 0.5f));
    const TexelSplit sy =
        df_texel(df_add_f(df_mul(pitch, df32{-kDfInvPi.hi * fh, -kDfInvPi.lo * fh}), 0.5f * fh - 0.5f));
    auto tx = [&](int i, int j) {
      i = std::min(std::max(i, 0), in.env_w - 1);
      j = std::min(std::max(j, 0), in.env_h - 1);
      return texel3(in.env + 4 * (static_cast<size_t>(j) * in.env_w + i));
    };
    const f3 c = weighted4(tx(sx.i, sy.i), tx(sx.i + 1, sy.i), tx(sx.i, sy.i + 1), tx(sx.i + 1, sy.i + 1), sx, sy);
    return in.tonemap ? tonemap(c) : c;
  }
  const float* nr = in.normal + 4 * idx;
  const float* al = in.albedo + 4 * idx;
  const float* mr = in.mro + 4 * idx;
  const f3 P = mk(pos[0], pos[1], pos[2]);
  const f3 N = normalize(mk(nr[0], nr[1], nr[2]));
  const float metallic = mr[0], roughness = mr[1];
  float ao = mr[2];
  const f3 I = normalize(dir);
  const float NI = dot(N, I);
  const float nd = 2.0f * NI;
  const f3 R = mk(std::fma(-nd, N.x, I.x), std::fma(-nd, N.y, I.y), std::fma(-nd, N.z, I.z));
  // reflection
  const float lod = roughness * static_cast<float>(in.pref_mips - 1);
  const float d = std::min(std::max(lod, 0.0f), static_cast<float>(in.pref_mips - 1));
  f3 refl;
  int face;
  float pc, pt;
  cube_select(R, face, pc, pt);
  if (d < 1.0f && !(in.fp32_steps & kFp32Reflection)) {
    const df3 v = view_ray_df(in, cxp, cyp);
    const f3 n = mk(nr[0], nr[1], nr[2]);
    const df32 nn = df_add(df_add(df_two_prod(n.x, n.x), df_two_prod(n.y, n.y)), df_two_prod(n.z, n.z));
    const df32 nv = df_add(df_add(df_mul_f(v.x, n.x), df_mul_f(v.y, n.y)), df_mul_f(v.z, n.z));
    const df32 two_nv{2.0f * nv.hi, 2.0f * nv.lo};
    const df3 Rd{df_sub(df_mul(v.x, nn), df_mul_f(two_nv, n.x)), df_sub(df_mul(v.y, nn), df_mul_f(two_nv, n.y)),
                 df_sub(df_mul(v.z, nn), df_mul_f(two_nv, n.z))};
    refl = cube_bilinear_df(in.prefiltered, in.pref_size, Rd);
    if (d > 0.0f) {
      const f3 b = cube_bilinear(in.prefiltered + 4 * cube_mip_offset(in.pref_size, 1), std::max(1, in.pref_size >> 1),
                                 face, pc, pt);
      const float w0 = 1.0f - d;
      refl = mk(std::fma(b.x, d, refl.x * w0), std::fma(b.y, d, refl.y * w0), std::fma(b.z, d, refl.z * w0));
    }
  } else {
    const float fl = std::floor(d);
    const int l0 = static_cast<int>(fl), l1 = std::min(l0 + 1, in.pref_mips - 1);
    refl = cube_bilinear(in.prefiltered + 4 * cube_mip_offset(in.pref_size, l0), std::max(1, in.pref_size >> l0),
                         face, pc, pt);
    if (d != fl)
      refl = lerp3(refl, cube_bilinear(in.prefiltered + 4 * cube_mip_offset(in.pref_size, l1),
                                       std::max(1, in.pref_size >> l1), face, pc, pt),
                   d - fl);
  }
  cube_select(N, face, pc, pt);
  const f3 irr = cube_bilinear(in.irradiance, in.irr_size, face, pc, pt);
  if (in.ssao_on) ao = ssao(in, px, py, P, N);
  // pbrMaterial: environment term
  const float n_dot_v = std::max(-NI, 0.0f);
  const f3 base = mk(al[0], al[1], al[2]);
  const f3 F0 = mk(std::fma(metallic, base.x - 0.04f, 0.04f), std::fma(metallic, base.y - 0.04f, 0.04f),
                   std::fma(metallic, base.z - 0.04f, 0.04f));
  const float a = roughness * roughness, a2 = a * a, kdir = (a + 1.0f) * (a + 1.0f) * 0.125f;
  const float om_r = 1.0f - roughness;
  const f3 dF = mk(std::max(om_r, F0.x) - F0.x, std::max(om_r, F0.y) - F0.y, std::max(om_r, F0.z) - F0.z);
  const f3 kd = scale(base, 1.0f - metallic);
  float A, B;
  lut_sample(in, n_dot_v, roughness, A, B);
  const float fx = std::pow(1.0f - n_dot_v, 5.0f);
  const f3 Fv = mk(std::fma(dF.x, fx, F0.x), std::fma(dF.y, fx, F0.y), std::fma(dF.z, fx, F0.z));
  f3 color = mk((irr.x * ((1.0f - Fv.x) * kd.x) + refl.x * std::fma(Fv.x, A, B)) * ao,
                (irr.y * ((1.0f - Fv.y) * kd.y) + refl.y * std::fma(Fv.y, A, B)) * ao,
                (irr.z * ((1.0f - Fv.z) * kd.z) + refl.z * std::fma(Fv.z, A, B)) * ao);
  // point lights
  const float gv = n_dot_v / std::fma(n_dot_v, 1.0f - kdir, kdir);
  const float c1 = a2 * gv * 0.318309886f;
  const f3 kdp = scale(kd, 0.318309886f);
  for (int i = 0; i < in.light_count; ++i) {
    const float* lt = in.lights + 8 * i;
    const f3 L = mk(lt[0] - P.x, lt[1] - P.y, lt[2] - P.z);
    const float nl_raw = dot(N, L);
    if (!(nl_raw > 0.0f)) continue;
    const float d2 = dot(L, L), inv_d = rsq(d2);
    const float h2 = std::fma(2.0f * dot(I, L), inv_d, 2.0f);
    const float n_dot_l = nl_raw * inv_d;
    float n_dot_h = std::max(std::fma(nl_raw, inv_d, NI) * rsq(h2), 0.0f);
    if (h2 < 1e-2f && !(in.fp32_steps & kFp32HalfVector)) {
      if (h2 >= 1e-4f) {
        const f3 h = mk(std::fma(L.x, inv_d, I.x), std::fma(L.y, inv_d, I.y), std::fma(L.z, inv_d, I.z));
        n_dot_h = std::max(dot(N, h) * rsq(dot(h, h)), 0.0f);
      } else {
        const df3 Lu = normalize_df({df_two_sum(lt[0], -P.x), df_two_sum(lt[1], -P.y), df_two_sum(lt[2], -P.z)});
        const df3 Vu = normalize_df(view_ray_df(in, cxp, cyp));
        const df3 h{df_add(Vu.x, Lu.x), df_add(Vu.y, Lu.y), df_add(Vu.z, Lu.z)};
        const df32 hh = df_add(df_add(df_mul(h.x, h.x), df_mul(h.y, h.y)), df_mul(h.z, h.z));
        const df32 nh = df_add(df_add(df_mul_f(h.x, N.x), df_mul_f(h.y, N.y)), df_mul_f(h.z, N.z));
        n_dot_h = std::max(df_div(nh, df_sqrt(hh)).hi, 0.0f);
      }
    }
    const float x5 = std::pow(1.0f - n_dot_h, 5.0f);
    const float t = std::fma(n_dot_h * n_dot_h, a2 - 1.0f, 1.0f);
    const float s = c1 * n_dot_l / ((t * t) * std::fma(n_dot_l, 1.0f - kdir, kdir) * std::fma(n_dot_l, 4.0f * n_dot_v, 1e-4f));
    const float w = (inv_d * inv_d) * n_dot_l;
    const f3 F = mk(std::fma(dF.x, x5, F0.x), std::fma(dF.y, x5, F0.y), std::fma(dF.z, x5, F0.z));
    color.x = std::fma(std::fma(F.x, s, std::fma(-F.x, kdp.x, kdp.x)) * lt[4], w, color.x);
    color.y = std::fma(std::fma(F.y, s, std::fma(-F.y, kdp.y, kdp.y)) * lt[5], w, color.y);
    color.z = std::fma(std::fma(F.z, s, std::fma(-F.z, kdp.z, kdp.z)) * lt[6], w, color.z);
  }
  return in.tonemap ? tonemap(color) : color;
}

}  // namespace port

extern "C" {

// The port's inputs: the oracle's resolve arguments (oracle_capi.cpp) plus
// the affine view ray in FP32 and as df32 pairs, projection * view and the
// env level 0 (the first texels of the pyramid).
struct port_resolve_args {
  const float *position, *normal, *albedo, *mro;
  int width, height;
  const float* env;
  int env_w, env_h;
  const float* prefiltered;
  int pref_size, pref_mips;
  const float* irradiance;
  int irr_size;
  const float* lut;
  int lut_size;
  const float* lights;
  int light_count;
  const float* view;
  const float* proj_view;
  const double* ray_affine;  // o[3], dx[3], dy[3] in FP64
  float exposure;
  int tonemap;
  int ssao_on, ssao_samples;
  float ssao_radius, ssao_bias;
  const float* ssao_kernel;
  const float* ssao_noise;
};

// Resolve pixels[:, (x, y)] (count of them; pixels NULL: the whole frame in
// row order) into out_rgba, with `threads` OpenMP threads (0: all).
// fp32_steps: 0 for the kernel's arithmetic; bits 1 / 2 / 4 put the
// background, the mip-0 reflection tap or the half vector back to FP32.
void port_resolve(const port_resolve_args* a, const int* pixels, int count, int threads, unsigned fp32_steps,
                  float* out_rgba) {
  port::Inputs in{};
  in.position = a->position;
  in.normal = a->normal;
  in.albedo = a->albedo;
  in.mro = a->mro;
  in.W = a->width;
  in.H = a->height;
  in.env = a->env;
  in.env_w = a->env_w;
  in.env_h = a->env_h;
  in.prefiltered = a->prefiltered;
  in.pref_size = a->pref_size;
  in.pref_mips = a->pref_mips;
  in.irradiance = a->irradiance;
  in.irr_size = a->irr_size;
  in.lut = a->lut;
  in.lut_size = a->lut_size;
  in.lights = a->lights;
  in.light_count = a->light_count;
  in.view = a->view;
  in.proj_view = a->proj_view;
  for (int r = 0; r < 3; ++r) {
    const double c[3] = {a->ray_affine[r], a->ray_affine[3 + r], a->ray_affine[6 + r]};
    float* f[3] = {in.ray0, in.rayx, in.rayy};
    for (int k = 0; k < 3; ++k) {
      f[k][r] = static_cast<float>(c[k]);
      in.ray_df[k][r][0] = static_cast<float>(c[k]);
      in.ray_df[k][r][1] = static_cast<float>(c[k] - static_cast<double>(in.ray_df[k][r][0]));
    }
  }
  in.exposure = a->exposure;
  in.tonemap = a->tonemap;
  in.ssao_on = a->ssao_on;
  in.ssao_samples = a->ssao_samples;
  in.ssao_radius = a->ssao_radius;
  in.ssao_bias = a->ssao_bias;
  in.ssao_kernel = a->ssao_kernel;
  in.ssao_noise = a->ssao_noise;
  in.fp32_steps = fp32_steps;
  const long n = pixels ? count : static_cast<long>(in.W) * in.H;
  const int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(nt)
  for (long i = 0; i < n; ++i) {
    const int px = pixels ? pixels[2 * i] : static_cast<int>(i % in.W);
    const int py = pixels ? pixels[2 * i + 1] : static_cast<int>(i / in.W);
    const port::f3 c = port::resolve_pixel(in, px, py);
    float* o = out_rgba + 4 * i;
    o[0] = c.x;
    o[1] = c.y;
    o[2] = c.z;
    o[3] = 1.0f;
  }
}

}  // extern "C"
