// glTF texture sampling shared by the G-buffer subpass and the alpha-tested
// shadow cubes: the Vulkan sampler rules in software, and the per-triangle
// alpha-test record.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "althea/althea.h"

namespace althea {

// Per-triangle alpha-test record of the glTF subpass, written by the setup
// kernel so that the raster's test waits on one record instead of the chain
// primitive -> material -> indices -> coordinates -> texture view: the base
// colour texture coordinates at the three vertices, the alpha factor and
// cut-off, and the base colour texture's view (null texels: no test, a
// pass-through primitive).
struct GbAlpha {
  float2 uv[3];
  float factor, cutoff;
  althea_texture_view tex;
};

// ---------------------------------------------------------------------------
// Texture sampling (Vulkan rules): wrap per integer texel coordinate; LINEAR
// = bilinear at (u w - 0.5, v h - 0.5), NEAREST = texel floor(u w); sRGB
// texels decoded to linear before filtering; the mip level from
// lambda = log2(max(|d(u w, v h)/dx|, |d(u w, v h)/dy|)): lambda <= 0
// magnifies (mag filter, level 0), else the min filter at
// d = min(lambda, mips - 1), the nearest level ceil(d + 0.5) - 1 or the
// linear blend of levels floor(d) and floor(d) + 1.
// ---------------------------------------------------------------------------

// The sRGB decode of every 8-bit value: the transfer function
// (c <= 0.04045 ? c / 12.92 : ((c + 0.055) / 1.055)^2.4, c = k / 255)
// evaluated in FP64 and rounded to FP32, a constant of the code object (no
// upload, so nothing here waits on a stream).
static __device__ const float gb_srgb_lut[256] = {
    0.000000000e+00f, 3.035269910e-04f, 6.070539821e-04f, 9.105809731e-04f, 1.214107964e-03f, 1.517634955e-03f,
    1.821161946e-03f, 2.124688821e-03f, 2.428215928e-03f, 2.731742803e-03f, 3.035269910e-03f, 3.346535843e-03f,
    3.676507389e-03f, 4.024717025e-03f, 4.391442053e-03f, 4.776953254e-03f, 5.181516521e-03f, 5.605391692e-03f,
    6.048833020e-03f, 6.512090564e-03f, 6.995410193e-03f, 7.499032188e-03f, 8.023193106e-03f, 8.568125777e-03f,
    9.134058841e-03f, 9.721217677e-03f, 1.032982301e-02f, 1.096009370e-02f, 1.161224488e-02f, 1.228648797e-02f,
    1.298303250e-02f, 1.370208338e-02f, 1.444384363e-02f, 1.520851441e-02f, 1.599629410e-02f, 1.680737548e-02f,
    1.764195412e-02f, 1.850022003e-02f, 1.938236132e-02f, 2.028856240e-02f, 2.121900953e-02f, 2.217388526e-02f,
    2.315336652e-02f, 2.415763214e-02f, 2.518685907e-02f, 2.624122240e-02f, 2.732089162e-02f, 2.842603996e-02f,
    2.955683507e-02f, 3.071344458e-02f, 3.189603239e-02f, 3.310476616e-02f, 3.433980793e-02f, 3.560131416e-02f,
    3.688944876e-02f, 3.820437193e-02f, 3.954623640e-02f, 4.091519862e-02f, 4.231141135e-02f, 4.373503104e-02f,
    4.518620297e-02f, 4.666508734e-02f, 4.817182571e-02f, 4.970656708e-02f, 5.126945674e-02f, 5.286064744e-02f,
    5.448027700e-02f, 5.612849072e-02f, 5.780543014e-02f, 5.951123685e-02f, 6.124605238e-02f, 6.301001459e-02f,
    6.480326504e-02f, 6.662593782e-02f, 6.847816706e-02f, 7.036009431e-02f, 7.227185369e-02f, 7.421357185e-02f,
    7.618538290e-02f, 7.818742096e-02f, 8.021982014e-02f, 8.228270710e-02f, 8.437620848e-02f, 8.650045842e-02f,
    8.865558356e-02f, 9.084171057e-02f, 9.305896610e-02f, 9.530746937e-02f, 9.758734703e-02f, 9.989872575e-02f,
    1.022417322e-01f, 1.046164855e-01f, 1.070231050e-01f, 1.094617099e-01f, 1.119324267e-01f, 1.144353747e-01f,
    1.169706658e-01f, 1.195384264e-01f, 1.221387759e-01f, 1.247718185e-01f, 1.274376810e-01f, 1.301364750e-01f,
    1.328683197e-01f, 1.356333345e-01f, 1.384316087e-01f, 1.412632912e-01f, 1.441284716e-01f, 1.470272690e-01f,
    1.499597877e-01f, 1.529261470e-01f, 1.559264660e-01f, 1.589608341e-01f, 1.620293707e-01f, 1.651321948e-01f,
    1.682693958e-01f, 1.714411080e-01f, 1.746474057e-01f, 1.778884232e-01f, 1.811642498e-01f, 1.844749898e-01f,
    1.878207773e-01f, 1.912016869e-01f, 1.946178377e-01f, 1.980693191e-01f, 2.015562505e-01f, 2.050787359e-01f,
    2.086368650e-01f, 2.122307569e-01f, 2.158605009e-01f, 2.195262015e-01f, 2.232279629e-01f, 2.269658744e-01f,
    2.307400554e-01f, 2.345505804e-01f, 2.383975685e-01f, 2.422811240e-01f, 2.462013215e-01f, 2.501582801e-01f,
    2.541520894e-01f, 2.581828535e-01f, 2.622506618e-01f, 2.663556039e-01f, 2.704977989e-01f, 2.746773064e-01f,
    2.788942754e-01f, 2.831487358e-01f, 2.874408364e-01f, 2.917706370e-01f, 2.961382568e-01f, 3.005437851e-01f,
    3.049873114e-01f, 3.094689250e-01f, 3.139887154e-01f, 3.185467720e-01f, 3.231432140e-01f, 3.277781010e-01f,
    3.324515224e-01f, 3.371636271e-01f, 3.419144154e-01f, 3.467040658e-01f, 3.515326083e-01f, 3.564001322e-01f,
    3.613067865e-01f, 3.662526011e-01f, 3.712376952e-01f, 3.762621284e-01f, 3.813260198e-01f, 3.864294291e-01f,
    3.915724754e-01f, 3.967552185e-01f, 4.019777775e-01f, 4.072402120e-01f, 4.125426114e-01f, 4.178850651e-01f,
    4.232676625e-01f, 4.286904931e-01f, 4.341536462e-01f, 4.396571815e-01f, 4.452011883e-01f, 4.507857859e-01f,
    4.564110339e-01f, 4.620769918e-01f, 4.677838087e-01f, 4.735314846e-01f, 4.793201685e-01f, 4.851499498e-01f,
    4.910208583e-01f, 4.969329834e-01f, 5.028864741e-01f, 5.088813305e-01f, 5.149176717e-01f, 5.209955573e-01f,
    5.271151066e-01f, 5.332763791e-01f, 5.394794941e-01f, 5.457244515e-01f, 5.520114303e-01f, 5.583403707e-01f,
    5.647115111e-01f, 5.711248517e-01f, 5.775804520e-01f, 5.840784311e-01f, 5.906188488e-01f, 5.972017646e-01f,
    6.038273573e-01f, 6.104955673e-01f, 6.172065735e-01f, 6.239603758e-01f, 6.307571530e-01f, 6.375968456e-01f,
    6.444796920e-01f, 6.514056325e-01f, 6.583748460e-01f, 6.653872728e-01f, 6.724431515e-01f, 6.795424819e-01f,
    6.866853237e-01f, 6.938717365e-01f, 7.011018991e-01f, 7.083757520e-01f, 7.156934738e-01f, 7.230551243e-01f,
    7.304607630e-01f, 7.379103899e-01f, 7.454041839e-01f, 7.529422045e-01f, 7.605245113e-01f, 7.681511641e-01f,
    7.758222222e-01f, 7.835378051e-01f, 7.912979126e-01f, 7.991027236e-01f, 8.069522381e-01f, 8.148465753e-01f,
    8.227857351e-01f, 8.307698965e-01f, 8.387989998e-01f, 8.468732238e-01f, 8.549926281e-01f, 8.631572127e-01f,
    8.713670969e-01f, 8.796223998e-01f, 8.879231215e-01f, 8.962693810e-01f, 9.046611786e-01f, 9.130986333e-01f,
    9.215818644e-01f, 9.301108718e-01f, 9.386857152e-01f, 9.473065138e-01f, 9.559733272e-01f, 9.646862745e-01f,
    9.734452963e-01f, 9.822505713e-01f, 9.911020994e-01f, 1.000000000e+00f,
};

// Power-of-two extents wrap with a mask (the same value as the modulo,
// negative coordinates included), others with an integer modulo.
__device__ __forceinline__ int gb_wrap(int i, int n, int mode) {
  if (mode == ALTHEA_WRAP_CLAMP_TO_EDGE) return min(max(i, 0), n - 1);
  const bool pow2 = (n & (n - 1)) == 0;
  if (mode == ALTHEA_WRAP_MIRRORED_REPEAT) {
    int m;
    if (pow2) {
      m = i & (2 * n - 1);
    } else {
      m = i % (2 * n);
      if (m < 0) m += 2 * n;
    }
    return m < n ? m : 2 * n - 1 - m;
  }
  if (pow2) return i & (n - 1);
  const int m = i % n;
  return m < 0 ? m + n : m;
}

template <bool kAlphaOnly>
__device__ __forceinline__ float4 gb_texel(const althea_texture_view& t, size_t off, int w, int i, int j) {
  const uchar4 c = reinterpret_cast<const uchar4*>(t.texels)[off + static_cast<size_t>(j) * w + i];
  if (kAlphaOnly) return make_float4(0.0f, 0.0f, 0.0f, c.w * (1.0f / 255.0f));
  if (t.srgb) return make_float4(gb_srgb_lut[c.x], gb_srgb_lut[c.y], gb_srgb_lut[c.z], c.w * (1.0f / 255.0f));
  return make_float4(c.x * (1.0f / 255.0f), c.y * (1.0f / 255.0f), c.z * (1.0f / 255.0f), c.w * (1.0f / 255.0f));
}

template <bool kAlphaOnly = false>
__device__ float4 gb_sample_level(const althea_texture_view& t, int level, float u, float v, int filter) {
  size_t off = 0;
  for (int l = 0; l < level; ++l) off += static_cast<size_t>(max(1, t.width >> l)) * max(1, t.height >> l);
  const int w = max(1, t.width >> level), h = max(1, t.height >> level);
  if (filter == ALTHEA_FILTER_NEAREST) {
    const int i = gb_wrap(static_cast<int>(floorf(u * w)), w, t.wrap_u);
    const int j = gb_wrap(static_cast<int>(floorf(v * h)), h, t.wrap_v);
    return gb_texel<kAlphaOnly>(t, off, w, i, j);
  }
  const float x = fmaf(u, static_cast<float>(w), -0.5f), y = fmaf(v, static_cast<float>(h), -0.5f);
  const float fx = floorf(x), fy = floorf(y);
  const float ax = x - fx, ay = y - fy;
  const int i0 = gb_wrap(static_cast<int>(fx), w, t.wrap_u), i1 = gb_wrap(static_cast<int>(fx) + 1, w, t.wrap_u);
  const int j0 = gb_wrap(static_cast<int>(fy), h, t.wrap_v), j1 = gb_wrap(static_cast<int>(fy) + 1, h, t.wrap_v);
  const float4 t00 = gb_texel<kAlphaOnly>(t, off, w, i0, j0), t10 = gb_texel<kAlphaOnly>(t, off, w, i1, j0);
  const float4 t01 = gb_texel<kAlphaOnly>(t, off, w, i0, j1), t11 = gb_texel<kAlphaOnly>(t, off, w, i1, j1);
  auto lerp2 = [&](float a, float b, float c, float d) {
    const float top = fmaf(b - a, ax, a), bot = fmaf(d - c, ax, c);
    return fmaf(bot - top, ay, top);
  };
  return make_float4(lerp2(t00.x, t10.x, t01.x, t11.x), lerp2(t00.y, t10.y, t01.y, t11.y),
                     lerp2(t00.z, t10.z, t01.z, t11.z), lerp2(t00.w, t10.w, t01.w, t11.w));
}

// Texture coordinate and its screen derivatives (forward differences to the
// next pixel to the right and below on the triangle's plane).
struct GbUv {
  float u, v, dudx, dvdx, dudy, dvdy;
};

template <bool kAlphaOnly = false>
__device__ float4 gb_sample(const althea_texture_view& t, const GbUv c) {  // by value: no scratch copy
  const float W = static_cast<float>(t.width), H = static_cast<float>(t.height);
  const float ax = c.dudx * W, bx = c.dvdx * H, ay = c.dudy * W, by = c.dvdy * H;
  const float rho2 = fmaxf(fmaf(ax, ax, bx * bx), fmaf(ay, ay, by * by));
  const float lambda = 0.5f * log2f(rho2);  // log2 of the larger scale factor
  if (!(lambda > 0.0f)) return gb_sample_level<kAlphaOnly>(t, 0, c.u, c.v, t.mag_filter);
  const int q = t.mips - 1;
  if (q == 0) return gb_sample_level<kAlphaOnly>(t, 0, c.u, c.v, t.min_filter);
  const float d = fminf(lambda, static_cast<float>(q));
  if (t.mipmap_mode == ALTHEA_FILTER_NEAREST) {
    const int level = min(max(static_cast<int>(ceilf(d + 0.5f)) - 1, 0), q);
    return gb_sample_level<kAlphaOnly>(t, level, c.u, c.v, t.min_filter);
  }
  const float fl = floorf(d);
  const int l0 = static_cast<int>(fl), l1 = min(l0 + 1, q);
  const float delta = d - fl;
  const float4 a = gb_sample_level<kAlphaOnly>(t, l0, c.u, c.v, t.min_filter);
  if (delta == 0.0f) return a;
  const float4 b = gb_sample_level<kAlphaOnly>(t, l1, c.u, c.v, t.min_filter);
  return make_float4(fmaf(b.x - a.x, delta, a.x), fmaf(b.y - a.y, delta, a.y), fmaf(b.z - a.z, delta, a.z),
                     fmaf(b.w - a.w, delta, a.w));
}

__device__ __forceinline__ GbUv gb_uv_of(float2 a, float2 b, float2 cc, const float l[3], const float lx[3],
                                         const float ly[3]) {
  GbUv c{};
  auto at = [&](const float w[3], float& u, float& v) {
    u = fmaf(w[2], cc.x, fmaf(w[1], b.x, w[0] * a.x));
    v = fmaf(w[2], cc.y, fmaf(w[1], b.y, w[0] * a.y));
  };
  float ux, vx, uy, vy;
  at(l, c.u, c.v);
  at(lx, ux, vx);
  at(ly, uy, vy);
  c.dudx = ux - c.u, c.dvdx = vx - c.v, c.dudy = uy - c.u, c.dvdy = vy - c.v;
  return c;
}

}  // namespace althea
