// C ABI over the FP64 oracle so Python tests can call it through ctypes.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp): loaded by tests/, the smoke check
// and the CPU-baseline leg of bench.py, never by the product library.
#include <omp.h>

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle.hpp"
#include "oracle_mlp.hpp"

using namespace oracle;

extern "C" {

// ----------------------------------------------------------------- core math
void oracle_hash_rng(uint32_t sx, uint32_t sy, int count, uint32_t* out_u, float* out_f) {
  HashRng a{sx, sy}, b{sx, sy};
  for (int i = 0; i < count; ++i) {
    if (out_u) out_u[i] = a.nextu();
    if (out_f) out_f[i] = b.nextf();
  }
}

void oracle_hammersley(uint32_t n, uint32_t first, uint32_t count, uint32_t* out_bits,
                       float* out_xi_f32, double* out_xi_f64) {
  for (uint32_t k = 0; k < count; ++k) {
    uint32_t i = first + k;
    if (out_bits) out_bits[k] = bitreverse32(i);
    if (out_xi_f32) hammersley_f32(i, n, out_xi_f32[2 * k], out_xi_f32[2 * k + 1]);
    if (out_xi_f64) hammersley(i, n, out_xi_f64[2 * k], out_xi_f64[2 * k + 1]);
  }
}

void oracle_frame(const double* raw, int count, double* out9) {
  for (int i = 0; i < count; ++i) {
    Frame f = local_to_world({raw[3 * i], raw[3 * i + 1], raw[3 * i + 2]});
    double* o = out9 + 9 * i;
    o[0] = f.t.x; o[1] = f.t.y; o[2] = f.t.z;
    o[3] = f.b.x; o[4] = f.b.y; o[5] = f.b.z;
    o[6] = f.n.x; o[7] = f.n.y; o[8] = f.n.z;
  }
}

void oracle_ggx_sample(const double* xi, int count, double a2, double* out3) {
  for (int i = 0; i < count; ++i) {
    dvec3 h = sample_ggx(xi[2 * i], xi[2 * i + 1], a2);
    out3[3 * i] = h.x; out3[3 * i + 1] = h.y; out3[3 * i + 2] = h.z;
  }
}

double oracle_ggx_d(double n_dot_h, double a2) { return distribution_ggx(n_dot_h, a2); }

int oracle_mip_count(int w, int h) { return mip_count(w, h); }

// ------------------------------------------------------------ environment
struct OracleEnv {
  Pyramid pyr;
  int w, h;
};

void* oracle_env_create(const float* level0, int w, int h) {
  OracleEnv* e = new OracleEnv;
  e->pyr = build_pyramid(level0, w, h);
  e->w = w;
  e->h = h;
  return e;
}
void oracle_env_destroy(void* env) { delete static_cast<OracleEnv*>(env); }
size_t oracle_env_texels(void* env) { return static_cast<OracleEnv*>(env)->pyr.rgba.size() / 4; }
void oracle_env_pyramid(void* env, double* out) {
  const Pyramid& p = static_cast<OracleEnv*>(env)->pyr;
  std::memcpy(out, p.rgba.data(), p.rgba.size() * sizeof(double));
}

// mode 0: precompute lookup (REPEAT, pole branch, trilinear at lod)
// mode 1: background lookup (CLAMP, atan2, level 0; lod ignored)
void oracle_env_lookup(void* env, const double* dirs, const double* lods, int count, int mode,
                       double* out3) {
  const Pyramid& p = static_cast<OracleEnv*>(env)->pyr;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < count; ++i) {
    dvec3 d{dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]};
    dvec3 r = mode == 0 ? sample_env_precompute(p, d, lods[i]) : sample_env_background(p, d);
    out3[3 * i] = r.x; out3[3 * i + 1] = r.y; out3[3 * i + 2] = r.z;
  }
}

// ------------------------------------------------------------------- cubes
void oracle_cube_texel_dirs(int size, const int* ids, int count, double* out3) {
  for (int i = 0; i < count; ++i) {
    dvec3 d = normalize(cube_texel_raw_dir(ids[3 * i], ids[3 * i + 1], ids[3 * i + 2], size));
    out3[3 * i] = d.x; out3[3 * i + 1] = d.y; out3[3 * i + 2] = d.z;
  }
}
void oracle_cube_select(const double* dirs, int count, int* face, double* st) {
  for (int i = 0; i < count; ++i) {
    cube_select({dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]}, face[i], st[2 * i], st[2 * i + 1]);
  }
}
void oracle_cube_fold(int face, int i, int j, int size, int* out3) {
  cube_fold(face, i, j, size, out3[0], out3[1], out3[2]);
}
void oracle_cube_sample(const float* cube, int size, int mips, const double* dirs, const double* lods,
                        int count, double* out3) {
  CubeView c{cube, size, mips};
#pragma omp parallel for schedule(static)
  for (int i = 0; i < count; ++i) {
    dvec3 r = cube_sample(c, {dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]}, lods[i]);
    out3[3 * i] = r.x; out3[3 * i + 1] = r.y; out3[3 * i + 2] = r.z;
  }
}

// -------------------------------------------------------------- precompute
// ids: (mip, face, x, y) per texel; roughness of mip m is m / (mips - 1)
// (Src/ImageBasedLighting.cpp:383 uses i / 4 for its 5 mips).
// out4: rgb and the NdotL weight sum.
void oracle_prefilter_texels(void* env, int face_size, int mips, int samples, const int* ids,
                             int count, double* out4) {
  const OracleEnv* e = static_cast<OracleEnv*>(env);
  auto rad = [&](dvec3 d, double lod) { return sample_env_precompute(e->pyr, d, lod); };
#pragma omp parallel for schedule(dynamic, 4)
  for (int i = 0; i < count; ++i) {
    int m = ids[4 * i], f = ids[4 * i + 1], x = ids[4 * i + 2], y = ids[4 * i + 3];
    int s = std::max(1, face_size >> m);
    double rough = mips > 1 ? static_cast<double>(m) / (mips - 1) : 0.0;
    double wsum;
    dvec3 r = prefilter_dir(rad, cube_texel_raw_dir(f, x, y, s), rough, samples, e->w, e->h, &wsum);
    out4[4 * i] = r.x; out4[4 * i + 1] = r.y; out4[4 * i + 2] = r.z; out4[4 * i + 3] = wsum;
  }
}

// dir_rough: float4 (raw dir xyz, roughness) per point.
void oracle_prefilter_points(void* env, int samples, const float* dir_rough, int count, double* out3) {
  const OracleEnv* e = static_cast<OracleEnv*>(env);
  auto rad = [&](dvec3 d, double lod) { return sample_env_precompute(e->pyr, d, lod); };
#pragma omp parallel for schedule(dynamic, 4)
  for (int i = 0; i < count; ++i) {
    const float* p = dir_rough + 4 * i;
    dvec3 r = prefilter_dir(rad, {p[0], p[1], p[2]}, p[3], samples, e->w, e->h);
    out3[3 * i] = r.x; out3[3 * i + 1] = r.y; out3[3 * i + 2] = r.z;
  }
}

// ids: (face, x, y) per irradiance texel.
void oracle_irradiance_texels(void* env, int face_size, int theta_samples, const int* ids, int count,
                              double* out3) {
  const OracleEnv* e = static_cast<OracleEnv*>(env);
  auto rad = [&](dvec3 d, double lod) { return sample_env_precompute(e->pyr, d, lod); };
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < count; ++i) {
    dvec3 r = irradiance_dir(rad, cube_texel_raw_dir(ids[3 * i], ids[3 * i + 1], ids[3 * i + 2], face_size),
                             e->w, e->h, theta_samples);
    out3[3 * i] = r.x; out3[3 * i + 1] = r.y; out3[3 * i + 2] = r.z;
  }
}

// Cross-check of the two prefilter estimators at the reference's equirect
// output texels (x, y) of a w x h mip (PreFilterEnvMap.comp:126-177): the
// reference's hash-RNG estimator with `ref_samples` (and its standard error)
// and the Hammersley estimator with `ham_samples` at the same direction.
void oracle_prefilter_crosscheck(void* env, const int* texels, int count, int w, int h, double roughness,
                                 int ref_samples, int ham_samples, double* out_ref3, double* out_se3,
                                 double* out_ham3) {
  const OracleEnv* e = static_cast<OracleEnv*>(env);
  auto rad = [&](dvec3 d, double lod) { return sample_env_precompute(e->pyr, d, lod); };
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < count; ++i) {
    const int x = texels[2 * i], y = texels[2 * i + 1];
    dvec3 N = reference_texel_dir(x, y, w, h);
    dvec3 r = prefilter_reference_estimator(rad, N, roughness, ref_samples, static_cast<uint32_t>(x),
                                            static_cast<uint32_t>(y), e->w, e->h, out_se3 + 3 * i);
    dvec3 q = prefilter_dir(rad, N, roughness, ham_samples, e->w, e->h);
    out_ref3[3 * i] = r.x; out_ref3[3 * i + 1] = r.y; out_ref3[3 * i + 2] = r.z;
    out_ham3[3 * i] = q.x; out_ham3[3 * i + 1] = q.y; out_ham3[3 * i + 2] = q.z;
  }
}

// Prefilter of a pyramid whose level l holds the constant 2^l in every
// channel (a W x H source's levels, filled directly rather than blitted), at
// raw directions: pins the source LOD of PreFilterEnvMap.comp:155-161 through
// the real trilinear lookup, since the output is then a function of the LOD
// alone.
void oracle_prefilter_level_constant(const double* raw_dirs, int count, double roughness, int samples, int env_w,
                                     int env_h, double* out3) {
  Pyramid p;
  p.levels = mip_count(env_w, env_h);
  size_t total = 0;
  for (int l = 0; l < p.levels; ++l) {
    p.w.push_back(std::max(1, env_w >> l));
    p.h.push_back(std::max(1, env_h >> l));
    p.offset.push_back(total);
    total += static_cast<size_t>(p.w[l]) * p.h[l];
  }
  p.rgba.assign(4 * total, 0.0);
  for (int l = 0; l < p.levels; ++l)
    for (size_t t = p.offset[l]; t < p.offset[l] + static_cast<size_t>(p.w[l]) * p.h[l]; ++t)
      for (int c = 0; c < 4; ++c) p.rgba[4 * t + c] = std::ldexp(1.0, l);
  auto rad = [&](dvec3 d, double lod) { return sample_env_precompute(p, d, lod); };
  for (int i = 0; i < count; ++i) {
    dvec3 r = prefilter_dir(rad, {raw_dirs[3 * i], raw_dirs[3 * i + 1], raw_dirs[3 * i + 2]}, roughness, samples,
                            env_w, env_h);
    out3[3 * i] = r.x; out3[3 * i + 1] = r.y; out3[3 * i + 2] = r.z;
  }
}

// Analytic radiance fields for closed-form pins: kind 0 constant (p0,p1,p2),
// kind 1 linear p . dir per channel (p[0..2] for r, p[3..5] g, p[6..8] b).
static dvec3 analytic(int kind, const double* p, dvec3 d) {
  if (kind == 0) return {p[0], p[1], p[2]};
  return {p[0] * d.x + p[1] * d.y + p[2] * d.z, p[3] * d.x + p[4] * d.y + p[5] * d.z,
          p[6] * d.x + p[7] * d.y + p[8] * d.z};
}
void oracle_prefilter_analytic(int kind, const double* params, const double* raw_dirs, int count,
                               double roughness, int samples, int env_w, int env_h, double* out3) {
  auto rad = [&](dvec3 d, double) { return analytic(kind, params, d); };
  for (int i = 0; i < count; ++i) {
    dvec3 r = prefilter_dir(rad, {raw_dirs[3 * i], raw_dirs[3 * i + 1], raw_dirs[3 * i + 2]}, roughness,
                            samples, env_w, env_h);
    out3[3 * i] = r.x; out3[3 * i + 1] = r.y; out3[3 * i + 2] = r.z;
  }
}
void oracle_irradiance_analytic(int kind, const double* params, const double* raw_dirs, int count,
                                int env_w, int env_h, int theta_samples, double* out3) {
  auto rad = [&](dvec3 d, double) { return analytic(kind, params, d); };
  for (int i = 0; i < count; ++i) {
    dvec3 r = irradiance_dir(rad, {raw_dirs[3 * i], raw_dirs[3 * i + 1], raw_dirs[3 * i + 2]}, env_w, env_h,
                             theta_samples);
    out3[3 * i] = r.x; out3[3 * i + 1] = r.y; out3[3 * i + 2] = r.z;
  }
}

// ids: (x, y) LUT texels; texel centre (NdotV, roughness) = ((x+.5)/size, (y+.5)/size).
void oracle_brdf_lut(int size, int samples, const int* ids, int count, double* out2) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < count; ++i) {
    double nv = (ids[2 * i] + 0.5) / size, r = (ids[2 * i + 1] + 0.5) / size;
    brdf_lut_point(nv, r, samples, out2[2 * i], out2[2 * i + 1]);
  }
}
void oracle_brdf_lut_point(double n_dot_v, double roughness, int samples, double* out2) {
  brdf_lut_point(n_dot_v, roughness, samples, out2[0], out2[1]);
}

// ------------------------------------------------------------------ resolve
struct oracle_resolve_args {
  const float* position;
  const float* normal;
  const float* albedo;
  const float* mro;
  int width, height;
  void* env;
  const float* prefiltered;
  int pref_size, pref_mips;
  const float* irradiance;
  int irr_size;
  const float* lut;
  int lut_size;
  const float* lights;  // [n][8]: position xyz, pad, emission xyz, pad
  int light_count;
  const float* projection;
  const float* inverse_projection;
  const float* view;
  const float* inverse_view;
  float exposure;
  int tonemap;
  int ssao_on;
  const float* ssao_kernel;  // [K][4]
  const float* ssao_noise;   // [16][4]
  int ssao_samples;
  float ssao_radius;
  float ssao_bias;
  const float* reflection;  // optional RGBA32F reflection buffer, all levels
  int refl_width, refl_height, refl_mips;
  const float* shadow_cubes;  // optional [light][face][y][x]
  int shadow_size;
};

static ResolveInputs resolve_inputs(const oracle_resolve_args* a, const std::vector<Light>& lights, ReflChain* chain,
                                    ShadowSet* shadows) {
  ResolveInputs in;
  in.position = a->position;
  in.normal = a->normal;
  in.albedo = a->albedo;
  in.mro = a->mro;
  in.W = a->width;
  in.H = a->height;
  in.env = a->env ? &static_cast<OracleEnv*>(a->env)->pyr : nullptr;
  in.prefiltered = {a->prefiltered, a->pref_size, a->pref_mips};
  in.irradiance = {a->irradiance, a->irr_size, 1};
  in.lut = {a->lut, a->lut_size};
  in.lights = lights.data();
  in.light_count = a->light_count;
  in.proj = dmat4::from_float(a->projection);
  in.inv_proj = dmat4::from_float(a->inverse_projection);
  in.view = dmat4::from_float(a->view);
  in.inv_view = dmat4::from_float(a->inverse_view);
  in.exposure = a->exposure;
  in.tonemap = a->tonemap != 0;
  in.ssao_on = a->ssao_on != 0;
  in.ssao = {a->ssao_kernel, a->ssao_noise, a->ssao_samples, a->ssao_radius, a->ssao_bias};
  if (a->reflection) {
    *chain = {a->reflection, a->refl_width, a->refl_height, a->refl_mips};
    in.reflection = chain;
  }
  if (a->shadow_cubes) {
    *shadows = {a->shadow_cubes, a->shadow_size};
    in.shadows = shadows;
  }
  return in;
}

static std::vector<Light> resolve_lights(const oracle_resolve_args* a) {
  std::vector<Light> lights(a->light_count);
  for (int i = 0; i < a->light_count; ++i) {
    const float* l = a->lights + 8 * i;
    lights[i].position = {l[0], l[1], l[2]};
    lights[i].emission = {l[4], l[5], l[6]};
  }
  return lights;
}

// refl_in (optional, 3 per pixel): prefiltered radiance replacing the cube
// lookup where refl_in[3 i] is not NaN; out_query (optional, 4 per pixel):
// reflection direction and roughness of each pixel's lookup.
void oracle_resolve_ex(const oracle_resolve_args* a, const int* pixels, int count, const double* refl_in,
                       double* out3, double* out_ao, double* out_margin, double* out_query) {
  std::vector<Light> lights = resolve_lights(a);
  ReflChain chain;
  ShadowSet shadows;
  ResolveInputs in = resolve_inputs(a, lights, &chain, &shadows);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < count; ++i) {
    double ao, margin;
    dvec3 c = resolve_pixel(in, pixels[2 * i], pixels[2 * i + 1], &ao, &margin, refl_in ? refl_in + 3 * i : nullptr,
                            out_query ? out_query + 4 * i : nullptr);
    out3[3 * i] = c.x; out3[3 * i + 1] = c.y; out3[3 * i + 2] = c.z;
    if (out_ao) out_ao[i] = ao;
    if (out_margin) out_margin[i] = margin;
  }
}

void oracle_resolve(const oracle_resolve_args* a, const int* pixels, int count, double* out3, double* out_ao,
                    double* out_margin) {
  oracle_resolve_ex(a, pixels, count, nullptr, out3, out_ao, out_margin, nullptr);
}

// The reference's SSAO (ssao_reference) and the hemisphere kernel's AO at
// covered pixels[:, (x, y)]; uncovered pixels give NaN in both.
void oracle_ssao_compare(const oracle_resolve_args* a, const int* pixels, int count, double* out_ref,
                         double* out_kernel) {
  std::vector<Light> lights = resolve_lights(a);
  ReflChain chain;
  ShadowSet shadows;
  ResolveInputs in = resolve_inputs(a, lights, &chain, &shadows);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < count; ++i) {
    const int px = pixels[2 * i], py = pixels[2 * i + 1];
    const size_t idx = static_cast<size_t>(py) * in.W + px;
    const float* p = in.position + 4 * idx;
    const float* n = in.normal + 4 * idx;
    if (p[3] == 0.0f) {
      out_ref[i] = out_kernel[i] = std::nan("");
      continue;
    }
    dvec3 P{p[0], p[1], p[2]};
    dvec3 N = normalize({n[0], n[1], n[2]});
    out_ref[i] = ssao_reference(in.position, in.normal, in.W, in.H, px, py, P, N, in.view, in.proj);
    double margin;
    out_kernel[i] = ssao(in.position, in.W, in.H, px, py, P, N, in.view, in.proj, in.ssao, &margin);
  }
}

// Screen-space reflection of pixels[:, (x, y)]: out4 rgba, margin, hit step
// (-1 miss, -2 uncovered), hit uv and hit position (optional, 2 and 3 per
// pixel, NaN without a hit).
void oracle_ssr(const oracle_resolve_args* a, const int* pixels, int count, double* out4, double* out_margin,
                int* out_step, double* out_uv, double* out_pos) {
  std::vector<Light> lights = resolve_lights(a);
  ReflChain chain;
  ShadowSet shadows;
  ResolveInputs in = resolve_inputs(a, lights, &chain, &shadows);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < count; ++i) {
    double margin, uv[2] = {std::nan(""), std::nan("")}, pos[3] = {std::nan(""), std::nan(""), std::nan("")};
    int step;
    ssr_pixel(in, pixels[2 * i], pixels[2 * i + 1], out4 + 4 * i, &margin, &step, uv, pos);
    if (out_pos) for (int c = 0; c < 3; ++c) out_pos[3 * i + c] = pos[c];
    if (out_margin) out_margin[i] = margin;
    if (out_step) out_step[i] = step;
    if (out_uv) { out_uv[2 * i] = uv[0]; out_uv[2 * i + 1] = uv[1]; }
  }
}

// One level of the glossy convolution: dst (dh x dw x 4 doubles) from the
// RGBA32F level src (sh x sw).
void oracle_convolve_level(const float* src, int sw, int sh, int dw, int dh, int vertical, double* dst) {
#pragma omp parallel for schedule(static)
  for (int y = 0; y < dh; ++y)
    for (int x = 0; x < dw; ++x)
      convolve_texel(src, sw, sh, dw, dh, x, y, vertical != 0, dst + 4 * (static_cast<size_t>(y) * dw + x));
}

// textureLod of an RGBA32F reflection buffer at the centres of pixels[:, (x, y)].
void oracle_reflection_sample(const float* data, int W, int H, int mips, const int* pixels, const double* lods,
                              int count, double* out4) {
  ReflChain c{data, W, H, mips};
  for (int i = 0; i < count; ++i) reflection_sample(c, pixels[2 * i], pixels[2 * i + 1], lods[i], out4 + 4 * i);
}

// G-buffer producer at pixels (x, y): out n x 16 doubles, margin, triangle
// (-1 uncovered).  camera: projection, inverse projection, view, inverse
// view (column-major float[16] each), then width and height.
void oracle_gbuffer(const float* const* cam, int width, int height, const float* pos, const float* nrm,
                    const float* col, const float* mro, const uint32_t* idx, int tris, const int* pixels, int count,
                    double* out, double* margin, int* tri) {
  ResolveInputs in{};
  in.W = width;
  in.H = height;
  in.proj = dmat4::from_float(cam[0]);
  in.inv_proj = dmat4::from_float(cam[1]);
  in.view = dmat4::from_float(cam[2]);
  in.inv_view = dmat4::from_float(cam[3]);
  const GbMesh mesh{pos, nrm, col, mro, idx, tris};
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < count; ++i)
    gbuffer_pixel(mesh, in, pixels[2 * i], pixels[2 * i + 1], out + 16 * static_cast<size_t>(i),
                  margin ? margin + i : nullptr, tri ? tri + i : nullptr);
}

// glTF subpass at pixels: prims / mats / texs as the C ABI's structs
// (althea_gltf_primitive, althea_gltf_material, althea_texture_view with host
// texel pointers).
struct OracleGltfPrim {
  const float *pos, *tan, *bit, *nrm;
  const float* uv[4];
  const uint32_t* idx;
  int vertex_count, tris, material;
};
struct OracleTex {
  const uint8_t* texels;
  int w, h, mips, srgb, mag, min, mipmode, wrap_u, wrap_v;
};
// The oracle's primitive and texture tables from the C views (every texture
// a material names).
static void gltf_tables(const OracleGltfPrim* prims, int count, const GltfMaterial* mats, const OracleTex* texs,
                        std::vector<GltfPrim>& ps, std::vector<TexView>& tv) {
  ps.resize(count);
  for (int k = 0; k < count; ++k)
    ps[k] = {prims[k].pos, prims[k].tan, prims[k].bit, prims[k].nrm,
             {prims[k].uv[0], prims[k].uv[1], prims[k].uv[2], prims[k].uv[3]}, prims[k].idx, prims[k].tris,
             prims[k].material};
  int nt = 0;
  for (int k = 0; k < count; ++k) nt = std::max(nt, ps[k].material + 1);
  int ntex = 0;
  for (int k = 0; k < nt; ++k)
    ntex = std::max({ntex, mats[k].base_texture + 1, mats[k].normal_texture + 1, mats[k].metallic_roughness_texture + 1});
  tv.resize(ntex);
  for (int k = 0; k < ntex; ++k)
    tv[k] = {texs[k].texels, texs[k].w, texs[k].h, texs[k].mips, texs[k].srgb, texs[k].mag, texs[k].min,
             texs[k].mipmode, texs[k].wrap_u, texs[k].wrap_v};
}

void oracle_gbuffer_gltf(const float* const* cam, int width, int height, const OracleGltfPrim* prims, int count,
                         const GltfMaterial* mats, const OracleTex* texs, const int* pixels, int n, double* out,
                         double* margin, int* tri) {
  ResolveInputs in{};
  in.W = width;
  in.H = height;
  in.proj = dmat4::from_float(cam[0]);
  in.inv_proj = dmat4::from_float(cam[1]);
  in.view = dmat4::from_float(cam[2]);
  in.inv_view = dmat4::from_float(cam[3]);
  std::vector<GltfPrim> ps;
  std::vector<TexView> tv;
  gltf_tables(prims, count, mats, texs, ps, tv);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; ++i)
    gltf_pixel(ps.data(), count, mats, tv.data(), in, pixels[2 * i], pixels[2 * i + 1],
               out + 16 * static_cast<size_t>(i), margin ? margin + i : nullptr, tri ? tri + i : nullptr);
}

// Texture sampling at n points (u, v, du/dx, dv/dx, du/dy, dv/dy): out n x 4,
// margin n.
void oracle_texture_sample(const OracleTex* t, const double* q, int n, double* out, double* margin) {
  const TexView tv{t->texels, t->w, t->h, t->mips, t->srgb, t->mag, t->min, t->mipmode, t->wrap_u, t->wrap_v};
  for (int i = 0; i < n; ++i) {
    double m = 1e30;
    tex_sample(tv, q[6 * i], q[6 * i + 1], q[6 * i + 2], q[6 * i + 3], q[6 * i + 4], q[6 * i + 5], out + 4 * i, &m);
    if (margin) margin[i] = m;
  }
}

void oracle_texture_mips(uint8_t* texels, int w, int h, int mips, int srgb) { texture_mips(texels, w, h, mips, srgb); }

// Shadow cubes of `light_count` lights ([n][8] floats) from a mesh:
// out [light][face][y][x] doubles.
void oracle_shadow_cubes(const float* positions, const uint32_t* indices, int triangles, const float* lights,
                         int light_count, int size, double* out) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int l = 0; l < light_count; ++l) {
    const float* lp = lights + 8 * l;
    shadow_cube(positions, indices, triangles, {lp[0], lp[1], lp[2]}, size,
                out + static_cast<size_t>(l) * 6 * size * size);
  }
}

// Alpha-tested shadow cubes of glTF primitives: out and margin
// [light][face][y][x] doubles (margin 1e30 where no candidate was tested).
void oracle_shadow_cubes_gltf(const OracleGltfPrim* prims, int count, const GltfMaterial* mats, const OracleTex* texs,
                              const float* lights, int light_count, int size, double* out, double* margin) {
  std::vector<GltfPrim> ps;
  std::vector<TexView> tv;
  gltf_tables(prims, count, mats, texs, ps, tv);
  const size_t per = static_cast<size_t>(6) * size * size;
  for (size_t k = 0; k < per * light_count; ++k) margin[k] = 1e30;
#pragma omp parallel for schedule(dynamic, 1)
  for (int l = 0; l < light_count; ++l) {
    const float* lp = lights + 8 * l;
    shadow_cube_gltf(ps.data(), count, mats, tv.data(), {lp[0], lp[1], lp[2]}, size, out + l * per, margin + l * per);
  }
}

// pbrMaterial for one shading point.  vec: pos, V, N, base, refl, irr (18
// doubles); lights [n][8] floats.
void oracle_pbr_material(const double* vec, double metallic, double roughness, double ao, const float* lut,
                         int lut_size, const float* lights_f, int light_count, double* out3) {
  std::vector<Light> lights(light_count);
  for (int i = 0; i < light_count; ++i) {
    const float* l = lights_f + 8 * i;
    lights[i].position = {l[0], l[1], l[2]};
    lights[i].emission = {l[4], l[5], l[6]};
  }
  auto v3 = [&](int k) { return dvec3{vec[3 * k], vec[3 * k + 1], vec[3 * k + 2]}; };
  dvec3 c = pbr_material(v3(0), v3(1), v3(2), v3(3), v3(4), v3(5), metallic, roughness, ao, Lut{lut, lut_size},
                         lights.data(), light_count);
  out3[0] = c.x; out3[1] = c.y; out3[2] = c.z;
}

// ---------------------------------------------------------------------- mlp
int oracle_mlp_param_count(void) { return mlp::kParams; }

void oracle_mlp_encode(const double* dir_rough, int count, double* out28) {
  for (int i = 0; i < count; ++i) mlp::encode(dir_rough + 4 * i, dir_rough[4 * i + 3], out28 + 28 * i);
}

// out3: log radiance per sample.
void oracle_mlp_forward(const float* params, const float* dir_rough, int count, int bf16, double* out3) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < count; ++i) {
    double dir[3] = {dir_rough[4 * i], dir_rough[4 * i + 1], dir_rough[4 * i + 2]};
    mlp::Activations a;
    mlp::forward(params, dir, dir_rough[4 * i + 3], bf16 != 0, a);
    for (int o = 0; o < 3; ++o) out3[3 * i + o] = a.y[o];
  }
}

}  // extern "C"

template <class P>
static double mlp_loss_grad_all(const P* params, const float* dir_rough, const float* targets, int count, int bf16,
                                double* grad) {
  int threads = omp_get_max_threads();
  std::vector<double> partial(static_cast<size_t>(threads) * mlp::kParams, 0.0);
  std::vector<double> losses(threads, 0.0);
#pragma omp parallel
  {
    int t = omp_get_thread_num(), nt = omp_get_num_threads();
    int lo = static_cast<int>(static_cast<long long>(count) * t / nt);
    int hi = static_cast<int>(static_cast<long long>(count) * (t + 1) / nt);
    if (hi > lo) {
      double* g = &partial[static_cast<size_t>(t) * mlp::kParams];
      losses[t] = mlp::loss_grad(params, dir_rough + 4 * lo, targets + 4 * lo, hi - lo, count,
                                 bf16 != 0, g);
    }
  }
  std::memset(grad, 0, sizeof(double) * mlp::kParams);
  double loss = 0.0;
  for (int t = 0; t < threads; ++t) {
    for (int k = 0; k < mlp::kParams; ++k) grad[k] += partial[static_cast<size_t>(t) * mlp::kParams + k];
    loss += losses[t];
  }
  return loss;
}

extern "C" {

void oracle_adam(double* params, double* m, double* v, const double* grad, int n, int step, double lr,
                 double b1, double b2, double eps) {
  mlp::adam(params, m, v, grad, n, step, lr, b1, b2, eps);
}

double oracle_mlp_loss_grad(const float* params, const float* dir_rough, const float* targets, int count,
                            int bf16, double* grad) {
  return mlp_loss_grad_all(params, dir_rough, targets, count, bf16, grad);
}

// FP64 parameters (exact mode), for finite-difference checks.
double oracle_mlp_loss_grad64(const double* params, const float* dir_rough, const float* targets, int count,
                              double* grad) {
  return mlp_loss_grad_all(params, dir_rough, targets, count, 0, grad);
}

float oracle_bf16_round(float x) { return mlp::bf16_round(x); }

int oracle_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
