// C ABI entry points: argument validation, handle management and workspace
// plumbing around the kernel launchers.  No exception crosses this boundary.
#include <cstddef>
#include <cstdint>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <new>
#include <string>
#include <vector>

#include "althea/althea.h"
#include "internal.hpp"
#include "mlp_view.hpp"

namespace althea {

// The uniform and light layouts are the reference's byte for byte, so a
// consumer passes its existing buffers: GlobalUniforms
// (Include/Althea/GlobalUniforms.h:15-31, six mat4 then the scalars, 416 B)
// and PointLight (Include/Althea/PointLight.h:31-34, two alignas(16) vec3).
static_assert(sizeof(althea_globals) == 416, "GlobalUniforms is 416 bytes");
static_assert(offsetof(althea_globals, inverse_projection) == 64 && offsetof(althea_globals, view) == 128 &&
                  offsetof(althea_globals, prev_view) == 192 && offsetof(althea_globals, inverse_view) == 256 &&
                  offsetof(althea_globals, prev_inverse_view) == 320 && offsetof(althea_globals, mouse_uv) == 384,
              "GlobalUniforms matrices");
static_assert(offsetof(althea_globals, light_count) == 392 && offsetof(althea_globals, light_buffer_handle) == 396 &&
                  offsetof(althea_globals, time) == 400 && offsetof(althea_globals, exposure) == 404 &&
                  offsetof(althea_globals, input_mask) == 408 && offsetof(althea_globals, frame_count) == 412,
              "GlobalUniforms scalars");
static_assert(sizeof(althea_point_light) == 32 && offsetof(althea_point_light, position) == 0 &&
                  offsetof(althea_point_light, emission) == 16,
              "PointLight: alignas(16) position and emission");

namespace {
thread_local char g_error[512] = "";
}

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_error, sizeof(g_error), fmt, ap);
  va_end(ap);
}

althea_status check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s launch failed: %s", what, hipGetErrorString(e));
    return ALTHEA_ERROR_HIP;
  }
  return ALTHEA_OK;
}

PyramidGeometry pyramid_geometry(int width, int height) {
  PyramidGeometry g{};
  g.levels = althea_env_levels(width, height);
  size_t total = 0;
  for (int l = 0; l < g.levels; ++l) {
    g.w[l] = std::max(1, width >> l);
    g.h[l] = std::max(1, height >> l);
    g.off[l] = static_cast<uint32_t>(total);
    total += static_cast<size_t>(g.w[l]) * g.h[l];
  }
  g.texels = total;
  return g;
}

size_t cube_mip_offset(int size, int mip) {
  size_t off = 0;
  for (int m = 0; m < mip; ++m) {
    size_t s = static_cast<size_t>(std::max(1, size >> m));
    off += 6 * s * s;
  }
  return off;
}

// Stream-ordered workspace: allocated and released on the caller's stream.
class Workspace {
 public:
  Workspace(size_t bytes, hipStream_t s) : s_(s) {
    if (bytes == 0) return;
    e_ = hipMallocAsync(&ptr_, bytes, s);
    if (e_ != hipSuccess) {
      ptr_ = nullptr;
      set_error("hipMallocAsync(%zu) failed: %s", bytes, hipGetErrorString(e_));
    }
  }
  ~Workspace() {
    if (ptr_) (void)hipFreeAsync(ptr_, s_);
  }
  bool ok() const { return e_ == hipSuccess; }
  void* get() const { return ptr_; }

 private:
  void* ptr_ = nullptr;
  hipStream_t s_;
  hipError_t e_ = hipSuccess;
};

// Power-of-two faces up to 16384 (texel offsets of a cube and of its bordered
// copy fit 32 bits).
static bool cube_ok(int size, int mips) {
  return is_pow2(size) && size <= 16384 && mips >= 1 && mips <= 16 && (size >> (mips - 1)) >= 1;
}

// Power-of-two extents up to 2^15, at most 2^27 texels (16384 x 8192): every
// level's byte offsets, padded or not, fit 32 bits.
static bool env_dims_ok(int w, int h) {
  return is_pow2(w) && is_pow2(h) && w <= (1 << 15) && h <= (1 << 15) &&
         static_cast<int64_t>(w) * h <= (int64_t{1} << 27);
}

}  // namespace althea

using namespace althea;

// ---------------------------------------------------------------- handles
struct althea_ibl {
  althea_ibl_desc desc{};
  int env_w = 0, env_h = 0;
  float* env = nullptr;  // pyramid
  size_t env_texels = 0;
  float* prefiltered = nullptr;
  float* irradiance = nullptr;
  float* lut = nullptr;
  // Seamless-bordered copies of the two cubes (althea_cube_border), rewritten
  // after every precompute or cache load; the resolve and the SSR capture read them.
  float* prefiltered_bordered = nullptr;
  float* irradiance_bordered = nullptr;
  void* workspace = nullptr;
  size_t workspace_bytes = 0;
};

static althea_status resolve_args_ok(const althea_gbuffer_view* g, const althea_ibl_views* ibl,
                                     const althea_point_light* d_lights, int light_count,
                                     const althea_globals* globals, const althea_ssao_desc* ssao, void* d_out) {
  ALTHEA_REQUIRE(g && ibl && globals && d_out, "deferred_resolve: null argument");
  ALTHEA_REQUIRE(g->position && g->normal && g->albedo && g->mro, "deferred_resolve: null G-buffer plane");
  ALTHEA_REQUIRE(g->width >= 1 && g->height >= 1 && g->width <= 32768 && g->height <= 32768,
                 "deferred_resolve: extent %dx%d", g->width, g->height);
  ALTHEA_REQUIRE(g->format == ALTHEA_GBUFFER_FP32 || g->format == ALTHEA_GBUFFER_FP16,
                 "deferred_resolve: G-buffer format %d", g->format);
  ALTHEA_REQUIRE(ibl->environment && ibl->prefiltered && ibl->irradiance && ibl->brdf_lut,
                 "deferred_resolve: incomplete IBL views");
  ALTHEA_REQUIRE(env_dims_ok(ibl->env_width, ibl->env_height), "deferred_resolve: environment %dx%d",
                 ibl->env_width, ibl->env_height);
  ALTHEA_REQUIRE(cube_ok(ibl->prefiltered_size, ibl->prefiltered_mips), "deferred_resolve: prefiltered %d / %d",
                 ibl->prefiltered_size, ibl->prefiltered_mips);
  ALTHEA_REQUIRE(ibl->irradiance_size >= 1 && ibl->brdf_lut_size >= 1, "deferred_resolve: IBL sizes");
  ALTHEA_REQUIRE(light_count >= 0 && (light_count == 0 || d_lights), "deferred_resolve: lights");
  if (ssao && ssao->enabled)
    ALTHEA_REQUIRE(ssao->samples >= 1 && ssao->samples <= ALTHEA_SSAO_MAX_SAMPLES && ssao->radius > 0.0f,
                   "deferred_resolve: ssao samples %d radius %g", ssao->samples, ssao->radius);
  return ALTHEA_OK;
}

static althea_status reflection_ok(const althea_reflection_view* r, const althea_gbuffer_view* g,
                                   const char* what) {
  ALTHEA_REQUIRE(r && r->data, "%s: null reflection buffer", what);
  ALTHEA_REQUIRE(r->format == ALTHEA_REFLECTION_FP16 || r->format == ALTHEA_REFLECTION_FP32,
                 "%s: reflection format %d", what, r->format);
  ALTHEA_REQUIRE(r->mips >= 1 && r->mips <= ALTHEA_REFLECTION_MAX_MIPS, "%s: reflection mips %d", what, r->mips);
  ALTHEA_REQUIRE(r->width >= 1 && r->height >= 1 && r->width <= 32768 && r->height <= 32768,
                 "%s: reflection extent %dx%d", what, r->width, r->height);
  if (g)
    ALTHEA_REQUIRE(r->width == g->width && r->height == g->height,
                   "%s: reflection buffer %dx%d, G-buffer %dx%d", what, r->width, r->height, g->width, g->height);
  return ALTHEA_OK;
}

extern "C" {

const char* althea_status_string(althea_status status) {
  switch (status) {
    case ALTHEA_OK: return "ok";
    case ALTHEA_ERROR_INVALID_ARGUMENT: return "invalid argument";
    case ALTHEA_ERROR_OUT_OF_MEMORY: return "out of device memory";
    case ALTHEA_ERROR_HIP: return "HIP runtime error";
    case ALTHEA_ERROR_NOT_FOUND: return "not found";
    case ALTHEA_ERROR_IO: return "I/O error";
    case ALTHEA_ERROR_UNSUPPORTED: return "unsupported";
  }
  return "unknown status";
}

const char* althea_last_error(void) { return g_error; }

int althea_version(void) { return 100; }

althea_status althea_hammersley(uint32_t n, uint32_t first, uint32_t count, uint32_t* d_bits, float* d_xi,
                                hipStream_t stream) {
  ALTHEA_REQUIRE(n > 0, "hammersley: n must be positive");
  ALTHEA_REQUIRE(d_bits || d_xi || count == 0, "hammersley: no output buffer");
  return launch_hammersley(n, first, count, d_bits, d_xi, stream);
}

int althea_env_levels(int width, int height) {
  if (width <= 0 || height <= 0) return 0;
  return 1 + static_cast<int>(std::floor(std::log2(static_cast<double>(std::max(width, height)))));
}

size_t althea_env_pyramid_texels(int width, int height) {
  if (width <= 0 || height <= 0) return 0;
  return pyramid_geometry(width, height).texels;
}

althea_status althea_env_pyramid_build(const float* d_level0, int width, int height, float* d_pyramid,
                                       hipStream_t stream) {
  ALTHEA_REQUIRE(d_level0 && d_pyramid, "env_pyramid_build: null buffer");
  ALTHEA_REQUIRE(env_dims_ok(width, height), "env_pyramid_build: %dx%d is not a power-of-two extent of at most 2^27 texels", width, height);
  return launch_pyramid(d_level0, width, height, d_pyramid, stream);
}

size_t althea_cube_texels(int size, int mips) {
  if (size <= 0 || mips <= 0) return 0;
  return cube_mip_offset(size, mips);
}

size_t althea_cube_bordered_texels(int size, int mips) {
  if (size < 1 || mips < 1 || mips > 16) return 0;
  return cube_bordered_texels(size, mips);
}

althea_status althea_cube_border(const float* d_cube, int size, int mips, float* d_bordered, hipStream_t stream) {
  ALTHEA_REQUIRE(d_cube && d_bordered, "cube_border: null buffer");
  ALTHEA_REQUIRE(size >= 1 && size <= 16384 && mips >= 1 && mips <= 16 && (size >> (mips - 1)) >= 1,
                 "cube_border: size %d / mips %d unsupported", size, mips);
  return launch_cube_border(d_cube, size, mips, d_bordered, stream);
}

althea_status althea_cube_sample(const float* d_cube, int size, int mips, const float* d_dir_lod, int count,
                                 float* d_out, hipStream_t stream) {
  ALTHEA_REQUIRE(count >= 0, "cube_sample: negative count");
  ALTHEA_REQUIRE(count == 0 || (d_cube && d_dir_lod && d_out), "cube_sample: null buffer");
  ALTHEA_REQUIRE(cube_ok(size, mips), "cube_sample: size %d / mips %d unsupported", size, mips);
  return launch_cube_sample(d_cube, size, mips, d_dir_lod, count, d_out, nullptr, stream);
}

althea_status althea_env_sample(const float* d_pyramid, int width, int height, const float* d_dir_lod, int count,
                                float* d_out, hipStream_t stream) {
  ALTHEA_REQUIRE(count >= 0, "env_sample: negative count");
  ALTHEA_REQUIRE(count == 0 || (d_pyramid && d_dir_lod && d_out), "env_sample: null buffer");
  ALTHEA_REQUIRE(env_dims_ok(width, height), "env_sample: %dx%d is not a power-of-two extent of at most 2^27 texels", width, height);
  return launch_env_sample(d_pyramid, width, height, d_dir_lod, count, d_out, nullptr, stream);
}

static bool samples_ok(int samples) { return samples >= 1 && samples <= (1 << 20); }

althea_status althea_prefilter_cube(const float* d_pyramid, int env_width, int env_height, int size, int mips,
                                    int samples, unsigned flags, float* d_cube, hipStream_t stream) {
  ALTHEA_REQUIRE(d_pyramid && d_cube, "prefilter_cube: null buffer");
  ALTHEA_REQUIRE(env_dims_ok(env_width, env_height), "prefilter_cube: %dx%d is not a power-of-two extent of at most 2^27 texels",
                 env_width, env_height);
  ALTHEA_REQUIRE(cube_ok(size, mips), "prefilter_cube: size %d / mips %d unsupported", size, mips);
  ALTHEA_REQUIRE(samples_ok(samples), "prefilter_cube: samples %d out of range", samples);
  Workspace ws(prefilter_workspace_bytes(env_width, env_height, size, mips, samples), stream);
  if (!ws.ok()) return ALTHEA_ERROR_OUT_OF_MEMORY;
  return launch_prefilter_cube(d_pyramid, env_width, env_height, size, mips, samples, flags, d_cube, ws.get(),
                               stream);
}

size_t althea_prefilter_cube_scratch_bytes(int mips, int samples) {
  if (mips <= 0 || samples <= 0) return 0;
  return prefilter_table_bytes(mips, samples);
}

althea_status althea_prefilter_cube_noalloc(const float* d_pyramid, int env_width, int env_height, int size,
                                            int mips, int samples, unsigned flags, float* d_cube,
                                            hipStream_t stream) {
  ALTHEA_REQUIRE(d_pyramid && d_cube, "prefilter_cube_noalloc: null buffer");
  ALTHEA_REQUIRE(env_dims_ok(env_width, env_height), "prefilter_cube_noalloc: %dx%d is not a power-of-two extent of at most 2^27 texels",
                 env_width, env_height);
  ALTHEA_REQUIRE(cube_ok(size, mips), "prefilter_cube_noalloc: size %d / mips %d unsupported", size, mips);
  ALTHEA_REQUIRE(samples_ok(samples), "prefilter_cube_noalloc: samples %d out of range", samples);
  const size_t need = prefilter_table_bytes(mips, samples);
  ALTHEA_REQUIRE(need <= ALTHEA_PREFILTER_SCRATCH_MAX_BYTES,
                 "prefilter_cube_noalloc: %d mips x %d samples need %zu bytes of sample tables, more than the "
                 "fixed %u-byte array (use althea_prefilter_cube)",
                 mips, samples, need, ALTHEA_PREFILTER_SCRATCH_MAX_BYTES);
  return launch_prefilter_cube_noalloc(d_pyramid, env_width, env_height, size, mips, samples, flags, d_cube,
                                       stream);
}

althea_status althea_prefilter_points(const float* d_pyramid, int env_width, int env_height, int samples,
                                      const float* d_dir_rough, int count, float* d_out, hipStream_t stream) {
  ALTHEA_REQUIRE(count >= 0, "prefilter_points: negative count");
  ALTHEA_REQUIRE(count == 0 || (d_pyramid && d_dir_rough && d_out), "prefilter_points: null buffer");
  ALTHEA_REQUIRE(env_dims_ok(env_width, env_height), "prefilter_points: %dx%d is not a power-of-two extent of at most 2^27 texels",
                 env_width, env_height);
  ALTHEA_REQUIRE(samples_ok(samples), "prefilter_points: samples %d out of range", samples);
  return launch_prefilter_points(d_pyramid, env_width, env_height, samples, d_dir_rough, count, d_out, nullptr,
                                 stream);
}

althea_status althea_irradiance_cube(const float* d_pyramid, int env_width, int env_height, int size,
                                     int theta_samples, unsigned flags, float* d_cube, hipStream_t stream) {
  ALTHEA_REQUIRE(d_pyramid && d_cube, "irradiance_cube: null buffer");
  ALTHEA_REQUIRE(env_dims_ok(env_width, env_height), "irradiance_cube: %dx%d is not a power-of-two extent of at most 2^27 texels",
                 env_width, env_height);
  ALTHEA_REQUIRE(size >= 1 && size <= 4096, "irradiance_cube: size %d out of range", size);
  ALTHEA_REQUIRE(theta_samples >= 1 && theta_samples <= 100000, "irradiance_cube: theta_samples %d out of range",
                 theta_samples);
  Workspace ws(irradiance_workspace_bytes(env_width, env_height, size, theta_samples), stream);
  if (!ws.ok()) return ALTHEA_ERROR_OUT_OF_MEMORY;
  return launch_irradiance(d_pyramid, env_width, env_height, size, theta_samples, flags, d_cube, ws.get(), stream);
}

althea_status althea_brdf_lut(int size, int samples, float* d_rg, hipStream_t stream) {
  ALTHEA_REQUIRE(d_rg, "brdf_lut: null buffer");
  ALTHEA_REQUIRE(size >= 1 && size <= 4096, "brdf_lut: size %d out of range", size);
  ALTHEA_REQUIRE(samples_ok(samples), "brdf_lut: samples %d out of range", samples);
  return launch_brdf_lut(size, samples, d_rg, stream);
}

// ------------------------------------------------------------- IBL handle
void althea_ibl_desc_default(althea_ibl_desc* desc) {
  if (!desc) return;
  desc->irradiance_size = 32;
  desc->irradiance_theta_samples = 300;
  desc->prefiltered_size = 128;
  desc->prefiltered_mips = 5;
  desc->prefilter_samples = 1024;
  desc->brdf_lut_size = 128;
  desc->brdf_lut_samples = 1024;
  desc->flags = 0;
}

althea_status althea_ibl_create(const althea_ibl_desc* desc, althea_ibl** out) {
  ALTHEA_REQUIRE(desc && out, "ibl_create: null argument");
  *out = nullptr;
  ALTHEA_REQUIRE(desc->irradiance_size >= 1 && desc->irradiance_size <= 4096, "ibl_create: irradiance_size %d",
                 desc->irradiance_size);
  ALTHEA_REQUIRE(desc->irradiance_theta_samples >= 1, "ibl_create: irradiance_theta_samples %d",
                 desc->irradiance_theta_samples);
  ALTHEA_REQUIRE(cube_ok(desc->prefiltered_size, desc->prefiltered_mips), "ibl_create: prefiltered %d / %d mips",
                 desc->prefiltered_size, desc->prefiltered_mips);
  ALTHEA_REQUIRE(samples_ok(desc->prefilter_samples), "ibl_create: prefilter_samples %d", desc->prefilter_samples);
  ALTHEA_REQUIRE(desc->brdf_lut_size >= 1 && desc->brdf_lut_size <= 4096, "ibl_create: brdf_lut_size %d",
                 desc->brdf_lut_size);
  ALTHEA_REQUIRE(samples_ok(desc->brdf_lut_samples), "ibl_create: brdf_lut_samples %d", desc->brdf_lut_samples);
  althea_ibl* ibl = new (std::nothrow) althea_ibl;
  if (!ibl) return ALTHEA_ERROR_OUT_OF_MEMORY;
  ibl->desc = *desc;
  size_t pref = sizeof(float) * 4 * althea_cube_texels(desc->prefiltered_size, desc->prefiltered_mips);
  size_t irr = sizeof(float) * 4 * 6 * static_cast<size_t>(desc->irradiance_size) * desc->irradiance_size;
  size_t lut = sizeof(float) * 2 * static_cast<size_t>(desc->brdf_lut_size) * desc->brdf_lut_size;
  size_t pref_b = sizeof(float) * 4 * cube_bordered_texels(desc->prefiltered_size, desc->prefiltered_mips);
  size_t irr_b = sizeof(float) * 4 * cube_bordered_texels(desc->irradiance_size, 1);
  if (hipMalloc(&ibl->prefiltered, pref) != hipSuccess || hipMalloc(&ibl->irradiance, irr) != hipSuccess ||
      hipMalloc(&ibl->lut, lut) != hipSuccess || hipMalloc(&ibl->prefiltered_bordered, pref_b) != hipSuccess ||
      hipMalloc(&ibl->irradiance_bordered, irr_b) != hipSuccess) {
    set_error("ibl_create: device allocation failed");
    althea_ibl_destroy(ibl);
    return ALTHEA_ERROR_OUT_OF_MEMORY;
  }
  *out = ibl;
  return ALTHEA_OK;
}

althea_status althea_ibl_destroy(althea_ibl* ibl) {
  if (!ibl) return ALTHEA_OK;
  (void)hipFree(ibl->env);
  (void)hipFree(ibl->prefiltered);
  (void)hipFree(ibl->irradiance);
  (void)hipFree(ibl->lut);
  (void)hipFree(ibl->prefiltered_bordered);
  (void)hipFree(ibl->irradiance_bordered);
  (void)hipFree(ibl->workspace);
  delete ibl;
  return ALTHEA_OK;
}

}  // extern "C"

// The handle's bordered cubes from its prefiltered and irradiance cubes.
static althea_status ibl_border(althea_ibl* ibl, hipStream_t stream) {
  const althea_ibl_desc& d = ibl->desc;
  althea_status st =
      launch_cube_border(ibl->prefiltered, d.prefiltered_size, d.prefiltered_mips, ibl->prefiltered_bordered, stream);
  if (st != ALTHEA_OK) return st;
  return launch_cube_border(ibl->irradiance, d.irradiance_size, 1, ibl->irradiance_bordered, stream);
}

// Sizes the handle's pyramid and workspace for a width x height source.
static althea_status ibl_prepare(althea_ibl* ibl, int width, int height, hipStream_t stream) {
  const althea_ibl_desc& d = ibl->desc;
  size_t texels = althea_env_pyramid_texels(width, height);
  if (texels != ibl->env_texels) {
    // The pyramid is reallocated only when the source extent changes.
    ALTHEA_TRY_HIP(hipStreamSynchronize(stream));
    (void)hipFree(ibl->env);
    ibl->env = nullptr;
    ibl->env_texels = 0;
    ALTHEA_TRY_HIP(hipMalloc(&ibl->env, sizeof(float) * 4 * texels));
    ibl->env_texels = texels;
  }
  ibl->env_w = width;
  ibl->env_h = height;
  size_t ws_bytes = std::max(prefilter_workspace_bytes(width, height, d.prefiltered_size, d.prefiltered_mips,
                                                       d.prefilter_samples),
                             irradiance_workspace_bytes(width, height, d.irradiance_size, d.irradiance_theta_samples));
  if (ws_bytes > ibl->workspace_bytes) {
    ALTHEA_TRY_HIP(hipStreamSynchronize(stream));
    (void)hipFree(ibl->workspace);
    ibl->workspace = nullptr;
    ibl->workspace_bytes = 0;
    ALTHEA_TRY_HIP(hipMalloc(&ibl->workspace, ws_bytes));
    ibl->workspace_bytes = ws_bytes;
  }
  return ALTHEA_OK;
}

extern "C" {

althea_status althea_ibl_precompute(althea_ibl* ibl, const float* d_equirect, int width, int height,
                                    hipStream_t stream) {
  ALTHEA_REQUIRE(ibl && d_equirect, "ibl_precompute: null argument");
  ALTHEA_REQUIRE(env_dims_ok(width, height), "ibl_precompute: %dx%d is not a power-of-two extent of at most 2^27 texels", width, height);
  const althea_ibl_desc& d = ibl->desc;
  althea_status st = ibl_prepare(ibl, width, height, stream);
  if (st != ALTHEA_OK) return st;
  st = launch_pyramid(d_equirect, width, height, ibl->env, stream);
  if (st != ALTHEA_OK) return st;
  st = launch_irradiance(ibl->env, width, height, d.irradiance_size, d.irradiance_theta_samples, d.flags,
                         ibl->irradiance, ibl->workspace, stream);
  if (st != ALTHEA_OK) return st;
  st = launch_prefilter_cube(ibl->env, width, height, d.prefiltered_size, d.prefiltered_mips, d.prefilter_samples,
                             d.flags, ibl->prefiltered, ibl->workspace, stream);
  if (st != ALTHEA_OK) return st;
  st = ibl_border(ibl, stream);
  if (st != ALTHEA_OK) return st;
  return launch_brdf_lut(d.brdf_lut_size, d.brdf_lut_samples, ibl->lut, stream);
}

althea_status althea_ibl_precompute_batched(althea_ibl* const* ibls, const althea_env_view* envs, int count,
                                            const althea_tuning* tuning, althea_tuning* used, hipStream_t stream) {
  ALTHEA_REQUIRE(ibls && envs && count >= 1, "ibl_precompute_batched: invalid argument");
  for (int i = 0; i < count; ++i) {
    ALTHEA_REQUIRE(ibls[i] && envs[i].rgba, "ibl_precompute_batched: null handle or map %d", i);
    ALTHEA_REQUIRE(env_dims_ok(envs[i].width, envs[i].height), "ibl_precompute_batched: map %d is %dx%d", i,
                   envs[i].width, envs[i].height);
    const althea_ibl_desc& a = ibls[0]->desc;
    const althea_ibl_desc& b = ibls[i]->desc;
    ALTHEA_REQUIRE(b.prefiltered_mips == a.prefiltered_mips && b.prefilter_samples == a.prefilter_samples &&
                       (b.flags & ALTHEA_PRECOMPUTE_FP32_SUMS) == (a.flags & ALTHEA_PRECOMPUTE_FP32_SUMS),
                   "ibl_precompute_batched: handle %d differs in mips, samples or precision from handle 0", i);
    for (int j = 0; j < i; ++j)
      ALTHEA_REQUIRE(ibls[j] != ibls[i], "ibl_precompute_batched: handle %d repeats handle %d", i, j);
  }
  std::vector<float*> pyr(count), cubes(count);
  std::vector<int> sizes(count);
  for (int i = 0; i < count; ++i) {
    althea_status st = ibl_prepare(ibls[i], envs[i].width, envs[i].height, stream);
    if (st != ALTHEA_OK) return st;
    pyr[i] = ibls[i]->env;
    cubes[i] = ibls[i]->prefiltered;
    sizes[i] = ibls[i]->desc.prefiltered_size;
  }
  const althea_ibl_desc& d0 = ibls[0]->desc;
  althea_status st = prefilter_batch_into(envs, count, sizes.data(), d0.prefiltered_mips, d0.prefilter_samples,
                                          d0.flags, pyr.data(), cubes.data(), tuning, used, stream);
  if (st != ALTHEA_OK) return st;
  for (int i = 0; i < count; ++i) {
    const althea_ibl_desc& d = ibls[i]->desc;
    st = launch_irradiance(ibls[i]->env, envs[i].width, envs[i].height, d.irradiance_size, d.irradiance_theta_samples,
                           d.flags, ibls[i]->irradiance, ibls[i]->workspace, stream);
    if (st != ALTHEA_OK) return st;
    st = launch_brdf_lut(d.brdf_lut_size, d.brdf_lut_samples, ibls[i]->lut, stream);
    if (st != ALTHEA_OK) return st;
    st = ibl_border(ibls[i], stream);
    if (st != ALTHEA_OK) return st;
  }
  return ALTHEA_OK;
}

althea_status althea_ibl_get_views(const althea_ibl* ibl, althea_ibl_views* v) {
  ALTHEA_REQUIRE(ibl && v, "ibl_get_views: null argument");
  v->environment = ibl->env;
  v->env_width = ibl->env_w;
  v->env_height = ibl->env_h;
  v->env_levels = ibl->env ? althea_env_levels(ibl->env_w, ibl->env_h) : 0;
  v->prefiltered = ibl->prefiltered;
  v->prefiltered_size = ibl->desc.prefiltered_size;
  v->prefiltered_mips = ibl->desc.prefiltered_mips;
  v->irradiance = ibl->irradiance;
  v->irradiance_size = ibl->desc.irradiance_size;
  v->brdf_lut = ibl->lut;
  v->brdf_lut_size = ibl->desc.brdf_lut_size;
  v->prefiltered_bordered = ibl->prefiltered_bordered;
  v->irradiance_bordered = ibl->irradiance_bordered;
  return ALTHEA_OK;
}

size_t althea_ibl_state_bytes(const althea_ibl* ibl) {
  if (!ibl) return 0;
  const althea_ibl_desc& d = ibl->desc;
  return sizeof(float) * 4 * althea_cube_texels(d.prefiltered_size, d.prefiltered_mips) +
         sizeof(float) * 4 * 6 * static_cast<size_t>(d.irradiance_size) * d.irradiance_size +
         sizeof(float) * 2 * static_cast<size_t>(d.brdf_lut_size) * d.brdf_lut_size;
}

// ---------------------------------------------------------------- resolve
althea_status althea_deferred_resolve(const althea_gbuffer_view* g, const althea_ibl_views* ibl,
                                      const althea_point_light* d_lights, int light_count,
                                      const althea_globals* globals, const althea_ssao_desc* ssao, unsigned flags,
                                      void* d_out, float* d_ao_out, hipStream_t stream) {
  althea_status st = resolve_args_ok(g, ibl, d_lights, light_count, globals, ssao, d_out);
  if (st != ALTHEA_OK) return st;
  return launch_resolve(g, ibl, d_lights, light_count, globals, ssao, flags, d_out, d_ao_out, nullptr, nullptr,
                        stream);
}

althea_status althea_deferred_resolve_mlp(const althea_gbuffer_view* g, const althea_ibl_views* ibl,
                                          const althea_mlp* mlp, const althea_point_light* d_lights,
                                          int light_count, const althea_globals* globals,
                                          const althea_ssao_desc* ssao, unsigned flags, void* d_out,
                                          float* d_ao_out, uint8_t* d_mlp_used, hipStream_t stream) {
  althea_status st = resolve_args_ok(g, ibl, d_lights, light_count, globals, ssao, d_out);
  if (st != ALTHEA_OK) return st;
  ALTHEA_REQUIRE(mlp, "deferred_resolve_mlp: null MLP");
  MlpInferenceView v{};
  // One queue region per resolve wave: 64 x 4 pixel tiles, four waves each.
  st = mlp_resolve_queue(mlp, static_cast<size_t>((g->width + 63) / 64) * ((g->height + 3) / 4) * 4, &v);
  if (st != ALTHEA_OK) return st;
  ALTHEA_REQUIRE(v.mon_mips > 0 && v.cell_ok, "deferred_resolve_mlp: the MLP's error monitor is not built");
  ALTHEA_REQUIRE(v.mon_size == ibl->prefiltered_size && v.mon_mips == ibl->prefiltered_mips,
                 "deferred_resolve_mlp: monitor built for a %d / %d cube, IBL has %d / %d", v.mon_size, v.mon_mips,
                 ibl->prefiltered_size, ibl->prefiltered_mips);
  return launch_resolve(g, ibl, d_lights, light_count, globals, ssao, flags, d_out, d_ao_out, &v, d_mlp_used,
                        stream);
}

// ------------------------------------------------------------ IBL cache
// The reference's cache layout (Src/ImageBasedLighting.cpp:428-446):
// <dir>/Prefiltered<m+1>.hdr per roughness level and <dir>/IrradianceMap.hdr,
// RGBE through stb (Src/Utilities.cpp:216-255).  The images here are the
// cubes: each level's six faces stacked top to bottom in face order
// (+X, -X, +Y, -Y, +Z, -Z), S wide and 6 S tall, as [face][y][x] stores them.
static std::string cache_file(const char* dir, int prefiltered_level) {
  std::string d(dir);
  if (prefiltered_level < 0) return d + "/IrradianceMap.hdr";
  return d + "/Prefiltered" + std::to_string(prefiltered_level + 1) + ".hdr";
}

size_t althea_reflection_mip_offset(int width, int height, int level) {
  if (width < 1 || height < 1 || level < 0) return 0;
  size_t off = 0;
  for (int m = 0; m < level; ++m)
    off += static_cast<size_t>(std::max(1, width >> m)) * static_cast<size_t>(std::max(1, height >> m));
  return off;
}

size_t althea_reflection_texels(int width, int height, int mips) {
  if (mips < 1 || mips > ALTHEA_REFLECTION_MAX_MIPS) return 0;
  return althea_reflection_mip_offset(width, height, mips);
}

size_t althea_ssr_workspace_bytes(int width, int height) {
  if (width < 1 || height < 1) return 0;
  return ssr_workspace_bytes(width, height);
}

althea_status althea_ibl_save_cache(const althea_ibl* ibl, const char* dir, hipStream_t stream) {
  ALTHEA_REQUIRE(ibl && dir && dir[0], "ibl_save_cache: null argument");
  ALTHEA_REQUIRE(ibl->env_w > 0, "ibl_save_cache: the handle holds no precomputed maps");
  std::error_code ec;
  std::filesystem::create_directories(dir, ec);
  if (ec) {
    set_error("ibl_save_cache: cannot create %s: %s", dir, ec.message().c_str());
    return ALTHEA_ERROR_IO;
  }
  ALTHEA_TRY_HIP(hipStreamSynchronize(stream));
  const althea_ibl_desc& d = ibl->desc;
  std::vector<float> host;
  auto save = [&](const float* dev, int S, int level) -> althea_status {
    host.resize(static_cast<size_t>(6) * S * S * 4);
    ALTHEA_TRY_HIP(hipMemcpy(host.data(), dev, host.size() * sizeof(float), hipMemcpyDeviceToHost));
    return althea_hdr_write(cache_file(dir, level).c_str(), host.data(), S, 6 * S);
  };
  for (int m = 0; m < d.prefiltered_mips; ++m) {
    althea_status st = save(ibl->prefiltered + 4 * cube_mip_offset(d.prefiltered_size, m),
                            std::max(1, d.prefiltered_size >> m), m);
    if (st != ALTHEA_OK) return st;
  }
  return save(ibl->irradiance, d.irradiance_size, -1);
}

althea_status althea_ibl_load_cache(althea_ibl* ibl, const float* d_equirect, int width, int height, const char* dir,
                                    hipStream_t stream) {
  ALTHEA_REQUIRE(ibl && d_equirect && dir && dir[0], "ibl_load_cache: null argument");
  ALTHEA_REQUIRE(env_dims_ok(width, height), "ibl_load_cache: %dx%d is not a power-of-two extent of at most 2^27 texels", width, height);
  const althea_ibl_desc& d = ibl->desc;
  // Every file present, each with the handle's extent, before any device work.
  for (int m = -1; m < d.prefiltered_mips; ++m) {
    const int S = m < 0 ? d.irradiance_size : std::max(1, d.prefiltered_size >> m);
    const std::string f = cache_file(dir, m);
    int w = 0, h = 0;
    althea_status st = althea_hdr_info(f.c_str(), &w, &h);
    if (st != ALTHEA_OK) return st;
    ALTHEA_REQUIRE(w == S && h == 6 * S, "ibl_load_cache: %s is %dx%d, the handle needs %dx%d", f.c_str(), w, h, S,
                   6 * S);
  }
  althea_status st = ibl_prepare(ibl, width, height, stream);
  if (st != ALTHEA_OK) return st;
  st = launch_pyramid(d_equirect, width, height, ibl->env, stream);
  if (st != ALTHEA_OK) return st;
  std::vector<float> host;
  auto load = [&](float* dev, int S, int level) -> althea_status {
    host.resize(static_cast<size_t>(6) * S * S * 4);
    althea_status s2 = althea_hdr_read(cache_file(dir, level).c_str(), host.data(), host.size());
    if (s2 != ALTHEA_OK) return s2;
    // Synchronous copy: `host` is reused for the next file.
    ALTHEA_TRY_HIP(hipMemcpyAsync(dev, host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice, stream));
    ALTHEA_TRY_HIP(hipStreamSynchronize(stream));
    return ALTHEA_OK;
  };
  for (int m = 0; m < d.prefiltered_mips; ++m) {
    st = load(ibl->prefiltered + 4 * cube_mip_offset(d.prefiltered_size, m), std::max(1, d.prefiltered_size >> m), m);
    if (st != ALTHEA_OK) return st;
  }
  st = load(ibl->irradiance, d.irradiance_size, -1);
  if (st != ALTHEA_OK) return st;
  st = ibl_border(ibl, stream);
  if (st != ALTHEA_OK) return st;
  return launch_brdf_lut(d.brdf_lut_size, d.brdf_lut_samples, ibl->lut, stream);
}

size_t althea_gbuffer_workspace_bytes(int width, int height, int triangle_count) {
  return gbuffer_workspace_bytes(width, height, triangle_count);
}

// Workspaces are carved into 256-byte aligned parts that the kernels access
// with 16-byte stores and 64-bit atomics: the base must be 256-byte aligned,
// as hipMalloc and torch allocations are.
static bool workspace_aligned(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 255) == 0; }

althea_status althea_gbuffer_draw(const althea_gbuffer_mesh* mesh, const althea_globals* globals,
                                  const althea_gbuffer_view* g, void* d_workspace, hipStream_t stream) {
  ALTHEA_REQUIRE(mesh && globals && g && d_workspace, "gbuffer_draw: null argument");
  ALTHEA_REQUIRE(workspace_aligned(d_workspace), "gbuffer_draw: workspace not 256-byte aligned");
  ALTHEA_REQUIRE(g->width >= 1 && g->height >= 1 && static_cast<long long>(g->width) * g->height < (1ll << 31),
                 "gbuffer_draw: extent %dx%d", g->width, g->height);
  ALTHEA_REQUIRE(g->position && g->normal && g->albedo && g->mro, "gbuffer_draw: null G-buffer plane");
  ALTHEA_REQUIRE(g->format == ALTHEA_GBUFFER_FP32 || g->format == ALTHEA_GBUFFER_FP16, "gbuffer_draw: format %d",
                 g->format);
  ALTHEA_REQUIRE(mesh->triangle_count >= 0 && mesh->triangle_count < (1 << 28) &&
                     (mesh->triangle_count == 0 ||
                      (mesh->positions && mesh->colors && mesh->mro && mesh->indices &&
                                                mesh->vertex_count >= 1)),
                 "gbuffer_draw: mesh with %d triangles, %d vertices", mesh->triangle_count, mesh->vertex_count);
  ALTHEA_REQUIRE(globals->projection[10] != 0.0f && globals->projection[10] != -1.0f && globals->projection[14] != 0.0f,
                 "gbuffer_draw: projection is not a [0, 1]-depth perspective");
  return launch_gbuffer_draw(mesh, globals, g, d_workspace, stream);
}

size_t althea_texture_mip_offset(int width, int height, int level) {
  if (width < 1 || height < 1 || level < 0) return 0;
  size_t off = 0;
  for (int l = 0; l < level; ++l)
    off += static_cast<size_t>(std::max(1, width >> l)) * static_cast<size_t>(std::max(1, height >> l));
  return off;
}

size_t althea_texture_texels(int width, int height, int mips) {
  if (width < 1 || height < 1 || mips < 1 || mips > 16) return 0;
  return althea_texture_mip_offset(width, height, mips);
}

static int texture_max_mips(int w, int h) {
  int m = 1;
  while ((std::max(w, h) >> m) > 0) ++m;
  return m;
}

althea_status althea_texture_build_mips(uint8_t* d_texels, int width, int height, int mips, int srgb,
                                        hipStream_t stream) {
  ALTHEA_REQUIRE(d_texels && width >= 1 && height >= 1 && width <= 16384 && height <= 16384,
                 "texture_build_mips: texels %p, %dx%d", static_cast<void*>(d_texels), width, height);
  ALTHEA_REQUIRE(mips >= 1 && mips <= texture_max_mips(width, height), "texture_build_mips: %d levels for %dx%d",
                 mips, width, height);
  return launch_texture_mips(d_texels, width, height, mips, srgb ? 1 : 0, stream);
}

static althea_status texture_ok(const althea_texture_view& t, int k) {
  ALTHEA_REQUIRE(t.texels && t.width >= 1 && t.height >= 1 && t.width <= 16384 && t.height <= 16384,
                 "texture %d: texels %p, %dx%d", k, static_cast<const void*>(t.texels), t.width, t.height);
  ALTHEA_REQUIRE(t.mips >= 1 && t.mips <= texture_max_mips(t.width, t.height), "texture %d: %d levels for %dx%d", k,
                 t.mips, t.width, t.height);
  auto filt = [](int f) { return f == ALTHEA_FILTER_NEAREST || f == ALTHEA_FILTER_LINEAR; };
  auto wrap = [](int w) { return w >= ALTHEA_WRAP_REPEAT && w <= ALTHEA_WRAP_CLAMP_TO_EDGE; };
  ALTHEA_REQUIRE(filt(t.mag_filter) && filt(t.min_filter) && filt(t.mipmap_mode) && wrap(t.wrap_u) && wrap(t.wrap_v),
                 "texture %d: sampler state", k);
  return ALTHEA_OK;
}

static int gltf_triangles(const althea_gltf_primitive* prims, int count) {
  long long total = 0;
  for (int k = 0; k < count; ++k) total += prims[k].triangle_count;
  return total < (1ll << 28) ? static_cast<int>(total) : -1;
}

size_t althea_gbuffer_gltf_workspace_bytes(int width, int height, const althea_gltf_primitive* primitives,
                                           int primitive_count, int material_count, int texture_count) {
  if (primitive_count < 0 || (primitive_count > 0 && !primitives) || material_count < 0 || texture_count < 0)
    return 0;
  const int total = gltf_triangles(primitives, primitive_count);
  if (total < 0) return 0;
  return gbuffer_gltf_workspace_bytes(width, height, total, primitive_count, material_count, texture_count);
}

althea_status althea_gbuffer_draw_gltf(const althea_gltf_primitive* primitives, int primitive_count,
                                       const althea_gltf_material* materials, int material_count,
                                       const althea_texture_view* textures, int texture_count,
                                       const althea_globals* globals, const althea_gbuffer_view* g, void* d_workspace,
                                       hipStream_t stream) {
  ALTHEA_REQUIRE(globals && g && d_workspace, "gbuffer_draw_gltf: null argument");
  ALTHEA_REQUIRE(workspace_aligned(d_workspace), "gbuffer_draw_gltf: workspace not 256-byte aligned");
  ALTHEA_REQUIRE(primitive_count >= 0 && (primitive_count == 0 || primitives), "gbuffer_draw_gltf: %d primitives",
                 primitive_count);
  ALTHEA_REQUIRE(material_count >= 0 && (material_count == 0 || materials) && texture_count >= 0 &&
                     (texture_count == 0 || textures),
                 "gbuffer_draw_gltf: %d materials, %d textures", material_count, texture_count);
  ALTHEA_REQUIRE(g->width >= 1 && g->height >= 1 && static_cast<long long>(g->width) * g->height < (1ll << 31),
                 "gbuffer_draw_gltf: extent %dx%d", g->width, g->height);
  ALTHEA_REQUIRE(g->position && g->normal && g->albedo && g->mro, "gbuffer_draw_gltf: null G-buffer plane");
  ALTHEA_REQUIRE(g->format == ALTHEA_GBUFFER_FP32 || g->format == ALTHEA_GBUFFER_FP16,
                 "gbuffer_draw_gltf: format %d", g->format);
  ALTHEA_REQUIRE(globals->projection[10] != 0.0f && globals->projection[10] != -1.0f && globals->projection[14] != 0.0f,
                 "gbuffer_draw_gltf: projection is not a [0, 1]-depth perspective");
  ALTHEA_REQUIRE(gltf_triangles(primitives, primitive_count) >= 0, "gbuffer_draw_gltf: too many triangles");
  for (int k = 0; k < texture_count; ++k) {
    const althea_status st = texture_ok(textures[k], k);
    if (st != ALTHEA_OK) return st;
  }
  for (int k = 0; k < material_count; ++k) {
    const althea_gltf_material& m = materials[k];
    auto tex = [&](int t) { return t >= 0 && t < texture_count; };
    ALTHEA_REQUIRE(tex(m.base_texture) && tex(m.normal_texture) && tex(m.metallic_roughness_texture),
                   "gbuffer_draw_gltf: material %d names a texture outside 0..%d", k, texture_count - 1);
    ALTHEA_REQUIRE(m.base_uv >= 0 && m.base_uv < 4 && m.metallic_roughness_uv >= 0 && m.metallic_roughness_uv < 4,
                   "gbuffer_draw_gltf: material %d texture-coordinate sets", k);
  }
  for (int k = 0; k < primitive_count; ++k) {
    const althea_gltf_primitive& q = primitives[k];
    ALTHEA_REQUIRE(q.triangle_count >= 0 && q.material >= 0 && q.material < material_count,
                   "gbuffer_draw_gltf: primitive %d: %d triangles, material %d", k, q.triangle_count, q.material);
    if (q.triangle_count == 0) continue;
    ALTHEA_REQUIRE(q.positions && q.tangents && q.bitangents && q.normals && q.indices && q.vertex_count >= 1,
                   "gbuffer_draw_gltf: primitive %d: null vertex data", k);
    const althea_gltf_material& m = materials[q.material];
    ALTHEA_REQUIRE(q.uvs[m.base_uv] && q.uvs[m.metallic_roughness_uv],
                   "gbuffer_draw_gltf: primitive %d lacks the texture coordinates its material reads", k);
  }
  return launch_gbuffer_draw_gltf(primitives, primitive_count, materials, material_count, textures, texture_count,
                                  globals, g, d_workspace, stream);
}

size_t althea_shadow_cube_floats(int size) {
  return size < 1 ? 0 : static_cast<size_t>(6) * (size + 2) * (size + 2);
}

althea_status althea_shadow_cubes(const althea_mesh_view* mesh, const althea_point_light* d_lights, int light_count,
                                  int size, float* d_cubes, hipStream_t stream) {
  ALTHEA_REQUIRE(mesh && d_cubes && light_count >= 1 && d_lights, "shadow_cubes: null argument or no light");
  ALTHEA_REQUIRE(size >= 1 && size <= 4096, "shadow_cubes: size %d", size);
  ALTHEA_REQUIRE(mesh->triangle_count >= 0 && (mesh->triangle_count == 0 || (mesh->positions && mesh->indices &&
                                                                               mesh->vertex_count >= 1)),
                 "shadow_cubes: mesh with %d triangles, %d vertices", mesh->triangle_count, mesh->vertex_count);
  ALTHEA_REQUIRE(static_cast<long long>(light_count) * mesh->triangle_count < (1ll << 31),
                 "shadow_cubes: %d lights x %d triangles", light_count, mesh->triangle_count);
  return launch_shadow_cubes(mesh, d_lights, light_count, size, d_cubes, stream);
}

size_t althea_shadow_cubes_workspace_bytes(int light_count, int triangle_count) {
  if (light_count < 1 || triangle_count < 0 || static_cast<long long>(light_count) * triangle_count >= (1ll << 31))
    return 0;
  return shadow_workspace_bytes(light_count, triangle_count);
}

althea_status althea_shadow_cubes_ws(const althea_mesh_view* mesh, const althea_point_light* d_lights,
                                     int light_count, int size, float* d_cubes, void* d_workspace,
                                     hipStream_t stream) {
  ALTHEA_REQUIRE(mesh && d_cubes && light_count >= 1 && d_lights && d_workspace,
                 "shadow_cubes_ws: null argument or no light");
  ALTHEA_REQUIRE(workspace_aligned(d_workspace), "shadow_cubes_ws: workspace not 256-byte aligned");
  ALTHEA_REQUIRE(size >= 1 && size <= 4096, "shadow_cubes_ws: size %d", size);
  ALTHEA_REQUIRE(mesh->triangle_count >= 0 && (mesh->triangle_count == 0 || (mesh->positions && mesh->indices &&
                                                                               mesh->vertex_count >= 1)),
                 "shadow_cubes_ws: mesh with %d triangles, %d vertices", mesh->triangle_count, mesh->vertex_count);
  ALTHEA_REQUIRE(static_cast<long long>(light_count) * mesh->triangle_count < (1ll << 31),
                 "shadow_cubes_ws: %d lights x %d triangles", light_count, mesh->triangle_count);
  return launch_shadow_cubes(mesh, d_lights, light_count, size, d_cubes, stream, d_workspace);
}

size_t althea_shadow_cubes_gltf_workspace_bytes(const althea_gltf_primitive* primitives, int primitive_count,
                                                int light_count) {
  if (primitive_count < 0 || (primitive_count > 0 && !primitives) || light_count < 1) return 0;
  const int total = gltf_triangles(primitives, primitive_count);
  if (total < 0 || static_cast<long long>(light_count) * total >= (1ll << 31)) return 0;
  return shadow_gltf_workspace_bytes(light_count, total, primitive_count);
}

althea_status althea_shadow_cubes_gltf(const althea_gltf_primitive* primitives, int primitive_count,
                                       const althea_gltf_material* materials, int material_count,
                                       const althea_texture_view* textures, int texture_count,
                                       const althea_point_light* d_lights, int light_count, int size, float* d_cubes,
                                       void* d_workspace, hipStream_t stream) {
  ALTHEA_REQUIRE(d_cubes && light_count >= 1 && d_lights && d_workspace,
                 "shadow_cubes_gltf: null argument or no light");
  ALTHEA_REQUIRE(workspace_aligned(d_workspace), "shadow_cubes_gltf: workspace not 256-byte aligned");
  ALTHEA_REQUIRE(size >= 1 && size <= 4096, "shadow_cubes_gltf: size %d", size);
  ALTHEA_REQUIRE(primitive_count >= 0 && (primitive_count == 0 || primitives), "shadow_cubes_gltf: %d primitives",
                 primitive_count);
  ALTHEA_REQUIRE(material_count >= 0 && (material_count == 0 || materials) && texture_count >= 0 &&
                     (texture_count == 0 || textures),
                 "shadow_cubes_gltf: %d materials, %d textures", material_count, texture_count);
  const int total = gltf_triangles(primitives, primitive_count);
  ALTHEA_REQUIRE(total >= 0 && static_cast<long long>(light_count) * total < (1ll << 31),
                 "shadow_cubes_gltf: %d lights x %d triangles", light_count, total);
  for (int k = 0; k < texture_count; ++k) {
    const althea_status st = texture_ok(textures[k], k);
    if (st != ALTHEA_OK) return st;
  }
  for (int k = 0; k < material_count; ++k) {
    const althea_gltf_material& m = materials[k];
    ALTHEA_REQUIRE(m.base_texture >= 0 && m.base_texture < texture_count && m.base_uv >= 0 && m.base_uv < 4,
                   "shadow_cubes_gltf: material %d: base colour texture %d, coordinate set %d", k, m.base_texture,
                   m.base_uv);
  }
  for (int k = 0; k < primitive_count; ++k) {
    const althea_gltf_primitive& q = primitives[k];
    ALTHEA_REQUIRE(q.triangle_count >= 0 && q.material >= 0 && q.material < material_count,
                   "shadow_cubes_gltf: primitive %d: %d triangles, material %d", k, q.triangle_count, q.material);
    if (q.triangle_count == 0) continue;
    ALTHEA_REQUIRE(q.positions && q.indices && q.vertex_count >= 1, "shadow_cubes_gltf: primitive %d: null vertex data",
                   k);
    ALTHEA_REQUIRE(q.uvs[materials[q.material].base_uv],
                   "shadow_cubes_gltf: primitive %d lacks the texture coordinates its material reads", k);
  }
  return launch_shadow_cubes_gltf(primitives, primitive_count, materials, textures, d_lights, light_count, size,
                                  d_cubes, d_workspace, stream);
}

althea_status althea_deferred_resolve_shadowed(const althea_gbuffer_view* g, const althea_ibl_views* ibl,
                                               const althea_point_light* d_lights, int light_count,
                                               const althea_globals* globals, const althea_ssao_desc* ssao,
                                               const althea_reflection_view* reflection, const float* d_shadow_cubes,
                                               int shadow_size, unsigned flags, void* d_out, float* d_ao_out,
                                               hipStream_t stream) {
  althea_status st = resolve_args_ok(g, ibl, d_lights, light_count, globals, ssao, d_out);
  if (st != ALTHEA_OK) return st;
  if (reflection) {
    st = reflection_ok(reflection, g, "deferred_resolve_shadowed");
    if (st != ALTHEA_OK) return st;
  }
  ALTHEA_REQUIRE(d_shadow_cubes && shadow_size >= 1 && shadow_size <= 4096,
                 "deferred_resolve_shadowed: shadow cubes %p, size %d", static_cast<const void*>(d_shadow_cubes),
                 shadow_size);
  return launch_resolve(g, ibl, d_lights, light_count, globals, ssao, flags, d_out, d_ao_out, nullptr, nullptr,
                        stream, reflection, d_shadow_cubes, shadow_size);
}

althea_status althea_ssr_capture(const althea_gbuffer_view* g, const althea_ibl_views* ibl,
                                 const althea_point_light* d_lights, int light_count, const althea_globals* globals,
                                 const althea_reflection_view* reflection, void* d_workspace, hipStream_t stream) {
  althea_status st =
      resolve_args_ok(g, ibl, d_lights, light_count, globals, nullptr, reflection ? reflection->data : nullptr);
  if (st != ALTHEA_OK) return st;
  st = reflection_ok(reflection, g, "ssr_capture");
  if (st != ALTHEA_OK) return st;
  ALTHEA_REQUIRE(d_workspace, "ssr_capture: null workspace");
  ALTHEA_REQUIRE(workspace_aligned(d_workspace), "ssr_capture: workspace not 256-byte aligned");
  return launch_ssr_capture(g, ibl, d_lights, light_count, globals, reflection, d_workspace, stream);
}

althea_status althea_ssr_capture_shadowed(const althea_gbuffer_view* g, const althea_ibl_views* ibl,
                                          const althea_point_light* d_lights, int light_count,
                                          const althea_globals* globals, const althea_reflection_view* reflection,
                                          const float* d_shadow_cubes, int shadow_size, void* d_workspace,
                                          hipStream_t stream) {
  althea_status st =
      resolve_args_ok(g, ibl, d_lights, light_count, globals, nullptr, reflection ? reflection->data : nullptr);
  if (st != ALTHEA_OK) return st;
  st = reflection_ok(reflection, g, "ssr_capture_shadowed");
  if (st != ALTHEA_OK) return st;
  ALTHEA_REQUIRE(d_workspace, "ssr_capture_shadowed: null workspace");
  ALTHEA_REQUIRE(workspace_aligned(d_workspace), "ssr_capture_shadowed: workspace not 256-byte aligned");
  ALTHEA_REQUIRE(d_shadow_cubes && shadow_size >= 1 && shadow_size <= 4096,
                 "ssr_capture_shadowed: shadow cubes %p, size %d", static_cast<const void*>(d_shadow_cubes),
                 shadow_size);
  return launch_ssr_capture(g, ibl, d_lights, light_count, globals, reflection, d_workspace, stream, d_shadow_cubes,
                            shadow_size);
}

althea_status althea_reflection_convolve(const althea_reflection_view* reflection, hipStream_t stream) {
  althea_status st = reflection_ok(reflection, nullptr, "reflection_convolve");
  if (st != ALTHEA_OK) return st;
  return launch_reflection_convolve(reflection, stream);
}

althea_status althea_deferred_resolve_reflect(const althea_gbuffer_view* g, const althea_ibl_views* ibl,
                                              const althea_point_light* d_lights, int light_count,
                                              const althea_globals* globals, const althea_ssao_desc* ssao,
                                              const althea_reflection_view* reflection, unsigned flags, void* d_out,
                                              float* d_ao_out, hipStream_t stream) {
  althea_status st = resolve_args_ok(g, ibl, d_lights, light_count, globals, ssao, d_out);
  if (st != ALTHEA_OK) return st;
  st = reflection_ok(reflection, g, "deferred_resolve_reflect");
  if (st != ALTHEA_OK) return st;
  return launch_resolve(g, ibl, d_lights, light_count, globals, ssao, flags, d_out, d_ao_out, nullptr, nullptr,
                        stream, reflection);
}

}  // extern "C"
