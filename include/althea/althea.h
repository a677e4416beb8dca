/*
 * althea.h - C ABI of the MI355X (gfx950) image-based-lighting precompute and
 * deferred-resolve library.
 *
 * It replaces the Vulkan compute/fragment path of nithinp7/Althea for:
 *   - ImageBasedLighting::createResources / precomputeResources
 *     (Include/Althea/ImageBasedLighting.h:45-49, Src/ImageBasedLighting.cpp:137-605)
 *     and the shaders it dispatches (Shaders/IBL_Precompute/GenIrradianceMap.comp,
 *     Shaders/IBL_Precompute/PreFilterEnvMap.comp), plus a computed BRDF LUT in
 *     place of Content/PrecomputedMaps/brdf_lut.png (Src/ImageBasedLighting.cpp:571-602);
 *   - the deferred lighting resolve DeferredPass.vert/DeferredPass.frag with
 *     pbrMaterial (Shaders/PBR/PBRMaterial.glsl:72-162) and SSAO (Shaders/SSAO.glsl),
 *     reading the G-buffer of GBufferResources (Include/Althea/DeferredRendering.h:38-102).
 *
 * Conventions
 *   - Every buffer is caller-owned device memory unless a handle says otherwise.
 *   - A d_workspace argument is 256-byte aligned device memory (hipMalloc and
 *     torch allocations are) of the size its *_workspace_bytes function gives;
 *     a misaligned one is rejected with ALTHEA_ERROR_INVALID_ARGUMENT.
 *   - Every call is stream-ordered and asynchronous on the given hipStream_t
 *     (0 = the legacy default stream); results are ready when the stream is.
 *   - Errors come back as althea_status; no C++ exception crosses this ABI.
 *     althea_last_error() gives a thread-local message for the last failure.
 *   - Handles are not thread-safe: calls on one handle need external
 *     synchronisation (the reference is single-threaded too,
 *     Src/SingleTimeCommandBuffer.cpp:28-54).
 *
 * Layouts in HBM
 *   - Equirect environment map: RGBA32F rows top-down; column u = yaw/2pi + 0.5
 *     with yaw = atan2(z, x); row v = 0.5 - atan2(y, |xz|)/pi (the shading
 *     convention of Shaders/Global/GlobalResources.glsl:48-54).  Power-of-two
 *     extents.  Its pyramid holds althea_env_levels(w,h) levels
 *     (Src/Utilities.cpp:97-100) concatenated, level l of max(1,w>>l) x max(1,h>>l).
 *     Sources have at most 2^27 texels (e.g. 16384 x 8192), so that every
 *     level's byte offsets fit 32 bits.
 *   - Cube map: RGBA32F [mip][face][y][x], faces +X,-X,+Y,-Y,+Z,-Z with the
 *     Vulkan cube-face-selection table; mip m has faces of max(1, size>>m).
 *   - BRDF LUT: RG32F [roughness row][NdotV column], texel centres at (i+0.5)/size.
 *   - G-buffer: four planes of vec4 per pixel (position.w = coverage, 0 means
 *     sky), FP32 or FP16, in the order of GBufferResources::buildMaterial
 *     (Src/DeferredRendering.cpp:23-35).
 */
#ifndef ALTHEA_ALTHEA_H_
#define ALTHEA_ALTHEA_H_

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define ALTHEA_API __attribute__((visibility("default")))
#else
#define ALTHEA_API
#endif

/* Status codes.  The reference throws std::runtime_error (Src/ImageBasedLighting.cpp:424,
 * Src/ComputePipeline.cpp:49); the ABI returns these instead. */
typedef enum althea_status {
  ALTHEA_OK = 0,
  ALTHEA_ERROR_INVALID_ARGUMENT = 1,
  ALTHEA_ERROR_OUT_OF_MEMORY = 2,
  ALTHEA_ERROR_HIP = 3,
  ALTHEA_ERROR_NOT_FOUND = 4, /* missing environment map (ImageBasedLighting.cpp:423-425) */
  ALTHEA_ERROR_IO = 5,
  ALTHEA_ERROR_UNSUPPORTED = 6
} althea_status;

ALTHEA_API const char* althea_status_string(althea_status status);
ALTHEA_API const char* althea_last_error(void);
ALTHEA_API int althea_version(void);

/* ------------------------------------------------------------------------ */
/* Sampling primitives                                                       */
/* ------------------------------------------------------------------------ */

/* Hammersley points i = first .. first+count-1 of an n-point set, computed by
 * the device function every sampling kernel uses: bits[k] = bitreverse(i) and
 * xi[2k..2k+1] = (i/n correctly rounded, float(bits) * 2^-32).  The reference
 * samples with a hash RNG instead (Shaders/IBL_Precompute/PreFilterEnvMap.comp:38-44). */
ALTHEA_API althea_status althea_hammersley(uint32_t n, uint32_t first, uint32_t count, uint32_t* d_bits,
                                           float* d_xi, hipStream_t stream);

/* Number of levels of an environment pyramid: 1 + floor(log2(max(w, h)))
 * (Src/Utilities.cpp:97-100), and the texel count of all levels together. */
ALTHEA_API int althea_env_levels(int width, int height);
ALTHEA_API size_t althea_env_pyramid_texels(int width, int height);

/* Builds the environment pyramid: level 0 is copied from d_level0 and level l
 * is the 2x2 box mean of level l-1, the LINEAR half-size blit chain of
 * Image::generateMipMaps (Src/Image.cpp:135-239). */
ALTHEA_API althea_status althea_env_pyramid_build(const float* d_level0, int width, int height,
                                                  float* d_pyramid, hipStream_t stream);

/* Texel count of a cube with `mips` levels and faces of `size` at mip 0. */
ALTHEA_API size_t althea_cube_texels(int size, int mips);

/* Seamless trilinear cube lookups: d_dir_lod holds float4 (dir.xyz, lod) per
 * query; d_out receives float4 (rgb, 1).  The software filter used by the
 * resolve for sampleEnvMap / sampleIrrMap (Shaders/PBR/PBRMaterial.glsl:4-19). */
ALTHEA_API althea_status althea_cube_sample(const float* d_cube, int size, int mips, const float* d_dir_lod,
                                            int count, float* d_out, hipStream_t stream);

/* Precompute lookups into an environment pyramid (REPEAT sampler, trilinear,
 * the pole branch of PreFilterEnvMap.comp:98-122): d_dir_lod float4 per
 * query, d_out float4.  The mirror-lookup path: angles and texel coordinates
 * in double-float (as the roughness-0 prefilter mip addresses its lookups),
 * each level's bilinear weights rounded once to FP32. */
ALTHEA_API althea_status althea_env_sample(const float* d_pyramid, int width, int height,
                                           const float* d_dir_lod, int count, float* d_out,
                                           hipStream_t stream);

/* ------------------------------------------------------------------------ */
/* IBL precompute kernels on caller buffers                                  */
/* ------------------------------------------------------------------------ */

/* Flags for the precompute kernels. */
#define ALTHEA_PRECOMPUTE_FP32_SUMS 0x1u /* all-FP32 reductions (precision study variant) */

/* GGX-prefiltered cube (PreFilterEnvMap.comp:126-177 per output texel, one
 * mip per roughness m/(mips-1) as ImageBasedLighting.cpp:352-402 does with i/4):
 * `samples` Hammersley half vectors, alpha = roughness, NdotL-weighted mean,
 * source LOD from the sample pdf.  d_cube holds althea_cube_texels(size, mips)
 * float4.  Sums over samples are FP64 unless ALTHEA_PRECOMPUTE_FP32_SUMS.  The
 * call's stream-ordered workspace holds the sample tables and a padded copy of
 * the pyramid (each level with its wrap-around neighbours stored around it,
 * about 1.34x the pyramid) that the sample loop reads. */
ALTHEA_API althea_status althea_prefilter_cube(const float* d_pyramid, int env_width, int env_height,
                                               int size, int mips, int samples, unsigned flags,
                                               float* d_cube, hipStream_t stream);

/* Largest sample-table scratch althea_prefilter_cube_noalloc accepts:
 * althea_prefilter_cube_scratch_bytes(mips, samples) must not exceed it
 * (4 MiB: about 32 B per mip and sample, e.g. 6 mips x 16384 samples). */
#define ALTHEA_PREFILTER_SCRATCH_MAX_BYTES (4u << 20)
ALTHEA_API size_t althea_prefilter_cube_scratch_bytes(int mips, int samples);

/* althea_prefilter_cube without any device allocation.  No counterpart in the
 * reference: this entry point is for callers that cannot allocate (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  The
 * per-roughness GGX sample tables live in a fixed 4 MiB device array that the
 * library holds on every GPU it runs on until the process ends; requests whose
 * tables exceed ALTHEA_PREFILTER_SCRATCH_MAX_BYTES are rejected with
 * ALTHEA_ERROR_INVALID_ARGUMENT before any device work.  It reads the pyramid
 * itself, where althea_prefilter_cube reads its padded copy; the cube written
 * is bit-identical to althea_prefilter_cube's.  Every call on one GPU shares that
 * array: calls that overlap on one GPU (other streams, other threads) race. */
ALTHEA_API althea_status althea_prefilter_cube_noalloc(const float* d_pyramid, int env_width, int env_height,
                                                       int size, int mips, int samples, unsigned flags,
                                                       float* d_cube, hipStream_t stream);

/* Prefiltered radiance at arbitrary points: d_dir_rough float4 (dir.xyz,
 * roughness) per point, d_out float4.  Used for the learned surrogate's labels. */
ALTHEA_API althea_status althea_prefilter_points(const float* d_pyramid, int env_width, int env_height,
                                                 int samples, const float* d_dir_rough, int count,
                                                 float* d_out, hipStream_t stream);

/* Diffuse irradiance cube (GenIrradianceMap.comp:106-155 per output texel):
 * theta_samples x int(H*theta_samples/W) Riemann sum (300 in the reference)
 * at source LOD log2(W/theta_samples).  d_cube holds 6*size*size float4.  The
 * call's stream-ordered workspace holds the angle tables and a padded copy of
 * the two source levels read. */
ALTHEA_API althea_status althea_irradiance_cube(const float* d_pyramid, int env_width, int env_height,
                                                int size, int theta_samples, unsigned flags, float* d_cube,
                                                hipStream_t stream);

/* Split-sum BRDF LUT (Karis 2013; consumed at PBRMaterial.glsl:110-111):
 * size x size RG32F, `samples` Hammersley points per texel, FP64 sums and
 * normalisation. */
ALTHEA_API althea_status althea_brdf_lut(int size, int samples, float* d_rg, hipStream_t stream);

/* ------------------------------------------------------------------------ */
/* Batched prefilter (no counterpart in the reference, which prefilters one  */
/* map with five dispatches, Src/ImageBasedLighting.cpp:352-402)             */
/* ------------------------------------------------------------------------ */

/* An equirect source map in device memory (RGBA32F, power-of-two extents). */
typedef struct althea_env_view {
  const float* rgba;
  int width, height;
} althea_env_view;

/* Launch shape of the persistent prefilter: a wave shades a tile_w x tile_h
 * texel tile (8x8, 16x4 or 4x16); the grid holds waves_per_cu waves per CU. */
typedef struct althea_tuning {
  int tile_w, tile_h;
  int waves_per_cu;
  int reserved;
} althea_tuning;

typedef struct althea_prefilter_batch althea_prefilter_batch;

/* Prefilters `count` maps (any mix of extents) into cubes of face_sizes[i]
 * with `mips` roughness levels m/(mips-1) and `samples` Hammersley samples,
 * the same estimator as althea_prefilter_cube.  The batch owns the pyramids
 * and output cubes; the sources must stay valid while the batch is used. */
ALTHEA_API althea_status althea_prefilter_batch_create(const althea_env_view* envs, int count, const int* face_sizes,
                                                       int mips, int samples, unsigned flags,
                                                       althea_prefilter_batch** out);
ALTHEA_API althea_status althea_prefilter_batch_destroy(althea_prefilter_batch* batch);
/* One run: a single hipGraph launch holding every map's pyramid chain, the
 * sample tables of each distinct source extent, a counter reset and one
 * persistent prefilter launch over all work tiles (largest first). */
ALTHEA_API althea_status althea_prefilter_batch_run(althea_prefilter_batch* batch, hipStream_t stream);
/* Times every candidate launch shape on a fixed subset of the work list and
 * keeps the fastest (synchronises `stream`). */
ALTHEA_API althea_status althea_prefilter_batch_autotune(althea_prefilter_batch* batch, althea_tuning* chosen,
                                                         hipStream_t stream);
ALTHEA_API althea_status althea_prefilter_batch_set_tuning(althea_prefilter_batch* batch, const althea_tuning* tuning);
ALTHEA_API althea_status althea_prefilter_batch_get_tuning(const althea_prefilter_batch* batch, althea_tuning* tuning);
/* Device cube of map `index` (layout of althea_prefilter_cube). */
ALTHEA_API althea_status althea_prefilter_batch_cube(const althea_prefilter_batch* batch, int index,
                                                     const float** d_cube, int* size, int* mips);
/* Output texels of all maps together. */
ALTHEA_API size_t althea_prefilter_batch_texels(const althea_prefilter_batch* batch);

/* ------------------------------------------------------------------------ */
/* IBL resources (ImageBasedLighting::createResources / IBLResources)        */
/* ------------------------------------------------------------------------ */

typedef struct althea_ibl althea_ibl;

typedef struct althea_ibl_desc {
  int irradiance_size;        /* cube face size of the irradiance map (32) */
  int irradiance_theta_samples; /* 300 in GenIrradianceMap.comp:133 */
  int prefiltered_size;       /* cube face size of prefiltered mip 0 (128, 512) */
  int prefiltered_mips;       /* roughness levels (5 in ImageBasedLighting.cpp:192) */
  int prefilter_samples;      /* Hammersley samples per texel (1024) */
  int brdf_lut_size;          /* 128 */
  int brdf_lut_samples;       /* 1024 */
  unsigned flags;             /* ALTHEA_PRECOMPUTE_* */
} althea_ibl_desc;

/* Device views of an IBL set: IBLResources::getHandles (ImageBasedLighting.h:36-42)
 * with device pointers in place of bindless indices, in the same order
 * (environment, prefiltered, irradiance, BRDF LUT). */
typedef struct althea_ibl_views {
  const float* environment; /* pyramid, level 0 first */
  int env_width, env_height, env_levels;
  const float* prefiltered;
  int prefiltered_size, prefiltered_mips;
  const float* irradiance;
  int irradiance_size;
  const float* brdf_lut;
  int brdf_lut_size;
  /* Optional seamless-bordered copies of the prefiltered and irradiance cubes
   * (althea_cube_border), which the deferred resolves and the SSR capture's
   * hit shading read; an IBL handle keeps them.  NULL: each such call builds
   * them in a stream-ordered allocation (a copy of the two cubes per call). */
  const float* prefiltered_bordered;
  const float* irradiance_bordered;
} althea_ibl_views;

/* Seamless-bordered copy of a cube (size x size faces, mips levels): every
 * face of mip m stored as (S + 2) x (S + 2) RGBA32F texels (S = size >> m),
 * the border holding what the seamless bilinear filter reads one texel beyond
 * each edge (the adjacent face's edge texel; at a corner the mean of the three
 * texels meeting there), so a lookup is four plain loads with the cube's own
 * result.  althea_cube_bordered_texels gives its size in float4. */
ALTHEA_API size_t althea_cube_bordered_texels(int size, int mips);
ALTHEA_API althea_status althea_cube_border(const float* d_cube, int size, int mips, float* d_bordered,
                                            hipStream_t stream);

ALTHEA_API void althea_ibl_desc_default(althea_ibl_desc* desc);
ALTHEA_API althea_status althea_ibl_create(const althea_ibl_desc* desc, althea_ibl** out);
ALTHEA_API althea_status althea_ibl_destroy(althea_ibl* ibl);
/* Builds the pyramid of d_equirect (copied into the handle), then irradiance,
 * prefiltered mips and the BRDF LUT (ImageBasedLighting.cpp:137-412). */
ALTHEA_API althea_status althea_ibl_precompute(althea_ibl* ibl, const float* d_equirect, int width, int height,
                                               hipStream_t stream);
/* althea_ibl_precompute for `count` handles at once: every handle's pyramid
 * and prefiltered mips come from one persistent, graph-captured batched
 * prefilter launch (as althea_prefilter_batch_run), then each handle's
 * irradiance and LUT.  The handles must share prefiltered_mips,
 * prefilter_samples and precision flags; face sizes and map extents may
 * differ.  tuning NULL runs the autotuner; `used` (optional) receives the
 * launch shape.  Synchronises `stream`.  (The reference precomputes one map
 * at a time, ImageBasedLighting.cpp:137-412.) */
ALTHEA_API althea_status althea_ibl_precompute_batched(althea_ibl* const* ibls, const althea_env_view* envs,
                                                       int count, const althea_tuning* tuning, althea_tuning* used,
                                                       hipStream_t stream);
ALTHEA_API althea_status althea_ibl_get_views(const althea_ibl* ibl, althea_ibl_views* views);
/* Device bytes the shading pass reads for this IBL set (prefiltered +
 * irradiance + LUT), the "IBL state per environment". */
ALTHEA_API size_t althea_ibl_state_bytes(const althea_ibl* ibl);

/* On-disk cache of the precomputed maps in the reference's layout
 * (Src/ImageBasedLighting.cpp:428-446, written at :345-349 and :397-401):
 * <dir>/Prefiltered<m+1>.hdr for each roughness level and
 * <dir>/IrradianceMap.hdr, RGBE files.  The images are this build's cubes,
 * each level's six faces stacked top to bottom in face order (S x 6S).
 * Save: after a precompute; creates `dir`; synchronises `stream`.  Load:
 * builds the environment pyramid from d_equirect, reads every level and the
 * irradiance cube (RGBE precision, as the reference's own reload), computes
 * the LUT; ALTHEA_ERROR_NOT_FOUND when a file is missing (then precompute and
 * save, as the reference does), ALTHEA_ERROR_INVALID_ARGUMENT when a file's
 * extent does not match the handle.  Synchronises `stream`. */
ALTHEA_API althea_status althea_ibl_save_cache(const althea_ibl* ibl, const char* dir, hipStream_t stream);
ALTHEA_API althea_status althea_ibl_load_cache(althea_ibl* ibl, const float* d_equirect, int width, int height,
                                               const char* dir, hipStream_t stream);

/* ------------------------------------------------------------------------ */
/* Deferred resolve (DeferredPass.frag + pbrMaterial + SSAO)                 */
/* ------------------------------------------------------------------------ */

#define ALTHEA_GBUFFER_FP32 0
#define ALTHEA_GBUFFER_FP16 1

typedef struct althea_gbuffer_view {
  const void* position; /* xyz world position, w = coverage (0 -> background) */
  const void* normal;   /* xyz normal (normalised by the resolve) */
  const void* albedo;   /* rgb base colour */
  const void* mro;      /* metallic, roughness, occlusion */
  int width, height;
  int format;           /* ALTHEA_GBUFFER_FP32 or ALTHEA_GBUFFER_FP16 */
} althea_gbuffer_view;

/* PointLight of Include/Althea/PointLight.h:31-34 (alignas(16) vec3 pairs). */
typedef struct althea_point_light {
  float position[3];
  float pad0;
  float emission[3];
  float pad1;
} althea_point_light;

/* GlobalUniforms of Include/Althea/GlobalUniforms.h:15-31: column-major mat4s
 * as glm stores them. */
typedef struct althea_globals {
  float projection[16];
  float inverse_projection[16];
  float view[16];
  float prev_view[16];
  float inverse_view[16];
  float prev_inverse_view[16];
  float mouse_uv[2];
  int32_t light_count;
  uint32_t light_buffer_handle;
  float time;
  float exposure;
  uint32_t input_mask;
  uint32_t frame_count;
} althea_globals;

#define ALTHEA_SSAO_MAX_SAMPLES 64

/* Hemisphere-kernel SSAO (replaces Shaders/SSAO.glsl's 24x12 ray march). */
typedef struct althea_ssao_desc {
  int enabled;    /* 0: keep the G-buffer occlusion channel */
  int samples;    /* <= ALTHEA_SSAO_MAX_SAMPLES */
  float radius;   /* world units (0.5) */
  float bias;     /* view-depth bias */
  float kernel[ALTHEA_SSAO_MAX_SAMPLES][4]; /* tangent-space offsets, z >= 0 */
  float noise[16][4];                       /* 4x4 rotation tile, indexed (x&3) + 4 (y&3) */
} althea_ssao_desc;

#define ALTHEA_RESOLVE_TONEMAP 0x1u  /* 1 - exp(-c * exposure), DeferredPass.frag:86-88 */
#define ALTHEA_RESOLVE_OUT_FP16 0x2u /* RGBA16F output instead of RGBA32F */

/* One frame of the deferred resolve: per pixel the view ray of
 * DeferredPass.vert:10-22, background environment when position.w == 0,
 * reflection from the prefiltered cube at LOD roughness*(mips-1), irradiance,
 * SSAO into the occlusion channel, pbrMaterial with `light_count` point lights,
 * optional tonemap.  d_out: width*height RGBA (FP32 or FP16).  d_ao_out
 * (optional, may be NULL): width*height floats of the occlusion used.
 * With SSAO enabled, every resolve entry point first writes the frame's view
 * depth (width*height floats) into a stream-ordered allocation
 * (hipMallocAsync / hipFreeAsync on `stream`) that the SSAO taps read. */
ALTHEA_API althea_status althea_deferred_resolve(const althea_gbuffer_view* gbuffer, const althea_ibl_views* ibl,
                                                 const althea_point_light* d_lights, int light_count,
                                                 const althea_globals* globals, const althea_ssao_desc* ssao,
                                                 unsigned flags, void* d_out, float* d_ao_out,
                                                 hipStream_t stream);

/* ------------------------------------------------------------------------ */
/* Screen-space reflection + glossy reflection buffer                        */
/* ------------------------------------------------------------------------ */

/* The reflection buffer of Src/ReflectionBuffer.cpp:20-94: one frame-sized
 * image with `mips` levels (the reference uses 5), level m of extent
 * max(1, width >> m) x max(1, height >> m), RGBA per texel, stored level
 * after level ([mip][y][x], no padding) in caller-owned device memory of
 * althea_reflection_texels() texels.  FP16 is the reference's
 * R16G16B16A16_SFLOAT (:37); FP32 serves precision studies. */
#define ALTHEA_REFLECTION_MAX_MIPS 8
#define ALTHEA_REFLECTION_FP16 0
#define ALTHEA_REFLECTION_FP32 1

typedef struct althea_reflection_view {
  void* data;
  int width, height; /* level 0: the frame */
  int mips;          /* 1 .. ALTHEA_REFLECTION_MAX_MIPS */
  int format;        /* ALTHEA_REFLECTION_FP16 or ALTHEA_REFLECTION_FP32 */
} althea_reflection_view;

ALTHEA_API size_t althea_reflection_texels(int width, int height, int mips);
ALTHEA_API size_t althea_reflection_mip_offset(int width, int height, int level);

/* Level 0 of the reflection buffer: ScreenSpaceReflection::captureReflection
 * (Src/ScreenSpaceReflection.cpp:59-83) drawing Shaders/SSR.{vert,frag}.  Per
 * pixel with G-buffer coverage (position.w != 0): reflect the view ray about
 * the normal, march 128 steps of 0.005 in uv towards the projection of
 * P + 10000 R, rebuilding each step's position from the bilinearly filtered
 * (CLAMP_TO_EDGE) depth of the G-buffer (far plane where uncovered); the
 * first step (after the first) where the ray crosses the plane through R
 * perpendicular to cross(R, N), within acos(0.999) of R, with a hit normal
 * facing the ray, is shaded by pbrMaterial with the IBL state and
 * `light_count` point lights (ambient occlusion 1) and stored as (rgb, 1).
 * Everything else stores (0, 0, 0, 0).  Linear HDR, no tonemap.  The
 * buffer's extent must equal the G-buffer's.  d_workspace:
 * althea_ssr_workspace_bytes(width, height) bytes of device memory for the
 * depth plane, overwritten. */
ALTHEA_API size_t althea_ssr_workspace_bytes(int width, int height);
ALTHEA_API althea_status althea_ssr_capture(const althea_gbuffer_view* gbuffer, const althea_ibl_views* ibl,
                                            const althea_point_light* d_lights, int light_count,
                                            const althea_globals* globals, const althea_reflection_view* reflection,
                                            void* d_workspace, hipStream_t stream);

/* althea_ssr_capture with the point-light shadow test in the hit shading: the
 * reference's pbrMaterial reads the shadow cubes for every caller
 * (PBRMaterial.glsl:119-140), SSR.frag's hit shading included.
 * d_shadow_cubes: light_count cubes as althea_shadow_cubes writes them. */
ALTHEA_API althea_status althea_ssr_capture_shadowed(const althea_gbuffer_view* gbuffer, const althea_ibl_views* ibl,
                                                     const althea_point_light* d_lights, int light_count,
                                                     const althea_globals* globals,
                                                     const althea_reflection_view* reflection,
                                                     const float* d_shadow_cubes, int shadow_size, void* d_workspace,
                                                     hipStream_t stream);

/* Levels 1 .. mips-1 of the reflection buffer:
 * ReflectionBuffer::convolveReflectionBuffer (Src/ReflectionBuffer.cpp:163-277)
 * running Shaders/SSRGlossyConvolve.comp per level.  Level m samples level
 * m-1 bilinearly (CLAMP_TO_EDGE) at uv = texel / extent(m) (no half-texel
 * offset), blurred along y for odd m and along x for even m by the 7-tap
 * linear-sampled Gaussian: uv offsets of 1.4118, 3.2941 and 5.1765 divided by
 * the level's width, along either axis (SSRGlossyConvolve.comp:42-52).
 * Levels are written in order on `stream`. */
ALTHEA_API althea_status althea_reflection_convolve(const althea_reflection_view* reflection, hipStream_t stream);

/* althea_deferred_resolve with the reflection buffer blended into the
 * prefiltered term (DeferredPass.frag:59-68): the buffer is read trilinearly
 * at the pixel centre at LOD roughness*(mips-1); where its alpha a >= 0.01
 * the term becomes mix(prefiltered, rgb / a, a). */
ALTHEA_API althea_status althea_deferred_resolve_reflect(const althea_gbuffer_view* gbuffer,
                                                         const althea_ibl_views* ibl,
                                                         const althea_point_light* d_lights, int light_count,
                                                         const althea_globals* globals, const althea_ssao_desc* ssao,
                                                         const althea_reflection_view* reflection, unsigned flags,
                                                         void* d_out, float* d_ao_out, hipStream_t stream);

/* ------------------------------------------------------------------------ */
/* G-buffer producer                                                          */
/* ------------------------------------------------------------------------ */

/* A triangle mesh with the per-vertex inputs of the pass-through G-buffer
 * shader (Shaders/GBuffer/GBufferPassThrough.frag: inPosition, inNormal,
 * inColor, inMetallicRoughnessOcclusion), float3 each, world space; normals
 * may be NULL (all zero). */
typedef struct althea_gbuffer_mesh {
  const float* positions;
  const float* normals;
  const float* colors;
  const float* mro;
  const uint32_t* indices; /* 3 per triangle */
  int vertex_count, triangle_count;
} althea_gbuffer_mesh;

/* Device workspace althea_gbuffer_draw needs: the per-pixel visibility keys
 * (8 B/px) and per-triangle setup (72 B/triangle), 256-byte aligned parts. */
ALTHEA_API size_t althea_gbuffer_workspace_bytes(int width, int height, int triangle_count);

/* SceneToGBufferPass (Src/DeferredRendering.cpp:268-332) with the
 * pass-through subpass (GBufferPassThrough.frag, as Src/Scene/Floor.cpp
 * draws it): rasterizes `mesh` with the camera of `globals` (projection *
 * view; the [0, 1]-depth perspective's near and far planes clip) into the
 * planes of `gbuffer` (written, FP32 or FP16).  Per pixel: the nearest
 * front-facing triangle (counter-clockwise on screen, back faces culled;
 * equal depths keep the earlier triangle, as a LESS depth test does) whose
 * interior holds the pixel centre, attributes interpolated perspective-
 * correctly there.  Planes: position (pos, 1), normal (the interpolated
 * normal, or where it is exactly zero normalize(cross(dFdx(pos), dFdy(pos)))
 * with the framebuffer's y growing downwards, 1), albedo (color, 1), MRO
 * (mro, 1); uncovered pixels (0, 0, 0, 1) and position (0, 0, 0, 0).
 * d_workspace: althea_gbuffer_workspace_bytes bytes. */
ALTHEA_API althea_status althea_gbuffer_draw(const althea_gbuffer_mesh* mesh, const althea_globals* globals,
                                             const althea_gbuffer_view* gbuffer, void* d_workspace,
                                             hipStream_t stream);

/* glTF material subpass (Shaders/Gltf/Gltf.{vert,frag} with fetchMaterial of
 * Shaders/InstanceData/InstanceData.glsl:27-63). */

/* Sampler state of a glTF texture (Src/Sampler.cpp:11-114). */
#define ALTHEA_WRAP_REPEAT 0
#define ALTHEA_WRAP_MIRRORED_REPEAT 1
#define ALTHEA_WRAP_CLAMP_TO_EDGE 2
#define ALTHEA_FILTER_NEAREST 0
#define ALTHEA_FILTER_LINEAR 1

/* An RGBA8 texture with its mip chain (Src/Texture.cpp:50-100): levels back
 * to back, level m of max(1, w >> m) x max(1, h >> m) texels
 * (althea_texture_mip_offset); srgb = 1 for R8G8B8A8_SRGB (decoded to linear
 * before filtering), 0 for UNORM.  Filtering is the Vulkan LINEAR / NEAREST
 * rule per axis with the mip level from the screen derivatives of the
 * texture coordinate, isotropic (the reference enables anisotropic filtering,
 * whose footprint Vulkan leaves to the implementation). */
typedef struct althea_texture_view {
  const uint8_t* texels;
  int width, height, mips;
  int srgb;
  int mag_filter, min_filter, mipmap_mode; /* ALTHEA_FILTER_*; mipmap_mode NEAREST or LINEAR */
  int wrap_u, wrap_v;                      /* ALTHEA_WRAP_* */
} althea_texture_view;

ALTHEA_API size_t althea_texture_texels(int width, int height, int mips);
ALTHEA_API size_t althea_texture_mip_offset(int width, int height, int level);
/* Image::generateMipMaps (Src/Image.cpp:135-230): level m from level m - 1
 * by a linear blit to its extent (the bilinear value at each destination
 * texel centre mapped into the source, clamped to the edge), in linear space
 * for sRGB textures, rounded to 8 bits.  Level 0 is the caller's. */
ALTHEA_API althea_status althea_texture_build_mips(uint8_t* d_texels, int width, int height, int mips, int srgb,
                                                   hipStream_t stream);

/* MaterialConstants as Src/Material.cpp fills them: factors, the texture of
 * each slot (an index into the texture table; the reference binds its 1x1
 * white / flat-normal defaults for absent textures, which the caller passes
 * as textures too) and the texture-coordinate set of the base colour and
 * metallic-roughness slots (the normal map is read with the base colour's
 * set, InstanceData.glsl:38-40). */
typedef struct althea_gltf_material {
  float base_color_factor[4];
  float metallic_factor, roughness_factor;
  float normal_scale;
  float alpha_cutoff; /* fragments with base colour alpha below it are discarded (0.5 unless set) */
  int base_texture, normal_texture, metallic_roughness_texture;
  int base_uv, metallic_roughness_uv; /* 0..3 */
} althea_gltf_material;

/* One drawn glTF primitive in world space: positions and the TBN frame
 * after the model transform (Gltf.vert: worldPosition, mat3(model) * tbn),
 * up to four texture-coordinate sets (float2, NULL when absent), triangles,
 * and its material. */
typedef struct althea_gltf_primitive {
  const float* positions;
  const float* tangents;
  const float* bitangents;
  const float* normals;
  const float* uvs[4];
  const uint32_t* indices;
  int vertex_count, triangle_count;
  int material;
} althea_gltf_primitive;

ALTHEA_API size_t althea_gbuffer_gltf_workspace_bytes(int width, int height, const althea_gltf_primitive* primitives,
                                                      int primitive_count, int material_count, int texture_count);

/* SceneToGBufferPass with the glTF subpass: every primitive drawn in order
 * into one depth-tested G-buffer (earlier primitives win equal depths).  Per
 * pixel the nearest front-facing triangle whose fragment survives the alpha
 * test, then Gltf.frag's outputs: albedo = baseColor (texture x factor),
 * normal = (normalize(t.x T + t.y B + t.z N), alpha) with t the normal-map
 * sample mapped to [-1, 1] and scaled by (normal_scale, normal_scale, 1),
 * MRO = (metallic, roughness, 0, alpha) from the metallic-roughness texture's
 * blue and green channels times the factors; position (worldPosition, 1).
 * Tables are host arrays; the texture texels and vertex data device memory. */
ALTHEA_API althea_status althea_gbuffer_draw_gltf(const althea_gltf_primitive* primitives, int primitive_count,
                                                  const althea_gltf_material* materials, int material_count,
                                                  const althea_texture_view* textures, int texture_count,
                                                  const althea_globals* globals, const althea_gbuffer_view* gbuffer,
                                                  void* d_workspace, hipStream_t stream);

/* ------------------------------------------------------------------------ */
/* Point-light shadow cubes                                                  */
/* ------------------------------------------------------------------------ */

/* A triangle mesh in world space: float3 positions, 3 uint32 indices per
 * triangle (the geometry the reference's shadow pass draws). */
typedef struct althea_mesh_view {
  const float* positions;
  const uint32_t* indices;
  int vertex_count, triangle_count;
} althea_mesh_view;

#define ALTHEA_SHADOW_NEAR 0.01f  /* PointLight.cpp:84, ShadowMap.frag */
#define ALTHEA_SHADOW_FAR 1000.0f

/* The shadow cubes of PointLightCollection::drawShadowMaps
 * (Src/PointLight.cpp:52-140, :235-282; Shaders/ShadowMapBindless.{vert,frag}):
 * per light a size x size x 6 cube (256 in the reference) of FP32 depth =
 * min(1, distance / 1000) of the nearest triangle along each texel-centre ray
 * from the light, 1 where nothing is hit (depth cleared to 1).  As the
 * reference's pass draws them: back faces culled (counter-clockwise front
 * faces, GraphicsPipeline.h:171-172), only hits between the view-depth planes
 * 0.01 and 1000 of the 90-degree face camera.  The face images are the
 * reference's: faces in Vulkan order, face f's texel (x, y) looks along R_f v,
 * where v is the Vulkan cube direction of that texel and R_f = diag(-1, 1, 1)
 * for the +-X and +-Z faces, diag(1, 1, -1) for +-Y (what its yaw / pitch face
 * cameras, Src/PointLight.cpp:73-131 with Src/Camera.cpp:49-66, render).
 * Storage: [light][face][y + 1][x + 1] with a one-texel border around each
 * face, (size + 2)^2 floats per face (althea_shadow_cube_floats): the border
 * holds what the seamless filter reads beyond the face's edge (the adjacent
 * face's edge texel; at a corner the mean of the three texels meeting there),
 * filled by this call.  d_mesh positions / indices and d_lights are device
 * memory; d_cubes holds light_count * althea_shadow_cube_floats(size) floats.
 * Every triangle of the mesh is opaque: alpha-masked geometry goes through
 * althea_shadow_cubes_gltf, which runs the shadow shaders' alpha test. */
ALTHEA_API size_t althea_shadow_cube_floats(int size);
ALTHEA_API althea_status althea_shadow_cubes(const althea_mesh_view* mesh, const althea_point_light* d_lights,
                                             int light_count, int size, float* d_cubes, hipStream_t stream);

/* althea_shadow_cubes with a caller-provided workspace
 * (althea_shadow_cubes_workspace_bytes: per light-triangle pair its six
 * texel boxes and slot offset, 56 B) that deals the boxes' texels out evenly
 * over the GPU; the same cubes, bit for bit, faster where box sizes vary. */
ALTHEA_API size_t althea_shadow_cubes_workspace_bytes(int light_count, int triangle_count);
ALTHEA_API althea_status althea_shadow_cubes_ws(const althea_mesh_view* mesh, const althea_point_light* d_lights,
                                                int light_count, int size, float* d_cubes, void* d_workspace,
                                                hipStream_t stream);

/* The shadow cubes of glTF primitives, as the reference's shadow pass draws
 * a model's primitives with their materials: ShadowMapBindless.frag and
 * ShadowMap.frag discard a fragment whose base colour alpha (the base colour
 * texture's alpha times the factor's) is below the material's cut-off, before
 * the depth test.  Here a texel's candidate hit is discarded when that alpha,
 * sampled at the hit (the material's base colour coordinate set,
 * interpolated perspective-correctly) with the mip level of the coordinate's
 * derivatives across the face's texels (forward differences to the next
 * texel right and below), is below the cut-off.  Everything else as
 * althea_shadow_cubes_ws, the same cubes for opaque materials.  Tables are
 * host arrays as for althea_gbuffer_draw_gltf (positions, indices and the
 * base colour coordinates of each primitive, and the texels, in device
 * memory); d_workspace: althea_shadow_cubes_gltf_workspace_bytes bytes. */
ALTHEA_API size_t althea_shadow_cubes_gltf_workspace_bytes(const althea_gltf_primitive* primitives,
                                                           int primitive_count, int light_count);
ALTHEA_API althea_status althea_shadow_cubes_gltf(const althea_gltf_primitive* primitives, int primitive_count,
                                                  const althea_gltf_material* materials, int material_count,
                                                  const althea_texture_view* textures, int texture_count,
                                                  const althea_point_light* d_lights, int light_count, int size,
                                                  float* d_cubes, void* d_workspace, hipStream_t stream);

/* althea_deferred_resolve (optionally reading a reflection buffer, which may
 * be NULL) with the point-light shadow test of PBRMaterial.glsl:119-140: light
 * i is skipped at a pixel when 1000 * cube_i(L.x, -L.y, -L.z) < |Lp - P| - 0.5,
 * L the unit vector from the pixel to the light, the cube read with the
 * seamless bilinear filter.  d_shadow_cubes: light_count bordered cubes as
 * althea_shadow_cubes writes them. */
ALTHEA_API althea_status althea_deferred_resolve_shadowed(
    const althea_gbuffer_view* gbuffer, const althea_ibl_views* ibl, const althea_point_light* d_lights,
    int light_count, const althea_globals* globals, const althea_ssao_desc* ssao,
    const althea_reflection_view* reflection, const float* d_shadow_cubes, int shadow_size, unsigned flags,
    void* d_out, float* d_ao_out, hipStream_t stream);

/* ------------------------------------------------------------------------ */
/* RGBE (.hdr) files                                                         */
/* ------------------------------------------------------------------------ */

/* The format of the reference's environment maps (Content/HDRI_Skybox) and of
 * its prefiltered-map cache (Content/PrecomputedMaps), read and written there
 * through stb_image (Src/Utilities.cpp:216-255): Radiance header, flat or
 * run-length coded RGBE scanlines, value = byte * 2^(E - 136).  Host memory,
 * RGBA32F rows top-down (alpha 1 on read, ignored on write).  A missing file
 * is ALTHEA_ERROR_NOT_FOUND (ImageBasedLighting.cpp:423-425 throws there), a
 * malformed one ALTHEA_ERROR_IO, another pixel format or orientation
 * ALTHEA_ERROR_UNSUPPORTED. */
ALTHEA_API althea_status althea_hdr_info(const char* path, int* width, int* height);
ALTHEA_API althea_status althea_hdr_read(const char* path, float* host_rgba, size_t capacity_floats);
ALTHEA_API althea_status althea_hdr_write(const char* path, const float* host_rgba, int width, int height);

/* ------------------------------------------------------------------------ */
/* Learned prefiltered radiance (no counterpart in the reference)            */
/* ------------------------------------------------------------------------ */

/* A 28 -> 64 -> 64 -> 64 -> 3 MLP standing in for the prefiltered cube:
 * input (reflection dir, sin/cos(2^k pi dir) for k = 0..3, roughness), SiLU
 * hidden layers, log-radiance output.  BF16 weights and activations, FP32
 * accumulation (gfx950 MFMA), FP32 master weights trained with Adam on the
 * mean squared log-radiance error.  Parameters in the order W1[64][28] b1[64]
 * W2[64][64] b2[64] W3[64][64] b3[64] W4[3][64] b4[3] (10371 floats). */
typedef struct althea_mlp althea_mlp;

typedef struct althea_mlp_desc {
  uint64_t seed;          /* weight initialisation */
  float learning_rate;    /* Adam (1e-3) */
  float beta1, beta2, epsilon;
  float init_output_bias; /* initial log-radiance bias */
} althea_mlp_desc;

ALTHEA_API void althea_mlp_desc_default(althea_mlp_desc* desc);
ALTHEA_API int althea_mlp_param_count(void);
ALTHEA_API althea_status althea_mlp_create(const althea_mlp_desc* desc, althea_mlp** out);
ALTHEA_API althea_status althea_mlp_destroy(althea_mlp* mlp);
/* Host copies of the FP32 master parameters (synchronous on `stream`);
 * set_params also resets the Adam moments and step. */
ALTHEA_API althea_status althea_mlp_set_params(althea_mlp* mlp, const float* host_params, hipStream_t stream);
ALTHEA_API althea_status althea_mlp_get_params(const althea_mlp* mlp, float* host_params, hipStream_t stream);
/* One Adam step on a batch: d_dir_rough float4 (dir.xyz, roughness) and
 * d_targets float4 (linear rgb radiance > 0) per sample; d_loss (optional)
 * receives the batch's mean squared log error before the step. */
ALTHEA_API althea_status althea_mlp_train_step(althea_mlp* mlp, const float* d_dir_rough, const float* d_targets,
                                               int count, float* d_loss, hipStream_t stream);
/* Adam step size for the following train steps (a schedule sets it between
 * steps; the moments and step count are kept). */
ALTHEA_API althea_status althea_mlp_set_learning_rate(althea_mlp* mlp, float learning_rate);
/* The batch gradient (10371 floats, parameter order) without an update. */
ALTHEA_API althea_status althea_mlp_gradients(althea_mlp* mlp, const float* d_dir_rough, const float* d_targets,
                                              int count, float* d_grad, float* d_loss, hipStream_t stream);
/* Radiance exp(y) per query: d_out float4 (rgb, 1). */
ALTHEA_API althea_status althea_mlp_infer(const althea_mlp* mlp, const float* d_dir_rough, int count, float* d_out,
                                          hipStream_t stream);
/* Bytes of inference state (BF16 weights + FP32 biases). */
ALTHEA_API size_t althea_mlp_state_bytes(const althea_mlp* mlp);

/* Error monitor.  The prefiltered cube is the reference: the MLP is evaluated at
 * every texel centre of every mip m of d_prefiltered (size, mips; roughness
 * m/(mips-1), the roughness the cube was prefiltered at) and per texel
 *   e = mean over rgb of |mlp - cube| / max(|cube|, ALTHEA_MLP_MONITOR_FLOOR).
 * Each face is split into an 8 x 8 grid of cells (texel x of a size-S face
 * is in cell floor((x + 0.5) * 8 / S)); a cell's error is the mean e of its
 * texels and the cell is accepted when that is <= threshold.  The handle
 * keeps the acceptance for althea_deferred_resolve_mlp.  d_cell_error
 * (optional, device) receives mips * 6 * 64 floats [mip][face][cy][cx];
 * accepted (optional, host) the number of accepted cells, which synchronises
 * `stream`. */
#define ALTHEA_MLP_MONITOR_CELLS 64
#define ALTHEA_MLP_MONITOR_FLOOR 1e-2f
ALTHEA_API althea_status althea_mlp_build_monitor(althea_mlp* mlp, const float* d_prefiltered, int size, int mips,
                                                  float threshold, float* d_cell_error, int* accepted,
                                                  hipStream_t stream);

/* althea_deferred_resolve with the prefiltered-radiance term taken from the
 * MLP (BF16 MFMA inside the resolve, weights in LDS) where the monitor
 * accepted the cell of the reflection direction at both mips bracketing the
 * pixel's LOD roughness*(mips-1), and from the cube everywhere else.  The
 * monitor must have been built for ibl->prefiltered_size / prefiltered_mips.
 * d_mlp_used (optional): one byte per pixel, 1 where the MLP supplied the
 * term. */
ALTHEA_API althea_status althea_deferred_resolve_mlp(const althea_gbuffer_view* gbuffer, const althea_ibl_views* ibl,
                                                     const althea_mlp* mlp, const althea_point_light* d_lights,
                                                     int light_count, const althea_globals* globals,
                                                     const althea_ssao_desc* ssao, unsigned flags, void* d_out,
                                                     float* d_ao_out, uint8_t* d_mlp_used, hipStream_t stream);

/* Diagnostic: one 16x16x32 BF16 MFMA through the lane layout the MLP kernels
 * assume.  d_a 16x32, d_b 32x16 (row-major FP32, rounded to BF16 on load),
 * d_d 16x16 row-major FP32 = A . B when the assumed layout holds. */
ALTHEA_API althea_status althea_mfma_bf16_probe(const float* d_a, const float* d_b, float* d_d, hipStream_t stream);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* ALTHEA_ALTHEA_H_ */
