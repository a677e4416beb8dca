// Host-side internals shared by the C ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>

#include "althea/althea.h"

namespace althea {

void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));

#define ALTHEA_TRY_HIP(expr)                                                       \
  do {                                                                             \
    hipError_t althea_e_ = (expr);                                                 \
    if (althea_e_ != hipSuccess) {                                                 \
      ::althea::set_error("%s failed: %s", #expr, hipGetErrorString(althea_e_));    \
      return ALTHEA_ERROR_HIP;                                                     \
    }                                                                              \
  } while (0)

#define ALTHEA_REQUIRE(cond, ...)                 \
  do {                                            \
    if (!(cond)) {                                \
      ::althea::set_error(__VA_ARGS__);           \
      return ALTHEA_ERROR_INVALID_ARGUMENT;       \
    }                                             \
  } while (0)

inline bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

// Launch-error check after a kernel launch.
althea_status check_launch(const char* what);

// Pyramid geometry: per level (w, h, texel offset) plus the total texel count.
struct PyramidGeometry {
  int levels;
  int w[16], h[16];
  uint32_t off[16];
  size_t texels;
};
PyramidGeometry pyramid_geometry(int width, int height);

// Cube mip offsets (texels).
size_t cube_mip_offset(int size, int mip);

// Seamless-bordered copies of cubes (BorderedCube in device_math.hpp): mip
// offsets, total texels, and the kernel that writes one from a cube.
struct BorderedCube;
BorderedCube bordered_cube_of(int size, int mips);
size_t cube_bordered_texels(int size, int mips);
althea_status launch_cube_border(const float* cube, int size, int mips, float* bordered, hipStream_t s);

struct GgxEntry;
struct LevelC;
struct GatherLevel;

// Gather copy of pyramid levels first..last (each padded with its REPEAT
// neighbours, (w + 2) x (h + 2) texels, concatenated): texel offset of each
// level and the total.
struct GatherGeometry {
  int first, last;
  uint32_t off[16];
  size_t texels;
};
GatherGeometry gather_geometry(int w, int h, int first, int last);
size_t gather_bytes(int w, int h, int first, int last);
void gather_levels(int w, int h, const GatherGeometry& gg, GatherLevel* out);
althea_status launch_gather(const float* pyr, int w, int h, const GatherGeometry& gg, float* gpyr, hipStream_t s);
// The split form (SplitLevel): the same padded texels as an RG plane of
// float2 followed by a B plane of floats.
struct SplitLevel;
size_t split_gather_bytes(int w, int h, int first, int last);
SplitLevel split_level(int w, int h, const GatherGeometry& gg, int l);
althea_status launch_split_gather(const float* pyr, int w, int h, const GatherGeometry& gg, float* gpyr,
                                  hipStream_t s);
// Sample tables of a prefilter (the part of its workspace the allocation-free
// entry point keeps in the library's fixed array).
size_t prefilter_table_bytes(int mips, int samples);
// Waves per block of the single-map prefilter for a cube and the table
// entries per partial sum (`chunk`), which the batched kernel uses too.
int prefilter_split(int size, int mips, int samples, int& chunk);

// Kernel launchers (ibl_kernels.hip).
althea_status launch_ggx_tables(int samples, int mips, int env_w, int env_h, GgxEntry* tables, int* counts,
                                double* wsum, float* wsum32, hipStream_t s);
void pyramid_levels(int w, int h, LevelC* out);
althea_status launch_hammersley(uint32_t n, uint32_t first, uint32_t count, uint32_t* bits, float* xi,
                                hipStream_t s);
althea_status launch_pyramid(const float* level0, int w, int h, float* pyr, hipStream_t s);
althea_status launch_prefilter_cube(const float* pyr, int env_w, int env_h, int size, int mips, int samples,
                                    unsigned flags, float* cube, void* workspace, hipStream_t s);
size_t prefilter_workspace_bytes(int env_w, int env_h, int size, int mips, int samples);
// The same prefilter with its tables in the library's fixed device array.
althea_status launch_prefilter_cube_noalloc(const float* pyr, int env_w, int env_h, int size, int mips,
                                            int samples, unsigned flags, float* cube, hipStream_t s);
althea_status launch_prefilter_points(const float* pyr, int env_w, int env_h, int samples,
                                      const float* dir_rough, int count, float* out, void* workspace,
                                      hipStream_t s);
althea_status launch_irradiance(const float* pyr, int env_w, int env_h, int size, int theta_samples,
                                unsigned flags, float* cube, void* workspace, hipStream_t s);
size_t irradiance_workspace_bytes(int env_w, int env_h, int size, int theta_samples);
althea_status launch_brdf_lut(int size, int samples, float* rg, hipStream_t s);
althea_status launch_cube_sample(const float* cube, int size, int mips, const float* dir_lod, int count,
                                 float* out, void* workspace, hipStream_t s);
althea_status launch_env_sample(const float* pyr, int w, int h, const float* dir_lod, int count, float* out,
                                void* workspace, hipStream_t s);

// Batched prefilter into caller buffers (batched_kernels.hip): pyramids of
// `envs` into `pyramids[i]`, prefiltered cubes into `cubes[i]`, one graph
// launch; tuning NULL runs the autotuner.  Synchronises `stream`.
althea_status prefilter_batch_into(const althea_env_view* envs, int count, const int* face_sizes, int mips,
                                   int samples, unsigned flags, float* const* pyramids, float* const* cubes,
                                   const althea_tuning* tuning, althea_tuning* used, hipStream_t stream);

// Host: the view ray of DeferredPass.vert:10-22 / SSR.vert:15-22 is linear
// in NDC, so affine in the pixel centre (p + 0.5): ray0 + rayx x + rayy y;
// and projection * view, column-major (resolve_kernels.hip).
struct ViewRayDf;
void view_ray_affine(const althea_globals& g, int W, int H, float ray0[3], float rayx[3], float rayy[3],
                     ViewRayDf* df = nullptr);
void view_projection(const althea_globals& g, float vp[16]);

// Deferred resolve (resolve_kernels.hip); with `mlp` the prefiltered term
// comes from the MLP where its monitor accepts.
struct MlpInferenceView;
althea_status launch_resolve(const althea_gbuffer_view* g, const althea_ibl_views* ibl,
                             const althea_point_light* lights, int light_count, const althea_globals* globals,
                             const althea_ssao_desc* ssao, unsigned flags, void* out, float* ao_out,
                             const MlpInferenceView* mlp, uint8_t* mlp_used, hipStream_t s,
                             const althea_reflection_view* reflection = nullptr,
                             const float* shadow_cubes = nullptr, int shadow_size = 0);

// Point-light shadow cubes (shadow_kernels.hip).
size_t gbuffer_workspace_bytes(int width, int height, int triangle_count);
size_t gbuffer_gltf_workspace_bytes(int width, int height, int triangle_count, int primitive_count,
                                    int material_count, int texture_count);
althea_status launch_gbuffer_draw_gltf(const althea_gltf_primitive* prims, int prim_count,
                                       const althea_gltf_material* materials, int material_count,
                                       const althea_texture_view* textures, int texture_count,
                                       const althea_globals* g, const althea_gbuffer_view* out, void* workspace,
                                       hipStream_t s);
althea_status launch_texture_mips(uint8_t* texels, int w, int h, int mips, int srgb, hipStream_t s);
althea_status launch_gbuffer_draw(const althea_gbuffer_mesh* mesh, const althea_globals* g,
                                  const althea_gbuffer_view* out, void* workspace, hipStream_t s);
size_t shadow_workspace_bytes(int light_count, int triangle_count);
althea_status launch_shadow_cubes(const althea_mesh_view* mesh, const althea_point_light* lights, int light_count,
                                  int size, float* cubes, hipStream_t s,
                                  void* workspace = nullptr);
// The glTF shadow pass (alpha-tested): the mesh path's workspace for the
// draw's triangles, then the triangle soup, the alpha-test records and the
// primitive table.
size_t shadow_gltf_workspace_bytes(int light_count, int triangles, int prim_count);
althea_status launch_shadow_cubes_gltf(const althea_gltf_primitive* prims, int prim_count,
                                       const althea_gltf_material* materials, const althea_texture_view* textures,
                                       const althea_point_light* lights, int light_count, int size, float* cubes,
                                       void* workspace, hipStream_t s);

// Screen-space reflection (ssr_kernels.hip): level 0 of the reflection
// buffer, then the glossy convolution of levels 1 .. mips-1.
struct ReflChain;
void refl_chain(const althea_reflection_view& v, ReflChain& c);
size_t ssr_workspace_bytes(int width, int height);
althea_status launch_ssr_capture(const althea_gbuffer_view* g, const althea_ibl_views* ibl,
                                 const althea_point_light* lights, int light_count, const althea_globals* globals,
                                 const althea_reflection_view* reflection, void* workspace, hipStream_t s,
                                 const float* shadow_cubes = nullptr, int shadow_size = 0);
althea_status launch_reflection_convolve(const althea_reflection_view* reflection, hipStream_t s);

}  // namespace althea
