// C++ host facade over the C ABI, shaped like nithinp7/Althea's
// Include/Althea/ImageBasedLighting.h and Include/Althea/DeferredRendering.h:
// the same type and function names, device buffers in place of VkImages,
// std::runtime_error on failure as the reference throws
// (Src/ImageBasedLighting.cpp:424).  Header-only; link libalthea_amd.so.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "althea/althea.h"

namespace AltheaEngine {

inline void throwIfFailed(althea_status status, const char* where) {
  if (status != ALTHEA_OK)
    throw std::runtime_error(std::string(where) + ": " + althea_status_string(status) + ": " + althea_last_error());
}

// IBLHandles (ImageBasedLighting.h:17-22): device pointers in the same order
// (environment, prefiltered, irradiance, BRDF LUT) instead of bindless indices.
struct IBLHandles {
  const float* environmentMap;
  const float* prefilteredMap;
  const float* irradianceMap;
  const float* brdfLut;
};

// IBLResources (ImageBasedLighting.h:24-43): move-only owner of the device
// images of one environment.
class IBLResources {
 public:
  IBLResources() = default;
  explicit IBLResources(althea_ibl* ibl) : _ibl(ibl) {}

  IBLHandles getHandles() const {
    althea_ibl_views v = views();
    return {v.environment, v.prefiltered, v.irradiance, v.brdf_lut};
  }
  althea_ibl_views views() const {
    althea_ibl_views v{};
    throwIfFailed(althea_ibl_get_views(_ibl.get(), &v), "IBLResources::views");
    return v;
  }
  // Device bytes the shading pass reads (prefiltered + irradiance + LUT).
  size_t stateBytes() const { return althea_ibl_state_bytes(_ibl.get()); }
  althea_ibl* get() const { return _ibl.get(); }
  explicit operator bool() const { return static_cast<bool>(_ibl); }

 private:
  struct Deleter {
    void operator()(althea_ibl* p) const { althea_ibl_destroy(p); }
  };
  std::unique_ptr<althea_ibl, Deleter> _ibl;
};

namespace ImageBasedLighting {

inline althea_ibl_desc defaultDesc() {
  althea_ibl_desc d{};
  althea_ibl_desc_default(&d);
  return d;
}

// ImageBasedLighting::createResources (ImageBasedLighting.h:46-49): the
// environment arrives as a device equirect map instead of a file name; the
// irradiance cube, prefiltered mip chain and BRDF LUT are computed on
// `stream` (the reference precomputes equirect images, :137-412).
inline IBLResources createResources(const float* d_equirect, int width, int height,
                                    const althea_ibl_desc& desc = defaultDesc(), hipStream_t stream = nullptr) {
  althea_ibl* ibl = nullptr;
  throwIfFailed(althea_ibl_create(&desc, &ibl), "ImageBasedLighting::createResources");
  IBLResources res(ibl);
  throwIfFailed(althea_ibl_precompute(ibl, d_equirect, width, height, stream), "ImageBasedLighting::createResources");
  return res;
}

// createResources from an RGBE file, the reference's form (it loads
// Content/HDRI_Skybox/<name>.hdr and throws "Specified environment map does
// not exist!" when the file is missing, ImageBasedLighting.cpp:420-425).  The
// map must have power-of-two extents.  Synchronises `stream` before the
// staging buffer is released.
namespace detail {
// An RGBE file in device memory, freed with the object.
struct DeviceHdr {
  float* data = nullptr;
  int width = 0, height = 0;
  DeviceHdr(const std::string& path, hipStream_t stream) {
    althea_status st = althea_hdr_info(path.c_str(), &width, &height);
    if (st == ALTHEA_ERROR_NOT_FOUND) throw std::runtime_error("Specified environment map does not exist!");
    throwIfFailed(st, "ImageBasedLighting::createResources");
    std::vector<float> host(static_cast<size_t>(width) * height * 4);
    throwIfFailed(althea_hdr_read(path.c_str(), host.data(), host.size()), "ImageBasedLighting::createResources");
    if (hipMalloc(&data, host.size() * sizeof(float)) != hipSuccess)
      throw std::runtime_error("ImageBasedLighting::createResources: device allocation failed");
    if (hipMemcpyAsync(data, host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess) {
      (void)hipFree(data);
      throw std::runtime_error("ImageBasedLighting::createResources: upload failed");
    }
  }
  ~DeviceHdr() { (void)hipFree(data); }
  DeviceHdr(const DeviceHdr&) = delete;
  DeviceHdr& operator=(const DeviceHdr&) = delete;
};
}  // namespace detail

inline IBLResources createResources(const std::string& hdrPath, const althea_ibl_desc& desc = defaultDesc(),
                                    hipStream_t stream = nullptr) {
  detail::DeviceHdr env(hdrPath, stream);
  IBLResources res = createResources(env.data, env.width, env.height, desc, stream);
  if (hipStreamSynchronize(stream) != hipSuccess)
    throw std::runtime_error("ImageBasedLighting::createResources: precompute failed");
  return res;
}

// The reference's whole createResources(app, cmd, envMapName)
// (ImageBasedLighting.cpp:415-446) over a content directory: the source is
// <content>/HDRI_Skybox/<name>.hdr (missing: throws the reference's
// message); the maps are read from the cache <content>/PrecomputedMaps/<name>/
// when every file is there, else precomputed and written to it
// (althea_ibl_load_cache / althea_ibl_save_cache).
inline IBLResources createResources(const std::string& contentDir, const std::string& envMapName,
                                    const althea_ibl_desc& desc = defaultDesc(), hipStream_t stream = nullptr) {
  detail::DeviceHdr env(contentDir + "/HDRI_Skybox/" + envMapName + ".hdr", stream);
  const std::string cache = contentDir + "/PrecomputedMaps/" + envMapName;
  althea_ibl* ibl = nullptr;
  throwIfFailed(althea_ibl_create(&desc, &ibl), "ImageBasedLighting::createResources");
  IBLResources res(ibl);
  althea_status st = althea_ibl_load_cache(ibl, env.data, env.width, env.height, cache.c_str(), stream);
  if (st == ALTHEA_ERROR_NOT_FOUND) {
    throwIfFailed(althea_ibl_precompute(ibl, env.data, env.width, env.height, stream),
                  "ImageBasedLighting::createResources");
    st = althea_ibl_save_cache(ibl, cache.c_str(), stream);
  }
  throwIfFailed(st, "ImageBasedLighting::createResources");
  if (hipStreamSynchronize(stream) != hipSuccess)
    throw std::runtime_error("ImageBasedLighting::createResources: precompute failed");
  return res;
}

}  // namespace ImageBasedLighting

// GBufferHandles: the four planes in GBufferResources::buildMaterial order
// (Src/DeferredRendering.cpp:23-35).
struct GBufferHandles {
  const void* position;
  const void* normal;
  const void* albedo;
  const void* metallicRoughnessOcclusion;
};

// GBufferResources (DeferredRendering.h:38-102) over caller-owned device
// planes of vec4 per pixel (FP32 or FP16).
class GBufferResources {
 public:
  GBufferResources() = default;
  GBufferResources(const void* position, const void* normal, const void* albedo, const void* mro, int width,
                   int height, int format = ALTHEA_GBUFFER_FP32)
      : _view{position, normal, albedo, mro, width, height, format} {}

  GBufferHandles getHandles() const { return {_view.position, _view.normal, _view.albedo, _view.mro}; }
  const althea_gbuffer_view& view() const { return _view; }

 private:
  althea_gbuffer_view _view{};
};

// ReflectionBuffer (Src/ReflectionBuffer.cpp:20-94): the frame-sized
// reflection image with `mips` levels (5 in the reference), RGBA16F by
// default, plus the capture's depth-plane workspace.  Owns its device memory.
class ReflectionBuffer {
 public:
  ReflectionBuffer() = default;
  ReflectionBuffer(int width, int height, int mips = 5, int format = ALTHEA_REFLECTION_FP16) {
    const size_t texels = althea_reflection_texels(width, height, mips);
    if (texels == 0) throw std::runtime_error("ReflectionBuffer: invalid extent or level count");
    const size_t bytes = texels * (format == ALTHEA_REFLECTION_FP16 ? 8 : 16);
    void* data = nullptr;
    void* ws = nullptr;
    if (hipMalloc(&data, bytes) != hipSuccess) throw std::runtime_error("ReflectionBuffer: hipMalloc failed");
    _data.reset(data);
    if (hipMalloc(&ws, althea_ssr_workspace_bytes(width, height)) != hipSuccess)
      throw std::runtime_error("ReflectionBuffer: hipMalloc failed");
    _workspace.reset(ws);
    if (hipMemset(data, 0, bytes) != hipSuccess) throw std::runtime_error("ReflectionBuffer: hipMemset failed");
    _view = {data, width, height, mips, format};
  }

  const althea_reflection_view& view() const { return _view; }
  void* workspace() const { return _workspace.get(); }

 private:
  struct Free {
    void operator()(void* p) const { (void)hipFree(p); }
  };
  std::unique_ptr<void, Free> _data, _workspace;
  althea_reflection_view _view{};
};

// ScreenSpaceReflection (Src/ScreenSpaceReflection.cpp): captureReflection
// draws Shaders/SSR.frag into level 0, convolveReflectionBuffer fills the
// glossy levels (Src/ReflectionBuffer.cpp:163-277).
namespace ScreenSpaceReflection {

inline void captureReflection(const GBufferResources& gbuffer, const IBLResources& ibl,
                              const althea_point_light* d_lights, int lightCount, const althea_globals& globals,
                              ReflectionBuffer& reflection, hipStream_t stream = nullptr) {
  althea_ibl_views v = ibl.views();
  throwIfFailed(althea_ssr_capture(&gbuffer.view(), &v, d_lights, lightCount, &globals, &reflection.view(),
                                   reflection.workspace(), stream),
                "ScreenSpaceReflection::captureReflection");
}

inline void convolveReflectionBuffer(ReflectionBuffer& reflection, hipStream_t stream = nullptr) {
  throwIfFailed(althea_reflection_convolve(&reflection.view(), stream),
                "ScreenSpaceReflection::convolveReflectionBuffer");
}

}  // namespace ScreenSpaceReflection

// SceneToGBufferPass (Src/DeferredRendering.cpp:268-332) with the
// pass-through subpass (GBufferPassThrough.frag): begin() rasterizes a mesh
// with the camera of `globals` into the caller's G-buffer planes.  Owns the
// workspace (visibility keys and per-triangle setup), grown on demand.
class SceneToGBufferPass {
 public:
  SceneToGBufferPass() = default;
  explicit SceneToGBufferPass(const GBufferResources& gbuffer) : _gbuffer(gbuffer.view()) {}

  void begin(const althea_gbuffer_mesh& mesh, const althea_globals& globals, hipStream_t stream = nullptr) {
    const size_t need = althea_gbuffer_workspace_bytes(_gbuffer.width, _gbuffer.height, mesh.triangle_count);
    if (need == 0) throw std::runtime_error("SceneToGBufferPass: invalid extent or mesh");
    reserve(need);
    throwIfFailed(althea_gbuffer_draw(&mesh, &globals, &_gbuffer, _workspace.get(), stream),
                  "SceneToGBufferPass::begin");
  }

  // The glTF subpass: primitives drawn in order with their materials and
  // textures (host tables; vertex data and texels in device memory).
  void beginGltf(const althea_gltf_primitive* primitives, int primitiveCount, const althea_gltf_material* materials,
                 int materialCount, const althea_texture_view* textures, int textureCount,
                 const althea_globals& globals, hipStream_t stream = nullptr) {
    const size_t need = althea_gbuffer_gltf_workspace_bytes(_gbuffer.width, _gbuffer.height, primitives,
                                                            primitiveCount, materialCount, textureCount);
    if (need == 0) throw std::runtime_error("SceneToGBufferPass: invalid extent or primitives");
    reserve(need);
    throwIfFailed(althea_gbuffer_draw_gltf(primitives, primitiveCount, materials, materialCount, textures,
                                           textureCount, &globals, &_gbuffer, _workspace.get(), stream),
                  "SceneToGBufferPass::beginGltf");
  }

 private:
  void reserve(size_t need) {
    if (need <= _bytes) return;
    void* ws = nullptr;
    if (hipMalloc(&ws, need) != hipSuccess) throw std::runtime_error("SceneToGBufferPass: hipMalloc failed");
    _workspace.reset(ws);
    _bytes = need;
  }
  struct Free {
    void operator()(void* p) const { (void)hipFree(p); }
  };
  althea_gbuffer_view _gbuffer{};
  std::unique_ptr<void, Free> _workspace;
  size_t _bytes = 0;
};

// PointLightCollection (Include/Althea/PointLight.h:39-126): the point lights
// in device memory (PointLight layout) and, when created with shadow maps,
// their cubes (shadowSize^2 x 6 per light, 256 in the reference, stored with
// the one-texel face border althea_shadow_cubes writes).  setLight changes
// the host copy; updateResource uploads it.  Owns its device memory.
class PointLightCollection {
 public:
  PointLightCollection() = default;
  PointLightCollection(size_t lightCount, bool createShadowMap, int shadowSize = 256)
      : _host(lightCount), _shadowSize(shadowSize) {
    if (lightCount == 0) throw std::runtime_error("PointLightCollection: no lights");
    void* lights = nullptr;
    if (hipMalloc(&lights, lightCount * sizeof(althea_point_light)) != hipSuccess)
      throw std::runtime_error("PointLightCollection: hipMalloc failed");
    _lights.reset(lights);
    if (createShadowMap) {
      const size_t floats = althea_shadow_cube_floats(shadowSize) * lightCount;
      if (floats == 0) throw std::runtime_error("PointLightCollection: invalid shadow map size");
      void* cubes = nullptr;
      if (hipMalloc(&cubes, floats * sizeof(float)) != hipSuccess)
        throw std::runtime_error("PointLightCollection: hipMalloc failed");
      _cubes.reset(cubes);
    }
  }

  void setLight(uint32_t lightId, const althea_point_light& light) { _host.at(lightId) = light; }
  const althea_point_light& getLight(uint32_t lightId) const { return _host.at(lightId); }
  size_t getCount() const { return _host.size(); }

  void updateResource(hipStream_t stream = nullptr) {
    throwIfHip(hipMemcpyAsync(_lights.get(), _host.data(), _host.size() * sizeof(althea_point_light),
                              hipMemcpyHostToDevice, stream),
               "PointLightCollection::updateResource");
  }

  // The shadow pass (Src/PointLight.cpp:235-282) over a world-space mesh,
  // through the workspace variant (the workspace grows to the largest mesh
  // drawn and is kept).
  void drawShadowMaps(const althea_mesh_view& mesh, hipStream_t stream = nullptr) {
    if (!_cubes) throw std::runtime_error("PointLightCollection: created without shadow maps");
    const int n = static_cast<int>(getCount());
    const size_t need = althea_shadow_cubes_workspace_bytes(n, mesh.triangle_count);
    if (need == 0) throw std::runtime_error("PointLightCollection::drawShadowMaps: too many light-triangle pairs");
    reserveWorkspace(need);
    throwIfFailed(althea_shadow_cubes_ws(&mesh, lights(), n, _shadowSize, static_cast<float*>(_cubes.get()),
                                         _workspace.get(), stream),
                  "PointLightCollection::drawShadowMaps");
  }

  // The shadow pass over a model's glTF primitives with their materials: the
  // shadow shaders' alpha test (ShadowMapBindless.frag) discards fragments
  // whose base colour alpha is below the material's cut-off.
  void drawShadowMaps(const althea_gltf_primitive* primitives, int primitiveCount,
                      const althea_gltf_material* materials, int materialCount, const althea_texture_view* textures,
                      int textureCount, hipStream_t stream = nullptr) {
    if (!_cubes) throw std::runtime_error("PointLightCollection: created without shadow maps");
    const int n = static_cast<int>(getCount());
    const size_t need = althea_shadow_cubes_gltf_workspace_bytes(primitives, primitiveCount, n);
    if (need == 0) throw std::runtime_error("PointLightCollection::drawShadowMaps: invalid glTF draw");
    reserveWorkspace(need);
    throwIfFailed(althea_shadow_cubes_gltf(primitives, primitiveCount, materials, materialCount, textures,
                                           textureCount, lights(), n, _shadowSize, static_cast<float*>(_cubes.get()),
                                           _workspace.get(), stream),
                  "PointLightCollection::drawShadowMaps");
  }

  const althea_point_light* lights() const { return static_cast<const althea_point_light*>(_lights.get()); }
  const float* shadowMaps() const { return static_cast<const float*>(_cubes.get()); }
  int getShadowMapExtent() const { return _shadowSize; }

 private:
  static void throwIfHip(hipError_t e, const char* where) {
    if (e != hipSuccess) throw std::runtime_error(std::string(where) + ": " + hipGetErrorString(e));
  }
  // Grows the kept workspace to `need` bytes (hipFree waits for the work
  // that used the old one).
  void reserveWorkspace(size_t need) {
    if (need <= _workspaceBytes) return;
    void* ws = nullptr;
    if (hipMalloc(&ws, need) != hipSuccess)
      throw std::runtime_error("PointLightCollection::drawShadowMaps: hipMalloc failed");
    _workspace.reset(ws);
    _workspaceBytes = need;
  }
  struct Free {
    void operator()(void* p) const { (void)hipFree(p); }
  };
  std::vector<althea_point_light> _host;
  std::unique_ptr<void, Free> _lights, _cubes, _workspace;
  size_t _workspaceBytes = 0;
  int _shadowSize = 0;
};

namespace ScreenSpaceReflection {

// captureReflection with the lights of `lights` and, in the hit shading,
// their shadow test (pbrMaterial runs it for every caller).
inline void captureReflection(const GBufferResources& gbuffer, const IBLResources& ibl,
                              const PointLightCollection& lights, const althea_globals& globals,
                              ReflectionBuffer& reflection, hipStream_t stream = nullptr) {
  if (!lights.shadowMaps()) throw std::runtime_error("ScreenSpaceReflection: lights without shadow maps");
  althea_ibl_views v = ibl.views();
  throwIfFailed(althea_ssr_capture_shadowed(&gbuffer.view(), &v, lights.lights(),
                                            static_cast<int>(lights.getCount()), &globals, &reflection.view(),
                                            lights.shadowMaps(), lights.getShadowMapExtent(), reflection.workspace(),
                                            stream),
                "ScreenSpaceReflection::captureReflection");
}

}  // namespace ScreenSpaceReflection

namespace DeferredRendering {

// The deferred lighting resolve of DeferredPass.frag (one frame).
inline void resolve(const GBufferResources& gbuffer, const IBLResources& ibl, const althea_point_light* d_lights,
                    int lightCount, const althea_globals& globals, const althea_ssao_desc* ssao, void* d_out,
                    unsigned flags = 0, float* d_ao = nullptr, hipStream_t stream = nullptr) {
  althea_ibl_views v = ibl.views();
  throwIfFailed(althea_deferred_resolve(&gbuffer.view(), &v, d_lights, lightCount, &globals, ssao, flags, d_out, d_ao,
                                        stream),
                "DeferredRendering::resolve");
}

// The same resolve reading a reflection buffer (DeferredPass.frag:59-68).
inline void resolve(const GBufferResources& gbuffer, const IBLResources& ibl, const althea_point_light* d_lights,
                    int lightCount, const althea_globals& globals, const althea_ssao_desc* ssao,
                    const ReflectionBuffer& reflection, void* d_out, unsigned flags = 0, float* d_ao = nullptr,
                    hipStream_t stream = nullptr) {
  althea_ibl_views v = ibl.views();
  throwIfFailed(althea_deferred_resolve_reflect(&gbuffer.view(), &v, d_lights, lightCount, &globals, ssao,
                                                &reflection.view(), flags, d_out, d_ao, stream),
                "DeferredRendering::resolve");
}

// The resolve with the lights of `lights` and their shadow test
// (PBRMaterial.glsl:119-140); `reflection` may be null.
inline void resolve(const GBufferResources& gbuffer, const IBLResources& ibl, const PointLightCollection& lights,
                    const althea_globals& globals, const althea_ssao_desc* ssao, const ReflectionBuffer* reflection,
                    void* d_out, unsigned flags = 0, float* d_ao = nullptr, hipStream_t stream = nullptr) {
  if (!lights.shadowMaps()) throw std::runtime_error("DeferredRendering::resolve: lights without shadow maps");
  althea_ibl_views v = ibl.views();
  throwIfFailed(althea_deferred_resolve_shadowed(&gbuffer.view(), &v, lights.lights(),
                                                 static_cast<int>(lights.getCount()), &globals, ssao,
                                                 reflection ? &reflection->view() : nullptr, lights.shadowMaps(),
                                                 lights.getShadowMapExtent(), flags, d_out, d_ao, stream),
                "DeferredRendering::resolve");
}

}  // namespace DeferredRendering
}  // namespace AltheaEngine
