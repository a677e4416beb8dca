// End-to-end use of the C++ facade (include/althea/ImageBasedLighting.hpp),
// written as a consumer of ImageBasedLighting::createResources and
// DeferredRendering would: constant environment -> closed-form checks on the
// prefiltered and irradiance cubes, then one resolve of a small G-buffer.
#include <hip/hip_runtime.h>

#include <cmath>
#include <filesystem>
#include <cstdio>
#include <string>
#include <unistd.h>
#include <vector>

#include "althea/ImageBasedLighting.hpp"

using namespace AltheaEngine;

#define HIP_OK(x)                                                     \
  do {                                                                \
    if ((x) != hipSuccess) {                                          \
      std::fprintf(stderr, "HIP failure %s at %d\n", #x, __LINE__); \
      return 1;                                                       \
    }                                                                 \
  } while (0)

int main() {
  const int W = 64, H = 32;
  std::vector<float> env(4 * W * H);
  for (int i = 0; i < W * H; ++i) {
    env[4 * i] = 2.0f;
    env[4 * i + 1] = 3.0f;
    env[4 * i + 2] = 5.0f;
    env[4 * i + 3] = 1.0f;
  }
  float* d_env = nullptr;
  HIP_OK(hipMalloc(&d_env, env.size() * 4));
  HIP_OK(hipMemcpy(d_env, env.data(), env.size() * 4, hipMemcpyHostToDevice));
  althea_ibl_desc desc = ImageBasedLighting::defaultDesc();
  desc.irradiance_size = 4;
  desc.prefiltered_size = 16;
  desc.prefiltered_mips = 5;
  desc.prefilter_samples = 64;
  desc.brdf_lut_size = 16;
  desc.brdf_lut_samples = 64;
  try {
    IBLResources ibl = ImageBasedLighting::createResources(d_env, W, H, desc);
    HIP_OK(hipDeviceSynchronize());
    IBLHandles h = ibl.getHandles();
    althea_ibl_views v = ibl.views();
    std::vector<float> pref(4 * althea_cube_texels(v.prefiltered_size, v.prefiltered_mips));
    std::vector<float> irr(4 * 6 * v.irradiance_size * v.irradiance_size);
    HIP_OK(hipMemcpy(pref.data(), h.prefilteredMap, pref.size() * 4, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(irr.data(), h.irradianceMap, irr.size() * 4, hipMemcpyDeviceToHost));
    double max_pref = 0.0, max_irr = 0.0;
    // Constant radiance: prefilter is the constant; irradiance is the
    // left Riemann sum (pi / 2P) cot(pi / 2P) of the constant, P = 150.
    const double P = 150.0, k = (M_PI / (2 * P)) / std::tan(M_PI / (2 * P));
    for (size_t t = 0; t < pref.size() / 4; ++t)
      for (int c = 0; c < 3; ++c)
        max_pref = std::fmax(max_pref, std::fabs(pref[4 * t + c] - env[c]) / env[c]);
    for (size_t t = 0; t < irr.size() / 4; ++t)
      for (int c = 0; c < 3; ++c)
        max_irr = std::fmax(max_irr, std::fabs(irr[4 * t + c] - env[c] * k) / (env[c] * k));
    // A 4x4 G-buffer facing the camera, one light.
    const int GW = 4, GH = 4;
    std::vector<float> pos(4 * GW * GH), nrm(4 * GW * GH), alb(4 * GW * GH), mro(4 * GW * GH);
    for (int i = 0; i < GW * GH; ++i) {
      pos[4 * i + 2] = -3.0f;
      pos[4 * i + 3] = 1.0f;
      nrm[4 * i + 2] = 1.0f;
      alb[4 * i] = alb[4 * i + 1] = alb[4 * i + 2] = 0.5f;
      mro[4 * i + 1] = 0.5f;
      mro[4 * i + 2] = 1.0f;
    }
    float *d_pos, *d_nrm, *d_alb, *d_mro, *d_out;
    althea_point_light light{{0.0f, 0.0f, 1.0f}, 0.0f, {10.0f, 10.0f, 10.0f}, 0.0f};
    althea_point_light* d_light;
    HIP_OK(hipMalloc(&d_pos, pos.size() * 4));
    HIP_OK(hipMalloc(&d_nrm, pos.size() * 4));
    HIP_OK(hipMalloc(&d_alb, pos.size() * 4));
    HIP_OK(hipMalloc(&d_mro, pos.size() * 4));
    HIP_OK(hipMalloc(&d_out, pos.size() * 4));
    HIP_OK(hipMalloc(&d_light, sizeof(light)));
    HIP_OK(hipMemcpy(d_pos, pos.data(), pos.size() * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d_nrm, nrm.data(), pos.size() * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d_alb, alb.data(), pos.size() * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d_mro, mro.data(), pos.size() * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d_light, &light, sizeof(light), hipMemcpyHostToDevice));
    althea_globals g{};
    for (int i = 0; i < 16; i += 5) g.projection[i] = g.inverse_projection[i] = g.view[i] = g.inverse_view[i] = 1.0f;
    g.inverse_projection[10] = 0.0f;
    g.inverse_projection[14] = -1.0f;  // view ray (ndc.x, ndc.y, -1)
    g.exposure = 1.0f;
    GBufferResources gbuf(d_pos, d_nrm, d_alb, d_mro, GW, GH);
    DeferredRendering::resolve(gbuf, ibl, d_light, 1, g, nullptr, d_out, ALTHEA_RESOLVE_TONEMAP);
    std::vector<float> out(pos.size());
    HIP_OK(hipMemcpy(out.data(), d_out, out.size() * 4, hipMemcpyDeviceToHost));
    bool finite = true;
    for (float f : out) finite = finite && std::isfinite(f) && f >= 0.0f && f <= 1.0f;
    std::printf("prefiltered max rel %.3g irradiance max rel %.3g resolve finite %d state bytes %zu\n", max_pref,
                max_irr, finite ? 1 : 0, ibl.stateBytes());
    // Screen-space reflections: capture, glossy levels, the resolve reading them.
    ReflectionBuffer refl(GW, GH, 5);
    ScreenSpaceReflection::captureReflection(gbuf, ibl, d_light, 1, g, refl);
    ScreenSpaceReflection::convolveReflectionBuffer(refl);
    DeferredRendering::resolve(gbuf, ibl, d_light, 1, g, nullptr, refl, d_out, ALTHEA_RESOLVE_TONEMAP);
    HIP_OK(hipMemcpy(out.data(), d_out, out.size() * 4, hipMemcpyDeviceToHost));
    bool refl_finite = true;
    for (float f : out) refl_finite = refl_finite && std::isfinite(f) && f >= 0.0f && f <= 1.0f;
    bool extent_threw = false;
    try {
      ReflectionBuffer wrong(GW + 1, GH, 5);
      ScreenSpaceReflection::captureReflection(gbuf, ibl, d_light, 1, g, wrong);
    } catch (const std::runtime_error& e) {
      extent_threw = true;
    }
    std::printf("reflections finite %d, extent mismatch threw %d\n", refl_finite ? 1 : 0, extent_threw ? 1 : 0);
    finite = finite && refl_finite && extent_threw;
    // Point-light shadows: a square between the light and the surface,
    // facing the light, hides it; the frame is then the frame without lights.
    PointLightCollection lights(1, true, 64);
    lights.setLight(0, light);
    lights.updateResource();
    const float quad[12] = {-2.0f, -2.0f, -1.0f, 2.0f, -2.0f, -1.0f, 2.0f, 2.0f, -1.0f, -2.0f, 2.0f, -1.0f};
    const uint32_t tris[6] = {0, 1, 2, 0, 2, 3};  // counter-clockwise seen from the light at z = 1
    float* d_quad;
    uint32_t* d_tris;
    HIP_OK(hipMalloc(&d_quad, sizeof(quad)));
    HIP_OK(hipMalloc(&d_tris, sizeof(tris)));
    HIP_OK(hipMemcpy(d_quad, quad, sizeof(quad), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d_tris, tris, sizeof(tris), hipMemcpyHostToDevice));
    lights.drawShadowMaps(althea_mesh_view{d_quad, d_tris, 4, 2});
    DeferredRendering::resolve(gbuf, ibl, lights, g, nullptr, nullptr, d_out, ALTHEA_RESOLVE_TONEMAP);
    std::vector<float> shadowed(out.size()), unlit(out.size());
    HIP_OK(hipMemcpy(shadowed.data(), d_out, out.size() * 4, hipMemcpyDeviceToHost));
    DeferredRendering::resolve(gbuf, ibl, nullptr, 0, g, nullptr, d_out, ALTHEA_RESOLVE_TONEMAP);
    HIP_OK(hipMemcpy(unlit.data(), d_out, out.size() * 4, hipMemcpyDeviceToHost));
    float shadow_diff = 0.0f;
    for (size_t i = 0; i < out.size(); ++i) shadow_diff = std::fmax(shadow_diff, std::fabs(shadowed[i] - unlit[i]));
    std::printf("shadowed frame vs unlit frame: max diff %.3g\n", shadow_diff);
    finite = finite && shadow_diff <= 1e-6f;
    // The G-buffer from a mesh: the occluder square drawn by the camera looking
    // down -z covers the whole 4x4 frame at z = -1, facing it.
    {
      const float white[12] = {0.8f, 0.8f, 0.8f, 0.8f, 0.8f, 0.8f, 0.8f, 0.8f, 0.8f, 0.8f, 0.8f, 0.8f};
      const float mro4[12] = {0.0f, 0.5f, 1.0f, 0.0f, 0.5f, 1.0f, 0.0f, 0.5f, 1.0f, 0.0f, 0.5f, 1.0f};
      float *d_white, *d_mro4;
      HIP_OK(hipMalloc(&d_white, sizeof(white)));
      HIP_OK(hipMalloc(&d_mro4, sizeof(mro4)));
      HIP_OK(hipMemcpy(d_white, white, sizeof(white), hipMemcpyHostToDevice));
      HIP_OK(hipMemcpy(d_mro4, mro4, sizeof(mro4), hipMemcpyHostToDevice));
      HIP_OK(hipMemcpy(d_quad, quad, sizeof(quad), hipMemcpyHostToDevice));
      althea_globals cam = g;  // a [0, 1]-depth perspective, 90 degrees, near 0.01, far 1000
      for (float& v : cam.projection) v = 0.0f;
      for (float& v : cam.inverse_projection) v = 0.0f;
      const float n = 0.01f, f = 1000.0f;
      cam.projection[0] = 1.0f;
      cam.projection[5] = -1.0f;
      cam.projection[10] = f / (n - f);
      cam.projection[11] = -1.0f;
      cam.projection[14] = -(f * n) / (f - n);
      cam.inverse_projection[0] = 1.0f;
      cam.inverse_projection[5] = -1.0f;
      cam.inverse_projection[11] = 1.0f / cam.projection[14];
      cam.inverse_projection[14] = -1.0f;
      cam.inverse_projection[15] = cam.projection[10] / cam.projection[14];
      float *gp, *gn, *ga, *gm;
      HIP_OK(hipMalloc(&gp, pos.size() * 4));
      HIP_OK(hipMalloc(&gn, pos.size() * 4));
      HIP_OK(hipMalloc(&ga, pos.size() * 4));
      HIP_OK(hipMalloc(&gm, pos.size() * 4));
      HIP_OK(hipMemcpy(d_tris, tris, sizeof(tris), hipMemcpyHostToDevice));
      GBufferResources produced(gp, gn, ga, gm, GW, GH);
      SceneToGBufferPass pass(produced);
      pass.begin(althea_gbuffer_mesh{d_quad, nullptr, d_white, d_mro4, d_tris, 4, 2}, cam);
      std::vector<float> gpos(pos.size()), gnrm(pos.size());
      HIP_OK(hipMemcpy(gpos.data(), gp, gpos.size() * 4, hipMemcpyDeviceToHost));
      HIP_OK(hipMemcpy(gnrm.data(), gn, gnrm.size() * 4, hipMemcpyDeviceToHost));
      bool gb_ok = true;
      for (int i = 0; i < GW * GH; ++i)
        gb_ok = gb_ok && gpos[4 * i + 3] == 1.0f && std::fabs(gpos[4 * i + 2] + 1.0f) < 1e-5f &&
                std::fabs(gnrm[4 * i + 2] + 1.0f) < 1e-5f;  // derivative normal: away from the eye
      std::printf("G-buffer from a mesh: every pixel covered at z = -1, normal away from the eye: %d\n",
                  gb_ok ? 1 : 0);
      finite = finite && gb_ok;
      for (float* q : {gp, gn, ga, gm, d_white, d_mro4}) HIP_OK(hipFree(q));
    }
    HIP_OK(hipFree(d_quad));
    HIP_OK(hipFree(d_tris));
    // Errors come back as exceptions, as the reference throws.
    bool threw = false;
    try {
      ImageBasedLighting::createResources(d_env, 60, 32, desc);
    } catch (const std::runtime_error& e) {
      threw = true;
      std::printf("rejected non-power-of-two source: %s\n", e.what());
    }
    // The reference's file form: a missing .hdr throws its message; an .hdr
    // written here precomputes like the device buffer did.
    bool missing_threw = false;
    try {
      ImageBasedLighting::createResources(std::string("/nonexistent/Skybox.hdr"), desc);
    } catch (const std::runtime_error& e) {
      missing_threw = std::string(e.what()) == "Specified environment map does not exist!";
    }
    const std::string path = "/tmp/althea_facade_demo_" + std::to_string(getpid()) + ".hdr";
    throwIfFailed(althea_hdr_write(path.c_str(), env.data(), W, H), "althea_hdr_write");
    IBLResources from_file = ImageBasedLighting::createResources(path, desc);
    std::remove(path.c_str());
    std::vector<float> pref2(pref.size());
    HIP_OK(hipMemcpy(pref2.data(), from_file.getHandles().prefilteredMap, pref2.size() * 4, hipMemcpyDeviceToHost));
    const bool same = pref2 == pref;  // 2, 3 and 5 are exact in RGBE
    std::printf("missing file threw %d, file route identical %d\n", missing_threw ? 1 : 0, same ? 1 : 0);
    // The reference's content-directory form with its cache: the first call
    // precomputes and writes PrecomputedMaps/Demo/, the second reads it.
    const std::string content = "/tmp/althea_facade_content_" + std::to_string(getpid());
    std::filesystem::create_directories(content + "/HDRI_Skybox");
    throwIfFailed(althea_hdr_write((content + "/HDRI_Skybox/Demo.hdr").c_str(), env.data(), W, H), "althea_hdr_write");
    IBLResources computed = ImageBasedLighting::createResources(content, std::string("Demo"), desc);
    const bool cached = std::filesystem::exists(content + "/PrecomputedMaps/Demo/Prefiltered1.hdr") &&
                        std::filesystem::exists(content + "/PrecomputedMaps/Demo/IrradianceMap.hdr");
    IBLResources loaded = ImageBasedLighting::createResources(content, std::string("Demo"), desc);
    std::vector<float> pref3(pref.size());
    HIP_OK(hipMemcpy(pref3.data(), loaded.getHandles().prefilteredMap, pref3.size() * 4, hipMemcpyDeviceToHost));
    double max_cache = 0.0;  // RGBE: 8 bits under the pixel's shared exponent
    for (size_t i = 0; i < pref.size(); i += 4) {
      const double big = std::max(pref[i], std::max(pref[i + 1], pref[i + 2]));
      for (int c = 0; c < 3; ++c) max_cache = std::max(max_cache, std::fabs(pref3[i + c] - pref[i + c]) / big);
    }
    std::filesystem::remove_all(content);
    std::printf("cache written %d, cache reload max rel %.3g\n", cached ? 1 : 0, max_cache);
    const bool ok = max_pref < 1e-5 && max_irr < 1e-5 && finite && threw && missing_threw && same && cached &&
                    max_cache < 1.0 / 128;
    std::printf("facade %s\n", ok ? "ok" : "FAILED");
    return ok ? 0 : 2;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "exception: %s\n", e.what());
    return 3;
  }
}
