// Seeded synthetic inputs for the benchmark and test workloads (host C++,
// OpenMP).  Every value comes from a counter-based SplitMix64 stream keyed by
// (seed, stream, index), so the bytes are independent of the thread count.
//
//   environment  equirect RGB: per-texel log-normal(0, sigma) per channel plus
//                Gaussian sun lobes peak * exp(-theta^2 / (2 s^2)) at uniform
//                sphere directions (peaks log-uniform in [peak_min, peak_max]);
//                orientation of the shading convention
//                (Shaders/Global/GlobalResources.glsl:48-54)
//   camera       glm::lookAt view and glm::perspective (depth [0,1]) with the
//                Vulkan y flip of Src/Camera.cpp:94-103, near 0.01, far 1000
//   G-buffer     rays of the resolve's view-ray formula (DeferredPass.vert:10-22)
//                cast against a 4-octave value-noise fBm heightfield; normals
//                from its analytic gradient; albedo U[0,1]^3, metallic
//                Bernoulli(0.3), roughness U[0.05,1], AO 1; sky pixels w = 0
//   lights       uniform in the scene box, white, intensity U[1,50]
//   SSAO kernel  cosine-weighted hemisphere directions scaled by
//                0.1 + 0.9 (k/K)^2; noise tile 16 x (U[-1,1], U[-1,1], 0)
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

namespace {

constexpr double kPi = 3.14159265358979323846;

inline uint64_t splitmix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
// Uniform in [0, 1) from (seed, stream, index, draw).
inline double uni(uint64_t seed, uint64_t stream, uint64_t index, uint64_t draw) {
  uint64_t h = splitmix(seed ^ splitmix(stream * 0x100000001B3ull ^ splitmix(index * 8 + draw)));
  return static_cast<double>(h >> 11) * (1.0 / 9007199254740992.0);
}
inline double normal01(uint64_t seed, uint64_t stream, uint64_t index, uint64_t draw) {
  double u1 = uni(seed, stream, index, 2 * draw), u2 = uni(seed, stream, index, 2 * draw + 1);
  return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(2.0 * kPi * u2);
}

struct V3 {
  double x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 scl(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 norm(V3 a) { return scl(a, 1.0 / std::sqrt(dot(a, a))); }

inline V3 sphere_dir(double u1, double u2) {
  double z = 1.0 - 2.0 * u1, r = std::sqrt(std::max(0.0, 1.0 - z * z)), phi = 2.0 * kPi * u2;
  return {r * std::cos(phi), r * std::sin(phi), z};
}

// ------------------------------------------------------------ heightfield
struct Terrain {
  uint64_t seed;
  double amplitude = 1.5, scale = 4.0;
  int octaves = 4;

  double lattice(int64_t ix, int64_t iz, int oct) const {
    uint64_t h = splitmix(seed ^ splitmix(static_cast<uint64_t>(ix) * 0x9E3779B1ull ^
                                          splitmix(static_cast<uint64_t>(iz) * 0x85EBCA77ull + oct)));
    return static_cast<double>(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  }
  // Value noise with quintic fade; returns value and d/dx, d/dz.
  void noise(double x, double z, int oct, double& v, double& dx, double& dz) const {
    double fx = std::floor(x), fz = std::floor(z);
    double tx = x - fx, tz = z - fz;
    int64_t ix = static_cast<int64_t>(fx), iz = static_cast<int64_t>(fz);
    double a = lattice(ix, iz, oct), b = lattice(ix + 1, iz, oct);
    double c = lattice(ix, iz + 1, oct), d = lattice(ix + 1, iz + 1, oct);
    double ux = tx * tx * tx * (tx * (tx * 6 - 15) + 10), uz = tz * tz * tz * (tz * (tz * 6 - 15) + 10);
    double dux = 30 * tx * tx * (tx * (tx - 2) + 1), duz = 30 * tz * tz * (tz * (tz - 2) + 1);
    v = a + (b - a) * ux + (c - a) * uz + (a - b - c + d) * ux * uz;
    dx = dux * ((b - a) + (a - b - c + d) * uz);
    dz = duz * ((c - a) + (a - b - c + d) * ux);
  }
  void height(double x, double z, double& h, double& hx, double& hz) const {
    h = hx = hz = 0.0;
    double freq = 1.0 / scale, amp = amplitude;
    for (int o = 0; o < octaves; ++o) {
      double v, dx, dz;
      noise(x * freq, z * freq, o, v, dx, dz);
      h += amp * v;
      hx += amp * dx * freq;
      hz += amp * dz * freq;
      freq *= 2.0;
      amp *= 0.5;
    }
  }
  double h(double x, double z) const {
    double v, a, b;
    height(x, z, v, a, b);
    return v;
  }
};

}  // namespace

extern "C" {

// Equirect environment map, RGBA32F (alpha 1).  suns_out (optional) receives
// [suns][4] = (direction, peak).
void althea_synth_env(uint64_t seed, int w, int h, double sigma, int suns, double peak_min, double peak_max,
                      double sun_sigma_deg, float* out, float* suns_out) {
  const int kMaxSuns = 64;
  suns = std::min(suns, kMaxSuns);
  V3 dirs[kMaxSuns];
  double peaks[kMaxSuns];
  for (int k = 0; k < suns; ++k) {
    dirs[k] = sphere_dir(uni(seed, 1, k, 0), uni(seed, 1, k, 1));
    double t = uni(seed, 1, k, 2);
    peaks[k] = peak_min * std::pow(peak_max / peak_min, t);
    if (suns_out) {
      suns_out[4 * k] = static_cast<float>(dirs[k].x);
      suns_out[4 * k + 1] = static_cast<float>(dirs[k].y);
      suns_out[4 * k + 2] = static_cast<float>(dirs[k].z);
      suns_out[4 * k + 3] = static_cast<float>(peaks[k]);
    }
  }
  const double s = sun_sigma_deg * kPi / 180.0;
  const double inv2s2 = 1.0 / (2.0 * s * s);
#pragma omp parallel for schedule(static)
  for (int y = 0; y < h; ++y) {
    double pitch = (0.5 - (y + 0.5) / h) * kPi;
    for (int x = 0; x < w; ++x) {
      double yaw = ((x + 0.5) / w - 0.5) * 2.0 * kPi;
      V3 d{std::cos(pitch) * std::cos(yaw), std::sin(pitch), std::cos(pitch) * std::sin(yaw)};
      double sun = 0.0;
      for (int k = 0; k < suns; ++k) {
        double c = std::min(1.0, std::max(-1.0, dot(d, dirs[k])));
        double th = std::acos(c);
        sun += peaks[k] * std::exp(-th * th * inv2s2);
      }
      uint64_t idx = static_cast<uint64_t>(y) * w + x;
      float* o = out + 4 * idx;
      for (int c = 0; c < 3; ++c) o[c] = static_cast<float>(std::exp(sigma * normal01(seed, 2, idx, c)) + sun);
      o[3] = 1.0f;
    }
  }
}

// Camera matrices into the GlobalUniforms layout (6 column-major mat4 +
// mouse_uv, light_count, handle, time, exposure, input_mask, frame_count),
// 104 floats.  eye/center in world units, fovy in degrees.
void althea_synth_camera(const double* eye, const double* center, double fovy_deg, int width, int height,
                         int light_count, float exposure, float* globals) {
  V3 e{eye[0], eye[1], eye[2]}, c{center[0], center[1], center[2]};
  V3 f = norm(sub(c, e));
  V3 s = norm(cross(f, {0.0, 1.0, 0.0}));
  V3 u = cross(s, f);
  double view[16] = {0}, proj[16] = {0}, iview[16] = {0}, iproj[16] = {0};
  // glm::lookAtRH
  view[0] = s.x; view[4] = s.y; view[8] = s.z;
  view[1] = u.x; view[5] = u.y; view[9] = u.z;
  view[2] = -f.x; view[6] = -f.y; view[10] = -f.z;
  view[12] = -dot(s, e); view[13] = -dot(u, e); view[14] = dot(f, e);
  view[15] = 1.0;
  // glm::perspectiveRH_ZO then [1][1] *= -1 (Src/Camera.cpp:95-102)
  const double near = 0.01, far = 1000.0;
  double aspect = static_cast<double>(width) / height;
  double th = std::tan(0.5 * fovy_deg * kPi / 180.0);
  proj[0] = 1.0 / (aspect * th);
  proj[5] = -1.0 / th;
  proj[10] = far / (near - far);
  proj[11] = -1.0;
  proj[14] = -(far * near) / (far - near);
  // Inverses: rigid view, and the closed-form inverse of the perspective.
  iview[0] = s.x; iview[1] = s.y; iview[2] = s.z;
  iview[4] = u.x; iview[5] = u.y; iview[6] = u.z;
  iview[8] = -f.x; iview[9] = -f.y; iview[10] = -f.z;
  iview[12] = e.x; iview[13] = e.y; iview[14] = e.z; iview[15] = 1.0;
  iproj[0] = 1.0 / proj[0];
  iproj[5] = 1.0 / proj[5];
  iproj[11] = 1.0 / proj[14];
  iproj[14] = -1.0;
  iproj[15] = proj[10] / proj[14];
  std::memset(globals, 0, sizeof(float) * 104);
  for (int i = 0; i < 16; ++i) {
    globals[i] = static_cast<float>(proj[i]);
    globals[16 + i] = static_cast<float>(iproj[i]);
    globals[32 + i] = static_cast<float>(view[i]);
    globals[48 + i] = static_cast<float>(view[i]);
    globals[64 + i] = static_cast<float>(iview[i]);
    globals[80 + i] = static_cast<float>(iview[i]);
  }
  std::memcpy(&globals[98], &light_count, 4);
  globals[101] = exposure;
}

// G-buffer planes (RGBA32F each) for a camera produced by althea_synth_camera.
// amplitude 0 gives the flat plane y = 0.
void althea_synth_gbuffer(uint64_t seed, const float* globals, int width, int height, double amplitude,
                          float* position, float* normal, float* albedo, float* mro) {
  Terrain terrain{seed};
  terrain.amplitude = amplitude;
  const float* ip = globals + 16;
  const float* iv = globals + 64;
  V3 eye{iv[12], iv[13], iv[14]};
#pragma omp parallel for schedule(dynamic, 4)
  for (int py = 0; py < height; ++py) {
    for (int px = 0; px < width; ++px) {
      double nx = (px + 0.5) / width * 2.0 - 1.0, ny = (py + 0.5) / height * 2.0 - 1.0;
      double ev[3];
      for (int r = 0; r < 3; ++r) ev[r] = ip[0 * 4 + r] * nx + ip[1 * 4 + r] * ny + ip[3 * 4 + r];
      V3 d{iv[0] * ev[0] + iv[4] * ev[1] + iv[8] * ev[2], iv[1] * ev[0] + iv[5] * ev[1] + iv[9] * ev[2],
           iv[2] * ev[0] + iv[6] * ev[1] + iv[10] * ev[2]};
      d = norm(d);
      // March to the first sign change of ray height - terrain, then bisect.
      double t0 = 0.02, t = t0;
      double f0 = eye.y + d.y * t0 - terrain.h(eye.x + d.x * t0, eye.z + d.z * t0);
      bool hit = false;
      double lo = 0, hi = 0;
      // |h| <= amplitude * (1 + 1/2 + 1/4 + 1/8): a ray above that, not
      // descending, cannot hit.
      const double hmax = 1.875 * terrain.amplitude + 1e-6;
      while (t < 80.0 && !(d.y >= 0.0 && eye.y + d.y * t > hmax)) {
        double step = 0.03 + 0.015 * t;
        double tn = t + step;
        double fn = eye.y + d.y * tn - terrain.h(eye.x + d.x * tn, eye.z + d.z * tn);
        if (f0 > 0.0 && fn <= 0.0) {
          hit = true;
          lo = t;
          hi = tn;
          break;
        }
        t = tn;
        f0 = fn;
      }
      size_t idx = static_cast<size_t>(py) * width + px;
      float* P = position + 4 * idx;
      float* N = normal + 4 * idx;
      float* A = albedo + 4 * idx;
      float* M = mro + 4 * idx;
      if (!hit) {
        for (int c = 0; c < 4; ++c) P[c] = N[c] = A[c] = M[c] = 0.0f;
        continue;
      }
      for (int it = 0; it < 40; ++it) {
        double mid = 0.5 * (lo + hi);
        double fm = eye.y + d.y * mid - terrain.h(eye.x + d.x * mid, eye.z + d.z * mid);
        if (fm > 0.0) lo = mid; else hi = mid;
      }
      double th = 0.5 * (lo + hi);
      V3 hp = add(eye, scl(d, th));
      double hv, hx, hz;
      terrain.height(hp.x, hp.z, hv, hx, hz);
      V3 n = norm({-hx, 1.0, -hz});
      P[0] = static_cast<float>(hp.x); P[1] = static_cast<float>(hv); P[2] = static_cast<float>(hp.z); P[3] = 1.0f;
      N[0] = static_cast<float>(n.x); N[1] = static_cast<float>(n.y); N[2] = static_cast<float>(n.z); N[3] = 0.0f;
      for (int c = 0; c < 3; ++c) A[c] = static_cast<float>(uni(seed, 3, idx, c));
      A[3] = 1.0f;
      M[0] = uni(seed, 4, idx, 0) < 0.3 ? 1.0f : 0.0f;
      M[1] = static_cast<float>(0.05 + 0.95 * uni(seed, 4, idx, 1));
      M[2] = 1.0f;
      M[3] = 0.0f;
    }
  }
}

// The heightfield of althea_synth_gbuffer as a triangle mesh: an n x n grid
// of cells over [-half, half]^2 in x and z, vertices (n+1)^2 at the terrain
// height, two triangles per cell (indices 6 n^2).
void althea_synth_heightfield_mesh(uint64_t seed, double amplitude, int n, double half, float* positions,
                                   uint32_t* indices) {
  Terrain terrain{seed};
  terrain.amplitude = amplitude;
  for (int j = 0; j <= n; ++j)
    for (int i = 0; i <= n; ++i) {
      double x = -half + 2.0 * half * i / n, z = -half + 2.0 * half * j / n;
      double h, hx, hz;
      terrain.height(x, z, h, hx, hz);
      float* p = positions + 3 * (static_cast<size_t>(j) * (n + 1) + i);
      p[0] = static_cast<float>(x);
      p[1] = static_cast<float>(h);
      p[2] = static_cast<float>(z);
    }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      uint32_t a = static_cast<uint32_t>(j * (n + 1) + i), b = a + 1, c = a + (n + 1), d = c + 1;
      uint32_t* t = indices + 6 * (static_cast<size_t>(j) * n + i);
      t[0] = a; t[1] = c; t[2] = b;
      t[3] = b; t[4] = c; t[5] = d;
    }
}

// Point lights [n][8] (position xyz, 0, emission xyz, 0) in the box
// [-half, half] x [ymin, ymax] x [-half, half].
void althea_synth_lights(uint64_t seed, int n, double half, double ymin, double ymax, double imin, double imax,
                         float* out) {
  for (int i = 0; i < n; ++i) {
    float* o = out + 8 * i;
    o[0] = static_cast<float>(-half + 2.0 * half * uni(seed, 5, i, 0));
    o[1] = static_cast<float>(ymin + (ymax - ymin) * uni(seed, 5, i, 1));
    o[2] = static_cast<float>(-half + 2.0 * half * uni(seed, 5, i, 2));
    o[3] = 0.0f;
    float e = static_cast<float>(imin + (imax - imin) * uni(seed, 5, i, 3));
    o[4] = o[5] = o[6] = e;
    o[7] = 0.0f;
  }
}

// SSAO kernel [n][4] and noise [16][4].
void althea_synth_ssao(uint64_t seed, int n, float* kernel, float* noise) {
  for (int k = 0; k < n; ++k) {
    double u1 = uni(seed, 6, k, 0), u2 = uni(seed, 6, k, 1);
    double r = std::sqrt(u1), phi = 2.0 * kPi * u2;
    double t = static_cast<double>(k) / n;
    double s = 0.1 + 0.9 * t * t;
    kernel[4 * k] = static_cast<float>(r * std::cos(phi) * s);
    kernel[4 * k + 1] = static_cast<float>(r * std::sin(phi) * s);
    kernel[4 * k + 2] = static_cast<float>(std::sqrt(std::max(0.0, 1.0 - u1)) * s);
    kernel[4 * k + 3] = 0.0f;
  }
  for (int k = 0; k < 16; ++k) {
    noise[4 * k] = static_cast<float>(2.0 * uni(seed, 7, k, 0) - 1.0);
    noise[4 * k + 1] = static_cast<float>(2.0 * uni(seed, 7, k, 1) - 1.0);
    noise[4 * k + 2] = 0.0f;
    noise[4 * k + 3] = 0.0f;
  }
}

// Uniform-sphere directions with U[rmin, rmax] roughness: [n][4].
void althea_synth_dir_rough(uint64_t seed, int64_t n, double rmin, double rmax, float* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    V3 d = sphere_dir(uni(seed, 8, i, 0), uni(seed, 8, i, 1));
    float* o = out + 4 * i;
    o[0] = static_cast<float>(d.x);
    o[1] = static_cast<float>(d.y);
    o[2] = static_cast<float>(d.z);
    o[3] = static_cast<float>(rmin + (rmax - rmin) * uni(seed, 8, i, 2));
  }
}

}  // extern "C"
