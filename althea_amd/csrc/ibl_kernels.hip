// IBL precompute kernels for gfx950: Hammersley, the environment pyramid, the
// per-roughness GGX sample tables, the prefiltered cube, the irradiance cube,
// the split-sum BRDF LUT, point prefiltering and sampler probes.
//
// The GGX prefilter (Shaders/IBL_Precompute/PreFilterEnvMap.comp:126-177)
// evaluated at a cube texel with N = V = R has every per-sample quantity except
// the frame rotation independent of the texel: with h the local half vector,
// V.H = h.z, so L = frame * (2 h.z h.x, 2 h.z h.y, 2 h.z^2 - 1), NdotL = the z
// term and the source LOD depends on h.z only.  A one-block kernel per
// roughness builds these entries in FP64 (rounded once to FP32), drops samples
// with NdotL <= 0 in sample order, and sums the weights in FP64.  The texel
// kernel then reads one wave-uniform entry per sample (scalar loads), rotates
// it by its frame and filters the source in software.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>

#include "device_math.hpp"
#include "internal.hpp"
#include "precise_lookup.hpp"
#include "prefilter_common.hpp"

namespace althea {

// -------------------------------------------------------------- Hammersley
__global__ void hammersley_kernel(uint32_t n, uint32_t first, uint32_t count, uint32_t* __restrict__ bits,
                                  float2* __restrict__ xi) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  uint32_t i = first + k;
  if (bits) bits[k] = vdc_bits(i);
  if (xi) {
    float x1, x2;
    hammersley(i, n, x1, x2);
    xi[k] = make_float2(x1, x2);
  }
}

althea_status launch_hammersley(uint32_t n, uint32_t first, uint32_t count, uint32_t* bits, float* xi,
                                hipStream_t s) {
  if (count == 0) return ALTHEA_OK;
  dim3 grid((count + 255) / 256);
  hipLaunchKernelGGL(hammersley_kernel, grid, dim3(256), 0, s, n, first, count, bits,
                     reinterpret_cast<float2*>(xi));
  return check_launch("hammersley_kernel");
}

// ----------------------------------------------------------------- pyramid
// Level l = mean of the 2x2 (2x1 / 1x2 at unit extents) block of level l-1,
// the LINEAR half-size blit of Image::generateMipMaps (Src/Image.cpp:135-239).
__global__ void pyramid_level_kernel(const float4* __restrict__ src, int sw, int sh, float4* __restrict__ dst,
                                     int dw, int dh) {
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  int y = blockIdx.y;
  if (x >= dw) return;
  int fx = sw > dw ? 2 : 1, fy = sh > dh ? 2 : 1;
  const float4* r0 = src + (size_t)(y * fy) * sw + x * fx;
  float4 a = r0[0];
  float4 acc = a;
  float scale = 1.0f;
  if (fx == 2 && fy == 2) {
    float4 b = r0[1], c = r0[sw], d = r0[sw + 1];
    acc = make_float4(((a.x + b.x) + (c.x + d.x)), ((a.y + b.y) + (c.y + d.y)), ((a.z + b.z) + (c.z + d.z)),
                      ((a.w + b.w) + (c.w + d.w)));
    scale = 0.25f;
  } else if (fx == 2) {
    float4 b = r0[1];
    acc = make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
    scale = 0.5f;
  } else if (fy == 2) {
    float4 b = r0[sw];
    acc = make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
    scale = 0.5f;
  }
  dst[(size_t)y * dw + x] = make_float4(acc.x * scale, acc.y * scale, acc.z * scale, acc.w * scale);
}

althea_status launch_pyramid(const float* level0, int w, int h, float* pyr, hipStream_t s) {
  PyramidGeometry g = pyramid_geometry(w, h);
  if (level0 != pyr)
    ALTHEA_TRY_HIP(hipMemcpyAsync(pyr, level0, sizeof(float4) * (size_t)w * h, hipMemcpyDeviceToDevice, s));
  float4* p = reinterpret_cast<float4*>(pyr);
  for (int l = 1; l < g.levels; ++l) {
    dim3 grid((g.w[l] + 255) / 256, g.h[l]);
    hipLaunchKernelGGL(pyramid_level_kernel, grid, dim3(256), 0, s, p + g.off[l - 1], g.w[l - 1], g.h[l - 1],
                       p + g.off[l], g.w[l], g.h[l]);
    althea_status st = check_launch("pyramid_level_kernel");
    if (st != ALTHEA_OK) return st;
  }
  return ALTHEA_OK;
}

// ------------------------------------------------------- GGX sample tables
// One block per roughness level.  Entry e (in sample order, NdotL > 0 only):
// L in the frame (T, B, N), and the source LOD clamped to the pyramid's
// [0, levels-1] split into (l0, l1 = min(l0+1, levels-1), delta).
// counts[m] = kept entries, wsum[m] = FP64 sum of L.n, wsum32[m] = the same
// sum accumulated in FP32 in sample order.
struct TableParams {
  int samples;
  int mips;
  int env_levels;
  double sa_texel;  // 4 pi / (6 W H)  (PreFilterEnvMap.comp:158)
};

constexpr int kTableThreads = 256;

__global__ __launch_bounds__(kTableThreads) void ggx_table_kernel(TableParams p, GgxEntry* __restrict__ tables,
                                                                  int* __restrict__ counts,
                                                                  double* __restrict__ wsum,
                                                                  float* __restrict__ wsum32) {
  const int m = blockIdx.x;
  const int tid = threadIdx.x;
  const double rough = p.mips > 1 ? static_cast<double>(m) / (p.mips - 1) : 0.0;
  const double a2 = rough * rough;  // alpha = roughness (PreFilterEnvMap.comp:139-140)
  const double two_pi = 6.283185307179586476925;
  __shared__ int s_scan[kTableThreads];
  __shared__ double s_sum[kTableThreads];
  __shared__ float s_keep_w[kTableThreads];
  __shared__ int s_base;
  __shared__ float s_w32;
  if (tid == 0) {
    s_base = 0;
    s_w32 = 0.0f;
  }
  double my_sum = 0.0;
  GgxEntry* tab = tables + static_cast<size_t>(m) * p.samples;
  __syncthreads();
  for (int chunk = 0; chunk < p.samples; chunk += kTableThreads) {
    int s = chunk + tid;
    bool keep = false;
    GgxEntry entry{};
    float w32 = 0.0f;
    if (s < p.samples) {
      double x1 = static_cast<double>(s) / p.samples;
      double x2 = static_cast<double>(__builtin_bitreverse32(static_cast<uint32_t>(s))) * (1.0 / 4294967296.0);
      double den = 1.0 + (a2 - 1.0) * x2;
      double cos_t = sqrt((1.0 - x2) / den);
      double sin_t = sqrt(a2 * x2 / den);  // = sqrt(1 - cos_t^2) without cancellation
      double phi = two_pi * x1;
      double hx = cos(phi) * sin_t, hy = sin(phi) * sin_t, hz = cos_t;
      double lz = 2.0 * hz * hz - 1.0;
      if (lz > 0.0) {
        keep = true;
        double lod = 0.0;
        if (rough != 0.0) {
          double t = hz * hz * (a2 - 1.0) + 1.0;
          double D = a2 / (3.14159265358979323846 * t * t);
          double pdf = D * hz / (4.0 * hz + 0.00001);
          double sa_sample = 1.0 / (p.samples * pdf + 0.0001);
          lod = 0.5 * log2(sa_sample / p.sa_texel);
        }
        lod = fmin(fmax(lod, 0.0), static_cast<double>(p.env_levels - 1));
        // Rounded to FP32 as a sampler LOD would be, then split.
        const float lodf = static_cast<float>(lod);
        const float fl = floorf(lodf);
        entry.lx = static_cast<float>(2.0 * hz * hx);
        entry.ly = static_cast<float>(2.0 * hz * hy);
        entry.lz = static_cast<float>(lz);
        const double delta = static_cast<double>(lodf - fl);
        entry.l0 = static_cast<int>(fl);
        entry.l1 = min(entry.l0 + 1, p.env_levels - 1);
        entry.w0 = static_cast<float>(lz * (1.0 - delta));
        entry.w1 = static_cast<float>(lz * delta);
        entry.weight = static_cast<float>(lz);
        my_sum += lz;
        w32 = static_cast<float>(lz);
      }
    }
    // Block-wide exclusive scan of keep flags (Hillis-Steele in LDS).
    s_scan[tid] = keep ? 1 : 0;
    s_keep_w[tid] = w32;
    __syncthreads();
    for (int off = 1; off < kTableThreads; off <<= 1) {
      int v = tid >= off ? s_scan[tid - off] : 0;
      __syncthreads();
      s_scan[tid] += v;
      __syncthreads();
    }
    int incl = s_scan[tid];
    if (keep) tab[s_base + incl - 1] = entry;
    if (tid == 0) {
      // FP32 weight sum in sample order, the reference's totalWeight += NdotL.
      float acc = s_w32;
      for (int k = 0; k < kTableThreads; ++k) acc += s_keep_w[k];
      s_w32 = acc;
    }
    __syncthreads();
    if (tid == kTableThreads - 1) s_base += incl;
    __syncthreads();
  }
  s_sum[tid] = my_sum;
  __syncthreads();
  for (int off = kTableThreads / 2; off > 0; off >>= 1) {
    if (tid < off) s_sum[tid] += s_sum[tid + off];
    __syncthreads();
  }
  if (tid == 0) {
    counts[m] = s_base;
    wsum[m] = s_sum[0];
    wsum32[m] = s_w32;
    // Roughness 0 (alpha = 0): every half vector is N, so every kept sample
    // is L = N at LOD 0 with NdotL = 1 and the entries are bit-identical.  One
    // entry carrying their summed weight gives the same weighted mean (v * n /
    // n, exact for n <= 2^24) with one lookup instead of n.
    if (rough == 0.0 && s_base > 1) {
      tab[0].weight = tab[0].w0 = static_cast<float>(s_base);
      counts[m] = 1;
    }
  }
}

// ------------------------------------------------------------ prefilter
struct PrefilterParams {
  LevelC lv[kMaxLevels];        // source pyramid levels
  GatherLevel glv[kMaxLevels];  // the same levels in the gather copy
  CubeMips cube;
  int samples;
  int chunk;                        // table entries per partial sum (prefilter_sums)
  int block_start[kMaxLevels + 1];  // blocks before each mip
};

constexpr int kPrefilterTile = 8;

// kGather: the sample loop reads the padded gather copy `gpyr`; otherwise the
// pyramid (bit-identical sums, more address arithmetic per lookup).
// kSplit waves per block.  A rough mip's block is one 8x8 tile whose table
// is cut into chunks of p.chunk entries, wave k summing chunk k; wave 0 adds
// the partials in chunk order (prefilter_sums: the bits of one pass).  A
// roughness-0 block holds kSplit tiles, one per wave.  kSplit > 1 fills the
// GPU when the rough mips have few tiles (a 128^2 cube: 510 tiles, half a
// wave per SIMD).
template <bool kF64Sums, bool kGather, int kSplit>
__global__ __launch_bounds__(64 * kSplit) void prefilter_cube_kernel(const char* __restrict__ pyr,
                                                                     const char* __restrict__ gpyr,
                                                                     const GgxEntry* __restrict__ tables,
                                                                     const int* __restrict__ counts,
                                                                     const double* __restrict__ wsum,
                                                                     const PrefilterParams p,
                                                                     float4* __restrict__ cube) {
  const int b = blockIdx.x;
  int m = 0;
  while (m + 1 < p.cube.mips && b >= p.block_start[m + 1]) ++m;
  const int S = max(1, p.cube.size >> m);
  const int tpf = (S + kPrefilterTile - 1) / kPrefilterTile;
  const auto level = [&](int l) -> const LevelC& { return p.lv[l]; };
  const bool mirror = m == 0 || p.cube.mips == 1;  // roughness m / (mips - 1) == 0 (block-uniform)
  if (kSplit == 1) {
    const int local = b - p.block_start[m];
    const int face = local / (tpf * tpf);
    const int t = local - face * tpf * tpf;
    const int x = (t % tpf) * kPrefilterTile + (threadIdx.x & 7);
    const int y = (t / tpf) * kPrefilterTile + (threadIdx.x >> 3);
    const bool active = x < S && y < S;
    float4 out;
    if (mirror) {
      // the mirror lookup
      out = prefilter_texel_mirror(pyr, level, face, min(x, S - 1), min(y, S - 1), S);
    } else {
      const Frame fr = local_to_world(cube_texel_raw_dir(face, min(x, S - 1), min(y, S - 1), S));
      const GgxEntry* __restrict__ tab = tables + static_cast<size_t>(m) * p.samples;
      if (kGather)
        out = prefilter_value<kF64Sums>(prefilter_sums_whole<kF64Sums>(GatherAcc{gpyr, p.glv}, tab, 0, counts[m], fr),
                                        wsum[m]);
      else
        out = prefilter_value<kF64Sums>(prefilter_sums_whole<kF64Sums>(PyramidAcc{pyr, p.lv}, tab, 0, counts[m], fr),
                                        wsum[m]);
    }
    if (active) cube[p.cube.off[m] + (static_cast<size_t>(face) * S + y) * S + x] = out;
    return;
  }
  using T = typename PrefilterSums<kF64Sums>::T;
  __shared__ T s_part[kSplit][4][64];
  const int tiles = 6 * tpf * tpf;
  const int wave = static_cast<int>(threadIdx.x >> 6), lane = static_cast<int>(threadIdx.x & 63);
  const int local = mirror ? (b - p.block_start[m]) * kSplit + wave : b - p.block_start[m];
  const int lt = min(local, tiles - 1);
  const int face = lt / (tpf * tpf);
  const int t = lt - face * tpf * tpf;
  const int x = (t % tpf) * kPrefilterTile + (lane & 7);
  const int y = (t / tpf) * kPrefilterTile + (lane >> 3);
  const bool active = local < tiles && x < S && y < S;
  const size_t dst = p.cube.off[m] + (static_cast<size_t>(face) * S + min(y, S - 1)) * S + min(x, S - 1);
  if (mirror) {
    const float4 out = prefilter_texel_mirror(pyr, level, face, min(x, S - 1), min(y, S - 1), S);
    if (active) cube[dst] = out;
    return;
  }
  const Frame fr = local_to_world(cube_texel_raw_dir(face, min(x, S - 1), min(y, S - 1), S));
  const GgxEntry* __restrict__ tab = tables + static_cast<size_t>(m) * p.samples;
  const int count = counts[m];
  const int lo = min(count, wave * p.chunk), hi = min(count, lo + p.chunk);
  // A wave's range lies inside one chunk: its sums are one pass.
  const PrefilterSums<kF64Sums> part =
      kGather ? prefilter_sums_whole<kF64Sums>(GatherAcc{gpyr, p.glv}, tab, lo, hi, fr)
              : prefilter_sums_whole<kF64Sums>(PyramidAcc{pyr, p.lv}, tab, lo, hi, fr);
  s_part[wave][0][lane] = part.r;
  s_part[wave][1][lane] = part.g;
  s_part[wave][2][lane] = part.b;
  s_part[wave][3][lane] = static_cast<T>(part.w);
  __syncthreads();
  if (wave == 0) {
    PrefilterSums<kF64Sums> tot;
    for (int k = 0; k < kSplit; ++k) {
      tot.r += s_part[k][0][lane];
      tot.g += s_part[k][1][lane];
      tot.b += s_part[k][2][lane];
      tot.w += static_cast<float>(s_part[k][3][lane]);
    }
    if (active) cube[dst] = prefilter_value<kF64Sums>(tot, wsum[m]);
  }
}

// Waves per block of the single-map prefilter and the table entries per
// partial sum: the rough mips' tiles times the waves reach 4096 (four waves
// per SIMD) with at most 8 waves and 64 entries per chunk; chunks are
// multiples of the 32-entry FP32 blocks.  One wave: a chunk longer than any
// table.  The batched kernel cuts each map's sums the same way.
int prefilter_split(int size, int mips, int samples, int& chunk) {
  long long rough = 0;
  for (int m = 1; m < mips; ++m) {
    const int S = std::max(1, size >> m);
    const int tpf = (S + kPrefilterTile - 1) / kPrefilterTile;
    rough += 6ll * tpf * tpf;
  }
  int k = 1;
  while (k < 8 && rough > 0 && rough * k < 4096 && samples / (2 * k) >= 64) k *= 2;
  chunk = k == 1 ? (1 << 30) : ((samples + k - 1) / k + 31) / 32 * 32;
  return k;
}

static size_t align256(size_t v) { return (v + 255) & ~static_cast<size_t>(255); }

void pyramid_levels(int w, int h, LevelC* out) {
  PyramidGeometry g = pyramid_geometry(w, h);
  for (int l = 0; l < g.levels; ++l) out[l] = make_level(g.w[l], g.h[l], g.off[l]);
}

// ------------------------------------------------------------ gather copy
// Levels first..last of a pyramid, each padded with its REPEAT neighbours
// ((w + 2) x (h + 2) texels, GatherLevel), concatenated from texel 0.
GatherGeometry gather_geometry(int w, int h, int first, int last) {
  const PyramidGeometry g = pyramid_geometry(w, h);
  GatherGeometry out{};
  out.first = first;
  out.last = last;
  size_t total = 0;
  for (int l = first; l <= last; ++l) {
    out.off[l] = static_cast<uint32_t>(total);
    total += static_cast<size_t>(g.w[l] + 2) * (g.h[l] + 2);
  }
  out.texels = total;
  return out;
}

size_t gather_bytes(int w, int h, int first, int last) {
  if (w <= 0 || h <= 0) return 0;
  return align256(16 * gather_geometry(w, h, first, last).texels);
}

void gather_levels(int w, int h, const GatherGeometry& gg, GatherLevel* out) {
  const PyramidGeometry g = pyramid_geometry(w, h);
  for (int l = gg.first; l <= gg.last; ++l) out[l] = make_gather_level(g.w[l], g.h[l], gg.off[l]);
}

struct PadParams {
  int first, last;
  int w[kMaxLevels], h[kMaxLevels];
  uint32_t src[kMaxLevels];      // pyramid texel offset of each level
  uint32_t dst[kMaxLevels + 1];  // gather texel offset of each level; dst[last + 1] = total
};

// Texels of a split gather copy's RG plane (the B plane starts 8 times this
// many bytes in): the total rounded up to 64, so both planes start 256-byte
// aligned.
static size_t split_plane_texels(const GatherGeometry& gg) { return (gg.texels + 63) & ~static_cast<size_t>(63); }

size_t split_gather_bytes(int w, int h, int first, int last) {
  if (w <= 0 || h <= 0) return 0;
  return 12 * split_plane_texels(gather_geometry(w, h, first, last));
}

SplitLevel split_level(int w, int h, const GatherGeometry& gg, int l) {
  const PyramidGeometry g = pyramid_geometry(w, h);
  const LevelC c = make_level(g.w[l], g.h[l], 0);
  SplitLevel s{};
  s.kx = c.kx;
  s.cx = c.cx;
  s.ky = c.ky;
  s.cy = c.cy;
  s.rowbytes = static_cast<uint32_t>(g.w[l] + 2) * 8u;
  s.rg = 8u * gg.off[l];
  s.b = static_cast<uint32_t>(8 * split_plane_texels(gg)) + 4u * gg.off[l];
  return s;
}

// One thread per gather texel: padded (r, c) of level l is the pyramid texel
// ((r - 1) mod h, (c - 1) mod w).  kSplit writes it as the split copy's RG
// and B planes (plane_texels apart) instead of one float4.
template <bool kSplit>
__global__ __launch_bounds__(256) void gather_pad_kernel(const float4* __restrict__ pyr, const PadParams p,
                                                         float4* __restrict__ gpyr, size_t plane_texels) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= p.dst[p.last + 1]) return;
  int l = p.first;
  while (l < p.last && i >= p.dst[l + 1]) ++l;
  const int w = p.w[l], h = p.h[l];
  const uint32_t k = i - p.dst[l];
  const int r = static_cast<int>(k / static_cast<uint32_t>(w + 2));
  const int c = static_cast<int>(k) - r * (w + 2);
  const int sr = r == 0 ? h - 1 : (r == h + 1 ? 0 : r - 1);
  const int sc = c == 0 ? w - 1 : (c == w + 1 ? 0 : c - 1);
  const float4 v = pyr[p.src[l] + static_cast<size_t>(sr) * w + sc];
  if (kSplit) {
    reinterpret_cast<float2*>(gpyr)[i] = make_float2(v.x, v.y);
    reinterpret_cast<float*>(reinterpret_cast<float2*>(gpyr) + plane_texels)[i] = v.z;
  } else {
    gpyr[i] = v;
  }
}

static althea_status launch_gather_as(const float* pyr, int w, int h, const GatherGeometry& gg, float* gpyr,
                                      bool split, hipStream_t s) {
  const PyramidGeometry g = pyramid_geometry(w, h);
  PadParams p{};
  p.first = gg.first;
  p.last = gg.last;
  for (int l = gg.first; l <= gg.last; ++l) {
    p.w[l] = g.w[l];
    p.h[l] = g.h[l];
    p.src[l] = g.off[l];
    p.dst[l] = gg.off[l];
  }
  p.dst[gg.last + 1] = static_cast<uint32_t>(gg.texels);
  const uint32_t blocks = static_cast<uint32_t>((gg.texels + 255) / 256);
  if (split)
    hipLaunchKernelGGL(gather_pad_kernel<true>, dim3(blocks), dim3(256), 0, s, reinterpret_cast<const float4*>(pyr),
                       p, reinterpret_cast<float4*>(gpyr), split_plane_texels(gg));
  else
    hipLaunchKernelGGL(gather_pad_kernel<false>, dim3(blocks), dim3(256), 0, s, reinterpret_cast<const float4*>(pyr),
                       p, reinterpret_cast<float4*>(gpyr), size_t{0});
  return check_launch("gather_pad_kernel");
}

althea_status launch_gather(const float* pyr, int w, int h, const GatherGeometry& gg, float* gpyr, hipStream_t s) {
  return launch_gather_as(pyr, w, h, gg, gpyr, false, s);
}

althea_status launch_split_gather(const float* pyr, int w, int h, const GatherGeometry& gg, float* gpyr,
                                  hipStream_t s) {
  return launch_gather_as(pyr, w, h, gg, gpyr, true, s);
}

static CubeMips cube_mips_of(int size, int mips) {
  CubeMips c{};
  c.size = size;
  c.mips = mips;
  for (int m = 0; m < mips; ++m) c.off[m] = static_cast<uint32_t>(cube_mip_offset(size, m));
  return c;
}


althea_status launch_ggx_tables(int samples, int mips, int env_w, int env_h, GgxEntry* tables, int* counts,
                                double* wsum, float* wsum32, hipStream_t s) {
  TableParams tp;
  tp.samples = samples;
  tp.mips = mips;
  tp.env_levels = althea_env_levels(env_w, env_h);
  tp.sa_texel = 4.0 * M_PI / (6.0 * static_cast<double>(env_w) * env_h);
  hipLaunchKernelGGL(ggx_table_kernel, dim3(mips), dim3(kTableThreads), 0, s, tp, tables, counts, wsum, wsum32);
  return check_launch("ggx_table_kernel");
}

size_t prefilter_table_bytes(int mips, int samples) {
  return align256(sizeof(GgxEntry) * static_cast<size_t>(mips) * samples) + align256(sizeof(int) * mips) +
         align256(sizeof(double) * mips) + align256(sizeof(float) * mips);
}

// The sample tables, then the gather copy of every level.
size_t prefilter_workspace_bytes(int env_w, int env_h, int, int mips, int samples) {
  return prefilter_table_bytes(mips, samples) + gather_bytes(env_w, env_h, 0, althea_env_levels(env_w, env_h) - 1);
}

// ------------------------------------------- allocation-free prefilter
// The sample tables of althea_prefilter_cube_noalloc: a fixed device array,
// one copy per GPU, held until the process ends.  In host code `scratch`
// names a host-side stand-in, so its device address is looked up once per
// device with hipGetSymbolAddress and only that address is handed to kernels.
static_assert(ALTHEA_PREFILTER_SCRATCH_MAX_BYTES % 256 == 0, "scratch size");
static __device__ __align__(256) unsigned char scratch[ALTHEA_PREFILTER_SCRATCH_MAX_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Device address of `scratch` on the current device.
static althea_status prefilter_scratch(void** out) {
  int dev = 0;
  ALTHEA_TRY_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("prefilter_cube_noalloc: device %d beyond the %d supported", dev, kMaxDevices);
    return ALTHEA_ERROR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    ALTHEA_TRY_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return ALTHEA_OK;
}

// Every table entry, count and weight sum the prefilter reads is written by
// the table kernel of the same call before the texel kernel runs, so nothing
// left in the array by an earlier call is read and nothing needs clearing.
static althea_status prefilter_cube_tables_in(const float* pyr, int env_w, int env_h, int size, int mips,
                                              int samples, unsigned flags, float* cube, void* workspace,
                                              float* gather, hipStream_t s);

althea_status launch_prefilter_cube_noalloc(const float* pyr, int env_w, int env_h, int size, int mips,
                                            int samples, unsigned flags, float* cube, hipStream_t s) {
  void* ws = nullptr;
  althea_status st = prefilter_scratch(&ws);
  if (st != ALTHEA_OK) return st;
  return prefilter_cube_tables_in(pyr, env_w, env_h, size, mips, samples, flags, cube, ws, nullptr, s);
}

// Tables in `workspace`; the sample loop reads the gather copy `gather` when
// one is given (written here first), else the pyramid.
static althea_status prefilter_cube_tables_in(const float* pyr, int env_w, int env_h, int size, int mips,
                                              int samples, unsigned flags, float* cube, void* workspace,
                                              float* gather, hipStream_t s) {
  char* ws = static_cast<char*>(workspace);
  GgxEntry* tables = reinterpret_cast<GgxEntry*>(ws);
  ws += align256(sizeof(GgxEntry) * static_cast<size_t>(mips) * samples);
  int* counts = reinterpret_cast<int*>(ws);
  ws += align256(sizeof(int) * mips);
  double* wsum = reinterpret_cast<double*>(ws);
  ws += align256(sizeof(double) * mips);
  float* wsum32 = reinterpret_cast<float*>(ws);

  althea_status st = launch_ggx_tables(samples, mips, env_w, env_h, tables, counts, wsum, wsum32, s);
  if (st != ALTHEA_OK) return st;

  PrefilterParams p{};
  pyramid_levels(env_w, env_h, p.lv);
  if (gather) {
    const GatherGeometry gg = gather_geometry(env_w, env_h, 0, althea_env_levels(env_w, env_h) - 1);
    st = launch_gather(pyr, env_w, env_h, gg, gather, s);
    if (st != ALTHEA_OK) return st;
    gather_levels(env_w, env_h, gg, p.glv);
  }
  p.cube = cube_mips_of(size, mips);
  p.samples = samples;
  const int split = prefilter_split(size, mips, samples, p.chunk);
  int blocks = 0;
  for (int m = 0; m < mips; ++m) {
    p.block_start[m] = blocks;
    int S = std::max(1, size >> m);
    int tpf = (S + kPrefilterTile - 1) / kPrefilterTile;
    const int tiles = 6 * tpf * tpf;
    blocks += (m == 0 || mips == 1) ? (tiles + split - 1) / split : tiles;
  }
  p.block_start[mips] = blocks;
  const char* pyr_b = reinterpret_cast<const char*>(pyr);
  const char* g_b = reinterpret_cast<const char*>(gather);
  float4* cube4 = reinterpret_cast<float4*>(cube);
  const bool f64 = !(flags & ALTHEA_PRECOMPUTE_FP32_SUMS);
  const int variant = (f64 ? 1 : 0) | (gather ? 2 : 0);
  switch (split) {
#define ALTHEA_PREFILTER_LAUNCH(k)                                                                                   \
  case k: {                                                                                                          \
    auto kernel = variant == 3   ? prefilter_cube_kernel<true, true, k>                                              \
                  : variant == 1 ? prefilter_cube_kernel<true, false, k>                                             \
                  : variant == 2 ? prefilter_cube_kernel<false, true, k>                                             \
                                 : prefilter_cube_kernel<false, false, k>;                                           \
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(64 * k), 0, s, pyr_b, g_b, tables, counts, wsum, p, cube4);        \
    break;                                                                                                           \
  }
    ALTHEA_PREFILTER_LAUNCH(1)
    ALTHEA_PREFILTER_LAUNCH(2)
    ALTHEA_PREFILTER_LAUNCH(4)
    ALTHEA_PREFILTER_LAUNCH(8)
#undef ALTHEA_PREFILTER_LAUNCH
    default:
      break;
  }
  return check_launch("prefilter_cube_kernel");
}

althea_status launch_prefilter_cube(const float* pyr, int env_w, int env_h, int size, int mips, int samples,
                                    unsigned flags, float* cube, void* workspace, hipStream_t s) {
  float* gather = reinterpret_cast<float*>(static_cast<char*>(workspace) + prefilter_table_bytes(mips, samples));
  return prefilter_cube_tables_in(pyr, env_w, env_h, size, mips, samples, flags, cube, workspace, gather, s);
}

// -------------------------------------------------------- point prefilter
// Same estimator at per-lane (direction, roughness): the sample quantities are
// evaluated per lane in FP32 (sin^2 theta formed without cancellation), sums
// in FP32 blocks flushed to FP64.
struct PointParams {
  LevelC lv[kMaxLevels];
  int levels;
  int samples;
  int count;
  float log2_sa_texel;
};

// Trilinear lookup with per-lane LOD: level constants from LDS.
__device__ __forceinline__ f3 lds_trilinear(const char* __restrict__ pyr, const LevelC* s_lv, int levels, float yaw,
                                            float pitch, float lod) {
  const float d = fminf(fmaxf(lod, 0.0f), static_cast<float>(levels - 1));
  const float fl = floorf(d);
  const int l0 = static_cast<int>(fl);
  const float delta = d - fl;
  f3 v = level_bilinear(pyr, s_lv[l0], yaw, pitch);
  if (delta != 0.0f) v = lerp3(v, level_bilinear(pyr, s_lv[min(l0 + 1, levels - 1)], yaw, pitch), delta);
  return v;
}

__global__ __launch_bounds__(256) void prefilter_points_kernel(const char* __restrict__ pyr, const PointParams p,
                                                               const float4* __restrict__ dir_rough,
                                                               float4* __restrict__ out) {
  __shared__ LevelC s_lv[kMaxLevels];
  if (threadIdx.x < kMaxLevels) s_lv[threadIdx.x] = p.lv[threadIdx.x];
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.count) return;
  const float4 q = dir_rough[i];
  const Frame fr = local_to_world(mk(q.x, q.y, q.z));
  const float rough = q.w;
  const float a2 = rough * rough;
  const float fn = static_cast<float>(p.samples);
  const float max_lod = static_cast<float>(p.levels - 1);
  double ar = 0.0, ag = 0.0, ab = 0.0, aw = 0.0;
  float sr = 0.0f, sg = 0.0f, sb = 0.0f, sw = 0.0f;
  for (int s = 0; s < p.samples; ++s) {
    float x1, x2;
    hammersley(static_cast<uint32_t>(s), static_cast<uint32_t>(p.samples), x1, x2);
    const float den = fmaf(a2 - 1.0f, x2, 1.0f);
    const float cos_t = sqrtf((1.0f - x2) / den);
    const float sin_t = sqrtf(a2 * x2 / den);
    float sp, cp;
    sincospif(2.0f * x1, &sp, &cp);
    const float hz = cos_t;
    const float lz = fmaf(2.0f * hz, hz, -1.0f);
    if (lz > 0.0f) {
      const float lx = 2.0f * hz * (cp * sin_t), ly = 2.0f * hz * (sp * sin_t);
      float lod = 0.0f;
      if (rough != 0.0f) {
        const float t = fmaf(hz * hz, a2 - 1.0f, 1.0f);
        const float D = a2 / (kPi * t * t);
        const float pdf = D * hz / fmaf(4.0f, hz, 0.00001f);
        lod = 0.5f * (-log2f(fmaf(fn, pdf, 0.0001f)) - p.log2_sa_texel);
      }
      lod = fminf(fmaxf(lod, 0.0f), max_lod);
      f3 L = fr.t * lx + fr.b * ly + fr.n * lz;
      float yaw, pitch;
      equirect_angles(L, yaw, pitch);
      f3 v = lds_trilinear(pyr, s_lv, p.levels, yaw, pitch, lod);
      sr = fmaf(v.x, lz, sr);
      sg = fmaf(v.y, lz, sg);
      sb = fmaf(v.z, lz, sb);
      sw += lz;
    }
    if ((s & (kBlockSums - 1)) == kBlockSums - 1) {
      ar += sr; ag += sg; ab += sb; aw += sw;
      sr = sg = sb = sw = 0.0f;
    }
  }
  ar += sr; ag += sg; ab += sb; aw += sw;
  const double inv = 1.0 / aw;
  out[i] = make_float4(static_cast<float>(ar * inv), static_cast<float>(ag * inv), static_cast<float>(ab * inv),
                       1.0f);
}

althea_status launch_prefilter_points(const float* pyr, int env_w, int env_h, int samples, const float* dir_rough,
                                      int count, float* out, void*, hipStream_t s) {
  if (count == 0) return ALTHEA_OK;
  PointParams p{};
  pyramid_levels(env_w, env_h, p.lv);
  p.levels = althea_env_levels(env_w, env_h);
  p.samples = samples;
  p.count = count;
  p.log2_sa_texel = static_cast<float>(std::log2(4.0 * M_PI / (6.0 * static_cast<double>(env_w) * env_h)));
  hipLaunchKernelGGL(prefilter_points_kernel, dim3((count + 255) / 256), dim3(256), 0, s,
                     reinterpret_cast<const char*>(pyr), p, reinterpret_cast<const float4*>(dir_rough),
                     reinterpret_cast<float4*>(out));
  return check_launch("prefilter_points_kernel");
}

// ------------------------------------------------------------ irradiance
// GenIrradianceMap.comp:106-155 per cube texel: the Riemann sum over
// theta_samples azimuths x phi_samples polar angles of radiance(d) cos sin,
// read at the fixed source LOD log2(W / theta_samples), times
// pi / (theta_samples phi_samples).
struct IrradianceParams {
  SplitLevel lv0, lv1;  // the two source levels around LOD log2(W / theta_samples), split gather copies
  SplitLevel grid;      // their blend grid (blend)
  int blend;
  int levels;
  int size;
  int theta_samples;
  int phi_samples;
  int l0, l1;
  float delta;
  double scale;  // pi / theta_samples / phi_samples
};

// phi_tab[j] = (cos, sin, cos sin (1 - delta), cos sin delta): the sample
// weight split over the two source levels (delta = the LOD's fraction);
// with the blend grid, whose values carry the split, (cos, sin, cos sin, 0).
__global__ void irradiance_tables_kernel(int theta_samples, int phi_samples, float delta, int blend,
                                         float2* __restrict__ theta_tab, float4* __restrict__ phi_tab) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double pi = 3.14159265358979323846;
  if (i < theta_samples) {
    double th = i * 2.0 * pi / theta_samples;
    theta_tab[i] = make_float2(static_cast<float>(cos(th)), static_cast<float>(sin(th)));
  }
  if (i < phi_samples) {
    double ph = i * 0.5 * pi / phi_samples;
    double c = cos(ph), s = sin(ph), d = static_cast<double>(delta);
    phi_tab[i] = blend ? make_float4(static_cast<float>(c), static_cast<float>(s), static_cast<float>(c * s), 0.0f)
                       : make_float4(static_cast<float>(c), static_cast<float>(s),
                                     static_cast<float>(c * s * (1.0 - d)), static_cast<float>(c * s * d));
  }
}

// One wave per output texel.  A lane owns one polar column (phi index
// 8 k + lane / 8) per pass and walks the azimuth rows theta = 8 s + lane % 8,
// so for a fixed step the wave's 64 lookups cover a compact patch of 8 polar
// by 8 azimuth samples (4.8 degrees of polar angle by at most 9.6 degrees,
// times sin(phi), of azimuth at the configured 300 x 150) and the gathers of
// a step share cache lines; a ring of 64 azimuths reaches 77 degrees.  Of
// the patches 4 x 16, 8 x 8, 16 x 4 and 32 x 2 (polar x azimuth), 8 x 8
// wastes the fewest lanes on the configured 150 x 300 samples: 19 passes of
// 38 steps (722 wave-steps) against 10 of 75 (750) for 16 x 4.  Each lane sums its column's samples over theta in FP32 in theta
// order and flushes the column into FP64 (kF64Sums) or FP32; the lanes are
// combined by a fixed xor tree across the wave.
constexpr int kIrrPhiLanes = 8, kIrrThetaLanes = 8;
static_assert(kIrrPhiLanes * kIrrThetaLanes == 64, "one wave per texel");

// kLevels: 1 (one source level), 2 (two levels, each its own lookup) or 3
// (the two levels' blend grid, one lookup of weight w0).
template <int kLevels>
__device__ __forceinline__ void irradiance_column(const char* __restrict__ gpyr, const float2* __restrict__ theta_tab,
                                                  const IrradianceParams& p, f3 u, f3 v, f3 w, float w0, float w1,
                                                  int a, float& sr, float& sg, float& sb) {
  for (int i = a; i < p.theta_samples; i += kIrrThetaLanes) {
    const float2 th = theta_tab[i];
    // n cos(phi) + (t cos(theta) + b sin(theta)) sin(phi), with u = t sin(phi),
    // v = b sin(phi), w = n cos(phi).
    const f3 d = mk(fmaf(u.x, th.x, fmaf(v.x, th.y, w.x)), fmaf(u.y, th.x, fmaf(v.y, th.y, w.y)),
                    fmaf(u.z, th.x, fmaf(v.z, th.y, w.z)));
    float yaw, pitch;
    equirect_angles(d, yaw, pitch);
    if (kLevels == 3) {
      blend_bilinear_acc(gpyr, p.grid, yaw, pitch, w0, sr, sg, sb);
    } else {
      split_bilinear_acc(gpyr, p.lv0, yaw, pitch, w0, sr, sg, sb);
      if (kLevels == 2) split_bilinear_acc(gpyr, p.lv1, yaw, pitch, w1, sr, sg, sb);
    }
  }
}

// Azimuth tables of up to this many rows are read from LDS (kLds): the
// table's loads then leave the texture data return, which bounds the kernel.
constexpr int kIrrThetaLdsMax = 4096;

template <bool kF64Sums, bool kLds>
__global__ __launch_bounds__(256) void irradiance_cube_kernel(const char* __restrict__ gpyr,
                                                              const float2* __restrict__ theta_glob,
                                                              const float4* __restrict__ phi_tab,
                                                              const IrradianceParams p, float4* __restrict__ cube) {
  extern __shared__ float2 s_theta[];  // theta_samples rows when kLds
  if (kLds) {
    for (int i = threadIdx.x; i < p.theta_samples; i += 256) s_theta[i] = theta_glob[i];
    __syncthreads();
  }
  const float2* theta_tab = kLds ? s_theta : theta_glob;
  const int S = p.size;
  const int texel = blockIdx.x * 4 + static_cast<int>(threadIdx.x >> 6);
  if (texel >= 6 * S * S) return;  // whole waves; no block barrier follows
  const int lane = threadIdx.x & 63;
  const int face = texel / (S * S);
  const int rem = texel - face * S * S;
  const int y = rem / S, x = rem - y * S;
  const Frame fr = local_to_world(cube_texel_raw_dir(face, x, y, S));
  const int a = lane % kIrrThetaLanes, b = lane / kIrrThetaLanes;
  double ar = 0.0, ag = 0.0, ab = 0.0;
  float fr32 = 0.0f, fg32 = 0.0f, fb32 = 0.0f;
  for (int j = b; j < p.phi_samples; j += kIrrPhiLanes) {
    const float4 ph = phi_tab[j];
    const f3 u = fr.t * ph.y, v = fr.b * ph.y, w = fr.n * ph.x;
    float sr = 0.0f, sg = 0.0f, sb = 0.0f;
    if (p.blend)
      irradiance_column<3>(gpyr, theta_tab, p, u, v, w, ph.z, ph.w, a, sr, sg, sb);
    else if (p.delta != 0.0f)
      irradiance_column<2>(gpyr, theta_tab, p, u, v, w, ph.z, ph.w, a, sr, sg, sb);
    else
      irradiance_column<1>(gpyr, theta_tab, p, u, v, w, ph.z, ph.w, a, sr, sg, sb);
    if (kF64Sums) {
      ar += sr;
      ag += sg;
      ab += sb;
    } else {
      fr32 += sr;
      fg32 += sg;
      fb32 += sb;
    }
  }
  if (kF64Sums) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      ar += __shfl_xor(ar, o, 64);
      ag += __shfl_xor(ag, o, 64);
      ab += __shfl_xor(ab, o, 64);
    }
    if (lane == 0)
      cube[texel] = make_float4(static_cast<float>(ar * p.scale), static_cast<float>(ag * p.scale),
                                static_cast<float>(ab * p.scale), 1.0f);
  } else {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      fr32 += __shfl_xor(fr32, o, 64);
      fg32 += __shfl_xor(fg32, o, 64);
      fb32 += __shfl_xor(fb32, o, 64);
    }
    const float sc = static_cast<float>(p.scale);
    if (lane == 0) cube[texel] = make_float4(fr32 * sc, fg32 * sc, fb32 * sc, 1.0f);
  }
}

static int irradiance_phi_samples(int w, int h, int theta) {
  // int(pushConstants.height * thetaSamples / pushConstants.width) in FP32
  // (GenIrradianceMap.comp:135-136).
  return static_cast<int>(static_cast<float>(h) * static_cast<float>(theta) / static_cast<float>(w));
}

// Source LOD log2(W / thetaSamples) (GenIrradianceMap.comp:139) in FP32 as
// the shader computes it, clamped to the pyramid: levels l0, l1 and the
// fraction delta.
struct IrradianceLod {
  int l0, l1;
  float delta;
};
static IrradianceLod irradiance_lod(int w, int h, int theta) {
  const int levels = althea_env_levels(w, h);
  float lod = std::log2(static_cast<float>(w) / static_cast<float>(theta));
  lod = std::min(std::max(lod, 0.0f), static_cast<float>(levels - 1));
  IrradianceLod r;
  r.l0 = static_cast<int>(std::floor(lod));
  r.l1 = std::min(r.l0 + 1, levels - 1);
  r.delta = lod - static_cast<float>(r.l0);
  return r;
}

// Blend grid of levels l0 and l1 = l0 + 1 (two levels, each extent of l0
// even and l1 exactly half of it): nodes per axis 3 w0 / 2, plus the REPEAT
// border; RG plane texels rounded up to 64.
struct BlendGrid {
  bool ok;
  int w0, h0, nx, ny;
  size_t plane;
};
static BlendGrid blend_grid(int w, int h, const IrradianceLod& lod) {
  const PyramidGeometry g = pyramid_geometry(w, h);
  BlendGrid b{};
  if (lod.delta == 0.0f || lod.l1 != lod.l0 + 1) return b;
  b.w0 = g.w[lod.l0];
  b.h0 = g.h[lod.l0];
  b.ok = b.w0 >= 2 && b.h0 >= 2 && b.w0 % 2 == 0 && b.h0 % 2 == 0 && g.w[lod.l1] == b.w0 / 2 &&
         g.h[lod.l1] == b.h0 / 2;
  b.nx = 3 * b.w0 / 2;
  b.ny = 3 * b.h0 / 2;
  b.plane = (static_cast<size_t>(b.nx + 2) * (b.ny + 2) + 63) & ~static_cast<size_t>(63);
  return b;
}

// The theta and phi tables, then the blend grid of levels l0, l1 or their
// split gather copy.
size_t irradiance_workspace_bytes(int w, int h, int, int theta) {
  int phi = irradiance_phi_samples(w, h, theta);
  const size_t tables = align256(sizeof(float2) * theta) + align256(sizeof(float4) * std::max(phi, 1));
  if (w <= 0 || h <= 0 || theta <= 0) return tables;
  const IrradianceLod lod = irradiance_lod(w, h, theta);
  const BlendGrid b = blend_grid(w, h, lod);
  if (b.ok) return tables + 12 * b.plane;
  return tables + split_gather_bytes(w, h, lod.l0, lod.l1);
}

struct BlendGridParams {
  int w0, h0, nx, ny;
  uint32_t off0, off1;  // pyramid texel offsets of the two levels
  double delta;
  size_t plane;
};

// One thread per padded node (r, c): node (c - 1, r - 1) wrapped to the
// grid, at fine coordinates (2 (n / 3) + (n % 3) / 2) per axis; its value
// (1 - delta) B0 + delta B1 in FP64 (B1 at x0 / 2 - 1/4, REPEAT), rounded
// once, into the RG and B planes.
__global__ __launch_bounds__(256) void blend_grid_kernel(const float4* __restrict__ pyr, const BlendGridParams p,
                                                         float* __restrict__ out) {
  const size_t W = static_cast<size_t>(p.nx) + 2, i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= W * (p.ny + 2)) return;
  const int r = static_cast<int>(i / W), c = static_cast<int>(i - r * W);
  auto wrap = [](int k, int n) { return k < 0 ? k + n : (k >= n ? k - n : k); };
  const int n = wrap(c - 1, p.nx), m = wrap(r - 1, p.ny);
  const double x0 = 2.0 * (n / 3) + 0.5 * (n % 3), y0 = 2.0 * (m / 3) + 0.5 * (m % 3);
  auto bilinear = [&](uint32_t off, int w, int h, double x, double y, double* v) {
    const double fx = floor(x), fy = floor(y), ax = x - fx, ay = y - fy;
    const int i0 = static_cast<int>(fx), j0 = static_cast<int>(fy);
    auto at = [&](int i, int j) {
      return pyr[off + static_cast<size_t>((j % h + h) % h) * w + (i % w + w) % w];
    };
    const float4 t00 = at(i0, j0), t10 = at(i0 + 1, j0), t01 = at(i0, j0 + 1), t11 = at(i0 + 1, j0 + 1);
    const float a[4][3] = {{t00.x, t00.y, t00.z}, {t10.x, t10.y, t10.z}, {t01.x, t01.y, t01.z}, {t11.x, t11.y, t11.z}};
    for (int k = 0; k < 3; ++k)
      v[k] = (1.0 - ay) * ((1.0 - ax) * a[0][k] + ax * a[1][k]) + ay * ((1.0 - ax) * a[2][k] + ax * a[3][k]);
  };
  double b0[3], b1[3];
  bilinear(p.off0, p.w0, p.h0, x0, y0, b0);
  bilinear(p.off1, p.w0 / 2, p.h0 / 2, 0.5 * x0 - 0.25, 0.5 * y0 - 0.25, b1);
  float f[3];
  for (int k = 0; k < 3; ++k) f[k] = static_cast<float>((1.0 - p.delta) * b0[k] + p.delta * b1[k]);
  reinterpret_cast<float2*>(out)[i] = make_float2(f[0], f[1]);
  out[2 * p.plane + i] = f[2];
}

althea_status launch_irradiance(const float* pyr, int env_w, int env_h, int size, int theta_samples,
                                unsigned flags, float* cube, void* workspace, hipStream_t s) {
  const int phi = irradiance_phi_samples(env_w, env_h, theta_samples);
  ALTHEA_REQUIRE(phi >= 1, "irradiance: int(H*theta/W) = %d phi samples", phi);
  char* ws = static_cast<char*>(workspace);
  float2* theta_tab = reinterpret_cast<float2*>(ws);
  ws += align256(sizeof(float2) * theta_samples);
  float4* phi_tab = reinterpret_cast<float4*>(ws);
  ws += align256(sizeof(float4) * phi);
  float* gather = reinterpret_cast<float*>(ws);

  IrradianceParams p{};
  PyramidGeometry g = pyramid_geometry(env_w, env_h);
  p.levels = g.levels;
  p.size = size;
  p.theta_samples = theta_samples;
  p.phi_samples = phi;
  const IrradianceLod lod = irradiance_lod(env_w, env_h, theta_samples);
  p.l0 = lod.l0;
  p.l1 = lod.l1;
  p.delta = lod.delta;
  const BlendGrid bg = blend_grid(env_w, env_h, lod);
  althea_status st = ALTHEA_OK;
  if (bg.ok) {
    BlendGridParams q{};
    q.w0 = bg.w0;
    q.h0 = bg.h0;
    q.nx = bg.nx;
    q.ny = bg.ny;
    q.off0 = g.off[p.l0];
    q.off1 = g.off[p.l1];
    q.delta = static_cast<double>(lod.delta);
    q.plane = bg.plane;
    const size_t nodes = static_cast<size_t>(bg.nx + 2) * (bg.ny + 2);
    hipLaunchKernelGGL(blend_grid_kernel, dim3(static_cast<unsigned>((nodes + 255) / 256)), dim3(256), 0, s,
                       reinterpret_cast<const float4*>(pyr), q, gather);
    st = check_launch("blend_grid_kernel");
    if (st != ALTHEA_OK) return st;
    const LevelC c0 = make_level(bg.w0, bg.h0, 0);
    p.grid.kx = c0.kx;
    p.grid.cx = c0.cx;
    p.grid.ky = c0.ky;
    p.grid.cy = c0.cy;
    p.grid.rowbytes = static_cast<uint32_t>(bg.nx + 2) * 8u;
    p.grid.rg = 0;
    p.grid.b = static_cast<uint32_t>(8 * bg.plane);
    p.blend = 1;
  } else {
    const GatherGeometry gg = gather_geometry(env_w, env_h, p.l0, p.l1);
    st = launch_split_gather(pyr, env_w, env_h, gg, gather, s);
    if (st != ALTHEA_OK) return st;
    p.lv0 = split_level(env_w, env_h, gg, p.l0);
    p.lv1 = split_level(env_w, env_h, gg, p.l1);
  }
  p.scale = M_PI / theta_samples / phi;
  const int n = std::max(theta_samples, phi);
  hipLaunchKernelGGL(irradiance_tables_kernel, dim3((n + 255) / 256), dim3(256), 0, s, theta_samples, phi, p.delta,
                     p.blend, theta_tab, phi_tab);
  st = check_launch("irradiance_tables_kernel");
  if (st != ALTHEA_OK) return st;
  const dim3 grid((6 * size * size + 3) / 4);
  const char* g4 = reinterpret_cast<const char*>(gather);
  float4* cube4 = reinterpret_cast<float4*>(cube);
  const bool f64 = !(flags & ALTHEA_PRECOMPUTE_FP32_SUMS);
  if (theta_samples <= kIrrThetaLdsMax) {
    const size_t lds = sizeof(float2) * static_cast<size_t>(theta_samples);
    if (f64)
      hipLaunchKernelGGL((irradiance_cube_kernel<true, true>), grid, dim3(256), lds, s, g4, theta_tab, phi_tab, p,
                         cube4);
    else
      hipLaunchKernelGGL((irradiance_cube_kernel<false, true>), grid, dim3(256), lds, s, g4, theta_tab, phi_tab, p,
                         cube4);
  } else if (f64) {
    hipLaunchKernelGGL((irradiance_cube_kernel<true, false>), grid, dim3(256), 0, s, g4, theta_tab, phi_tab, p, cube4);
  } else {
    hipLaunchKernelGGL((irradiance_cube_kernel<false, false>), grid, dim3(256), 0, s, g4, theta_tab, phi_tab, p,
                       cube4);
  }
  return check_launch("irradiance_cube_kernel");
}

// --------------------------------------------------------------- BRDF LUT
// Karis split-sum (A, B) at texel centre (NdotV, roughness).  A block covers
// 64 NdotV columns of one roughness row (the sample quantities are then
// wave-uniform) with kLutWaves waves, each owning a contiguous range of
// samples that is a whole number of 32-sample blocks.  FP32 per-sample
// arithmetic, 32-sample FP32 block sums flushed into FP64, the waves' FP64
// partials combined in wave order, FP64 normalisation.
constexpr int kLutWaves = 16;

__global__ __launch_bounds__(64 * kLutWaves) void brdf_lut_kernel(int size, int samples, int chunk,
                                                                  float2* __restrict__ rg) {
  __shared__ double s_a[kLutWaves][64], s_b[kLutWaves][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int x = blockIdx.x * 64 + lane;
  const int y = blockIdx.y;
  const float nv = (static_cast<float>(x) + 0.5f) / static_cast<float>(size);
  const float r = (static_cast<float>(y) + 0.5f) / static_cast<float>(size);
  const float alpha = r * r;
  const float a2 = alpha * alpha;
  const float k = 0.5f * alpha;
  const float vx = sqrtf(fmaxf(0.0f, fmaf(-nv, nv, 1.0f)));
  const float gv = nv / fmaf(nv, 1.0f - k, k);
  double sa = 0.0, sb = 0.0;
  float fa = 0.0f, fb = 0.0f;
  const int s0 = wave * chunk, s1 = min(samples, s0 + chunk);
  for (int s = s0; s < s1; ++s) {
    float x1, x2;
    hammersley(static_cast<uint32_t>(s), static_cast<uint32_t>(samples), x1, x2);
    const float den = fmaf(a2 - 1.0f, x2, 1.0f);
    const float cos_t = sqrtf((1.0f - x2) / den);
    const float sin_t = sqrtf(a2 * x2 / den);
    float sp, cp;
    sincospif(2.0f * x1, &sp, &cp);
    const float hx = cp * sin_t, hz = cos_t;
    const float vdh = fmaf(vx, hx, nv * hz);
    const float ndl = fmaf(2.0f * vdh, hz, -nv);
    if (ndl > 0.0f) {
      const float gl = ndl / fmaf(ndl, 1.0f - k, k);
      const float g_vis = gv * gl * vdh / (hz * nv);
      const float om = 1.0f - vdh;
      const float om2 = om * om;
      const float fc = om2 * om2 * om;
      fa = fmaf(1.0f - fc, g_vis, fa);
      fb = fmaf(fc, g_vis, fb);
    }
    if ((s & (kBlockSums - 1)) == kBlockSums - 1) {
      sa += fa;
      sb += fb;
      fa = fb = 0.0f;
    }
  }
  sa += fa;
  sb += fb;
  s_a[wave][lane] = sa;
  s_b[wave][lane] = sb;
  __syncthreads();
  if (wave != 0 || x >= size) return;
  double ta = 0.0, tb = 0.0;
#pragma unroll
  for (int w = 0; w < kLutWaves; ++w) {
    ta += s_a[w][lane];
    tb += s_b[w][lane];
  }
  rg[y * size + x] = make_float2(static_cast<float>(ta / samples), static_cast<float>(tb / samples));
}

althea_status launch_brdf_lut(int size, int samples, float* rg, hipStream_t s) {
  dim3 grid((size + 63) / 64, size);
  const int per_wave = (samples + kLutWaves - 1) / kLutWaves;
  const int chunk = (per_wave + kBlockSums - 1) / kBlockSums * kBlockSums;
  hipLaunchKernelGGL(brdf_lut_kernel, grid, dim3(64 * kLutWaves), 0, s, size, samples, chunk,
                     reinterpret_cast<float2*>(rg));
  return check_launch("brdf_lut_kernel");
}

// ---------------------------------------------------------- sampler probes
__global__ void cube_sample_kernel(const float4* __restrict__ cube, const CubeMips cm,
                                   const float4* __restrict__ dir_lod, int count, float4* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float4 q = dir_lod[i];
  f3 v = cube_sample(cube, cm, mk(q.x, q.y, q.z), q.w);
  out[i] = make_float4(v.x, v.y, v.z, 1.0f);
}

althea_status launch_cube_sample(const float* cube, int size, int mips, const float* dir_lod, int count,
                                 float* out, void*, hipStream_t s) {
  if (count == 0) return ALTHEA_OK;
  CubeMips cm = cube_mips_of(size, mips);
  hipLaunchKernelGGL(cube_sample_kernel, dim3((count + 255) / 256), dim3(256), 0, s,
                     reinterpret_cast<const float4*>(cube), cm, reinterpret_cast<const float4*>(dir_lod), count,
                     reinterpret_cast<float4*>(out));
  return check_launch("cube_sample_kernel");
}

// --------------------------------------------------- seamless-bordered cubes
BorderedCube bordered_cube_of(int size, int mips) {
  BorderedCube c{};
  c.size = size;
  c.mips = mips;
  size_t off = 0;
  for (int m = 0; m < mips; ++m) {
    c.off[m] = static_cast<uint32_t>(off);
    const size_t w = static_cast<size_t>(std::max(1, size >> m)) + 2;
    off += 6 * w * w;
  }
  return c;
}

size_t cube_bordered_texels(int size, int mips) {
  size_t n = 0;
  for (int m = 0; m < mips; ++m) {
    const size_t w = static_cast<size_t>(std::max(1, size >> m)) + 2;
    n += 6 * w * w;
  }
  return n;
}

// One thread per bordered texel: bordered (r, c) of face f at mip m is
// cube_tap(f, c - 1, r - 1), the tap a seamless bilinear lookup reads there.
__global__ __launch_bounds__(256) void cube_border_kernel(const float4* __restrict__ cube, const CubeMips cm,
                                                          const BorderedCube bc, uint32_t total,
                                                          float4* __restrict__ out) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= total) return;
  int m = 0;
  while (m + 1 < bc.mips && i >= bc.off[m + 1]) ++m;
  const int S = max(1, bc.size >> m), W = S + 2;
  const uint32_t k = i - bc.off[m];
  const int face = static_cast<int>(k / static_cast<uint32_t>(W * W));
  const int rem = static_cast<int>(k) - face * W * W;
  const int r = rem / W, c = rem - r * W;
  const f3 v = cube_tap(cube + cm.off[m], S, face, c - 1, r - 1);
  out[i] = make_float4(v.x, v.y, v.z, 1.0f);
}

althea_status launch_cube_border(const float* cube, int size, int mips, float* bordered, hipStream_t s) {
  const CubeMips cm = cube_mips_of(size, mips);
  const BorderedCube bc = bordered_cube_of(size, mips);
  const uint32_t total = static_cast<uint32_t>(cube_bordered_texels(size, mips));
  hipLaunchKernelGGL(cube_border_kernel, dim3((total + 255) / 256), dim3(256), 0, s,
                     reinterpret_cast<const float4*>(cube), cm, bc, total, reinterpret_cast<float4*>(bordered));
  return check_launch("cube_border_kernel");
}

struct EnvSampleParams {
  LevelC lv[kMaxLevels];
  int levels;
  int count;
};

__global__ __launch_bounds__(256) void env_sample_kernel(const char* __restrict__ pyr, const EnvSampleParams p,
                                                         const float4* __restrict__ dir_lod,
                                                         float4* __restrict__ out) {
  __shared__ LevelC s_lv[kMaxLevels];
  if (threadIdx.x < kMaxLevels) s_lv[threadIdx.x] = p.lv[threadIdx.x];
  __syncthreads();
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.count) return;
  // The mirror-lookup path: angles and texel coordinates in df32 (the
  // addressing of the roughness-0 prefilter), the two levels blended with
  // weights 1 - delta and delta.
  const float4 q = dir_lod[i];
  const float len_xz = sqrtf(fmaf(q.x, q.x, q.z * q.z));
  const int pole = len_xz > 0.001f ? 0 : (q.y > 0.0f ? 1 : -1);
  df32 yaw, pitch;
  equirect_angles_df(df3{df_of(q.x), df_of(q.y), df_of(q.z)}, pole, yaw, pitch);
  const float d = fminf(fmaxf(q.w, 0.0f), static_cast<float>(p.levels - 1));
  const float fl = floorf(d);
  const int l0 = static_cast<int>(fl);
  const float delta = d - fl;
  auto at = [&](int l) {
    const LevelC& c = s_lv[l];
    df32 x, y;
    equirect_texel_df(yaw, pitch, static_cast<int>(c.wmask + 1), static_cast<int>(c.hmask + 1), x, y);
    return level_bilinear_df(pyr, c, x, y);
  };
  f3 v = at(l0);
  if (delta != 0.0f) {
    const f3 b = at(min(l0 + 1, p.levels - 1));
    const float w0 = 1.0f - delta;
    v = mk(fmaf(b.x, delta, v.x * w0), fmaf(b.y, delta, v.y * w0), fmaf(b.z, delta, v.z * w0));
  }
  out[i] = make_float4(v.x, v.y, v.z, 1.0f);
}

althea_status launch_env_sample(const float* pyr, int w, int h, const float* dir_lod, int count, float* out, void*,
                                hipStream_t s) {
  if (count == 0) return ALTHEA_OK;
  EnvSampleParams p{};
  pyramid_levels(w, h, p.lv);
  p.levels = althea_env_levels(w, h);
  p.count = count;
  hipLaunchKernelGGL(env_sample_kernel, dim3((count + 255) / 256), dim3(256), 0, s,
                     reinterpret_cast<const char*>(pyr), p, reinterpret_cast<const float4*>(dir_lod),
                     reinterpret_cast<float4*>(out));
  return check_launch("env_sample_kernel");
}

}  // namespace althea
