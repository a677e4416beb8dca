// Per-texel GGX prefilter shared by the single-map and the batched
// (persistent) kernels, so both produce bit-identical texels.
#pragma once

#include <type_traits>

#include "device_math.hpp"
#include "precise_lookup.hpp"

namespace althea {

// Sample-table entry (one per Hammersley sample with NdotL > 0, in sample
// order): L in the texel frame (T, B, N), the source LOD split into levels
// (l0, l1) and fraction delta, and the sample's weight in the NdotL-weighted
// mean (NdotL = lz, or the summed weight of coinciding samples) with its
// shares of the two levels, w0 = weight (1 - delta) and w1 = weight delta
// (rounded once from FP64; w1 = 0 where the sample reads one level).
// 32 bytes, one scalar load.
struct GgxEntry {
  float lx, ly, lz, w0;
  int l0, l1;
  float w1, weight;
};
static_assert(sizeof(GgxEntry) == 32, "one scalar load per entry");

constexpr int kBlockSums = 32;  // FP32 partial sums flushed to FP64 every 32 samples

// Source lookups of the sample loop: acc(l, yaw, pitch, w, r, g, b) adds w
// times the bilinear value of level l.  GatherAcc reads the padded gather
// copy, PyramidAcc the pyramid itself; both read the same texels with the
// same coordinates and weights, so their sums are bit-identical.
struct GatherAcc {
  const char* __restrict__ gpyr;
  const GatherLevel* __restrict__ lv;
  __device__ __forceinline__ void operator()(int l, float yaw, float pitch, float w, float& r, float& g,
                                             float& b) const {
    gather_bilinear_acc(gpyr, lv[l], yaw, pitch, w, r, g, b);
  }
};
struct PyramidAcc {
  const char* __restrict__ pyr;
  const LevelC* __restrict__ lv;
  __device__ __forceinline__ void operator()(int l, float yaw, float pitch, float w, float& r, float& g,
                                             float& b) const {
    level_bilinear_acc(pyr, lv[l], yaw, pitch, w, r, g, b);
  }
};

// Sums of the NdotL-weighted source over table entries [lo, hi) for one
// texel (PreFilterEnvMap.comp:147-169): the FP32 per-sample body (each
// level's bilinear in weight form, added into running FP32 sums).  kF64Sums:
// the FP32 sums are flushed into FP64 every kBlockSums entries (blocks
// aligned at multiples of kBlockSums), into a partial that is added into the
// total t at every multiple of `chunk` (itself a multiple of kBlockSums) and
// at the end.  Otherwise all-FP32, the same chunking, and the weight sum in
// tw.  A texel's entries cut into chunks summed apart (by several waves) and
// their partials added in chunk order give the same bits as one pass.
template <bool kF64Sums>
struct PrefilterSums {
  using T = typename std::conditional<kF64Sums, double, float>::type;
  T r = 0, g = 0, b = 0;
  float w = 0.0f;
};

// One pass over [lo, hi) with no chunk boundary inside (chunk >= hi - lo):
// the running FP64 sum of the FP32 blocks, as t = 0 + that sum.
template <bool kF64Sums, class Acc>
__device__ __forceinline__ PrefilterSums<kF64Sums> prefilter_sums_whole(const Acc& acc,
                                                                        const GgxEntry* __restrict__ tab, int lo,
                                                                        int hi, const Frame& fr) {
  using T = typename PrefilterSums<kF64Sums>::T;
  PrefilterSums<kF64Sums> t;
  T ar = 0, ag = 0, ab = 0;
  float sr = 0.0f, sg = 0.0f, sb = 0.0f, sw = 0.0f;
  for (int e = lo; e < hi; ++e) {
    const GgxEntry en = tab[e];
    const f3 L = mk(fmaf(fr.t.x, en.lx, fmaf(fr.b.x, en.ly, fr.n.x * en.lz)),
                    fmaf(fr.t.y, en.lx, fmaf(fr.b.y, en.ly, fr.n.y * en.lz)),
                    fmaf(fr.t.z, en.lx, fmaf(fr.b.z, en.ly, fr.n.z * en.lz)));
    float yaw, pitch;
    equirect_angles(L, yaw, pitch);
    acc(en.l0, yaw, pitch, en.w0, sr, sg, sb);
    if (__float_as_uint(en.w1) != 0u) acc(en.l1, yaw, pitch, en.w1, sr, sg, sb);
    if (kF64Sums) {
      if ((e & (kBlockSums - 1)) == kBlockSums - 1) {
        ar += sr;
        ag += sg;
        ab += sb;
        sr = sg = sb = 0.0f;
      }
    } else {
      sw += en.weight;
    }
  }
  ar += sr;
  ag += sg;
  ab += sb;
  t.r = ar;
  t.g = ag;
  t.b = ab;
  t.w = sw;
  return t;
}

template <bool kF64Sums, class Acc>
__device__ __forceinline__ PrefilterSums<kF64Sums> prefilter_sums(const Acc& acc, const GgxEntry* __restrict__ tab,
                                                                  int lo, int hi, int chunk, const Frame& fr) {
  if (hi - lo <= chunk) return prefilter_sums_whole<kF64Sums>(acc, tab, lo, hi, fr);
  using T = typename PrefilterSums<kF64Sums>::T;
  PrefilterSums<kF64Sums> t;
  T ar = 0, ag = 0, ab = 0;  // the chunk's partial
  float sr = 0.0f, sg = 0.0f, sb = 0.0f, sw = 0.0f;
  int chunk_end = lo + chunk;
  for (int e = lo; e < hi; ++e) {
    const GgxEntry en = tab[e];
    const f3 L = mk(fmaf(fr.t.x, en.lx, fmaf(fr.b.x, en.ly, fr.n.x * en.lz)),
                    fmaf(fr.t.y, en.lx, fmaf(fr.b.y, en.ly, fr.n.y * en.lz)),
                    fmaf(fr.t.z, en.lx, fmaf(fr.b.z, en.ly, fr.n.z * en.lz)));
    float yaw, pitch;
    equirect_angles(L, yaw, pitch);
    acc(en.l0, yaw, pitch, en.w0, sr, sg, sb);
    if (__float_as_uint(en.w1) != 0u) acc(en.l1, yaw, pitch, en.w1, sr, sg, sb);
    if (!kF64Sums) sw += en.weight;
    if (kF64Sums ? (e & (kBlockSums - 1)) == kBlockSums - 1 : e + 1 == chunk_end) {
      ar += sr;
      ag += sg;
      ab += sb;
      sr = sg = sb = 0.0f;
    }
    if (e + 1 == chunk_end) {
      t.r += ar;
      t.g += ag;
      t.b += ab;
      t.w += sw;
      ar = ag = ab = 0;
      sw = 0.0f;
      chunk_end += chunk;
    }
  }
  ar += sr;
  ag += sg;
  ab += sb;
  t.r += ar;
  t.g += ag;
  t.b += ab;
  t.w += sw;
  return t;
}

// The texel's value from its sums: FP64 sums over the FP64 weight sum, or
// all-FP32 sums over their own weight sum.
template <bool kF64Sums>
__device__ __forceinline__ float4 prefilter_value(const PrefilterSums<kF64Sums>& t, double wsum) {
  if (kF64Sums) {
    const double inv = 1.0 / wsum;
    return make_float4(static_cast<float>(t.r * inv), static_cast<float>(t.g * inv), static_cast<float>(t.b * inv),
                       1.0f);
  }
  return make_float4(static_cast<float>(t.r) / t.w, static_cast<float>(t.g) / t.w, static_cast<float>(t.b) / t.w,
                     1.0f);
}

// The whole table in one pass, cut into chunks of `chunk` entries as above.
template <bool kF64Sums, class Acc>
__device__ __forceinline__ float4 prefilter_texel(const Acc& acc, const GgxEntry* __restrict__ tab, int count,
                                                  double wsum, const Frame& fr, int chunk) {
  return prefilter_value<kF64Sums>(prefilter_sums<kF64Sums>(acc, tab, 0, count, chunk, fr), wsum);
}

// Roughness 0 (PreFilterEnvMap.comp:139-161 with a = 0): every half vector
// is N, so every kept sample is the mirror lookup at N, LOD 0, and the
// NdotL-weighted mean is that one lookup.  It reads the full texel contrast
// of the source, so it is addressed in df32 from the texel's exact lattice
// direction (yaw and pitch do not depend on the direction's length); the pole
// branch (|xz| <= 0.001 of the unit direction, PreFilterEnvMap.comp:104-112)
// is decided in FP32 as the sampled path decides it.
template <class LevelOf>
__device__ __forceinline__ float4 prefilter_texel_mirror(const char* __restrict__ pyr, const LevelOf& level, int face,
                                                         int x, int y, int S) {
  const f3 raw = cube_texel_raw_dir(face, x, y, S);
  const f3 n = normalize(raw);
  const float len_xz = sqrtf(fmaf(n.x, n.x, n.z * n.z));
  const int pole = len_xz > 0.001f ? 0 : (n.y > 0.0f ? 1 : -1);
  const LevelC& c = level(0);
  df32 tx, ty;
  equirect_coords_df(df3{df_of(raw.x), df_of(raw.y), df_of(raw.z)}, static_cast<int>(c.wmask + 1),
                     static_cast<int>(c.hmask + 1), pole, tx, ty);
  const f3 v = level_bilinear_df(pyr, c, tx, ty);
  return make_float4(v.x, v.y, v.z, 1.0f);
}

}  // namespace althea
