// Batched GGX prefilter: many environment maps of uneven resolution in one
// persistent launch.  The reference prefilters one map with five sequential
// dispatches (Src/ImageBasedLighting.cpp:352-402); here every (map, mip,
// face, tile) of every map is one work item in a list sorted by estimated
// cost (largest first), and waves of a grid sized to the machine pull items
// from an atomic counter until the list is exhausted (greedy list scheduling
// of the longest items first).  Every texel runs prefilter_texel, the same
// function as the single-map kernel, so batched and single results are
// bit-identical.
//
// A run replays one hipGraph: the pyramid chains of all maps, the sample
// tables of every distinct source extent, a counter-reset kernel and the
// persistent kernel.  The tile shape and the grid are chosen by
// althea_prefilter_batch_autotune, which times every candidate on a fixed
// subset of the work.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <vector>

#include "device_math.hpp"
#include "internal.hpp"
#include "prefilter_common.hpp"

namespace althea {

struct BatchMap {
  const char* pyr;   // source pyramid (bytes)
  const char* gpyr;  // its gather copy (bytes)
  float4* cube;      // output cube, [mip][face][y][x]
  int size;          // face size at mip 0
  int table_set;     // sample-table set of this map's source extent
  int level_base;    // first LevelC / GatherLevel of this map
  int chunk;         // table entries per partial sum, as the single-map kernel cuts this map's sums
};

// Work item: map, mip, face, tile index within the face.
struct BatchParams {
  int num_tiles;
  int samples;
  int mips;
  int tile_w, tile_h;
};

__global__ void batch_reset_kernel(int* counter) {
  if (threadIdx.x == 0) counter[0] = 0;
}

template <bool kF64Sums, int kTW, int kTH>
__global__ __launch_bounds__(256) void prefilter_batch_kernel(const BatchMap* __restrict__ maps,
                                                              const LevelC* __restrict__ levels,
                                                              const GatherLevel* __restrict__ glevels,
                                                              const GgxEntry* __restrict__ tables,
                                                              const int* __restrict__ counts,
                                                              const double* __restrict__ wsum,
                                                              const int4* __restrict__ work, int* counter,
                                                              const BatchParams p) {
  static_assert(kTW * kTH == 64, "one wave per tile");
  const int lane = threadIdx.x & 63;
  for (;;) {
    int item = 0;
    if (lane == 0) item = atomicAdd(counter, 1);
    item = __builtin_amdgcn_readfirstlane(item);
    if (item >= p.num_tiles) break;  // every wave reaches this once the list is drained
    const int4 w = work[item];
    const BatchMap mp = maps[w.x];
    const int m = w.y, face = w.z, t = w.w;
    const int S = max(1, mp.size >> m);
    const int tpr = (S + kTW - 1) / kTW;
    const int x = (t % tpr) * kTW + (lane % kTW);
    const int y = (t / tpr) * kTH + (lane / kTW);
    const bool active = x < S && y < S;
    const int set_mip = mp.table_set * p.mips + m;
    const LevelC* lv = levels + mp.level_base;
    const auto level = [&](int l) -> const LevelC& { return lv[l]; };
    float4 out;
    if (m == 0 || p.mips == 1) {
      // roughness 0: the mirror lookup, as the single-map kernel
      out = prefilter_texel_mirror(mp.pyr, level, face, min(x, S - 1), min(y, S - 1), S);
    } else {
      const Frame fr = local_to_world(cube_texel_raw_dir(face, min(x, S - 1), min(y, S - 1), S));
      out = prefilter_texel<kF64Sums>(GatherAcc{mp.gpyr, glevels + mp.level_base},
                                      tables + static_cast<size_t>(set_mip) * p.samples, counts[set_mip],
                                      wsum[set_mip], fr, mp.chunk);
    }
    if (active) {
      size_t off = 0;
      for (int k = 0; k < m; ++k) {
        const size_t s = static_cast<size_t>(max(1, mp.size >> k));
        off += 6 * s * s;
      }
      mp.cube[off + (static_cast<size_t>(face) * S + y) * S + x] = out;
    }
  }
}

// Source-independent count of kept samples (NdotL > 0) at roughness r: the
// table kernel's test, evaluated on the host for cost estimates.
static int kept_samples(int samples, double rough) {
  const double a2 = rough * rough;
  int kept = 0;
  for (int s = 0; s < samples; ++s) {
    uint32_t v = static_cast<uint32_t>(s);
    v = ((v >> 1) & 0x55555555u) | ((v & 0x55555555u) << 1);
    v = ((v >> 2) & 0x33333333u) | ((v & 0x33333333u) << 2);
    v = ((v >> 4) & 0x0F0F0F0Fu) | ((v & 0x0F0F0F0Fu) << 4);
    v = ((v >> 8) & 0x00FF00FFu) | ((v & 0x00FF00FFu) << 8);
    v = (v >> 16) | (v << 16);
    const double x2 = static_cast<double>(v) * (1.0 / 4294967296.0);
    const double c2 = (1.0 - x2) / (1.0 + (a2 - 1.0) * x2);
    if (2.0 * c2 - 1.0 > 0.0) ++kept;
  }
  return kept;
}

struct TileShape {
  int w, h;
};
static const TileShape kShapes[] = {{8, 8}, {16, 4}, {4, 16}};

}  // namespace althea

using namespace althea;

struct althea_prefilter_batch {
  int count = 0, mips = 0, samples = 0;
  unsigned flags = 0;
  std::vector<althea_env_view> envs;
  std::vector<int> sizes;
  // Device state.
  std::vector<float*> pyramids;
  std::vector<float*> gathers;  // gather copies of the pyramids, owned by the batch
  std::vector<float*> cubes;
  bool owns_buffers = true;  // false when the pyramids and cubes belong to IBL handles
  BatchMap* d_maps = nullptr;
  LevelC* d_levels = nullptr;
  GatherLevel* d_glevels = nullptr;
  GgxEntry* d_tables = nullptr;
  int* d_counts = nullptr;
  double* d_wsum = nullptr;
  float* d_wsum32 = nullptr;
  int* d_counter = nullptr;
  std::vector<std::pair<int, int>> extents;  // distinct (w, h), one table set each
  // Work list of the current tile shape.
  int4* d_work = nullptr;
  int num_tiles = 0;
  size_t texels = 0;
  althea_tuning tuning{8, 8, 16, 0};
  int cus = 0;
  // Graph of one run.
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t capture = nullptr;
};

namespace {

int shape_index(const althea_tuning& t) {
  for (int i = 0; i < 3; ++i)
    if (kShapes[i].w == t.tile_w && kShapes[i].h == t.tile_h) return i;
  return -1;
}

// Work list for a tile shape, sorted by estimated cost, largest first.
std::vector<int4> build_work(const althea_prefilter_batch* b, int tw, int th, size_t* texels) {
  std::vector<double> mip_cost(b->mips);
  for (int m = 0; m < b->mips; ++m) {
    const double r = b->mips > 1 ? static_cast<double>(m) / (b->mips - 1) : 0.0;
    // Roughness 0 is one table entry (coinciding samples); rougher mips blend
    // two source levels for most samples.
    mip_cost[m] = r == 0.0 ? 1.0 : kept_samples(b->samples, r) * 1.45;
  }
  struct Item {
    double cost;
    int4 w;
  };
  std::vector<Item> items;
  size_t tex = 0;
  for (int i = 0; i < b->count; ++i)
    for (int m = 0; m < b->mips; ++m) {
      const int S = std::max(1, b->sizes[i] >> m);
      const int tpr = (S + tw - 1) / tw, tpc = (S + th - 1) / th;
      const int active = std::min(S, tw) * std::min(S, th);
      for (int f = 0; f < 6; ++f)
        for (int t = 0; t < tpr * tpc; ++t) items.push_back({mip_cost[m] * active, make_int4(i, m, f, t)});
      tex += 6ull * S * S;
    }
  std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& c) { return a.cost > c.cost; });
  std::vector<int4> out(items.size());
  for (size_t k = 0; k < items.size(); ++k) out[k] = items[k].w;
  if (texels) *texels = tex;
  return out;
}

althea_status upload_work(althea_prefilter_batch* b) {
  std::vector<int4> work = build_work(b, b->tuning.tile_w, b->tuning.tile_h, &b->texels);
  (void)hipFree(b->d_work);
  b->d_work = nullptr;
  ALTHEA_TRY_HIP(hipMalloc(&b->d_work, sizeof(int4) * work.size()));
  ALTHEA_TRY_HIP(hipMemcpy(b->d_work, work.data(), sizeof(int4) * work.size(), hipMemcpyHostToDevice));
  b->num_tiles = static_cast<int>(work.size());
  return ALTHEA_OK;
}

template <bool kF64>
void launch_persistent_t(const althea_prefilter_batch* b, int shape, int blocks, const int4* work, int tiles,
                         hipStream_t s) {
  BatchParams p{tiles, b->samples, b->mips, kShapes[shape].w, kShapes[shape].h};
  switch (shape) {
    case 0:
      hipLaunchKernelGGL((prefilter_batch_kernel<kF64, 8, 8>), dim3(blocks), dim3(256), 0, s, b->d_maps, b->d_levels,
                         b->d_glevels, b->d_tables, b->d_counts, b->d_wsum, work, b->d_counter, p);
      break;
    case 1:
      hipLaunchKernelGGL((prefilter_batch_kernel<kF64, 16, 4>), dim3(blocks), dim3(256), 0, s, b->d_maps,
                         b->d_levels, b->d_glevels, b->d_tables, b->d_counts, b->d_wsum, work, b->d_counter, p);
      break;
    default:
      hipLaunchKernelGGL((prefilter_batch_kernel<kF64, 4, 16>), dim3(blocks), dim3(256), 0, s, b->d_maps,
                         b->d_levels, b->d_glevels, b->d_tables, b->d_counts, b->d_wsum, work, b->d_counter, p);
      break;
  }
}

// Counter reset + persistent kernel over `tiles` items of `work`.
althea_status launch_persistent(const althea_prefilter_batch* b, int shape, int waves_per_cu, const int4* work,
                                int tiles, hipStream_t s) {
  hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, s, b->d_counter);
  althea_status st = check_launch("batch_reset_kernel");
  if (st != ALTHEA_OK) return st;
  const int blocks = std::max(1, b->cus * waves_per_cu / 4);
  if (b->flags & ALTHEA_PRECOMPUTE_FP32_SUMS)
    launch_persistent_t<false>(b, shape, blocks, work, tiles, s);
  else
    launch_persistent_t<true>(b, shape, blocks, work, tiles, s);
  return check_launch("prefilter_batch_kernel");
}

// Pyramids of all maps with their gather copies, and the sample tables of
// every source extent.
althea_status launch_prep(const althea_prefilter_batch* b, hipStream_t s) {
  for (int i = 0; i < b->count; ++i) {
    const int w = b->envs[i].width, h = b->envs[i].height;
    althea_status st = launch_pyramid(b->envs[i].rgba, w, h, b->pyramids[i], s);
    if (st != ALTHEA_OK) return st;
    st = launch_gather(b->pyramids[i], w, h, gather_geometry(w, h, 0, althea_env_levels(w, h) - 1), b->gathers[i], s);
    if (st != ALTHEA_OK) return st;
  }
  for (size_t k = 0; k < b->extents.size(); ++k) {
    const size_t set = k * b->mips;
    althea_status st = launch_ggx_tables(b->samples, b->mips, b->extents[k].first, b->extents[k].second,
                                         b->d_tables + set * b->samples, b->d_counts + set, b->d_wsum + set,
                                         b->d_wsum32 + set, s);
    if (st != ALTHEA_OK) return st;
  }
  return ALTHEA_OK;
}

void drop_graph(althea_prefilter_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

althea_status build_graph(althea_prefilter_batch* b) {
  drop_graph(b);
  ALTHEA_TRY_HIP(hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal));
  althea_status st = launch_prep(b, b->capture);
  if (st == ALTHEA_OK)
    st = launch_persistent(b, shape_index(b->tuning), b->tuning.waves_per_cu, b->d_work, b->num_tiles, b->capture);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->capture, &g);
  if (st != ALTHEA_OK) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  ALTHEA_TRY_HIP(e);
  b->graph = g;
  ALTHEA_TRY_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return ALTHEA_OK;
}

}  // namespace

// ext_pyramids / ext_cubes (optional): caller buffers the batch writes into
// instead of its own.
static althea_status batch_create(const althea_env_view* envs, int count, const int* face_sizes, int mips,
                                  int samples, unsigned flags, float* const* ext_pyramids, float* const* ext_cubes,
                                  althea_prefilter_batch** out) {
  ALTHEA_REQUIRE(envs && face_sizes && out && count >= 1, "prefilter_batch_create: invalid argument");
  *out = nullptr;
  ALTHEA_REQUIRE(samples >= 1 && samples <= (1 << 20), "prefilter_batch_create: samples %d", samples);
  ALTHEA_REQUIRE(mips >= 1 && mips <= 16, "prefilter_batch_create: mips %d", mips);
  for (int i = 0; i < count; ++i) {
    ALTHEA_REQUIRE(envs[i].rgba && is_pow2(envs[i].width) && is_pow2(envs[i].height) && envs[i].width <= (1 << 15) &&
                       envs[i].height <= (1 << 15) &&
                       static_cast<int64_t>(envs[i].width) * envs[i].height <= (int64_t{1} << 27),
                   "prefilter_batch_create: map %d extent %dx%d (powers of two, at most 2^27 texels)", i,
                   envs[i].width, envs[i].height);
    ALTHEA_REQUIRE(is_pow2(face_sizes[i]) && (face_sizes[i] >> (mips - 1)) >= 1,
                   "prefilter_batch_create: map %d face size %d with %d mips", i, face_sizes[i], mips);
  }
  althea_prefilter_batch* b = new (std::nothrow) althea_prefilter_batch;
  if (!b) return ALTHEA_ERROR_OUT_OF_MEMORY;
  auto fail = [&](althea_status st) {
    althea_prefilter_batch_destroy(b);
    return st;
  };
  b->count = count;
  b->mips = mips;
  b->samples = samples;
  b->flags = flags;
  b->envs.assign(envs, envs + count);
  b->sizes.assign(face_sizes, face_sizes + count);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    set_error("prefilter_batch_create: device query failed");
    return fail(ALTHEA_ERROR_HIP);
  }
  std::map<std::pair<int, int>, int> set_of;
  std::vector<BatchMap> maps(count);
  std::vector<LevelC> levels(static_cast<size_t>(count) * kMaxLevels);
  std::vector<GatherLevel> glevels(static_cast<size_t>(count) * kMaxLevels);
  b->pyramids.assign(count, nullptr);
  b->gathers.assign(count, nullptr);
  b->cubes.assign(count, nullptr);
  b->owns_buffers = ext_pyramids == nullptr;
  for (int i = 0; i < count; ++i) {
    const auto ext = std::make_pair(envs[i].width, envs[i].height);
    auto it = set_of.find(ext);
    if (it == set_of.end()) {
      it = set_of.emplace(ext, static_cast<int>(b->extents.size())).first;
      b->extents.push_back(ext);
    }
    if (!b->owns_buffers) {
      b->pyramids[i] = ext_pyramids[i];
      b->cubes[i] = ext_cubes[i];
    } else if (hipMalloc(&b->pyramids[i], 16 * althea_env_pyramid_texels(envs[i].width, envs[i].height)) !=
                   hipSuccess ||
               hipMalloc(&b->cubes[i], 16 * althea_cube_texels(face_sizes[i], mips)) != hipSuccess) {
      set_error("prefilter_batch_create: out of device memory at map %d", i);
      return fail(ALTHEA_ERROR_OUT_OF_MEMORY);
    }
    const int w = envs[i].width, h = envs[i].height;
    const GatherGeometry gg = gather_geometry(w, h, 0, althea_env_levels(w, h) - 1);
    if (hipMalloc(&b->gathers[i], 16 * gg.texels) != hipSuccess) {
      set_error("prefilter_batch_create: out of device memory at map %d", i);
      return fail(ALTHEA_ERROR_OUT_OF_MEMORY);
    }
    pyramid_levels(w, h, &levels[static_cast<size_t>(i) * kMaxLevels]);
    gather_levels(w, h, gg, &glevels[static_cast<size_t>(i) * kMaxLevels]);
    maps[i] = BatchMap{reinterpret_cast<const char*>(b->pyramids[i]), reinterpret_cast<const char*>(b->gathers[i]),
                       reinterpret_cast<float4*>(b->cubes[i]), face_sizes[i], it->second, i * kMaxLevels, 0};
    prefilter_split(face_sizes[i], mips, samples, maps[i].chunk);
  }
  const size_t sets = b->extents.size() * mips;
  if (hipMalloc(&b->d_maps, sizeof(BatchMap) * count) != hipSuccess ||
      hipMalloc(&b->d_levels, sizeof(LevelC) * levels.size()) != hipSuccess ||
      hipMalloc(&b->d_glevels, sizeof(GatherLevel) * glevels.size()) != hipSuccess ||
      hipMalloc(&b->d_tables, sizeof(GgxEntry) * sets * samples) != hipSuccess ||
      hipMalloc(&b->d_counts, sizeof(int) * sets) != hipSuccess ||
      hipMalloc(&b->d_wsum, sizeof(double) * sets) != hipSuccess ||
      hipMalloc(&b->d_wsum32, sizeof(float) * sets) != hipSuccess ||
      hipMalloc(&b->d_counter, 256) != hipSuccess) {
    set_error("prefilter_batch_create: out of device memory");
    return fail(ALTHEA_ERROR_OUT_OF_MEMORY);
  }
  if (hipMemcpy(b->d_maps, maps.data(), sizeof(BatchMap) * count, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_levels, levels.data(), sizeof(LevelC) * levels.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_glevels, glevels.data(), sizeof(GatherLevel) * glevels.size(), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking) != hipSuccess) {
    set_error("prefilter_batch_create: upload failed");
    return fail(ALTHEA_ERROR_HIP);
  }
  althea_status st = upload_work(b);
  if (st != ALTHEA_OK) return fail(st);
  *out = b;
  return ALTHEA_OK;
}

extern "C" {

althea_status althea_prefilter_batch_create(const althea_env_view* envs, int count, const int* face_sizes, int mips,
                                            int samples, unsigned flags, althea_prefilter_batch** out) {
  return batch_create(envs, count, face_sizes, mips, samples, flags, nullptr, nullptr, out);
}

althea_status althea_prefilter_batch_destroy(althea_prefilter_batch* b) {
  if (!b) return ALTHEA_OK;
  drop_graph(b);
  if (b->capture) (void)hipStreamDestroy(b->capture);
  if (b->owns_buffers) {
    for (float* p : b->pyramids) (void)hipFree(p);
    for (float* p : b->cubes) (void)hipFree(p);
  }
  for (float* p : b->gathers) (void)hipFree(p);
  (void)hipFree(b->d_maps);
  (void)hipFree(b->d_levels);
  (void)hipFree(b->d_glevels);
  (void)hipFree(b->d_tables);
  (void)hipFree(b->d_counts);
  (void)hipFree(b->d_wsum);
  (void)hipFree(b->d_wsum32);
  (void)hipFree(b->d_counter);
  (void)hipFree(b->d_work);
  delete b;
  return ALTHEA_OK;
}

althea_status althea_prefilter_batch_set_tuning(althea_prefilter_batch* b, const althea_tuning* t) {
  ALTHEA_REQUIRE(b && t, "prefilter_batch_set_tuning: null argument");
  ALTHEA_REQUIRE(shape_index(*t) >= 0, "prefilter_batch_set_tuning: tile %dx%d (8x8, 16x4 or 4x16)", t->tile_w,
                 t->tile_h);
  ALTHEA_REQUIRE(t->waves_per_cu >= 4 && t->waves_per_cu <= 32, "prefilter_batch_set_tuning: waves_per_cu %d",
                 t->waves_per_cu);
  const bool reshape = t->tile_w != b->tuning.tile_w || t->tile_h != b->tuning.tile_h;
  b->tuning = *t;
  drop_graph(b);
  return reshape ? upload_work(b) : ALTHEA_OK;
}

althea_status althea_prefilter_batch_get_tuning(const althea_prefilter_batch* b, althea_tuning* t) {
  ALTHEA_REQUIRE(b && t, "prefilter_batch_get_tuning: null argument");
  *t = b->tuning;
  return ALTHEA_OK;
}

althea_status althea_prefilter_batch_run(althea_prefilter_batch* b, hipStream_t stream) {
  ALTHEA_REQUIRE(b, "prefilter_batch_run: null batch");
  if (!b->exec) {
    althea_status st = build_graph(b);
    if (st != ALTHEA_OK) return st;
  }
  ALTHEA_TRY_HIP(hipGraphLaunch(b->exec, stream));
  return ALTHEA_OK;
}

althea_status althea_prefilter_batch_autotune(althea_prefilter_batch* b, althea_tuning* chosen,
                                              hipStream_t stream) {
  ALTHEA_REQUIRE(b, "prefilter_batch_autotune: null batch");
  althea_status st = launch_prep(b, stream);
  if (st != ALTHEA_OK) return st;
  hipEvent_t e0, e1;
  ALTHEA_TRY_HIP(hipEventCreate(&e0));
  ALTHEA_TRY_HIP(hipEventCreate(&e1));
  const int waves[] = {8, 16, 32};
  float best = 1e30f;
  althea_tuning best_t = b->tuning;
  int4* d_sub = nullptr;
  for (int shape = 0; shape < 3 && st == ALTHEA_OK; ++shape) {
    // A fixed subset: every k-th item of this shape's sorted work list, so the
    // mix of mips and extents matches the full run.
    std::vector<int4> work = build_work(b, kShapes[shape].w, kShapes[shape].h, nullptr);
    const int stride = std::max<int>(1, static_cast<int>(work.size() / 20000));
    std::vector<int4> sub;
    for (size_t k = 0; k < work.size(); k += stride) sub.push_back(work[k]);
    (void)hipFree(d_sub);
    d_sub = nullptr;
    if (hipMalloc(&d_sub, sizeof(int4) * sub.size()) != hipSuccess ||
        hipMemcpy(d_sub, sub.data(), sizeof(int4) * sub.size(), hipMemcpyHostToDevice) != hipSuccess) {
      set_error("prefilter_batch_autotune: work upload failed");
      st = ALTHEA_ERROR_HIP;
      break;
    }
    for (int wv : waves) {
      // Warm-up, then the timed launch.
      st = launch_persistent(b, shape, wv, d_sub, static_cast<int>(sub.size()), stream);
      if (st != ALTHEA_OK) break;
      (void)hipEventRecord(e0, stream);
      st = launch_persistent(b, shape, wv, d_sub, static_cast<int>(sub.size()), stream);
      (void)hipEventRecord(e1, stream);
      if (st != ALTHEA_OK || hipEventSynchronize(e1) != hipSuccess) {
        st = st != ALTHEA_OK ? st : ALTHEA_ERROR_HIP;
        break;
      }
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (ms < best) {
        best = ms;
        best_t = althea_tuning{kShapes[shape].w, kShapes[shape].h, wv, 0};
      }
    }
  }
  (void)hipFree(d_sub);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st != ALTHEA_OK) return st;
  st = althea_prefilter_batch_set_tuning(b, &best_t);
  if (st == ALTHEA_OK && chosen) *chosen = best_t;
  return st;
}

althea_status althea_prefilter_batch_cube(const althea_prefilter_batch* b, int index, const float** d_cube, int* size,
                                          int* mips) {
  ALTHEA_REQUIRE(b && index >= 0 && index < b->count, "prefilter_batch_cube: index %d", index);
  if (d_cube) *d_cube = b->cubes[index];
  if (size) *size = b->sizes[index];
  if (mips) *mips = b->mips;
  return ALTHEA_OK;
}

size_t althea_prefilter_batch_texels(const althea_prefilter_batch* b) { return b ? b->texels : 0; }

}  // extern "C"

namespace althea {

althea_status prefilter_batch_into(const althea_env_view* envs, int count, const int* face_sizes, int mips,
                                   int samples, unsigned flags, float* const* pyramids, float* const* cubes,
                                   const althea_tuning* tuning, althea_tuning* used, hipStream_t stream) {
  althea_prefilter_batch* b = nullptr;
  althea_status st = batch_create(envs, count, face_sizes, mips, samples, flags, pyramids, cubes, &b);
  if (st != ALTHEA_OK) return st;
  althea_tuning t{};
  st = tuning ? althea_prefilter_batch_set_tuning(b, tuning) : althea_prefilter_batch_autotune(b, &t, stream);
  if (st == ALTHEA_OK) st = althea_prefilter_batch_run(b, stream);
  if (st == ALTHEA_OK && used) st = althea_prefilter_batch_get_tuning(b, used);
  // The batch's tables and work list are in use until the launch completes.
  const hipError_t e = hipStreamSynchronize(stream);
  althea_prefilter_batch_destroy(b);
  if (st == ALTHEA_OK && e != hipSuccess) {
    set_error("prefilter_batch_into: %s", hipGetErrorString(e));
    st = ALTHEA_ERROR_HIP;
  }
  return st;
}

}  // namespace althea
