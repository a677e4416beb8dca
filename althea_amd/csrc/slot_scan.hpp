// Work dealt out as "slots": every item (a triangle, a light-triangle pair)
// owns a run of slots (the texels or pixels of its boxes), numbered over all
// items by an exclusive prefix sum of the run lengths.  A fixed grid of waves
// then takes equal contiguous ranges of slots, 64 per step, and finds each
// slot's item by searching the prefix sums; large and small items cost the
// same per slot.  Shared by the G-buffer and the shadow-cube rasterizers.
#pragma once

#include <hip/hip_runtime.h>

namespace althea {

constexpr int kSlotBlock = 256;   // items per block of the first scan level
constexpr int kSlotScan = 1024;   // threads of the block-total scan

// Exclusive prefix sum of `v` over a 256-thread block (wave shuffles, then
// the waves' totals in LDS); every thread of the block must call it.
__device__ __forceinline__ unsigned long long slot_block_scan(unsigned long long v,
                                                             unsigned long long* wave_total /* LDS [4] */,
                                                             unsigned long long& block_total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned long long x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned long long y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) wave_total[wave] = x;
  __syncthreads();
  unsigned long long before = 0;
  for (int k = 0; k < wave; ++k) before += wave_total[k];
  block_total = wave_total[0] + wave_total[1] + wave_total[2] + wave_total[3];
  return before + x - v;
}

// Exclusive prefix sum of the n block totals in place by one 1024-thread
// block (each thread a contiguous run, read and written eight at a time so
// the loads are in flight together), the total into *total.  The threads'
// run sums are scanned by wave shuffles, then the 16 wave totals by one wave.
__device__ __forceinline__ void slot_scan_blocks(unsigned long long* __restrict__ blocks, int n,
                                                 unsigned long long* __restrict__ total,
                                                 unsigned long long* sums /* LDS [kSlotScan / 64] */) {
  constexpr int kWaves = kSlotScan / 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int run = (n + kSlotScan - 1) / kSlotScan;
  const int b = min(tid * run, n), e = min(b + run, n);
  unsigned long long s = 0;
  for (int i0 = b; i0 < e; i0 += 8) {
    unsigned long long v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = i0 + k < e ? blocks[i0 + k] : 0ull;
#pragma unroll
    for (int k = 0; k < 8; ++k) s += v[k];
  }
  unsigned long long x = s;  // inclusive scan over the wave
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned long long y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) sums[wave] = x;
  __syncthreads();
  if (wave == 0) {  // inclusive scan of the wave totals
    unsigned long long w = lane < kWaves ? sums[lane] : 0ull;
#pragma unroll
    for (int d = 1; d < kWaves; d <<= 1) {
      const unsigned long long y = __shfl_up(w, d, 64);
      if (lane >= d) w += y;
    }
    if (lane < kWaves) sums[lane] = w;
  }
  __syncthreads();
  unsigned long long acc = (wave > 0 ? sums[wave - 1] : 0ull) + x - s;  // this thread's first exclusive prefix
  for (int i0 = b; i0 < e; i0 += 8) {
    unsigned long long v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = i0 + k < e ? blocks[i0 + k] : 0ull;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (i0 + k < e) blocks[i0 + k] = acc;
      acc += v[k];
    }
  }
  if (tid == 0) *total = sums[kWaves - 1];
}

// Largest t in [lo, hi) with offsets[t] <= slot, given offsets[lo] <= slot <
// offsets[hi].
__device__ __forceinline__ int slot_find(const unsigned long long* __restrict__ offsets, int lo, int hi,
                                         unsigned long long slot) {
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (offsets[mid] <= slot)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// The item of slot base + lane for every lane of a wave step [base, end)
// (-1 where base + lane >= end) and the low 32 bits of its first slot, from
// windows of 64 consecutive prefix sums (one coalesced load each) and a
// binary search across the lanes; usually one window.  w0 names an item with
// offsets[w0] <= base and becomes the item of the step's last slot, which
// serves the next step.  Every lane of the wave must call it.
__device__ __forceinline__ int slot_step_items(const unsigned long long* __restrict__ offsets, int n,
                                               unsigned long long base, unsigned long long end, int& w0,
                                               unsigned& item_start) {
  const int lane = threadIdx.x & 63;
  const unsigned long long slot = base + lane;
  int item = -1;
  item_start = 0;
  for (bool pending = slot < end; __builtin_amdgcn_ballot_w64(pending);) {
    const unsigned long long o = offsets[min(w0 + lane, n)];
    const unsigned long long o_end = offsets[min(w0 + 64, n)];
    const long long rel = static_cast<long long>(o - base);
    const int r = static_cast<int>(rel < -1 ? -1 : (rel > 64 ? 64 : rel));
    const bool here = pending && slot < o_end;
    int lo = 0;  // largest k with r_k <= lane (r_0 <= lane for every pending lane)
#pragma unroll
    for (int h = 32; h >= 1; h >>= 1)
      if (__shfl(r, lo + h) <= lane) lo += h;
    const unsigned start = static_cast<unsigned>(__shfl(static_cast<int>(static_cast<unsigned>(o)), lo));
    if (here) item = w0 + lo, item_start = start;
    pending = pending && !here;
    w0 = min(w0 + 64, n - 1);  // reached only while some slot lies past this window
  }
  w0 = __shfl(item, end - base > 64 ? 63 : static_cast<int>(end - base) - 1);
  return item;
}

__device__ __forceinline__ unsigned long long slot_uniform(unsigned long long v) {
  const unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(v));
  const unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(v >> 32));
  return (static_cast<unsigned long long>(hi) << 32) | lo;
}

// This wave's contiguous range [begin, end) of the total slots, a multiple
// of 64 long; begin >= end when the wave has none.
__device__ __forceinline__ void slot_range(unsigned long long total, unsigned long long& begin,
                                           unsigned long long& end) {
  const unsigned waves = gridDim.x * (blockDim.x >> 6);
  const unsigned wave = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  const unsigned long long per = ((total + waves - 1) / waves + 63) & ~63ull;
  begin = slot_uniform(wave * per);
  end = begin + per < total ? begin + per : total;
}

}  // namespace althea
