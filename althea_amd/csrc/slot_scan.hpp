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
// block (each thread a contiguous run), the total into *total.
__device__ __forceinline__ void slot_scan_blocks(unsigned long long* __restrict__ blocks, int n,
                                                 unsigned long long* __restrict__ total,
                                                 unsigned long long* sums /* LDS [kSlotScan] */) {
  const int tid = threadIdx.x;
  const int run = (n + kSlotScan - 1) / kSlotScan;
  const int b = min(tid * run, n), e = min(b + run, n);
  unsigned long long s = 0;
  for (int i = b; i < e; ++i) s += blocks[i];
  sums[tid] = s;
  __syncthreads();
  for (int d = 1; d < kSlotScan; d <<= 1) {  // inclusive Hillis-Steele scan
    const unsigned long long add = tid >= d ? sums[tid - d] : 0ull;
    __syncthreads();
    sums[tid] += add;
    __syncthreads();
  }
  unsigned long long acc = sums[tid] - s;
  for (int i = b; i < e; ++i) {
    const unsigned long long v = blocks[i];
    blocks[i] = acc;
    acc += v;
  }
  if (tid == kSlotScan - 1) *total = sums[tid];
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

// Largest t >= lo with offsets[t] <= slot (offsets[lo] <= slot < offsets[n]):
// probes lo + 1, lo + 3, lo + 7, ... then a binary search in the last step.
__device__ __forceinline__ int slot_gallop(const unsigned long long* __restrict__ offsets, int lo, int n,
                                           unsigned long long slot) {
  int step = 1;
  while (lo + step < n && offsets[lo + step] <= slot) {
    lo += step;
    step <<= 1;
  }
  return slot_find(offsets, lo, min(lo + step, n), slot);
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
