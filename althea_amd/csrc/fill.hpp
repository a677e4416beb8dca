// Fill of a device array with one value: 16-byte stores over its 16-byte
// aligned middle, single stores at the ends (the caller's pointer need only
// be aligned to the element).  A kernel, not hipMemsetAsync, so that callers
// may capture it in a graph.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace althea {

template <class T>
__global__ __launch_bounds__(256) void fill_kernel(T* __restrict__ p, size_t n, T v) {
  static_assert(16 % sizeof(T) == 0, "element size divides 16");
  constexpr size_t kPer = 16 / sizeof(T);
  const size_t mis = (16 - (reinterpret_cast<uintptr_t>(p) & 15)) & 15;
  const size_t head = n < mis / sizeof(T) ? n : mis / sizeof(T);
  const size_t quads = (n - head) / kPer;
  const size_t tail0 = head + quads * kPer;
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < quads) {
    union {
      T e[kPer];
      uint4 q;
    } u;
#pragma unroll
    for (size_t k = 0; k < kPer; ++k) u.e[k] = v;
    reinterpret_cast<uint4*>(p + head)[i] = u.q;
  }
  if (i < head) p[i] = v;
  if (i < n - tail0) p[tail0 + i] = v;
}

// Blocks of 256 for fill_kernel over n elements of size `elem`.
inline unsigned fill_blocks(size_t n, size_t elem) {
  const size_t quads = n / (16 / elem) + 1;  // at least one thread per head and tail element
  const size_t threads = quads > 16 / elem ? quads : 16 / elem;
  return static_cast<unsigned>((threads + 255) / 256);
}

}  // namespace althea
