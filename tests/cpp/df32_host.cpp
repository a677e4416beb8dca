// Host build of althea_amd/csrc/df32.hpp for tests/test_df32.py: the header's
// functions are __host__ __device__ and use only IEEE FP32 operations; the
// host results are the kernels' results except where df_rcp's hardware
// reciprocal (device) and division (host) differ, which df_div's remainder
// correction brings to the same bound.
#define __host__
#define __device__
#define __forceinline__ inline
#ifdef DF32_DEVICE_RCP
// The device code path: df_rcp takes the hardware reciprocal, accurate to
// 1 ulp.  Modelled as the correctly rounded 1 / x moved one ulp down or up
// (alternately), so every quotient starts from a wrong last bit.
#include <math.h>
#define __HIP_DEVICE_COMPILE__ 1
static unsigned g_rcp_flip = 0;
static inline float __builtin_amdgcn_rcpf(float x) {
  const float r = 1.0f / x;
  return nextafterf(r, (g_rcp_flip++ & 1u) ? INFINITY : -INFINITY);
}
#endif
#include "../../althea_amd/csrc/df32.hpp"

using namespace althea;

extern "C" {
// out[2 i] + out[2 i + 1] = atan2(y, x) for FP32 pairs (y_hi, y_lo), (x_hi, x_lo).
void df32_atan2_batch(const float* y, const float* x, int n, float* out) {
  for (int i = 0; i < n; ++i) {
    df32 r = df_atan2({y[2 * i], y[2 * i + 1]}, {x[2 * i], x[2 * i + 1]});
    out[2 * i] = r.hi;
    out[2 * i + 1] = r.lo;
  }
}
void df32_div_sqrt_batch(const float* a, const float* b, int n, float* q, float* s) {
  for (int i = 0; i < n; ++i) {
    df32 r = df_div({a[2 * i], a[2 * i + 1]}, {b[2 * i], b[2 * i + 1]});
    q[2 * i] = r.hi;
    q[2 * i + 1] = r.lo;
    df32 t = df_sqrt({a[2 * i] < 0 ? -a[2 * i] : a[2 * i], a[2 * i] < 0 ? -a[2 * i + 1] : a[2 * i + 1]});
    s[2 * i] = t.hi;
    s[2 * i + 1] = t.lo;
  }
}
// floor and weights of texel coordinates x = (hi, lo).
void df32_texel_batch(const float* x, int n, int* idx, float* a, float* b) {
  for (int i = 0; i < n; ++i) {
    TexelSplit t = df_texel({x[2 * i], x[2 * i + 1]});
    idx[i] = t.i;
    a[i] = t.a;
    b[i] = t.b;
  }
}
}
