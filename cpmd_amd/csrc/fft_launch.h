// Helpers shared by the transform kernels (wfft.hip, vofrho.hip): LDS
// carving, twiddle staging, wave reductions and the host-side choice of lines
// per workgroup and of the compile-time transform length.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "fft_device.h"
#include "mc_internal.h"

namespace mc {

__device__ __forceinline__ double warp_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

template <typename T>
__device__ __forceinline__ void load_tw(v2<T>* tw, const double2* src, int n) {
  for (int t = threadIdx.x; t < n; t += kThreads) {
    const double2 w = src[t];
    tw[t] = cmk<v2<T>>((T)w.x, (T)w.y);
  }
}

template <typename T>
__device__ __forceinline__ v2<T>* lds_base() {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  return reinterpret_cast<v2<T>*>(smem_raw);
}


constexpr int kLdsBudget = 65536;

// lines per workgroup for a transform of length n: engine limits, LDS (lines + table) and 16
template <typename T>
inline int lines_for(const FftPlan& pl) {
  const int a = fast_a(pl.n);
  int l = a > 0 ? kThreads / a : max_lines(pl);
  l = std::min(l, kLdsBudget / (int)(sizeof(v2<T>) * pl.n) - 1);
  l = std::min(l, 16);
  MC_CHECK(l >= 1, MC_ERR_UNSUPPORTED, "transform length too large for one workgroup");
  return l;
}


// calls f(std::integral_constant<int, N>) with N = n if n has a compile-time
// engine, N = 0 otherwise
template <class F>
inline void with_length(int n, F&& f) {
  switch (n) {
#define MC_CASE(x) \
  case x: f(std::integral_constant<int, x>()); return;
    MC_FAST_SIZES(MC_CASE)
#undef MC_CASE
    default: f(std::integral_constant<int, 0>()); return;
  }
}

#define LAUNCH(kid, kern, grid, block, shm, ...)                                  \
  do {                                                                          \
    timer_begin(ctx, kid);                                                      \
    hipLaunchKernelGGL(kern, grid, block, shm, ctx->stream, __VA_ARGS__);       \
    MC_HIP(hipGetLastError());                                                  \
    timer_end(ctx, kid);                                                        \
  } while (0)


}  // namespace mc
