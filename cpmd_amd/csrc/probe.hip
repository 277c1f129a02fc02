// MFMA issue-rate probe (gfx950): the rate one matrix instruction reaches on
// this GPU with independent accumulator chains, register operands and no
// memory traffic in the loop.  It is the ceiling the contraction kernels are
// measured against beside the published dense peaks.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mc_internal.h"

namespace mc {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef double d4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int kProbeGrid = 4096;   // workgroups of 256 (16 per CU on 256 CUs)

// KIND 0: v_mfma_f32_16x16x32_bf16 (8 chains), 1: v_mfma_f32_32x32x16_bf16
// (4 chains), 2: v_mfma_f64_16x16x4_f64 (8 chains), 3: v_mfma_f32_16x16x4_f32
// (8 chains).  One float per thread is stored so that nothing is eliminated.
template <int KIND>
__global__ __launch_bounds__(256) void k_probe_mfma(int iters, float* __restrict__ out) {
  const float x = 1e-3f * threadIdx.x;
  float sum = 0.f;
  if constexpr (KIND == 0) {
    bf8 a, b;
    for (int u = 0; u < 8; ++u) {
      a[u] = (__bf16)(x + u);
      b[u] = (__bf16)(x - u);
    }
    f4 acc[8];
    for (int q = 0; q < 8; ++q) acc[q] = f4{0.f, 0.f, 0.f, 0.f};
    for (int it = 0; it < iters; ++it)
#pragma unroll
      for (int q = 0; q < 8; ++q) acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[q], 0, 0, 0);
    for (int q = 0; q < 8; ++q) sum += acc[q][0] + acc[q][3];
  } else if constexpr (KIND == 1) {
    bf8 a, b;
    for (int u = 0; u < 8; ++u) {
      a[u] = (__bf16)(x + u);
      b[u] = (__bf16)(x - u);
    }
    f16v acc[4];
    for (int q = 0; q < 4; ++q)
      for (int r = 0; r < 16; ++r) acc[q][r] = 0.f;
    for (int it = 0; it < iters; ++it)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[q], 0, 0, 0);
    for (int q = 0; q < 4; ++q) sum += acc[q][0] + acc[q][15];
  } else if constexpr (KIND == 2) {
    const double a = x, b = 1.0 - x;
    d4 acc[8];
    for (int q = 0; q < 8; ++q) acc[q] = d4{0.0, 0.0, 0.0, 0.0};
    for (int it = 0; it < iters; ++it)
#pragma unroll
      for (int q = 0; q < 8; ++q) acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[q], 0, 0, 0);
    for (int q = 0; q < 8; ++q) sum += (float)(acc[q][0] + acc[q][3]);
  } else {
    const float a = x, b = 1.f - x;
    f4 acc[8];
    for (int q = 0; q < 8; ++q) acc[q] = f4{0.f, 0.f, 0.f, 0.f};
    for (int it = 0; it < iters; ++it)
#pragma unroll
      for (int q = 0; q < 8; ++q) acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[q], 0, 0, 0);
    for (int q = 0; q < 8; ++q) sum += acc[q][0] + acc[q][3];
  }
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = sum;
}

template <int KIND>
double probe_once(hipStream_t st, float* out, int iters, double flops_per_iter_wave) {
  hipLaunchKernelGGL(k_probe_mfma<KIND>, dim3(kProbeGrid), dim3(256), 0, st, 16, out);   // warm up
  MC_HIP(hipGetLastError());
  hipEvent_t e0, e1;
  MC_HIP(hipEventCreate(&e0));
  MC_HIP(hipEventCreate(&e1));
  MC_HIP(hipEventRecord(e0, st));
  hipLaunchKernelGGL(k_probe_mfma<KIND>, dim3(kProbeGrid), dim3(256), 0, st, iters, out);
  MC_HIP(hipGetLastError());
  MC_HIP(hipEventRecord(e1, st));
  MC_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  MC_HIP(hipEventElapsedTime(&ms, e0, e1));
  MC_HIP(hipEventDestroy(e0));
  MC_HIP(hipEventDestroy(e1));
  return (double)kProbeGrid * 4 * iters * flops_per_iter_wave / (ms * 1e-3) * 1e-12;
}

}  // namespace
}  // namespace mc

extern "C" int mc_probe_mfma(int kind, double* tflops) {
  MC_API_BEGIN_RO
  MC_CHECK(tflops && kind >= 0 && kind <= 3, MC_ERR_ARG, "probe_mfma: kind must be 0..3");
  *tflops = 0.0;
  mc::DevBuf buf;   // counted against the device memory limit like every library buffer
  buf.ensure(sizeof(float) * mc::kProbeGrid * 256);
  float* out = buf.as<float>();
  hipStream_t st;
  MC_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  try {
    // best of three timings of about 5 ms each
    for (int rep = 0; rep < 3; ++rep) {
      double r = 0.0;
      switch (kind) {
        case 0: r = mc::probe_once<0>(st, out, 2000, 8 * 16384.0); break;
        case 1: r = mc::probe_once<1>(st, out, 2000, 4 * 32768.0); break;
        case 2: r = mc::probe_once<2>(st, out, 1000, 8 * 2048.0); break;
        default: r = mc::probe_once<3>(st, out, 2000, 8 * 2048.0); break;
      }
      *tflops = std::max(*tflops, r);
    }
  } catch (...) {
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    throw;
  }
  MC_HIP(hipStreamDestroy(st));
  MC_API_END
}
