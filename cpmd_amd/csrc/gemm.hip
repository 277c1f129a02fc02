// FP64 MFMA contractions for gfx950 (v_mfma_f64_16x16x4_f64).
//
// Wavefunctions are handled in their real view: a complex (ldg x n) block is a
// (2*ldg x n) column-major real matrix, so the Gamma-point overlap
// 2 Re(C1^H C2) is 2 * C1r^T C2r (CPMD's dsyrk/dgemm on 2*NGW rows,
// src/ovlap_utils.mod.F90:150-163).
//
//  * gemm_tn: P = A^T B over K (= 2*ngw) with A (K x m), B (K x n); split-K
//    partials reduced in a fixed order (deterministic); optional upper-tile
//    only (symmetric products) and a fused "projector" A operand built on the
//    fly from twnl * eigr (rnlsm1).
//  * gemm_nn: D = alpha * A G + beta * D with A (M x k) tall (M = 2*ngw), G
//    (k x n) small; optional upper-triangular G (trmm) and a fused complex
//    epilogue for fnonloc.
//
// Each workgroup computes a 64x64 tile with 4 waves (2x2), each wave a 32x32
// block = 2x2 MFMA tiles of 16x16.  Operands are staged through LDS in
// K-slices of 32.  MFMA operand layout (16x16x4 f64, checked on gfx950 with a
// lane probe): lane l holds A[i = l%16][k = l/16], B[k = l/16][j = l%16] and,
// in accumulator register r, D[i = 4r + l/16][j = l%16].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "mc_internal.h"

namespace mc {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kGT = 64;    // tile edge
constexpr int kGK = 32;    // K slice
constexpr int kGPad = 2;   // LDS row padding (doubles)
constexpr int kGLd = kGT + kGPad;

__device__ __forceinline__ void mfma_block(const double* As, const double* Bs, int wr, int wc, d4 (&acc)[2][2]) {
  const int l = threadIdx.x & 63;
  const int li = l & 15, lk = l >> 4;
#pragma unroll
  for (int kk = 0; kk < kGK; kk += 4) {
    double a0 = As[(kk + lk) * kGLd + wr + li];
    double a1 = As[(kk + lk) * kGLd + wr + 16 + li];
    double b0 = Bs[(kk + lk) * kGLd + wc + li];
    double b1 = Bs[(kk + lk) * kGLd + wc + 16 + li];
    acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
  }
}

// A^T B partial over K-chunk blockIdx.z.
//   A: (K x m) column-major, lda (doubles); if proj.eigr != nullptr, A is the
//      projector operand instead (K = 2*ngw rows: 2g = Re, 2g+1 = Im).
//   B: (K x n) column-major, ldb.
//   P: partials [nsplit][n][m] (column-major m x n per split).
// upper: only tiles with tile_row <= tile_col.
__global__ __launch_bounds__(256) void k_gemm_tn(const double* __restrict__ A, int lda,
                                                 const double* __restrict__ B, int ldb, int K, int m, int n,
                                                 int kchunk, double* __restrict__ P, int upper, ProjA proj) {
  __shared__ double As[kGK * kGLd];
  __shared__ double Bs[kGK * kGLd];
  const int ti = blockIdx.x, tj = blockIdx.y, s = blockIdx.z;
  if (upper && ti > tj) return;
  const int i0 = ti * kGT, j0 = tj * kGT;
  const int kbeg = s * kchunk;
  const int kend = min(K, kbeg + kchunk);
  const int wave = threadIdx.x >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  d4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
  const int lc = threadIdx.x >> 2;        // column within the tile
  const int lq = (threadIdx.x & 3) * 8;   // k offset within the slice
  for (int kb = kbeg; kb < kend; kb += kGK) {
    // ---- A slice: 32 k x 64 columns
    {
      const int col = i0 + lc;
      double v[8];
      if (proj.eigr) {
        // rows kb+lq .. +8 = G vectors g0 .. g0+4 (re, im interleaved); kb, lq even
        const int g0 = (kb + lq) >> 1;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int g = g0 + u;
          double re = 0.0, im = 0.0;
          if (col < m && 2 * g < kend) {
            const double2 e = proj.eigr[(size_t)col * proj.ldg + g];
            double t = proj.twnl[g];
            if (proj.geq0 && g == 0) t *= 0.5;
            // (-i)^l * t * e
            switch (proj.lmod4) {
              case 0: re = t * e.x; im = t * e.y; break;
              case 1: re = t * e.y; im = -t * e.x; break;
              case 2: re = -t * e.x; im = -t * e.y; break;
              default: re = -t * e.y; im = t * e.x; break;
            }
          }
          v[2 * u] = re;
          v[2 * u + 1] = im;
        }
      } else {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int k = kb + lq + u;
          v[u] = (col < m && k < kend) ? A[(size_t)col * lda + k] : 0.0;
        }
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) As[(lq + u) * kGLd + lc] = v[u];
    }
    // ---- B slice
    {
      const int col = j0 + lc;
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int k = kb + lq + u;
        v[u] = (col < n && k < kend) ? B[(size_t)col * ldb + k] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) Bs[(lq + u) * kGLd + lc] = v[u];
    }
    __syncthreads();
    mfma_block(As, Bs, wr, wc, acc);
    __syncthreads();
  }
  const int l = threadIdx.x & 63;
  double* Ps = P + (size_t)s * m * n;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int j = j0 + wc + b * 16 + (l & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + wr + a * 16 + (l >> 4) + 4 * r;
        if (i < m && j < n) Ps[(size_t)j * m + i] = acc[a][b][r];
      }
    }
}

// out(i,j) = alpha * sum_s P[s](i,j) - g0 * a0(i) b0(j); symmetric: only the
// upper triangle was computed, mirrored into the lower one (dmatc).
__global__ void k_reduce_tn(const double* __restrict__ P, int nsplit, int m, int n, double alpha,
                            const double* __restrict__ a0, int lda, const double* __restrict__ b0, int ldb,
                            int g0, int symmetric, double* __restrict__ out, int ldo) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= m || j >= n) return;
  int si = i, sj = j;
  if (symmetric && i > j) { si = j; sj = i; }
  double s = 0.0;
  for (int q = 0; q < nsplit; ++q) s += P[(size_t)q * m * n + (size_t)sj * m + si];
  double v = alpha * s;
  if (g0) v -= a0[(size_t)si * lda] * b0[(size_t)sj * ldb];
  out[(size_t)j * ldo + i] = v;
}

// D = alpha A G + beta D.  A (M x k) column-major lda; G (k x n) column-major ldgm;
// D (M x n) column-major ldd.  upper_g: G is upper triangular (k-loop stops at
// the diagonal block).  With epi.twnl, D (real view of c2) += scaled A G.
__global__ __launch_bounds__(256) void k_gemm_nn(const double* __restrict__ A, int lda,
                                                 const double* __restrict__ G, int ldgm, int M, int k, int n,
                                                 double alpha, double beta, double* __restrict__ D, int ldd,
                                                 int upper_g, FnlEpi epi) {
  __shared__ double As[kGK * kGLd];
  __shared__ double Gs[kGK * kGLd];
  const int i0 = blockIdx.x * kGT, j0 = blockIdx.y * kGT;
  const int wave = threadIdx.x >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  d4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
  const int kmax = upper_g ? min(k, j0 + kGT) : k;
  const int ak = threadIdx.x >> 3, ai = (threadIdx.x & 7) * 8;   // A slice: 32 k x 64 rows
  const int gj = threadIdx.x >> 2, gq = (threadIdx.x & 3) * 8;   // G slice: 32 k x 64 cols
  for (int kb = 0; kb < kmax; kb += kGK) {
    {
      const int kk = kb + ak;
      const double* src = A + (size_t)kk * lda + i0 + ai;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int row = i0 + ai + u;
        As[ak * kGLd + ai + u] = (kk < kmax && row < M) ? src[u] : 0.0;
      }
    }
    {
      const int col = j0 + gj;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int kk = kb + gq + u;
        Gs[(gq + u) * kGLd + gj] = (kk < kmax && col < n) ? G[(size_t)col * ldgm + kk] : 0.0;
      }
    }
    __syncthreads();
    mfma_block(As, Gs, wr, wc, acc);
    __syncthreads();
  }
  const int l = threadIdx.x & 63;
  const int q = l >> 4;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int j = j0 + wc + b * 16 + (l & 15);
      const int ib = i0 + wr + a * 16 + q;          // rows ib + 4r; ib has the parity of q
      if (epi.twnl) {
        double ffi = (j < n) ? epi.f[j] : 1.0;
        if (ffi < 1e-5) ffi = 1.0;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          // the imaginary row 2g+1 of G vector g lives in lane l^16 (q odd)
          const double xr = acc[a][b][r];
          const double xi = __shfl_xor(xr, 16, 64);
          const int i = ib + 4 * r;
          if ((q & 1) == 0 && j < n && i + 1 < M) {
            const double t = epi.twnl[i >> 1];
            double yr, yi;
            // ctm = -ffi (-i)^l ; y = t * ctm * x
            switch (epi.lmod4) {
              case 0: yr = -ffi * t * xr; yi = -ffi * t * xi; break;
              case 1: yr = -ffi * t * xi; yi = ffi * t * xr; break;
              case 2: yr = ffi * t * xr; yi = ffi * t * xi; break;
              default: yr = ffi * t * xi; yi = -ffi * t * xr; break;
            }
            double* d = D + (size_t)j * ldd + i;
            d[0] += yr;
            d[1] += yi;
          }
        }
      } else if (j < n) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = ib + 4 * r;
          if (i < M) {
            double* d = D + (size_t)j * ldd + i;
            *d = (beta == 0.0 ? 0.0 : beta * *d) + alpha * acc[a][b][r];
          }
        }
      }
    }
}

// ---------------------------------------------------------------------------
// host wrappers
// ---------------------------------------------------------------------------

// out (m x n, ldo) = alpha A^T B [- a0 b0^T] with K rows; A may be the projector operand.
void gemm_tn(mc_ctx* ctx, const double* A, int lda, const double* B, int ldb, int K, int m, int n, double alpha,
             bool g0corr, bool symmetric, double* out, int ldo, const ProjA* proj) {
  if (m <= 0 || n <= 0) return;
  const int tm = (m + kGT - 1) / kGT, tn = (n + kGT - 1) / kGT;
  int tiles = symmetric ? (tm * (tm + 1)) / 2 : tm * tn;
  // aim at ~4 workgroups per CU (256 CUs) with K chunks of at least 1024 rows
  int nsplit = std::max(1, std::min((1024 + tiles - 1) / tiles, std::max(1, K / 1024)));
  int kchunk = (K + nsplit - 1) / nsplit;
  kchunk = (kchunk + kGK - 1) / kGK * kGK;
  nsplit = (K + kchunk - 1) / kchunk;
  ctx->w_gemm.ensure(sizeof(double) * (size_t)nsplit * m * n);
  ProjA pa{};
  if (proj) pa = *proj;
  timer_begin(ctx, MC_KID_GEMM);
  hipLaunchKernelGGL(k_gemm_tn, dim3(tm, tn, nsplit), dim3(256), 0, ctx->stream, A, lda, B, ldb, K, m, n, kchunk,
                     ctx->w_gemm.as<double>(), symmetric ? 1 : 0, pa);
  MC_HIP(hipGetLastError());
  timer_end(ctx, MC_KID_GEMM);
  timer_begin(ctx, MC_KID_MISC);
  hipLaunchKernelGGL(k_reduce_tn, dim3((m + 255) / 256, n), dim3(256), 0, ctx->stream, ctx->w_gemm.as<double>(),
                     nsplit, m, n, alpha, A, lda, B, ldb, g0corr ? 1 : 0, symmetric ? 1 : 0, out, ldo);
  MC_HIP(hipGetLastError());
  timer_end(ctx, MC_KID_MISC);
}

void gemm_nn(mc_ctx* ctx, const double* A, int lda, const double* G, int ldgm, int M, int k, int n, double alpha,
             double beta, double* D, int ldd, bool upper_g, const FnlEpi* epi) {
  if (M <= 0 || n <= 0) return;
  FnlEpi e{};
  if (epi) e = *epi;
  timer_begin(ctx, MC_KID_GEMM);
  hipLaunchKernelGGL(k_gemm_nn, dim3((M + kGT - 1) / kGT, (n + kGT - 1) / kGT), dim3(256), 0, ctx->stream, A, lda,
                     G, ldgm, M, k, n, alpha, beta, D, ldd, upper_g ? 1 : 0, e);
  MC_HIP(hipGetLastError());
  timer_end(ctx, MC_KID_GEMM);
}

}  // namespace mc
