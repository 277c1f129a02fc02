// Three-term BF16 split contractions on gfx950 MFMA (v_mfma_f32_16x16x32_bf16)
// with FP32 accumulation: the mixed-precision variants of the tall products of
// gemm.hip (overlap, projector, rotation and fnonloc GEMMs).
//
// Every FP64 operand value x is split as hi = bf16(x), lo = bf16(x - hi) (both
// round-to-nearest from the FP32 value of x: |x - hi| <= 2^-8 |x|, |x - hi - lo|
// <= 2^-16 |x|), and a product is accumulated as hi*hi + hi*lo + lo*hi; the
// dropped lo*lo term and the split residuals leave at most 3 * 2^-16 of |x y|.
// Each K chunk accumulates in FP32; split-K partials are written and reduced
// in FP64 (k_reduce_tn of gemm.hip), so the FP32 accumulation error is that of
// one chunk.
//
// Diagonal of square overlaps.  On a diagonal entry sum_k a_k b_k with
// correlated a, b (S = C^T C, or <c0|H c0>) every product loses the same-signed
// lo*lo, and the FP32 accumulation of a chunk shrinks the sum by a few 1e-7
// (measured); both bias the norms that orthonormalisation relies on.  The
// diagonal of every square wavefunction overlap is therefore recomputed in
// FP64 (k_diag_tn, one pass over the two operands), which leaves only the
// small, sign-random errors of the off-diagonal entries.
//
// Tiles are 128 x 128 per workgroup, 4 waves of 64 x 64 (4 x 4 MFMA tiles, 64
// FP32 accumulators per lane), K slices of 32 (one MFMA depth).  Operand
// slices are read from HBM as FP64, split once per workgroup and staged
// in LDS as bf16 rows of 32 k values (stride 40 bf16, 80 bytes: 16-byte lane
// reads without bank conflicts); the next slice is fetched into registers while
// the current one is multiplied.  Workgroup ids are dealt to XCDs as in
// gemm.hip, so tiles that share an operand block share an L2.
//
// MFMA operand layout (16x16x32 bf16, checked on gfx950 with a lane probe):
// lane l holds A[i = l%16][k = 8(l/16) .. 8(l/16)+7], B[k = 8(l/16) ..][j = l%16]
// and, in accumulator register r, D[i = 4(l/16) + r][j = l%16].
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mc_internal.h"

namespace mc {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

namespace {

constexpr int kT = 128;       // tile edge
constexpr int kK = 32;        // K slice
constexpr int kLd = 40;       // LDS row stride (bf16)
constexpr int kXcd = 8;

__device__ __forceinline__ int xcd_tile3(int total) {
  const int id = blockIdx.x;
  const int per = (total + kXcd - 1) / kXcd;
  return (id % kXcd) * per + id / kXcd;
}

// split 8 FP64 values into bf16 hi / lo parts
__device__ __forceinline__ void split8(const double* v, bf8& hi, bf8& lo) {
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const float f = (float)v[u];
    const __bf16 h = (__bf16)f;
    hi[u] = h;
    lo[u] = (__bf16)(f - (float)h);
  }
}

// one K slice of a 64 x 64 wave block: acc += A_s B_s with A_s rows
// [wr, wr+64) and B_s columns [wc, wc+64), both stored as [row][k] bf16:
// hi*hi + hi*lo + lo*hi.
__device__ __forceinline__ void x3_slice(const __bf16* Ahi, const __bf16* Alo, const __bf16* Bhi, const __bf16* Blo,
                                         int wr, int wc, f4 (&acc)[4][4]) {
  const int l = threadIdx.x & 63;
  const int off = (l & 15) * kLd + 8 * (l >> 4);
  bf8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    ah[a] = *reinterpret_cast<const bf8*>(Ahi + (wr + 16 * a) * kLd + off);
    al[a] = *reinterpret_cast<const bf8*>(Alo + (wr + 16 * a) * kLd + off);
  }
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    bh[b] = *reinterpret_cast<const bf8*>(Bhi + (wc + 16 * b) * kLd + off);
    bl[b] = *reinterpret_cast<const bf8*>(Blo + (wc + 16 * b) * kLd + off);
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
      acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
      acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
    }
}

// store 16 consecutive k values of one tile row (column of the source) as hi/lo
__device__ __forceinline__ void put16(__bf16* hi, __bf16* lo, int row, int kq, const double* v) {
  bf8 h, o;
  split8(v, h, o);
  *reinterpret_cast<bf8*>(hi + row * kLd + kq) = h;
  *reinterpret_cast<bf8*>(lo + row * kLd + kq) = o;
  split8(v + 8, h, o);
  *reinterpret_cast<bf8*>(hi + row * kLd + kq + 8) = h;
  *reinterpret_cast<bf8*>(lo + row * kLd + kq + 8) = o;
}

}  // namespace

// ---------------------------------------------------------------------------
// A^T B partials (as k_gemm_tn): launch ceil(total/8)*8 workgroups of 256,
// total = nsplit * tm * tn; A (K x m) column-major or the projector operand,
// B (K x n); P[s] (m x n) FP64 partials; upper: skip tiles below the diagonal.
// kchunk is a multiple of 32.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256, 2) void k_x3_tn(const double* __restrict__ A, int lda, const double* __restrict__ B,
                                                  int ldb, int K, int m, int n, int kchunk, double* __restrict__ P,
                                                  int upper, ProjA proj, int tm, int tn, int total) {
  __shared__ __attribute__((aligned(16))) __bf16 sm[4 * kT * kLd];
  __bf16* Ahi = sm;
  __bf16* Alo = sm + kT * kLd;
  __bf16* Bhi = sm + 2 * kT * kLd;
  __bf16* Blo = sm + 3 * kT * kLd;
  const int t = xcd_tile3(total);
  if (t >= total) return;
  const int s = t / (tm * tn);
  const int tile = t - s * tm * tn;
  const int ti = tile / tn, tj = tile - ti * tn;
  if (upper && ti > tj) return;
  const int i0 = ti * kT, j0 = tj * kT;
  const int kbeg = s * kchunk;
  const int kend = min(K, kbeg + kchunk);
  const int wave = threadIdx.x >> 6;
  const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;
  f4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};
  const int lc = threadIdx.x >> 1;          // tile column loaded by this thread
  const int lq = (threadIdx.x & 1) * 16;    // first k of its 16
  double ra[16], rb[16];
  auto fetch = [&](int kb) {
    const int acol = i0 + lc, bcol = j0 + lc;
    if (proj.eigr) {
      const int g0 = (kb + lq) >> 1;   // 8 G vectors, (re, im) rows
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int g = g0 + u;
        double re = 0.0, im = 0.0;
        if (acol < m && 2 * g < kend) {
          const double2 e = proj.eigr[(size_t)acol * proj.ldg + g];
          double tw = proj.twnl[g];
          if (proj.gkk) tw *= proj.gkk[g];
          if (proj.geq0 && g == 0) tw *= 0.5;
          switch (proj.lmod4) {   // (-i)^l * tw * e
            case 0: re = tw * e.x; im = tw * e.y; break;
            case 1: re = tw * e.y; im = -tw * e.x; break;
            case 2: re = -tw * e.x; im = -tw * e.y; break;
            default: re = -tw * e.y; im = tw * e.x; break;
          }
        }
        ra[2 * u] = re;
        ra[2 * u + 1] = im;
      }
    } else {
      const double* src = A + (size_t)acol * lda + kb + lq;
      if (acol < m && kb + lq + 16 <= kend) {   // 16-byte vector loads (lda even, chunk 32-aligned)
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const double2 v = reinterpret_cast<const double2*>(src)[u];
          ra[2 * u] = v.x;
          ra[2 * u + 1] = v.y;
        }
      } else {
#pragma unroll
        for (int u = 0; u < 16; ++u) ra[u] = (acol < m && kb + lq + u < kend) ? src[u] : 0.0;
      }
    }
    const double* src = B + (size_t)bcol * ldb + kb + lq;
    if (bcol < n && kb + lq + 16 <= kend) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const double2 v = reinterpret_cast<const double2*>(src)[u];
        rb[2 * u] = v.x;
        rb[2 * u + 1] = v.y;
      }
    } else {
#pragma unroll
      for (int u = 0; u < 16; ++u) rb[u] = (bcol < n && kb + lq + u < kend) ? src[u] : 0.0;
    }
  };
  if (kbeg < kend) fetch(kbeg);
  for (int kb = kbeg; kb < kend; kb += kK) {
    put16(Ahi, Alo, lc, lq, ra);
    put16(Bhi, Blo, lc, lq, rb);
    __syncthreads();
    if (kb + kK < kend) fetch(kb + kK);
    x3_slice(Ahi, Alo, Bhi, Blo, wr, wc, acc);
    __syncthreads();
  }
  const int l = threadIdx.x & 63;
  double* Ps = P + (size_t)s * m * n;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int j = j0 + wc + 16 * b + (l & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + wr + 16 * a + 4 * (l >> 4) + r;
        if (i < m && j < n) Ps[(size_t)j * m + i] = (double)acc[a][b][r];
      }
    }
}

// ---------------------------------------------------------------------------
// D = alpha A G + beta D (as k_gemm_nn): A (M x k) column-major lda, G (k x n)
// column-major ldgm, D (M x n) ldd; upper_g: G upper triangular; epi.twnl:
// fnonloc epilogue on (re, im) row pairs.  M, lda, ldd even and A 16-byte
// aligned (the host wrapper checks).  Launch ceil(total/8)*8 workgroups of
// 256, total = tm * tn (column tile fastest).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256, 2) void k_x3_nn(const double* __restrict__ A, int lda, const double* __restrict__ G,
                                                  int ldgm, int M, int k, int n, double alpha, double beta,
                                                  double* __restrict__ D, int ldd, int upper_g, FnlEpi epi, int tn,
                                                  int total) {
  __shared__ __attribute__((aligned(16))) __bf16 sm[4 * kT * kLd];
  __bf16* Ahi = sm;
  __bf16* Alo = sm + kT * kLd;
  __bf16* Ghi = sm + 2 * kT * kLd;
  __bf16* Glo = sm + 3 * kT * kLd;
  const int t = xcd_tile3(total);
  if (t >= total) return;
  const int rt = t / tn, ct = t - rt * tn;
  const int i0 = rt * kT, j0 = ct * kT;
  const int wave = threadIdx.x >> 6;
  const int l = threadIdx.x & 63;
  const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;
  f4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};
  const int kmax = upper_g ? min(k, j0 + kT) : k;
  const bool gvec = (ldgm % 2 == 0) && ((reinterpret_cast<uintptr_t>(G) & 15) == 0);
  // A slice (128 rows x 32 k): wave w loads k = 8w .. 8w+7 of rows i0+2l, i0+2l+1
  const int ar = 2 * l, ak = 8 * wave;
  // G slice (32 k x 128 columns): column gc, k = gq .. gq+15
  const int gc = threadIdx.x >> 1, gq = (threadIdx.x & 1) * 16;
  double2 ra[8];
  double rg[16];
  auto fetch = [&](int kb) {
    const int row = i0 + ar;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int kk = kb + ak + u;
      ra[u] = (kk < kmax && row < M) ? *reinterpret_cast<const double2*>(A + (size_t)kk * lda + row)
                                     : double2{0.0, 0.0};
    }
    const int col = j0 + gc;
    const double* src = G + (size_t)col * ldgm + kb + gq;
    if (gvec && col < n && kb + gq + 16 <= kmax) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const double2 v = reinterpret_cast<const double2*>(src)[u];
        rg[2 * u] = v.x;
        rg[2 * u + 1] = v.y;
      }
    } else {
#pragma unroll
      for (int u = 0; u < 16; ++u) rg[u] = (col < n && kb + gq + u < kmax) ? src[u] : 0.0;
    }
  };
  if (kmax > 0) fetch(0);
  for (int kb = 0; kb < kmax; kb += kK) {
    {
      double v0[8], v1[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        v0[u] = ra[u].x;
        v1[u] = ra[u].y;
      }
      bf8 h, o;
      split8(v0, h, o);
      *reinterpret_cast<bf8*>(Ahi + ar * kLd + ak) = h;
      *reinterpret_cast<bf8*>(Alo + ar * kLd + ak) = o;
      split8(v1, h, o);
      *reinterpret_cast<bf8*>(Ahi + (ar + 1) * kLd + ak) = h;
      *reinterpret_cast<bf8*>(Alo + (ar + 1) * kLd + ak) = o;
    }
    put16(Ghi, Glo, gc, gq, rg);
    __syncthreads();
    if (kb + kK < kmax) fetch(kb + kK);
    x3_slice(Ahi, Alo, Ghi, Glo, wr, wc, acc);
    __syncthreads();
  }
  // epilogue: lane holds rows i .. i+3 (i = 4(l/16) + 16a + wr + i0) of column j
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int j = j0 + wc + 16 * b + (l & 15);
    if (j >= n) continue;
    double f = 1.0;
    if (epi.twnl) {
      f = epi.f[j];
      if (f < 1e-5) f = 1.0;
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int i = i0 + wr + 16 * a + 4 * (l >> 4);
      if (i >= M) continue;
      double* d = D + (size_t)j * ldd + i;
      if (epi.twnl) {
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          const double x0 = acc[a][b][2 * p], x1 = acc[a][b][2 * p + 1];
          const double tw = epi.twnl[(i >> 1) + p];
          double yr, yi;
          switch (epi.lmod4) {   // ctm = -f (-i)^l ; y = tw * ctm * x
            case 0: yr = -f * tw * x0; yi = -f * tw * x1; break;
            case 1: yr = -f * tw * x1; yi = f * tw * x0; break;
            case 2: yr = f * tw * x0; yi = f * tw * x1; break;
            default: yr = f * tw * x1; yi = -f * tw * x0; break;
          }
          if (i + 2 * p < M) {
            d[2 * p] += yr;
            d[2 * p + 1] += yi;
          }
        }
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (i + r < M) d[r] = (beta == 0.0 ? 0.0 : beta * d[r]) + alpha * (double)acc[a][b][r];
      }
    }
  }
}

// out(i, i) = alpha sum_k A(k, i) B(k, i) [- A(0, i) B(0, i)] in FP64, one
// workgroup per column, fixed reduction order
__global__ __launch_bounds__(256) void k_diag_tn(const double* __restrict__ A, int lda, const double* __restrict__ B,
                                                 int ldb, int K, double alpha, int g0corr, double* __restrict__ out,
                                                 int ldo) {
  const int i = blockIdx.x;
  const double* a = A + (size_t)lda * i;
  const double* b = B + (size_t)ldb * i;
  double s = 0.0;
  for (int k = threadIdx.x; k < K; k += 256) s += a[k] * b[k];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  __shared__ double part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = alpha * (part[0] + part[1] + part[2] + part[3]);
    if (g0corr) v -= a[0] * b[0];
    out[(size_t)ldo * i + i] = v;
  }
}

// ---------------------------------------------------------------------------
// host wrappers
// ---------------------------------------------------------------------------
void gemm_tn_x3(mc_ctx* ctx, const double* A, int lda, const double* B, int ldb, int K, int m, int n, double alpha,
                bool g0corr, bool symmetric, double* out, int ldo, const ProjA* proj) {
  if (m <= 0 || n <= 0) return;
  MC_CHECK(ldb % 2 == 0 && (reinterpret_cast<uintptr_t>(B) & 15) == 0 &&
               (proj || (lda % 2 == 0 && (reinterpret_cast<uintptr_t>(A) & 15) == 0)),
           MC_ERR_INTERNAL, "split-bf16 gemm_tn needs even lda, ldb and 16-byte aligned operands");
  const int tm = (m + kT - 1) / kT, tn = (n + kT - 1) / kT;
  const int tiles = symmetric ? (tm * (tm + 1)) / 2 : tm * tn;
  // about 1024 busy workgroups, K chunks of at least 1024 rows
  int nsplit = std::max(1, std::min((1024 + tiles - 1) / tiles, std::max(1, K / 1024)));
  int kchunk = (K + nsplit - 1) / nsplit;
  kchunk = (kchunk + kK - 1) / kK * kK;
  nsplit = (K + kchunk - 1) / kchunk;
  ctx->w_gemm.ensure(sizeof(double) * (size_t)nsplit * m * n);
  ProjA pa{};
  if (proj) pa = *proj;
  const int total = nsplit * tm * tn;
  const int launch = (total + kXcd - 1) / kXcd * kXcd;
  timer_begin(ctx, MC_KID_GEMM);
  hipLaunchKernelGGL(k_x3_tn, dim3(launch), dim3(256), 0, ctx->stream, A, lda, B, ldb, K, m, n, kchunk,
                     ctx->w_gemm.as<double>(), symmetric ? 1 : 0, pa, tm, tn, total);
  MC_HIP(hipGetLastError());
  timer_end(ctx, MC_KID_GEMM);
  reduce_tn(ctx, nsplit, m, n, alpha, A, lda, B, ldb, g0corr, symmetric, out, ldo);
  if (!proj && m == n) {
    timer_begin(ctx, MC_KID_GEMM);
    hipLaunchKernelGGL(k_diag_tn, dim3(n), dim3(256), 0, ctx->stream, A, lda, B, ldb, K, alpha, g0corr ? 1 : 0, out,
                       ldo);
    MC_HIP(hipGetLastError());
    timer_end(ctx, MC_KID_GEMM);
  }
}

void gemm_nn_x3(mc_ctx* ctx, const double* A, int lda, const double* G, int ldgm, int M, int k, int n, double alpha,
                double beta, double* D, int ldd, bool upper_g, const FnlEpi* epi) {
  if (M <= 0 || n <= 0) return;
  MC_CHECK(M % 2 == 0 && lda % 2 == 0 && ldd % 2 == 0 && (reinterpret_cast<uintptr_t>(A) & 15) == 0, MC_ERR_INTERNAL,
           "split-bf16 gemm_nn needs even M, lda, ldd and a 16-byte aligned A");
  FnlEpi e{};
  if (epi) e = *epi;
  const int tm = (M + kT - 1) / kT, tn = (n + kT - 1) / kT;
  const int total = tm * tn;
  const int launch = (total + kXcd - 1) / kXcd * kXcd;
  timer_begin(ctx, MC_KID_GEMM);
  hipLaunchKernelGGL(k_x3_nn, dim3(launch), dim3(256), 0, ctx->stream, A, lda, G, ldgm, M, k, n, alpha, beta, D, ldd,
                     upper_g ? 1 : 0, e, tn, total);
  MC_HIP(hipGetLastError());
  timer_end(ctx, MC_KID_GEMM);
}

}  // namespace mc
