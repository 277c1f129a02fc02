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
// Each workgroup computes a 128x128 tile (64x64 for small problems) with 4
// waves (2x2), each wave a 64x64 block = 4x4 MFMA tiles of 16x16 (64 FP64
// accumulators per lane).  Operands
// are staged through LDS in K slices of 16; the next slice is fetched into
// registers while the current one is multiplied.  Workgroup ids are remapped so
// that tiles sharing an operand block run on the same XCD (ids are dealt to the
// 8 XCDs round robin): consecutive remapped ids share the K chunk (gemm_tn) or
// the row block of A (gemm_nn).  The tall output of gemm_nn is written through
// an LDS stage as 512-byte column segments.
//
// With ctx->gemm_x3 (mixed precision) the tall products (K or M of at least
// kX3Min rows, i.e. the wavefunction contractions) go to the split-bf16 kernels
// of gemm_x3.hip; the small n x n products stay FP64.
//
// MFMA operand layout (16x16x4 f64, checked on gfx950 with a lane probe):
// lane l holds A[i = l%16][k = l/16], B[k = l/16][j = l%16] and, in
// accumulator register r, D[i = 4r + l/16][j = l%16].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "mc_internal.h"

namespace mc {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kBK = 16;           // K slice
constexpr int kXcd = 8;           // XCDs of an MI355X

// Tile geometry: BT x BT per workgroup (BT = 128 for the tall products, 64
// for small n x n ones), 2 x 2 waves of WB = BT/2, NB = WB/16 MFMA tiles per
// wave edge, LDS row stride LD.
template <int BT>
struct Tile {
  static constexpr int WB = BT / 2, NB = WB / 16, LD = BT + 4;
  static constexpr int PER = BT / 16;          // doubles per thread and operand slice
  static constexpr int TPC = 256 / BT;         // threads per tile column in a slice
};

template <int BT>
__device__ __forceinline__ void mfma_tile(const double* As, const double* Bs, int wr, int wc,
                                          d4 (&acc)[Tile<BT>::NB][Tile<BT>::NB]) {
  constexpr int NB = Tile<BT>::NB, LD = Tile<BT>::LD;
  const int l = threadIdx.x & 63;
  const int li = l & 15, lk = l >> 4;
#pragma unroll
  for (int kk = 0; kk < kBK; kk += 4) {
    double a[NB], b[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) a[i] = As[(kk + lk) * LD + wr + 16 * i + li];
#pragma unroll
    for (int j = 0; j < NB; ++j) b[j] = Bs[(kk + lk) * LD + wc + 16 * j + li];
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
  }
}

// remapped workgroup index: XCD x = id % 8 gets the contiguous range [x*per, (x+1)*per)
__device__ __forceinline__ int xcd_tile(int total) {
  const int id = blockIdx.x;
  const int per = (total + kXcd - 1) / kXcd;
  return (id % kXcd) * per + id / kXcd;
}

// ---------------------------------------------------------------------------
// A^T B partials.  Launch ceil(total/8)*8 workgroups, total = nsplit * tile_count(upper, tm, tn).
//   A: (K x m) column-major, lda; if proj.eigr != nullptr, A is the projector
//      operand (K = 2*ngw rows: 2g = Re, 2g+1 = Im).  B: (K x n), ldb.
//   P: partials [nsplit][n][m] (column-major m x n per split).
// upper (m == n): only the tiles on or above the diagonal are launched.
// ---------------------------------------------------------------------------
template <int BT>
__global__ __launch_bounds__(256) void k_gemm_tn(const double* __restrict__ A, int lda,
                                                 const double* __restrict__ B, int ldb, int K, int m, int n,
                                                 int kchunk, double* __restrict__ P, int upper, ProjA proj, int tm,
                                                 int tn, int total) {
  using TL = Tile<BT>;
  constexpr int NB = TL::NB, LD = TL::LD, PER = TL::PER, WB = TL::WB;
  __shared__ double As[kBK * LD];
  __shared__ double Bs[kBK * LD];
  const int t = xcd_tile(total);
  if (t >= total) return;
  const int tiles = tile_count(upper, tm, tn);
  const int s = t / tiles;
  int ti, tj;
  tile_ij(t - s * tiles, upper, tm, tn, ti, tj);
  const int i0 = ti * BT, j0 = tj * BT;
  const int kbeg = s * kchunk;
  const int kend = min(K, kbeg + kchunk);
  const int wave = threadIdx.x >> 6;
  const int wr = (wave >> 1) * WB, wc = (wave & 1) * WB;
  d4 acc[NB][NB];
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
  const int lc = threadIdx.x / TL::TPC;            // column within the tile
  const int lq = (threadIdx.x % TL::TPC) * PER;    // k offset within the slice (even)
  double ra[PER], rb[PER];
  auto fetch = [&](int kb) {
    const int acol = i0 + lc, bcol = j0 + lc;
    if (proj.eigr) {
      const int g0 = (kb + lq) >> 1;   // rows kb+lq .. +PER = G vectors g0 .. (re, im)
#pragma unroll
      for (int u = 0; u < PER / 2; ++u) {
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
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        const int k = kb + lq + u;
        ra[u] = (acol < m && k < kend) ? A[(size_t)acol * lda + k] : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int k = kb + lq + u;
      rb[u] = (bcol < n && k < kend) ? B[(size_t)bcol * ldb + k] : 0.0;
    }
  };
  if (kbeg < kend) fetch(kbeg);
  for (int kb = kbeg; kb < kend; kb += kBK) {
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      As[(lq + u) * LD + lc] = ra[u];
      Bs[(lq + u) * LD + lc] = rb[u];
    }
    __syncthreads();
    if (kb + kBK < kend) fetch(kb + kBK);
    mfma_tile<BT>(As, Bs, wr, wc, acc);
    __syncthreads();
  }
  const int l = threadIdx.x & 63;
  double* Ps = P + (size_t)s * m * n;
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int j = j0 + wc + b * 16 + (l & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + wr + a * 16 + (l >> 4) + 4 * r;
        if (i < m && j < n) Ps[(size_t)j * m + i] = acc[a][b][r];
      }
    }
}

// out(i,j) = alpha * sum_s P[s](i,j) - g0 * a0(i) b0(j); symmetric: only the
// upper triangle is used and mirrored into the lower one (dmatc).
// out(i, j) = alpha sum_q P[q](i, j) [- A(0, i) B(0, j)], partials P[q] of
// ldp rows (the first m used), mirrored from the upper triangle if symmetric
// sum_q P[q * stride + o], q = 0 .. nsplit-1 in order; eight loads in flight
// per step (the adds keep the sequential order, so the sum is the same)
__device__ __forceinline__ double sum_partials(const double* __restrict__ P, int nsplit, size_t stride, size_t o) {
  double s = 0.0;
  int q = 0;
  for (; q + 8 <= nsplit; q += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = P[(size_t)(q + u) * stride + o];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += v[u];
  }
  for (; q < nsplit; ++q) s += P[(size_t)q * stride + o];
  return s;
}

__global__ void k_reduce_tn(const double* __restrict__ P, int nsplit, int ldp, int m, int n, double alpha,
                            const double* __restrict__ a0, int lda, const double* __restrict__ b0, int ldb,
                            int g0, int symmetric, double* __restrict__ out, int ldo) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= m || j >= n) return;
  int si = i, sj = j;
  if (symmetric && i > j) { si = j; sj = i; }
  const double s = sum_partials(P, nsplit, (size_t)ldp * n, (size_t)sj * ldp + si);
  double v = alpha * s;
  if (g0) v -= a0[(size_t)si * lda] * b0[(size_t)sj * ldb];
  out[(size_t)j * ldo + i] = v;
}

// The symmetric case of k_reduce_tn (m == n, only tiles on or above the
// diagonal computed): out(i, j) = out(j, i) = the value of k_reduce_tn at
// i <= j.  One workgroup per 32 x 32 block (bi <= bj): coalesced sums down the
// columns of P, the mirrored block written through LDS.
__global__ __launch_bounds__(256) void k_reduce_tn_sym(const double* __restrict__ P, int nsplit, int ldp, int n,
                                                       double alpha, const double* __restrict__ a0, int lda,
                                                       const double* __restrict__ b0, int ldb, int g0,
                                                       double* __restrict__ out, int ldo) {
  __shared__ double tile[32][33];
  const int nb = (n + 31) / 32;
  int bi = 0, bj = blockIdx.x;   // blockIdx.x -> (bi, bj), bi <= bj, row-major over the upper triangle
  while (bj >= nb - bi) {
    bj -= nb - bi;
    ++bi;
  }
  bj += bi;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int i = 32 * bi + tx;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int jl = ty + 8 * u, j = 32 * bj + jl;
    double v = 0.0;
    if (i < n && j < n && i <= j) {
      const double s = sum_partials(P, nsplit, (size_t)ldp * n, (size_t)j * ldp + i);
      v = alpha * s;
      if (g0) v -= a0[(size_t)i * lda] * b0[(size_t)j * ldb];
      out[(size_t)j * ldo + i] = v;
    }
    tile[tx][jl] = v;
  }
  __syncthreads();
  // mirror: out(j, i) = value(i, j) for i < j; thread tx runs along j
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int il = ty + 8 * u, ii = 32 * bi + il, jj = 32 * bj + tx;
    if (ii < n && jj < n && ii < jj) out[(size_t)ii * ldo + jj] = tile[il][tx];
  }
}

// ---------------------------------------------------------------------------
// D = alpha A G + beta D.  A (M x k) column-major lda; G (k x n) column-major
// ldgm; D (M x n) column-major ldd.  upper_g: G upper triangular (the k loop
// stops after the diagonal block).  With epi.twnl, D (real view of c2) +=
// twnl * ctm * (A G) on (re, im) row pairs (fnonloc).
// Launch ceil(total/8)*8 workgroups, total = tm * tn (column tile fastest).
// ---------------------------------------------------------------------------
template <int BT>
__global__ __launch_bounds__(256, 2) void k_gemm_nn(const double* __restrict__ A, int lda,
                                                    const double* __restrict__ G, int ldgm, int M, int k, int n,
                                                    double alpha, double beta, double* __restrict__ D, int ldd,
                                                    int upper_g, FnlEpi epi, int tn, int total) {
  using TL = Tile<BT>;
  constexpr int NB = TL::NB, LD = TL::LD, PER = TL::PER, WB = TL::WB;
  __shared__ double smem[2 * kBK * LD];
  double* As = smem;
  double* Gs = smem + kBK * LD;
  const int t = xcd_tile(total);
  if (t >= total) return;
  const int rt = t / tn, ct = t - rt * tn;
  const int i0 = rt * BT, j0 = ct * BT;
  const int wave = threadIdx.x >> 6;
  const int wr = (wave >> 1) * WB, wc = (wave & 1) * WB;
  d4 acc[NB][NB];
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
  const int kmax = upper_g ? min(k, j0 + BT) : k;
  const int ak = threadIdx.x >> 4, ai = (threadIdx.x & 15) * PER;       // A slice: 16 k x BT rows
  const int gj = threadIdx.x / TL::TPC, gq = (threadIdx.x % TL::TPC) * PER;   // G slice: 16 k x BT cols
  double ra[PER], rg[PER];
  auto fetch = [&](int kb) {
    const int kk = kb + ak;
    const double* src = A + (size_t)kk * lda + i0 + ai;
#pragma unroll
    for (int u = 0; u < PER; ++u) ra[u] = (kk < kmax && i0 + ai + u < M) ? src[u] : 0.0;
    const int col = j0 + gj;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int kq = kb + gq + u;
      rg[u] = (kq < kmax && col < n) ? G[(size_t)col * ldgm + kq] : 0.0;
    }
  };
  if (kmax > 0) fetch(0);
  for (int kb = 0; kb < kmax; kb += kBK) {
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      As[ak * LD + ai + u] = ra[u];
      Gs[(gq + u) * LD + gj] = rg[u];
    }
    __syncthreads();
    if (kb + kBK < kmax) fetch(kb + kBK);
    mfma_tile<BT>(As, Gs, wr, wc, acc);
    __syncthreads();
  }
  // epilogue through LDS: per wave, one WB-row x 16-column strip at a time
  // (WB*16 doubles per wave inside smem, free after the last barrier)
  const int l = threadIdx.x & 63;
  double* stage = smem + wave * (WB * 16);
  constexpr int LPC = WB / 2;          // lanes per column (one row pair each)
  constexpr int CPS = 64 / LPC;        // columns per sweep
#pragma unroll
  for (int b = 0; b < NB; ++b) {
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
      for (int r = 0; r < 4; ++r) stage[(l & 15) * WB + a * 16 + (l >> 4) + 4 * r] = acc[a][b][r];
    __syncthreads();
#pragma unroll 1
    for (int sw = 0; sw < 16 / CPS; ++sw) {
      const int c = CPS * sw + l / LPC;
      const int p = l % LPC;
      const int j = j0 + wc + 16 * b + c;
      const int i = i0 + wr + 2 * p;
      const double x0 = stage[c * WB + 2 * p], x1 = stage[c * WB + 2 * p + 1];
      if (j < n && i < M) {
        double* d = D + (size_t)j * ldd + i;
        if (epi.twnl) {
          double f = epi.f[j];
          if (f < 1e-5) f = 1.0;
          const double tw = epi.twnl[i >> 1];
          double yr, yi;
          switch (epi.lmod4) {   // ctm = -f (-i)^l ; y = tw * ctm * x
            case 0: yr = -f * tw * x0; yi = -f * tw * x1; break;
            case 1: yr = -f * tw * x1; yi = f * tw * x0; break;
            case 2: yr = f * tw * x0; yi = f * tw * x1; break;
            default: yr = f * tw * x1; yi = -f * tw * x0; break;
          }
          d[0] += yr;
          if (i + 1 < M) d[1] += yi;
        } else {
          const double y0 = (beta == 0.0 ? 0.0 : beta * d[0]) + alpha * x0;
          d[0] = y0;
          if (epi.vcm) epi.vcm[(size_t)j * ldd + i] = epi.vodt * (y0 - A[(size_t)j * lda + i]);
          if (i + 1 < M) {
            const double y1 = (beta == 0.0 ? 0.0 : beta * d[1]) + alpha * x1;
            d[1] = y1;
            if (epi.vcm) epi.vcm[(size_t)j * ldd + i + 1] = epi.vodt * (y1 - A[(size_t)j * lda + i + 1]);
          }
        }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// host wrappers
// ---------------------------------------------------------------------------

// out (m x n, ldo) = alpha A^T B [- a0 b0^T] with K rows; A may be the projector operand.
void gemm_tn(mc_ctx* ctx, const double* A, int lda, const double* B, int ldb, int K, int m, int n, double alpha,
             bool g0corr, bool symmetric, double* out, int ldo, const ProjA* proj) {
  if (m <= 0 || n <= 0) return;
  const auto al16 = [](const double* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  if (ctx->gemm_x3 && K >= kX3Min && ldb % 2 == 0 && al16(B) && (proj || (lda % 2 == 0 && al16(A)))) {
    gemm_tn_x3(ctx, A, lda, B, ldb, K, m, n, alpha, g0corr, symmetric, out, ldo, proj);
    return;
  }
  // 128-tiles for long K, 64-tiles (more workgroups) for short K
  const int BT = K >= 16384 ? 128 : 64;
  const int tm = (m + BT - 1) / BT, tn = (n + BT - 1) / BT;
  if (symmetric && m != n) throw Error{MC_ERR_INTERNAL, "symmetric A^T B with m != n"};
  const int tiles = tile_count(symmetric, tm, tn);
  // about 1024 busy workgroups (4 per CU), K chunks of at least 512 rows
  int nsplit = std::max(1, std::min((1024 + tiles - 1) / tiles, std::max(1, K / 512)));
  int kchunk = (K + nsplit - 1) / nsplit;
  kchunk = (kchunk + kBK - 1) / kBK * kBK;
  nsplit = (K + kchunk - 1) / kchunk;
  ctx->w_gemm.ensure(sizeof(double) * (size_t)nsplit * m * n);
  ProjA pa{};
  if (proj) pa = *proj;
  const int total = nsplit * tiles;
  const int launch = (total + kXcd - 1) / kXcd * kXcd;
  timer_begin(ctx, MC_KID_GEMM_TN);
  if (BT == 128)
    hipLaunchKernelGGL(k_gemm_tn<128>, dim3(launch), dim3(256), 0, ctx->stream, A, lda, B, ldb, K, m, n, kchunk,
                       ctx->w_gemm.as<double>(), symmetric ? 1 : 0, pa, tm, tn, total);
  else
    hipLaunchKernelGGL(k_gemm_tn<64>, dim3(launch), dim3(256), 0, ctx->stream, A, lda, B, ldb, K, m, n, kchunk,
                       ctx->w_gemm.as<double>(), symmetric ? 1 : 0, pa, tm, tn, total);
  MC_HIP(hipGetLastError());
  timer_end(ctx, MC_KID_GEMM_TN);
  reduce_tn(ctx, nsplit, m, n, alpha, A, lda, B, ldb, g0corr, symmetric, out, ldo);
}

void reduce_tn(mc_ctx* ctx, int nsplit, int m, int n, double alpha, const double* A, int lda, const double* B, int ldb,
               bool g0corr, bool symmetric, double* out, int ldo) {
  timer_begin(ctx, MC_KID_MISC);
  if (symmetric && m == n) {
    const int nb = (n + 31) / 32;
    hipLaunchKernelGGL(k_reduce_tn_sym, dim3(nb * (nb + 1) / 2), dim3(256), 0, ctx->stream, ctx->w_gemm.as<double>(),
                       nsplit, m, n, alpha, A, lda, B, ldb, g0corr ? 1 : 0, out, ldo);
    MC_HIP(hipGetLastError());
    timer_end(ctx, MC_KID_MISC);
    return;
  }
  hipLaunchKernelGGL(k_reduce_tn, dim3((m + 255) / 256, n), dim3(256), 0, ctx->stream, ctx->w_gemm.as<double>(),
                     nsplit, m, m, n, alpha, A, lda, B, ldb, g0corr ? 1 : 0, symmetric ? 1 : 0, out, ldo);
  MC_HIP(hipGetLastError());
  timer_end(ctx, MC_KID_MISC);
}

void reduce_tn_rows(mc_ctx* ctx, int nsplit, int ldp, int r0, int m, int n, double alpha, double* out, int ldo) {
  timer_begin(ctx, MC_KID_MISC);
  hipLaunchKernelGGL(k_reduce_tn, dim3((m + 255) / 256, n), dim3(256), 0, ctx->stream, ctx->w_gemm.as<double>() + r0,
                     nsplit, ldp, m, n, alpha, (const double*)nullptr, 0, (const double*)nullptr, 0, 0, 0, out, ldo);
  MC_HIP(hipGetLastError());
  timer_end(ctx, MC_KID_MISC);
}

bool nn_plane_form(const mc_ctx* ctx, const double* A, int lda, const double* G, int ldgm, int M, int ldd) {
  return ctx->gemm_x3 && M >= kX3Min && M % 2 == 0 && lda % 2 == 0 && ldd % 2 == 0 &&
         (reinterpret_cast<uintptr_t>(A) & 15) == 0 && ldgm % 2 == 0 && (reinterpret_cast<uintptr_t>(G) & 15) == 0;
}

void gemm_nn(mc_ctx* ctx, const double* A, int lda, const double* G, int ldgm, int M, int k, int n, double alpha,
             double beta, double* D, int ldd, bool upper_g, const FnlEpi* epi) {
  if (M <= 0 || n <= 0) return;
  if (ctx->gemm_x3 && M >= kX3Min && M % 2 == 0 && lda % 2 == 0 && ldd % 2 == 0 &&
      (reinterpret_cast<uintptr_t>(A) & 15) == 0) {
    gemm_nn_x3(ctx, A, lda, G, ldgm, M, k, n, alpha, beta, D, ldd, upper_g, epi);
    return;
  }
  MC_CHECK(!epi || (!epi->pl && !epi->colsq), MC_ERR_INTERNAL,
           "planes of D_new and column sums need the split-bf16 product");
  gemm_nn_f64(ctx, A, lda, G, ldgm, M, k, n, alpha, beta, D, ldd, upper_g, epi, MC_KID_GEMM);
}

void gemm_nn_f64(mc_ctx* ctx, const double* A, int lda, const double* G, int ldgm, int M, int k, int n, double alpha,
                 double beta, double* D, int ldd, bool upper_g, const FnlEpi* epi, int kid) {
  if (M <= 0 || n <= 0) return;
  FnlEpi e{};
  if (epi) e = *epi;
  // 128-tiles when they already give >= 1024 workgroups, 64-tiles otherwise
  const int t128 = ((M + 127) / 128) * ((n + 127) / 128);
  const int BT = t128 >= 1024 ? 128 : 64;
  const int tm = (M + BT - 1) / BT, tn = (n + BT - 1) / BT;
  const int total = tm * tn;
  const int launch = (total + kXcd - 1) / kXcd * kXcd;
  timer_begin(ctx, kid);
  if (BT == 128)
    hipLaunchKernelGGL(k_gemm_nn<128>, dim3(launch), dim3(256), 0, ctx->stream, A, lda, G, ldgm, M, k, n, alpha, beta,
                       D, ldd, upper_g ? 1 : 0, e, tn, total);
  else
    hipLaunchKernelGGL(k_gemm_nn<64>, dim3(launch), dim3(256), 0, ctx->stream, A, lda, G, ldgm, M, k, n, alpha, beta,
                       D, ldd, upper_g ? 1 : 0, e, tn, total);
  MC_HIP(hipGetLastError());
  timer_end(ctx, kid);
}

}  // namespace mc
