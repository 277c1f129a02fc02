// Split-bf16 contractions on gfx950 MFMA (v_mfma_f32_16x16x32_bf16): the
// mixed-precision variants of the tall products of gemm.hip.
//
// A G products (rotation c0 X, fnonloc): every FP64 operand value x is split
// as hi = bf16(x), lo = bf16(x - hi) (both round-to-nearest from the FP32
// value of x: |x - hi| <= 2^-8 |x|, |x - hi - lo| <= 2^-16 |x|), and a product
// is accumulated as hi*hi + hi*lo + lo*hi; the dropped lo*lo term and the
// split residuals leave at most 3 * 2^-16 of |x y|.  The three MFMAs of a 32-k
// slice start from a zero accumulator and their sum is added to the FP32
// running sum on the VALU.  Tiles are 128 x 128 per workgroup, 4 waves of
// 64 x 64 (4 x 4 MFMA tiles, 64 FP32 accumulators per lane); operand slices
// are read from HBM as FP64, split once per workgroup and staged in LDS as
// bf16 rows of 32 k values (stride 40 bf16, 80 bytes: 16-byte lane reads
// without bank conflicts); the next slice is fetched into registers while the
// current one is multiplied.  Workgroup ids are dealt to XCDs as in gemm.hip,
// so tiles that share an operand block share an L2.
//
// A^T B products (overlaps, fnl, dfnl) go through bf16 operand planes and
// k_tn_planes (gemm_planes.hip); this file keeps their dispatch and the plane
// reuse.
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
  bf8 ah[4], al[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    ah[a] = *reinterpret_cast<const bf8*>(Ahi + (wr + 16 * a) * kLd + off);
    al[a] = *reinterpret_cast<const bf8*>(Alo + (wr + 16 * a) * kLd + off);
  }
  // Each slice product starts from a zero accumulator and is added to the
  // running sum on the VALU (round to nearest): the MFMA aligns its products to
  // the exponent of the accumulator it is given and drops what falls below its
  // last bit (checked on gfx950: 32 products of 1/16 ulp each added to a large
  // C leave C unchanged), which on long sums loses ~1e-6 of them.
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const bf8 bh = *reinterpret_cast<const bf8*>(Bhi + (wc + 16 * b) * kLd + off);
    const bf8 bl = *reinterpret_cast<const bf8*>(Blo + (wc + 16 * b) * kLd + off);
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      f4 t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh, f4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
      t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl, t, 0, 0, 0);
      t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh, t, 0, 0, 0);
      acc[a][b] += t;
    }
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
          if (i + r < M) {
            const double y = (beta == 0.0 ? 0.0 : beta * d[r]) + alpha * (double)acc[a][b][r];
            d[r] = y;
            if (epi.vcm) epi.vcm[(size_t)j * ldd + i + r] = epi.vodt * (y - A[(size_t)j * lda + i + r]);
          }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// host wrappers
// ---------------------------------------------------------------------------
PlaneScope::PlaneScope(mc_ctx* ctx) : c(ctx) {
  c->planes.scoped = true;
  c->planes.src[0] = c->planes.src[1] = nullptr;
}
PlaneScope::~PlaneScope() {
  c->planes.scoped = false;
  c->planes.src[0] = c->planes.src[1] = nullptr;
}

// planes of the wavefunction operand X (K x n): reused inside a PlaneScope or
// from the kept slot, else split into a slot other than `keep` (the other
// operand's planes).  A read-only call with column reuse on splits into slot
// 0 and keeps it for later calls on the same wavefunctions.
static const DevBuf& planes_for(mc_ctx* ctx, const double* X, int ldx, int K, int n, int np, const DevBuf* keep) {
  auto& C = ctx->planes;
  for (int i = 0; i < 2; ++i)
    if (C.scoped && C.src[i] == X && C.np[i] == np && C.K[i] == K && C.n[i] >= n) return C.buf[i];
  if (keep != &C.buf[0] && kept_planes(ctx, X, K, n, np, 0)) return C.buf[0];
  const bool persist = ctx->col_reuse && in_ro_call() && keep != &C.buf[0];
  int i = persist ? 0 : (keep == &C.buf[0]) ? 1 : (keep == &C.buf[1]) ? 0 : C.next;
  C.next = 1 - i;
  if (i == 0) C.keep_src = nullptr;   // slot 0 is rewritten
  split_planes(ctx, X, ldx, K, n, np, C.buf[i]);
  if (persist) {
    C.keep_src = X;
    C.keep_np = np;
    C.keep_K = K;
    C.keep_n = n;
    C.keep_epoch = wfn_epoch();
  }
  C.src[i] = C.scoped ? X : nullptr;
  C.np[i] = np;
  C.K[i] = K;
  C.n[i] = n;
  return C.buf[i];
}

// fnl / dfnl of one projector group in mixed precision: the nv stacked
// projector variants [v0, v0 + nv) against the planes of B in one contraction,
// variant vv reduced into outs[vv] (na x n, leading dimension ldos[vv], scale
// alphas[vv])
void proj_contract_x3(mc_ctx* ctx, const ProjA& base, const double* gk, int v0, int nv, const double* B, int ldb,
                      int K, int na, int n, double* const* outs, const int* ldos, const double* alphas) {
  if (na <= 0 || n <= 0 || nv <= 0) return;
  MC_CHECK(ldb % 2 == 0 && (reinterpret_cast<uintptr_t>(B) & 15) == 0, MC_ERR_INTERNAL,
           "split-bf16 projector products need an even, 16-byte aligned wavefunction operand");
  const int np = ctx->x3_parts == 2 ? 2 : 3;
  const int mv = (na + kT - 1) / kT * kT;   // stacked rows per variant
  proj_planes_stacked(ctx, base, gk, v0, nv, K, na, np, ctx->w_plp);
  const DevBuf& Bp = planes_for(ctx, B, ldb, K, n, np, nullptr);
  const int nsplit = tn_planes_partials(ctx, ctx->w_plp, Bp, np, K, nv * mv, n, false);
  for (int vv = 0; vv < nv; ++vv) reduce_tn_rows(ctx, nsplit, nv * mv, vv * mv, na, n, alphas[vv], outs[vv], ldos[vv]);
}

// The A^T B products of the mixed-precision mode (overlaps: orthonormality;
// fnl, dfnl: E_NL and the nonlocal forces): both operands as bf16 planes
// (ctx->x3_parts parts each: 3 -> FP32-level products, 2 -> 3-term products)
// through k_tn_planes.
void gemm_tn_x3(mc_ctx* ctx, const double* A, int lda, const double* B, int ldb, int K, int m, int n, double alpha,
                bool g0corr, bool symmetric, double* out, int ldo, const ProjA* proj) {
  if (m <= 0 || n <= 0) return;
  MC_CHECK(ldb % 2 == 0 && (reinterpret_cast<uintptr_t>(B) & 15) == 0 &&
               (proj || (lda % 2 == 0 && (reinterpret_cast<uintptr_t>(A) & 15) == 0)),
           MC_ERR_INTERNAL, "split-bf16 gemm_tn needs even lda, ldb and 16-byte aligned operands");
  const int np = ctx->x3_parts == 2 ? 2 : 3;
  if (proj) {
    proj_planes(ctx, *proj, K, m, np, ctx->w_plp);
    const DevBuf& Bp = planes_for(ctx, B, ldb, K, n, np, nullptr);
    gemm_tn_planes(ctx, ctx->w_plp, Bp, np, K, m, n, alpha, nullptr, 0, B, ldb, false, false, out, ldo);
    return;
  }
  const DevBuf& Ap = planes_for(ctx, A, lda, K, m, np, nullptr);
  const DevBuf& Bp = (B == A && n == m) ? Ap : planes_for(ctx, B, ldb, K, n, np, &Ap);
  gemm_tn_planes(ctx, Ap, Bp, np, K, m, n, alpha, A, lda, B, ldb, g0corr, symmetric, out, ldo);
}

void gemm_nn_x3(mc_ctx* ctx, const double* A, int lda, const double* G, int ldgm, int M, int k, int n, double alpha,
                double beta, double* D, int ldd, bool upper_g, const FnlEpi* epi) {
  if (M <= 0 || n <= 0) return;
  MC_CHECK(M % 2 == 0 && lda % 2 == 0 && ldd % 2 == 0 && (reinterpret_cast<uintptr_t>(A) & 15) == 0, MC_ERR_INTERNAL,
           "split-bf16 gemm_nn needs even M, lda, ldd and a 16-byte aligned A");
  FnlEpi e{};
  if (epi) e = *epi;
  if (ldgm % 2 == 0 && (reinterpret_cast<uintptr_t>(G) & 15) == 0) {
    gemm_nn_planes(ctx, A, lda, G, ldgm, M, k, n, alpha, beta, D, ldd, upper_g, e);
    return;
  }
  MC_CHECK(!e.pl && !e.colsq, MC_ERR_INTERNAL, "planes of D_new and column sums need the plane form of the product");
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
