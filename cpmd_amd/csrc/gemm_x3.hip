// Three-term BF16 split contractions on gfx950 MFMA (v_mfma_f32_16x16x32_bf16)
// with FP32 accumulation: the mixed-precision variants of the tall products of
// gemm.hip (overlap, projector, rotation and fnonloc GEMMs).
//
// Every FP64 operand value x is split as hi = bf16(x), lo = bf16(x - hi) (both
// round-to-nearest from the FP32 value of x: |x - hi| <= 2^-8 |x|, |x - hi - lo|
// <= 2^-16 |x|), and a product is accumulated as hi*hi + hi*lo + lo*hi; the
// dropped lo*lo term and the split residuals leave at most 3 * 2^-16 of |x y|.
// The three MFMAs of a 32-k slice start from a zero accumulator and their sum
// is added to the FP32 running sum on the VALU; each K chunk is written as an
// FP32 partial and the partials are reduced in FP64 (k_reduce_tn of gemm.hip).
//
// Diagonal of square overlaps.  On a diagonal entry sum_k a_k b_k with
// correlated a, b (S = C^T C, or <c0|H c0>) every product loses the same-signed
// lo*lo (a 2.8e-6 relative bias), which would show up as a norm error after
// orthonormalisation.  The
// diagonal of every square wavefunction overlap is therefore summed in FP64 as
// well: the diagonal tiles accumulate sum_k a_k b_k of their columns from the
// FP64 operand registers (per K chunk, reduced in a fixed order), which leaves
// only the small, sign-random errors of the off-diagonal entries.
//
// Tiles are 128 x 128 per workgroup, 4 waves of 64 x 64 (4 x 4 MFMA tiles, 64
// FP32 accumulators per lane), K slices of 32 (one MFMA depth).  Operand
// slices are read from HBM as FP64 (in k_x3_tn 16 consecutive lanes read one
// column's 256 contiguous bytes per load), split once per workgroup and staged
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

// The same slice for operands split into NP bf16 parts (part p of A at
// A + p * kT * kLd): NP = 2 is the 3-term product above, NP = 3 (x = hi + mid +
// lo to 24 bits) sums the six products down to 2^-16 relative, mid*mid,
// lo*hi, hi*lo, mid*hi, hi*mid, hi*hi (FP32-level accuracy).
template <int NP>
__device__ __forceinline__ void xn_slice(const __bf16* As, const __bf16* Bs, int wr, int wc, f4 (&acc)[4][4]) {
  constexpr int PS = kT * kLd;   // part stride
  const int l = threadIdx.x & 63;
  const int off = (l & 15) * kLd + 8 * (l >> 4);
  bf8 a_[NP][4];
#pragma unroll
  for (int q = 0; q < NP; ++q)
#pragma unroll
    for (int a = 0; a < 4; ++a) a_[q][a] = *reinterpret_cast<const bf8*>(As + q * PS + (wr + 16 * a) * kLd + off);
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    bf8 b_[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) b_[q] = *reinterpret_cast<const bf8*>(Bs + q * PS + (wc + 16 * b) * kLd + off);
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      f4 t;
      if (NP == 2) {
        t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_[1][a], b_[0], f4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_[0][a], b_[1], t, 0, 0, 0);
      } else {
        t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_[1][a], b_[1], f4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_[NP - 1][a], b_[0], t, 0, 0, 0);
        t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_[0][a], b_[NP - 1], t, 0, 0, 0);
        t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_[1][a], b_[0], t, 0, 0, 0);
        t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_[0][a], b_[1], t, 0, 0, 0);
      }
      t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_[0][a], b_[0], t, 0, 0, 0);
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
// A^T B partials (as k_gemm_tn): launch ceil(total/8)*8 workgroups of 256,
// total = nsplit * tm * tn; A (K x m) column-major or the projector operand,
// B (K x n); P[s] (m x n) FP64 partials; upper: skip tiles below the diagonal.
// kchunk is a multiple of 32.
// ---------------------------------------------------------------------------
template <int NP>
__global__ __launch_bounds__(256, 2) void k_x3_tn(const double* __restrict__ A, int lda, const double* __restrict__ B,
                                                  int ldb, int K, int m, int n, int kchunk, double* __restrict__ P,
                                                  int upper, ProjA proj, int tm, int tn, int total,
                                                  double* __restrict__ Pd) {
  __shared__ __attribute__((aligned(16))) __bf16 sm[2 * NP * kT * kLd];
  __bf16* As = sm;                      // NP parts of the A slice
  __bf16* Bs = sm + NP * kT * kLd;      // NP parts of the B slice
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
  // Operand slices (128 columns x 32 k, FP64) are read so that 16 consecutive
  // lanes cover one column's 256 contiguous bytes per load instruction: thread
  // t owns the double pair p = t % 16 (k = 2p, 2p+1) of the 8 columns
  // u * 16 + t / 16, u = 0..7.
  const int lp = threadIdx.x & 15;          // double pair within the slice
  const int lcol = threadIdx.x >> 4;        // first owned column (0..15)
  double2 ra[8], rb[8];
  auto fetch = [&](int kb) {
    const int k0 = kb + 2 * lp;
    const bool kin = k0 + 2 <= kend;        // kend even (K = 2 ngw rows), k0 even
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int acol = i0 + u * 16 + lcol, bcol = j0 + u * 16 + lcol;
      if (proj.eigr) {
        const int g = k0 >> 1;               // one G vector: (re, im) rows
        double re = 0.0, im = 0.0;
        if (acol < m && kin) {
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
        ra[u] = make_double2(re, im);
      } else {
        ra[u] = (acol < m && kin) ? *reinterpret_cast<const double2*>(A + (size_t)acol * lda + k0)
                                  : make_double2(0.0, 0.0);
      }
      rb[u] = (bcol < n && kin) ? *reinterpret_cast<const double2*>(B + (size_t)bcol * ldb + k0)
                                : make_double2(0.0, 0.0);
    }
  };
  auto put = [&](__bf16* dst, const double2 (&r)[8]) {
    typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int row = u * 16 + lcol;
      float f0 = (float)r[u].x, f1 = (float)r[u].y;
#pragma unroll
      for (int q = 0; q < NP; ++q) {       // x = part0 + part1 (+ part2), each bf16
        bf2 h;
        h[0] = (__bf16)f0;
        h[1] = (__bf16)f1;
        *reinterpret_cast<bf2*>(dst + q * kT * kLd + row * kLd + 2 * lp) = h;
        f0 -= (float)h[0];
        f1 -= (float)h[1];
      }
    }
  };
  // diagonal tiles of a square overlap also sum a_k b_k of their columns in
  // FP64: per slice the 16 lanes of each owned column reduce their products and
  // lane lp = u keeps the running sum of column u * 16 + lcol
  const bool dtile = Pd != nullptr && ti == tj;
  double dsum = 0.0;
  if (kbeg < kend) fetch(kbeg);
  for (int kb = kbeg; kb < kend; kb += kK) {
    if (dtile) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        double v = ra[u].x * rb[u].x + ra[u].y * rb[u].y;
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) v += __shfl_xor(v, off, 64);
        if (lp == u) dsum += v;
      }
    }
    put(As, ra);
    put(Bs, rb);
    __syncthreads();
    if (kb + kK < kend) fetch(kb + kK);
    xn_slice<NP>(As, Bs, wr, wc, acc);
    __syncthreads();
  }
  if (dtile && lp < 8) {
    const int col = i0 + lp * 16 + lcol;
    if (col < n) Pd[(size_t)s * n + col] = dsum;
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

// out(i, i) = alpha sum_s Pd[s][i] [- A(0, i) B(0, i)]: the FP64 diagonal of a
// square overlap from the per-chunk sums of the diagonal tiles, fixed order
__global__ void k_reduce_diag(const double* __restrict__ Pd, int nsplit, int n, double alpha,
                              const double* __restrict__ A, int lda, const double* __restrict__ B, int ldb,
                              int g0corr, double* __restrict__ out, int ldo) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = 0.0;
  for (int q = 0; q < nsplit; ++q) v += Pd[(size_t)q * n + i];
  v *= alpha;
  if (g0corr) v -= A[(size_t)lda * i] * B[(size_t)ldb * i];
  out[(size_t)ldo * i + i] = v;
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
  // about 1024 busy workgroups, K chunks of at least 1024 rows.  Projector
  // contractions (fnl, dfnl: no FP64 diagonal) always use 1024-row chunks: the
  // FP32 accumulation bias grows with the chunk length (a few 1e-7 at 3000
  // rows), and their partials are small (nat x n)
  int nsplit = std::max(1, std::min((1024 + tiles - 1) / tiles, std::max(1, K / 1024)));
  if (proj) nsplit = std::max(1, K / 1024);
  int kchunk = (K + nsplit - 1) / nsplit;
  kchunk = (kchunk + kK - 1) / kK * kK;
  nsplit = (K + kchunk - 1) / kchunk;
  const bool diag = !proj && m == n;
  ctx->w_gemm.ensure(sizeof(double) * ((size_t)nsplit * m * n + (diag ? (size_t)nsplit * n : 0)));
  double* Pd = diag ? ctx->w_gemm.as<double>() + (size_t)nsplit * m * n : nullptr;
  ProjA pa{};
  if (proj) pa = *proj;
  const int total = nsplit * tm * tn;
  const int launch = (total + kXcd - 1) / kXcd * kXcd;
  timer_begin(ctx, MC_KID_GEMM);
  // the A^T B products (overlaps: orthonormality; fnl, dfnl: E_NL and the
  // nonlocal forces) split their operands in three parts (FP32-level
  // products); ctx->x3_parts = 2 selects the 3-term split
  if (ctx->x3_parts == 2)
    hipLaunchKernelGGL(k_x3_tn<2>, dim3(launch), dim3(256), 0, ctx->stream, A, lda, B, ldb, K, m, n, kchunk,
                       ctx->w_gemm.as<double>(), symmetric ? 1 : 0, pa, tm, tn, total, Pd);
  else
    hipLaunchKernelGGL(k_x3_tn<3>, dim3(launch), dim3(256), 0, ctx->stream, A, lda, B, ldb, K, m, n, kchunk,
                       ctx->w_gemm.as<double>(), symmetric ? 1 : 0, pa, tm, tn, total, Pd);
  MC_HIP(hipGetLastError());
  timer_end(ctx, MC_KID_GEMM);
  reduce_tn(ctx, nsplit, m, n, alpha, A, lda, B, ldb, g0corr, symmetric, out, ldo);
  if (diag) {
    timer_begin(ctx, MC_KID_MISC);
    hipLaunchKernelGGL(k_reduce_diag, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, Pd, nsplit, n, alpha, A, lda,
                       B, ldb, g0corr ? 1 : 0, out, ldo);
    MC_HIP(hipGetLastError());
    timer_end(ctx, MC_KID_MISC);
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
