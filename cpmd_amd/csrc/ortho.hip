// Orthonormalisation, overlap, rotation, Lagrange projection, the CP
// constraint and the nonlocal projector contractions on gfx950.
//
//   ovlap     src/ovlap_utils.mod.F90:57-190       a = 2 Re(C1^H C2) - c1_0 c2_0
//   rgs       src/rgs_utils.mod.F90:31-162         S = C^T C, S = U^T U, C <- C U^-1, zclean
//   lowdin    src/lowdin_utils.mod.F90:27-192      C <- C S^-1/2, G=0 imaginary part cleaned
//   rotate    src/rotate_utils.mod.F90:63-100      c2 = a c1 gam + b c2
//   lagrange  src/forces_driver.mod.F90:344-370    gam = ovlap(c2,c0) f_i f_j, c2 -= c0 gam, zclean
//   rortog    src/rortog_utils.mod.F90:315-498     CP constraint fixed point, cm += c0 gam
//   rnlsm1    src/rnlsm1_utils.mod.F90:42-193      fnl = 2 Re <(-i)^l twnl eigr | c>
//   fnonloc   src/fnonloc_utils.mod.F90:209-303    c2 += -f (-i)^l twnl sum_ia eigr dd
//
// Dense n x n factorisations stay in FP64 on the GPU:
//  * Cholesky (upper, LAPACK dpotrf semantics) is a left-looking column sweep
//    in one workgroup; the triangular inverse (dtrtri) solves the columns of
//    U X = I in parallel, one wavefront per column;
//  * Loewdin's S^-1/2 is obtained with the coupled Newton-Schulz iteration
//    (Higham, "Functions of Matrices", 2008, eq. 6.35) on MFMA, which converges
//    to the same unique SPD inverse square root that CPMD assembles from the
//    eigenpairs of dspev.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "mc_internal.h"

namespace mc {

// ---------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------

// Cholesky S = U^T U of the upper triangle of S (n x n, column-major lds) in
// place, blocked right-looking with 64-wide blocks; the lower triangle is set
// to zero.  info (device) = 0 or the order of the first non-positive leading
// minor (dpotrf); once set, the later launches return at once.
constexpr int kPb = 64;

// factor the diagonal block S(kb:kb+nb, kb:kb+nb) in LDS, right-looking
__global__ __launch_bounds__(256) void k_potrf_diag(double* __restrict__ S, int lds, int kb, int nb, int* info) {
  __shared__ double T[kPb][kPb + 1];
  __shared__ int bad;
  if (kb > 0 && *info) return;
  const int t = threadIdx.x;
  for (int e = t; e < nb * nb; e += 256) {
    const int r = e % nb, c = e / nb;
    T[r][c] = S[(kb + r) + (size_t)lds * (kb + c)];
  }
  if (t == 0) bad = 0;
  __syncthreads();
  for (int j = 0; j < nb; ++j) {
    if (t == 0) {
      const double d = T[j][j];
      if (!(d > 0.0)) bad = kb + j + 1;
      else T[j][j] = sqrt(d);
    }
    __syncthreads();
    if (bad) break;
    const double dj = T[j][j];
    for (int c = j + 1 + t; c < nb; c += 256) T[j][c] /= dj;
    __syncthreads();
    const int m = nb - j - 1;   // trailing (r, c), j < r <= c < nb
    for (int e = t; e < m * m; e += 256) {
      const int r = j + 1 + e % m, c = j + 1 + e / m;
      if (r <= c) T[r][c] -= T[j][r] * T[j][c];
    }
    __syncthreads();
  }
  for (int e = t; e < nb * nb; e += 256) {
    const int r = e % nb, c = e / nb;
    if (r <= c) S[(kb + r) + (size_t)lds * (kb + c)] = T[r][c];
  }
  if (t == 0 && (bad || kb == 0)) *info = bad;
}

// block row: U(kb:kb+nb, j) = U_kk^-T S(kb:kb+nb, j) for kb+nb <= j < n; one
// thread per column, forward substitution with the column staged in LDS
__global__ __launch_bounds__(64) void k_potrf_trsm(double* __restrict__ S, int lds, int n, int kb, int nb,
                                                   const int* info) {
  __shared__ double U[kPb][kPb + 1];
  __shared__ double y[kPb][kPb + 1];
  if (*info) return;
  const int t = threadIdx.x;
  for (int e = t; e < nb * nb; e += 64) {
    const int r = e % nb, c = e / nb;
    U[r][c] = S[(kb + r) + (size_t)lds * (kb + c)];
  }
  const int j = kb + nb + blockIdx.x * 64 + t;
  if (j < n)
    for (int r = 0; r < nb; ++r) y[r][t] = S[(kb + r) + (size_t)lds * j];
  __syncthreads();
  if (j >= n) return;
  for (int r = 0; r < nb; ++r) {
    double v = y[r][t];
    for (int k = 0; k < r; ++k) v -= U[k][r] * y[k][t];
    y[r][t] = v / U[r][r];
  }
  for (int r = 0; r < nb; ++r) S[(kb + r) + (size_t)lds * j] = y[r][t];
}

// trailing update S(i, j) -= sum_k U(k, i) U(k, j), k in the block row,
// kb+nb <= i <= j < n, 32 x 32 tiles (upper tiles only)
__global__ __launch_bounds__(256) void k_potrf_update(double* __restrict__ S, int lds, int n, int kb, int nb,
                                                      const int* info) {
  __shared__ double A[kPb][33];
  __shared__ double B[kPb][33];
  if (*info) return;
  const int ti = blockIdx.x, tj = blockIdx.y;
  if (ti > tj) return;
  const int i0 = kb + nb + 32 * ti, j0 = kb + nb + 32 * tj;
  const int t = threadIdx.x;
  for (int e = t; e < nb * 32; e += 256) {
    const int k = e % nb, c = e / nb;
    A[k][c] = i0 + c < n ? S[(kb + k) + (size_t)lds * (i0 + c)] : 0.0;
    B[k][c] = j0 + c < n ? S[(kb + k) + (size_t)lds * (j0 + c)] : 0.0;
  }
  __syncthreads();
  for (int e = t; e < 32 * 32; e += 256) {
    const int r = e % 32, c = e / 32;
    const int i = i0 + r, j = j0 + c;
    if (i >= n || j >= n || i > j) continue;
    double s = 0.0;
    for (int k = 0; k < nb; ++k) s += A[k][r] * B[k][c];
    S[i + (size_t)lds * j] -= s;
  }
}

__global__ void k_zero_lower(double* __restrict__ S, int n, int lds) {
  const int j = blockIdx.x;
  for (int i = j + 1 + threadIdx.x; i < n; i += blockDim.x) S[i + (size_t)lds * j] = 0.0;
}

static void potrf_upper_dev(mc_ctx* ctx, double* S, int n, int lds, int* info) {
  for (int kb = 0; kb < n; kb += kPb) {
    const int nb = std::min(kPb, n - kb), rest = n - kb - nb;
    hipLaunchKernelGGL(k_potrf_diag, dim3(1), dim3(256), 0, ctx->stream, S, lds, kb, nb, info);
    MC_HIP(hipGetLastError());
    if (rest <= 0) break;
    hipLaunchKernelGGL(k_potrf_trsm, dim3((rest + 63) / 64), dim3(64), 0, ctx->stream, S, lds, n, kb, nb, info);
    MC_HIP(hipGetLastError());
    const int nt = (rest + 31) / 32;
    hipLaunchKernelGGL(k_potrf_update, dim3(nt, nt), dim3(256), 0, ctx->stream, S, lds, n, kb, nb, info);
    MC_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_zero_lower, dim3(n), dim3(256), 0, ctx->stream, S, n, lds);
  MC_HIP(hipGetLastError());
}

// X = U^{-1} (upper): one 64-thread workgroup per column j solves U x = e_j
// by back substitution with x staged in LDS (n doubles).
__global__ __launch_bounds__(64) void k_trtri_upper(const double* __restrict__ U, int n, int ldu,
                                                    double* __restrict__ X, int ldx) {
  extern __shared__ double xs[];
  const int j = blockIdx.x;
  const int l = threadIdx.x;
  for (int i = l; i < n; i += 64) xs[i] = 0.0;
  __syncthreads();
  if (l == 0) xs[j] = 1.0 / U[j + (size_t)ldu * j];
  __syncthreads();
  for (int i = j - 1; i >= 0; --i) {
    double s = 0.0;
    for (int k = i + 1 + l; k <= j; k += 64) s += U[i + (size_t)ldu * k] * xs[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (l == 0) xs[i] = -s / U[i + (size_t)ldu * i];
    __syncthreads();
  }
  double* x = X + (size_t)ldx * j;
  for (int i = l; i < n; i += 64) x[i] = xs[i];
}

// a(i,j) *= f(i) f(j) with f = (occupation != 0) (gscal, forces_driver.mod.F90:502-518)
__global__ void k_gscal(double* __restrict__ a, int n, int lda, const double* __restrict__ focc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= n) return;
  const double fi = focc[i] != 0.0 ? 1.0 : 0.0, fj = focc[j] != 0.0 ? 1.0 : 0.0;
  a[i + (size_t)lda * j] *= fi * fj;
}

// Im c(G=0, i) = 0 (zclean, src/utils.mod.F90:116-134); G=0 is row 0 of the device layout
__global__ void k_zclean(double2* __restrict__ c, int ldg, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c[(size_t)ldg * i].y = 0.0;
}

// out = alpha * a + beta * I (+ gamma * b), n x n
__global__ void k_axpby_eye(const double* __restrict__ a, double alpha, double beta, const double* __restrict__ b,
                            double gamma, double* __restrict__ out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= n) return;
  const size_t e = i + (size_t)n * j;
  double v = alpha * a[e] + (i == j ? beta : 0.0);
  if (b) v += gamma * b[e];
  out[e] = v;
}

// max |a - b| over n*n into partial[blockIdx.x] (then reduced on the host)
__global__ void k_maxdiff(const double* __restrict__ a, const double* __restrict__ b, size_t nn,
                          double* __restrict__ part) {
  double m = 0.0;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nn; e += (size_t)gridDim.x * blockDim.x) {
    const double d = fabs(a[e] - (b ? b[e] : 0.0));
    m = d > m ? d : m;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_down(m, off, 64);
    m = o > m ? o : m;
  }
  __shared__ double w[4];
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) r = w[q] > r ? w[q] : r;
    part[blockIdx.x] = r;
  }
}

// symma: a(i,j) = a(j,i) = (a(i,j) + a(j,i))/2 for i < j (src/utils.mod.F90:138-153)
__global__ void k_symma(double* __restrict__ a, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= j || j >= n) return;
  const double v = 0.5 * (a[i + (size_t)n * j] + a[j + (size_t)n * i]);
  a[i + (size_t)n * j] = v;
  a[j + (size_t)n * i] = v;
}

// dd(ia, i) = sum_jv h(iv, jv) fnl(isa0+ia, jv, i)   (fnonloc Goedecker branch :219-256)
__global__ void k_fnl_dd(const double* __restrict__ fnl, int nat, int nhxs, int isa0, int na, int ngh,
                         const double* __restrict__ h, int iv, int nstate, double* __restrict__ dd) {
  const int ia = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  if (ia >= na || i >= nstate) return;
  double s = 0.0;
  for (int jv = 0; jv < ngh; ++jv) {
    const double hv = h[iv + nhxs * jv];
    if (hv != 0.0) s += fnl[(isa0 + ia) + (size_t)nat * (jv + (size_t)nhxs * i)] * hv;
  }
  dd[ia + (size_t)na * i] = s;
}

__global__ void k_permute_rows_in(const double2* __restrict__ src, int lds, double2* __restrict__ dst, int ldd,
                                  const int* __restrict__ perm, int ngw, int ncol) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
  if (j >= ldd || c >= ncol) return;
  dst[(size_t)ldd * c + j] = j < ngw ? src[(size_t)lds * c + perm[j]] : make_double2(0.0, 0.0);
}
__global__ void k_permute_real_in(const double* __restrict__ src, int lds, double* __restrict__ dst, int ldd,
                                  const int* __restrict__ perm, int ngw, int ncol) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
  if (j >= ldd || c >= ncol) return;
  dst[(size_t)ldd * c + j] = j < ngw ? src[(size_t)lds * c + perm[j]] : 0.0;
}

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
static inline int wfn_rows(const mc_ctx* ctx) { return 2 * ctx->ngw; }   // K of the real view
static inline int wfn_ld(const mc_ctx* ctx) { return 2 * ctx->ldg; }     // real leading dim

static double* small(mc_ctx* ctx, int which, int n) {
  DevBuf& b = ctx->w_small[which];
  b.ensure(sizeof(double) * (size_t)n * n);
  return b.as<double>();
}

static double max_abs_diff(mc_ctx* ctx, const double* a, const double* b, size_t nn) {
  const int nb = 64;
  ctx->w_red2.ensure(sizeof(double) * nb);
  hipLaunchKernelGGL(k_maxdiff, dim3(nb), dim3(256), 0, ctx->stream, a, b, nn, ctx->w_red2.as<double>());
  MC_HIP(hipGetLastError());
  double h[nb];
  dl_add(ctx, h, ctx->w_red2.p, sizeof(h));
  dl_finish(ctx);
  double m = 0.0;
  for (int q = 0; q < nb; ++q) m = std::max(m, h[q]);
  return m;
}

// rnlfor (src/rnlfor_utils.mod.F90:201-280, Gamma): one thread per (k, isa);
// fion(k, isa) = -sum_i 2 f_i sum_{iv,jv} h(iv,jv) dfnl(isa,iv,k,i) fnl(isa,jv,i),
// states with |f_i| <= 1e-12 skipped.  out (3, nat).
// one workgroup per (atom, direction): threads over states, fixed-order sum
__global__ __launch_bounds__(256) void k_rnlfor(const double* __restrict__ fnl, const double* __restrict__ dfnl,
                                                const double* __restrict__ f, int n, int nat, int nhxs,
                                                const int* __restrict__ sp_of, const int* __restrict__ ngh,
                                                const double* __restrict__ h, double* __restrict__ out) {
  __shared__ double red[256];
  const int isa = blockIdx.x, k = blockIdx.y;
  const int is = sp_of[isa];
  const int nh = ngh[is];
  const double* hs = h + (size_t)nhxs * nhxs * is;
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const double fi = f[i];
    if (fabs(fi) <= 1e-12) continue;
    double s = 0.0;
    for (int iv = 0; iv < nh; ++iv) {
      const double d = dfnl[isa + (size_t)nat * (iv + (size_t)nhxs * (k + 3 * (size_t)i))];
      double t = 0.0;
      for (int jv = 0; jv < nh; ++jv) t += hs[iv + nhxs * jv] * fnl[isa + (size_t)nat * (jv + (size_t)nhxs * i)];
      s += d * t;
    }
    acc += 2.0 * fi * s;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[k + 3 * (size_t)isa] = -red[0];
}

// E_NL of atom isa = sum_i f_i sum_{iv,jv} h(iv,jv) fnl(isa,iv,i) fnl(isa,jv,i)
// (rnlrh / rnlrhg, src/rnlrh_utils.mod.F90:43-200 and :271-343 with imagp = 1:
// the diagonal wsg terms and the symmetric hlsg pairs, iv != jv counted twice,
// are the full double sum over a symmetric h); one workgroup per atom, a fixed
// reduction tree
__global__ __launch_bounds__(256) void k_rnlrh(const double* __restrict__ fnl, const double* __restrict__ f, int n,
                                               int nat, int nhxs, const int* __restrict__ sp_of,
                                               const int* __restrict__ ngh, const double* __restrict__ h,
                                               double* __restrict__ out) {
  __shared__ double red[256];
  const int isa = blockIdx.x;
  const int is = sp_of[isa];
  const int nh = ngh[is];
  const double* hs = h + (size_t)nhxs * nhxs * is;
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const double* x = fnl + isa + (size_t)nat * nhxs * i;
    double s = 0.0;
    for (int iv = 0; iv < nh; ++iv) {
      double t = 0.0;
      for (int jv = 0; jv < nh; ++jv) t += hs[iv + nhxs * jv] * x[(size_t)nat * jv];
      s += x[(size_t)nat * iv] * t;
    }
    acc += f[i] * s;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[isa] = red[0];
}

// ---------------------------------------------------------------------------
// drivers (device-resident wavefunctions)
// ---------------------------------------------------------------------------
void ovlap_dev(mc_ctx* ctx, int n, double* a_dev, int lda, const double2* c1, const double2* c2) {
  gemm_tn(ctx, (const double*)c1, wfn_ld(ctx), (const double*)c2, wfn_ld(ctx), wfn_rows(ctx), n, n, 2.0,
          ctx->geq0 != 0, c1 == c2, a_dev, lda, nullptr);
}

// C <- C X with X (n x n) on the device; through the workspace copy of C.  With
// split-bf16 products C X is formed as C + C (X - I), so that the rounding of the
// product scales with |X - I| (small near orthonormality), not with |X|.
static void apply_right(mc_ctx* ctx, double2* c, const double* X, int n, bool upper) {
  const size_t bytes = sizeof(double2) * (size_t)ctx->ldg * n;
  ctx->w_wfn.ensure(bytes);
  MC_HIP(hipMemcpyAsync(ctx->w_wfn.p, c, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  if (ctx->gemm_x3) {
    double* XmI = small(ctx, 6, n);
    hipLaunchKernelGGL(k_axpby_eye, dim3((n + 255) / 256, n), dim3(256), 0, ctx->stream, X, 1.0, -1.0,
                       (const double*)nullptr, 0.0, XmI, n);
    MC_HIP(hipGetLastError());
    gemm_nn(ctx, ctx->w_wfn.as<double>(), wfn_ld(ctx), XmI, n, wfn_ld(ctx), n, n, 1.0, 1.0, (double*)c, wfn_ld(ctx),
            upper, nullptr);
    return;
  }
  gemm_nn(ctx, ctx->w_wfn.as<double>(), wfn_ld(ctx), X, n, wfn_ld(ctx), n, n, 1.0, 0.0, (double*)c, wfn_ld(ctx),
          upper, nullptr);
}

static void zclean(mc_ctx* ctx, double2* c, int n) {
  if (!ctx->geq0 || n <= 0) return;
  hipLaunchKernelGGL(k_zclean, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, c, ctx->ldg, n);
  MC_HIP(hipGetLastError());
}

void rgs_dev(mc_ctx* ctx, double2* c, int n, double* smat_host) {
  if (n <= 0) return;
  MC_CHECK(n <= 4096, MC_ERR_UNSUPPORTED, "rgs: nstate > 4096");
  double* S = small(ctx, 0, n);
  double* X = small(ctx, 1, n);
  ovlap_dev(ctx, n, S, n, c, c);
  ctx->w_info.ensure(sizeof(int));
  potrf_upper_dev(ctx, S, n, n, ctx->w_info.as<int>());
  int info = 0;
  MC_HIP(hipMemcpyAsync(&info, ctx->w_info.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  MC_HIP(hipStreamSynchronize(ctx->stream));
  MC_CHECK(info == 0, MC_ERR_NOTPOSDEF,
           "UINV| the leading minor of order " + std::to_string(info) + " is not positive definite");
  hipLaunchKernelGGL(k_trtri_upper, dim3(n), dim3(64), sizeof(double) * n, ctx->stream, S, n, n, X, n);
  MC_HIP(hipGetLastError());
  apply_right(ctx, c, X, n, true);
  zclean(ctx, c, n);
  if (smat_host) {
    MC_HIP(hipMemcpyAsync(smat_host, X, sizeof(double) * n * n, hipMemcpyDeviceToHost, ctx->stream));
    MC_HIP(hipStreamSynchronize(ctx->stream));
  }
}

// Z = S^{-1/2} by the coupled Newton-Schulz iteration on S/c, c = ||S||_F:
// T = (3I - Z Y)/2, Y <- Y T, Z <- T Z; Y -> (S/c)^{1/2}, Z -> (S/c)^{-1/2}.
static void inv_sqrt(mc_ctx* ctx, const double* S, int n, double* Zout, int* iters) {
  double* Y = small(ctx, 2, n);
  double* Z = small(ctx, 3, n);
  double* T = small(ctx, 4, n);
  double* W = small(ctx, 5, n);
  std::vector<double> hs((size_t)n * n);
  MC_HIP(hipMemcpyAsync(hs.data(), S, sizeof(double) * n * n, hipMemcpyDeviceToHost, ctx->stream));
  MC_HIP(hipStreamSynchronize(ctx->stream));
  double fro = 0.0;
  for (double v : hs) fro += v * v;
  fro = std::sqrt(fro);
  MC_CHECK(fro > 0.0 && std::isfinite(fro), MC_ERR_ARG, "lowdin: zero or non-finite overlap matrix");
  // scale so that ||I - S/c||_2 < 1 when S is close to I: use c = max(1, ||S||_2 bound)
  double cmax = 0.0;
  for (int j = 0; j < n; ++j) {
    double r = 0.0;
    for (int i = 0; i < n; ++i) r += std::fabs(hs[i + (size_t)n * j]);
    cmax = std::max(cmax, r);
  }
  const double c = cmax > 1.0 ? cmax : 1.0;   // ||S||_2 <= ||S||_1 (symmetric)
  dim3 g2((n + 255) / 256, n);
  hipLaunchKernelGGL(k_axpby_eye, g2, dim3(256), 0, ctx->stream, S, 1.0 / c, 0.0, (const double*)nullptr, 0.0, Y, n);
  hipLaunchKernelGGL(k_axpby_eye, g2, dim3(256), 0, ctx->stream, S, 0.0, 1.0, (const double*)nullptr, 0.0, Z, n);
  MC_HIP(hipGetLastError());
  int it = 0;
  const int maxit = 100;
  bool done = false;
  for (it = 1; it <= maxit; ++it) {
    gemm_nn(ctx, Z, n, Y, n, n, n, n, 1.0, 0.0, W, n, false, nullptr);                 // W = Z Y
    hipLaunchKernelGGL(k_axpby_eye, g2, dim3(256), 0, ctx->stream, W, -0.5, 1.5, (const double*)nullptr, 0.0, T, n);
    MC_HIP(hipGetLastError());
    gemm_nn(ctx, Y, n, T, n, n, n, n, 1.0, 0.0, W, n, false, nullptr);                 // W = Y T
    MC_HIP(hipMemcpyAsync(Y, W, sizeof(double) * n * n, hipMemcpyDeviceToDevice, ctx->stream));
    gemm_nn(ctx, T, n, Z, n, n, n, n, 1.0, 0.0, W, n, false, nullptr);                 // W = T Z
    MC_HIP(hipMemcpyAsync(Z, W, sizeof(double) * n * n, hipMemcpyDeviceToDevice, ctx->stream));
    if (done) break;
    // ||T - I||: converged once the correction is at rounding level; one more sweep then
    hipLaunchKernelGGL(k_axpby_eye, g2, dim3(256), 0, ctx->stream, T, 1.0, -1.0, (const double*)nullptr, 0.0, W, n);
    const double dev = max_abs_diff(ctx, W, nullptr, (size_t)n * n);
    MC_CHECK(std::isfinite(dev), MC_ERR_NOCONV, "lowdin: Newton-Schulz iteration diverged");
    if (dev < 1e-15 * n) done = true;
  }
  MC_CHECK(done, MC_ERR_NOCONV, "lowdin: Newton-Schulz iteration did not converge in 100 steps");
  hipLaunchKernelGGL(k_axpby_eye, g2, dim3(256), 0, ctx->stream, Z, 1.0 / std::sqrt(c), 0.0, (const double*)nullptr,
                     0.0, Zout, n);
  MC_HIP(hipGetLastError());
  if (iters) *iters = it;
}

void lowdin_dev(mc_ctx* ctx, double2* c, int n, int* iters) {
  if (n <= 0) return;
  MC_CHECK(n < kX3Min, MC_ERR_UNSUPPORTED, "lowdin: nstate beyond 16383");
  double* S = small(ctx, 0, n);
  double* A = small(ctx, 1, n);
  ovlap_dev(ctx, n, S, n, c, c);
  inv_sqrt(ctx, S, n, A, iters);
  apply_right(ctx, c, A, n, false);
  zclean(ctx, c, n);
}

void rotate_dev(mc_ctx* ctx, double a, const double2* c1, double b, double2* c2, const double* gam_dev, int n) {
  MC_CHECK(c1 != c2, MC_ERR_ARG, "rotate: c1 and c2 must not alias");
  gemm_nn(ctx, (const double*)c1, wfn_ld(ctx), gam_dev, n, wfn_ld(ctx), n, n, a, b, (double*)c2, wfn_ld(ctx), false,
          nullptr);
}

void lagrange_dev(mc_ctx* ctx, const double2* c0, double2* c2, const double* f_host, int n) {
  if (n <= 0) return;
  double* G = small(ctx, 0, n);
  ovlap_dev(ctx, n, G, n, c2, c0);
  ctx->w_f2.ensure(sizeof(double) * n);
  MC_HIP(hipMemcpyAsync(ctx->w_f2.p, f_host, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_gscal, dim3((n + 255) / 256, n), dim3(256), 0, ctx->stream, G, n, n, ctx->w_f2.as<double>());
  MC_HIP(hipGetLastError());
  rotate_dev(ctx, -1.0, c0, 1.0, c2, G, n);
  zclean(ctx, c2, n);
}

// CP orthonormality constraint, rortho (src/rortog_utils.mod.F90:315-498).
static void rortog_core(mc_ctx* ctx, const double2* c0, double2* cm, double* gam_host, int tnon, int n,
                        double epsog, int maxit, int* iters, double2* vel = nullptr, double vdt = 0.0,
                        __bf16* planes_out = nullptr);

void rortog_dev(mc_ctx* ctx, const double2* c0, double2* cm, double* gam_host, int tnon, int n, double epsog,
                int maxit, int* iters) {
  if (n <= 0) return;
  PlaneScope keep(ctx);   // c0 and cm are read-only until the final rotation
  rortog_core(ctx, c0, cm, gam_host, tnon, n, epsog, maxit, iters);
}

// posupa (src/posupa_utils.mod.F90:41-160) with velupa's kick folded in:
// cm += dtm c2 (c2 may be null), cp = c0 + dt cm, rortog(c0, cp) (cp += c0
// gamx), cm = (cp - c0) / dt.  In mixed precision the first pass also builds
// the bf16 planes of c0 and cp for the two overlaps of rortog and leaves cm
// alone (the kicked velocity is only an intermediate); the velocity is written
// by the epilogue of the final rotation.
void posupa_dev(mc_ctx* ctx, const double2* c0, const double2* c2, double2* cm, double2* cp, int n, double dtm,
                double dt, double epsog, int maxit, double* gam_host, int* iters) {
  if (n <= 0) return;
  PlaneScope keep(ctx);
  const int K = wfn_rows(ctx);
  const bool planes = ctx->gemm_x3 && K >= kX3Min;
  if (planes)
    posupa_planes(ctx, (const double*)c0, (const double*)c2, (const double*)cm, (double*)cp, wfn_ld(ctx), K, n, dtm,
                  dt);
  else
    cp_posupa_dev(ctx, c0, c2, cm, cp, n, dtm, dt);
  // With column reuse the final rotation also writes the planes of the new
  // coefficients into slot 0 (c0's planes are not needed after it), kept for
  // rnlsm and the rest of the next step on them; otherwise c0's planes stay
  // kept (c0 is unchanged: cp and cm were written).
  const bool alias = c0 == cp || c0 == cm;
  __bf16* pl = planes && ctx->col_reuse && !alias && n % 2 == 0 ? ctx->planes.buf[0].as<__bf16>() : nullptr;
  rortog_core(ctx, c0, cp, gam_host, 0, n, epsog, maxit, iters, cm, dt, pl);
  if (pl)
    keep_new_planes(ctx, cp, K, n);
  else if (planes && !alias)
    keep_c0_planes(ctx, c0, K, n);
}

// gamy = -sym(a): gamy(i,j) = gamy(j,i) = -(a(i,j) + a(j,i))/2 (rortv's
// symmetrisation, src/rortv_utils.mod.F90:151-158)
__global__ void k_neg_symm(double* __restrict__ a, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i > j || j >= n) return;
  const double v = -0.5 * (a[i + (size_t)n * j] + a[j + (size_t)n * i]);
  a[i + (size_t)n * j] = v;
  a[j + (size_t)n * i] = v;
}

// rortv (src/rortv_utils.mod.F90:39-166, overlap constraints, no tharm /
// tmass): gamy = ovlap(cm, c0), gamy <- -sym(gamy), cm += c0 gamy, so that
// <c0|cm> + <cm|c0> = 0 for orthonormal c0 (the velocity constraint of the
// velocity Verlet CP step).  With c2, velupa's kick cm += dtm c2 comes
// first; in mixed precision it is one pass with the bf16 split of c0 and of
// the kicked cm for the overlap.
bool rortv_dev(mc_ctx* ctx, const double2* c0, double2* cm, double* gamy_host, int n, const double2* c2,
               double dtm, double* dotp_dev) {
  if (n <= 0) return false;
  MC_CHECK(n < kX3Min, MC_ERR_UNSUPPORTED, "rortv: nstate beyond 16383");
  PlaneScope keep(ctx);   // c0 and cm are read-only between the kick and the rotation
  const int K = wfn_rows(ctx);
  if (c2) {
    if (ctx->gemm_x3 && K >= kX3Min) {
      kick_planes(ctx, (const double*)c0, (const double*)c2, (double*)cm, wfn_ld(ctx), K, n, dtm);
    } else {
      timer_begin(ctx, MC_KID_MISC);
      hipLaunchKernelGGL(k_axpy, dim3(2048), dim3(256), 0, ctx->stream, dtm, (const double*)c2, (double*)cm,
                         (size_t)ctx->ldg * n * 2);
      MC_HIP(hipGetLastError());
      timer_end(ctx, MC_KID_MISC);
    }
  }
  double* G = small(ctx, 7, n);
  ovlap_dev(ctx, n, G, n, cm, c0);
  hipLaunchKernelGGL(k_neg_symm, dim3((n + 255) / 256, n), dim3(256), 0, ctx->stream, G, n);
  MC_HIP(hipGetLastError());
  const int ld = wfn_ld(ctx);
  bool fused = false;
  if (dotp_dev && c0 != cm && nn_plane_form(ctx, (const double*)c0, ld, G, n, ld, ld)) {
    // cm += c0 gamy with rekine's column norms of the result in its epilogue
    const int nrb = (ld + 127) / 128 * (128 / kColsqRows);
    ctx->w_colsq.ensure(sizeof(double) * (size_t)nrb * n);
    FnlEpi e{};
    e.colsq = ctx->w_colsq.as<double>();
    e.sq_g0 = ctx->geq0;
    e.sq_nrb = nrb;
    gemm_nn(ctx, (const double*)c0, ld, G, n, ld, n, n, 1.0, 1.0, (double*)cm, ld, false, &e);
    colsq_finish(ctx, e.colsq, nrb, n, (const double*)cm, ld, ctx->geq0, dotp_dev);
    fused = true;
  } else {
    rotate_dev(ctx, 1.0, c0, 1.0, cm, G, n);
  }
  if (gamy_host) dl_add(ctx, gamy_host, G, sizeof(double) * (size_t)n * n);
  // c0 is unchanged (cm was written): the planes of c0 the kick pass left in
  // the cache stay kept for the next call on it
  if (c2 && ctx->gemm_x3 && K >= kX3Min && c0 != cm) keep_c0_planes(ctx, c0, K, n);
  return fused;
}

static void rortog_core(mc_ctx* ctx, const double2* c0, double2* cm, double* gam_host, int tnon, int n,
                        double epsog, int maxit, int* iters, double2* vel, double vdt, __bf16* planes_out) {
  // the n x n products below go through gemm_tn / gemm_nn, which send
  // products with K or M >= kX3Min to the split-bf16 kernels in mixed precision
  MC_CHECK(n < kX3Min, MC_ERR_UNSUPPORTED, "rortog: nstate beyond 16383");
  double* rho = small(ctx, 0, n);
  double* sig = small(ctx, 1, n);
  double* del = small(ctx, 2, n);
  double* gam = small(ctx, 3, n);
  double* gamn = small(ctx, 4, n);
  double* scr = small(ctx, 5, n);
  double* scr2 = small(ctx, 6, n);
  dim3 g2((n + 255) / 256, n);
  // rho = 1 - symma(ovlap(c0, cm)); sig = 1 - ovlap(cm, cm); del = -(ovlap(c0,c0) - 1)/2
  ovlap_dev(ctx, n, rho, n, c0, cm);
  hipLaunchKernelGGL(k_symma, g2, dim3(256), 0, ctx->stream, rho, n);
  hipLaunchKernelGGL(k_axpby_eye, g2, dim3(256), 0, ctx->stream, rho, -1.0, 1.0, (const double*)nullptr, 0.0, rho, n);
  ovlap_dev(ctx, n, sig, n, cm, cm);
  hipLaunchKernelGGL(k_axpby_eye, g2, dim3(256), 0, ctx->stream, sig, -1.0, 1.0, (const double*)nullptr, 0.0, sig, n);
  if (tnon) {
    ovlap_dev(ctx, n, del, n, c0, c0);
    hipLaunchKernelGGL(k_axpby_eye, g2, dim3(256), 0, ctx->stream, del, -0.5, 0.5, (const double*)nullptr, 0.0, del, n);
  }
  MC_HIP(hipGetLastError());
  // gam = sig/2; sig <- rho^T rho / 2 + sig / 2
  hipLaunchKernelGGL(k_axpby_eye, g2, dim3(256), 0, ctx->stream, sig, 0.5, 0.0, (const double*)nullptr, 0.0, gam, n);
  gemm_tn(ctx, rho, n, rho, n, n, n, n, 0.5, false, false, scr, n, nullptr);
  hipLaunchKernelGGL(k_axpby_eye, g2, dim3(256), 0, ctx->stream, scr, 1.0, 0.0, sig, 0.5, sig, n);
  MC_HIP(hipGetLastError());
  int it = 0;
  double difgam = 0.0;
  while (true) {
    ++it;
    // scr = rho - gam; gamn = sig - scr^T scr / 2 [+ gam^T (del gam)]
    hipLaunchKernelGGL(k_axpby_eye, g2, dim3(256), 0, ctx->stream, rho, 1.0, 0.0, gam, -1.0, scr, n);
    gemm_tn(ctx, scr, n, scr, n, n, n, n, -0.5, false, false, gamn, n, nullptr);
    hipLaunchKernelGGL(k_axpby_eye, g2, dim3(256), 0, ctx->stream, gamn, 1.0, 0.0, sig, 1.0, gamn, n);
    if (tnon) {
      gemm_nn(ctx, del, n, gam, n, n, n, n, 1.0, 0.0, scr2, n, false, nullptr);       // scr2 = del gam
      gemm_tn(ctx, gam, n, scr2, n, n, n, n, 1.0, false, false, scr, n, nullptr);    // scr = gam^T scr2
      hipLaunchKernelGGL(k_axpby_eye, g2, dim3(256), 0, ctx->stream, gamn, 1.0, 0.0, scr, 1.0, gamn, n);
    }
    MC_HIP(hipGetLastError());
    difgam = max_abs_diff(ctx, gamn, gam, (size_t)n * n);
    MC_CHECK(std::isfinite(difgam), MC_ERR_NOCONV, "RORTHO: iteration diverged (non-finite gamma)");
    MC_HIP(hipMemcpyAsync(gam, gamn, sizeof(double) * n * n, hipMemcpyDeviceToDevice, ctx->stream));
    if (!(difgam > epsog * 0.5 && it <= maxit)) break;
  }
  MC_CHECK(it <= maxit, MC_ERR_NOCONV,
           "RORTHO: DIFGAM=" + std::to_string(difgam) + " ITER=" + std::to_string(it) + " > " + std::to_string(maxit));
  hipLaunchKernelGGL(k_symma, g2, dim3(256), 0, ctx->stream, gam, n);
  MC_HIP(hipGetLastError());
  if (vel) {
    // cm += c0 gam with the velocity vel = (cm - c0) / vdt of the result in the same pass
    FnlEpi e{};
    e.vcm = (double*)vel;
    e.vodt = 1.0 / vdt;
    e.pl = planes_out;
    e.pl_rows = wfn_rows(ctx);   // the K of the planes the contractions read
    gemm_nn(ctx, (const double*)c0, wfn_ld(ctx), gam, n, wfn_ld(ctx), n, n, 1.0, 1.0, (double*)cm, wfn_ld(ctx), false,
            &e);
  } else {
    rotate_dev(ctx, 1.0, c0, 1.0, cm, gam, n);
  }
  if (gam_host) {
    MC_HIP(hipMemcpyAsync(gam_host, gam, sizeof(double) * n * n, hipMemcpyDeviceToHost, ctx->stream));
    MC_HIP(hipStreamSynchronize(ctx->stream));
  }
  if (iters) *iters = it;
}

// ---------------------------------------------------------------------------
// nonlocal projectors
// ---------------------------------------------------------------------------
void set_projectors(mc_ctx* ctx, int nsp, const int* na, int nhxs, const int* ngh, const int* lang,
                    const double* hmat, const double* twnl, int ldtw) {
  MC_CHECK(nsp >= 0 && nhxs >= 0, MC_ERR_ARG, "set_projectors: bad sizes");
  auto& P = ctx->proj;
  P.nsp = nsp;
  P.nhxs = nhxs;
  P.na.assign(na, na + nsp);
  P.ngh.assign(ngh, ngh + nsp);
  P.lang.assign(lang, lang + (size_t)nhxs * nsp);
  P.nat = 0;
  for (int is = 0; is < nsp; ++is) {
    MC_CHECK(na[is] >= 0 && ngh[is] >= 0 && ngh[is] <= nhxs, MC_ERR_ARG, "set_projectors: bad na/ngh");
    P.nat += na[is];
  }
  const size_t nh = (size_t)nhxs * nhxs * nsp;
  P.d_h.ensure(sizeof(double) * std::max<size_t>(nh, 1));
  if (nh) MC_HIP(hipMemcpy(P.d_h.p, hmat, sizeof(double) * nh, hipMemcpyHostToDevice));
  P.h_h.assign(hmat, hmat + nh);
  // twnl (ngw, nhxs, nsp) in CPMD order -> device rows (ldg) in stick order
  const int ncol = nhxs * nsp;
  P.d_twnl.ensure(sizeof(double) * (size_t)ctx->ldg * std::max(ncol, 1));
  if (ncol) {
    DevBuf stage;
    stage.ensure(sizeof(double) * (size_t)ldtw * ncol);
    MC_HIP(hipMemcpy(stage.p, twnl, sizeof(double) * (size_t)ldtw * ncol, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_permute_real_in, dim3((ctx->ldg + 255) / 256, ncol), dim3(256), 0, ctx->stream,
                       stage.as<double>(), ldtw, P.d_twnl.as<double>(), ctx->ldg, ctx->d_perm.as<int>(), ctx->ngw,
                       ncol);
    MC_HIP(hipGetLastError());
    MC_HIP(hipStreamSynchronize(ctx->stream));
  }
  P.ready = true;
}

void set_eigr(mc_ctx* ctx, const double* eigr, int lde, int nat) {
  MC_CHECK(lde >= ctx->ngw && nat >= 0, MC_ERR_ARG, "set_eigr: lde < ngw");
  auto& P = ctx->proj;
  P.d_eigr.ensure(sizeof(double2) * (size_t)ctx->ldg * std::max(nat, 1));
  P.eigr_nat = nat;
  P.eigr_serial = -1;   // set from the host: not the phase tables' product
  if (!nat) return;
  DevBuf stage;
  stage.ensure(sizeof(double2) * (size_t)lde * nat);
  MC_HIP(hipMemcpy(stage.p, eigr, sizeof(double2) * (size_t)lde * nat, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_permute_rows_in, dim3((ctx->ldg + 255) / 256, nat), dim3(256), 0, ctx->stream,
                     stage.as<double2>(), lde, P.d_eigr.as<double2>(), ctx->ldg, ctx->d_perm.as<int>(), ctx->ngw, nat);
  MC_HIP(hipGetLastError());
  MC_HIP(hipStreamSynchronize(ctx->stream));
}

// fnl(isa, iv, i) = 2 Re sum_G conj((-i)^l twnl eigr) c (rnlsm1,
// src/rnlsm1_utils.mod.F90:42-193) and dfnl(isa, iv, k, i) = 2 tpiba Re sum_G
// conj(gk_k twnl (-i)^(l+1) eigr) c (rnlsm2, src/rnlsm2_utils.mod.F90:321-482),
// G = 0 halved, per (species, projector) group.  In mixed precision the
// requested variants of a group are one stacked split-bf16 contraction over
// the planes of c0.
static void rnlsm_groups(mc_ctx* ctx, const double2* c0, int n, bool do_fnl, bool do_dfnl, double* fnl_host,
                         double* dfnl_host) {
  auto& P = ctx->proj;
  const auto& D = ctx->dens;
  MC_CHECK(P.ready, MC_ERR_ARG, "rnlsm: projectors not set (mc_set_projectors)");
  MC_CHECK(P.eigr_nat == P.nat, MC_ERR_ARG, "rnlsm: structure factors not set for all atoms (mc_set_eigr)");
  MC_CHECK(!do_dfnl || P.gk_ready, MC_ERR_ARG, "rnlsm2: G vectors not set (mc_set_gk)");
  const size_t nf = (size_t)P.nat * P.nhxs * std::max(n, 1);
  if (do_fnl) {
    P.d_fnl.ensure(sizeof(double) * nf);
    P.fnl_nstate = n;
    hipLaunchKernelGGL(k_zero, dim3(64), dim3(256), 0, ctx->stream, P.d_fnl.as<double>(), nf);
    MC_HIP(hipGetLastError());
  }
  if (do_dfnl) {
    P.d_dfnl.ensure(sizeof(double) * nf * 3);
    P.dfnl_nstate = n;
    hipLaunchKernelGGL(k_zero, dim3(64), dim3(256), 0, ctx->stream, P.d_dfnl.as<double>(), nf * 3);
    MC_HIP(hipGetLastError());
  }
  const double tpiba = std::sqrt(ctx->tpiba2);
  const int K = wfn_rows(ctx), ldb = wfn_ld(ctx);
  const double* B = (const double*)c0;
  const bool x3 = ctx->gemm_x3 && K >= kX3Min;
  const int ldf = P.nat * P.nhxs, ldd = P.nat * P.nhxs * 3;
  int isa0 = 0;
  for (int is = 0; is < P.nsp; ++is) {
    for (int iv = 0; iv < P.ngh[is]; ++iv) {
      if (P.na[is] == 0 || n == 0) continue;
      ProjA pa;
      pa.eigr = P.d_eigr.as<double2>() + (size_t)isa0 * ctx->ldg;
      pa.twnl = P.d_twnl.as<double>() + (size_t)(is * P.nhxs + iv) * ctx->ldg;
      pa.ldg = ctx->ldg;
      pa.lmod4 = ((P.lang[iv + (size_t)P.nhxs * is] % 4) + 4) % 4;
      pa.geq0 = ctx->geq0;
      if (x3 && D.ions_ready && P.eigr_serial == D.ions_serial && ctx->d_gmw.p) {
        // the plane build forms eigr from the phase tables it was made from
        pa.gmw = ctx->d_gmw.as<int>();
        pa.ph1 = D.d_ph1.as<double2>();
        pa.ph2 = D.d_ph2.as<double2>();
        pa.ph3 = D.d_ph3.as<double2>();
        pa.n1 = ctx->n1;
        pa.n2 = ctx->n2;
        pa.n3 = ctx->n3;
        pa.isa0 = isa0;
      }
      double* fout = P.d_fnl.as<double>() + isa0 + (size_t)P.nat * iv;
      double* dout[3];
      for (int k = 0; k < 3; ++k)
        dout[k] = do_dfnl ? P.d_dfnl.as<double>() + isa0 + (size_t)P.nat * (iv + (size_t)P.nhxs * k) : nullptr;
      if (x3) {
        double* outs[4];
        int ldos[4];
        double alphas[4];
        int nv = 0;
        if (do_fnl) { outs[nv] = fout; ldos[nv] = ldf; alphas[nv] = 2.0; ++nv; }
        if (do_dfnl)
          for (int k = 0; k < 3; ++k) { outs[nv] = dout[k]; ldos[nv] = ldd; alphas[nv] = 2.0 * tpiba; ++nv; }
        proj_contract_x3(ctx, pa, do_dfnl ? P.d_gk.as<double>() : nullptr, do_fnl ? 0 : 1, nv, B, ldb, K, P.na[is],
                         n, outs, ldos, alphas);
      } else {
        if (do_fnl) gemm_tn(ctx, nullptr, 0, B, ldb, K, P.na[is], n, 2.0, false, false, fout, ldf, &pa);
        if (do_dfnl) {
          ProjA pd = pa;
          pd.lmod4 = (pa.lmod4 + 1) % 4;
          for (int k = 0; k < 3; ++k) {
            pd.gkk = P.d_gk.as<double>() + (size_t)k * ctx->ldg;
            gemm_tn(ctx, nullptr, 0, B, ldb, K, P.na[is], n, 2.0 * tpiba, false, false, dout[k], ldd, &pd);
          }
        }
      }
    }
    isa0 += P.na[is];
  }
  if (fnl_host && do_fnl) dl_add(ctx, fnl_host, P.d_fnl.p, sizeof(double) * (size_t)P.nat * P.nhxs * n);
  if (dfnl_host && do_dfnl) dl_add(ctx, dfnl_host, P.d_dfnl.p, sizeof(double) * (size_t)P.nat * P.nhxs * 3 * n);
  if ((fnl_host && do_fnl) || (dfnl_host && do_dfnl)) dl_finish(ctx);
}

void rnlsm1_dev(mc_ctx* ctx, const double2* c0, int n, double* fnl_host) {
  rnlsm_groups(ctx, c0, n, true, false, fnl_host, nullptr);
}

void rnlsm2_dev(mc_ctx* ctx, const double2* c0, int n, double* dfnl_host) {
  rnlsm_groups(ctx, c0, n, false, true, nullptr, dfnl_host);
}

// rnlsm (src/rnlsm_utils.mod.F90:23-49): rnlsm1 and, with tfor, rnlsm2 of the
// same coefficients, in one pass over the projector groups
void rnlsm_dev(mc_ctx* ctx, const double2* c0, int n, bool tfor, double* fnl_host, double* dfnl_host) {
  rnlsm_groups(ctx, c0, n, true, tfor, fnl_host, dfnl_host);
}

void set_gk(mc_ctx* ctx, const double* gk) {
  auto& P = ctx->proj;
  std::vector<double> t((size_t)ctx->ngw * 3);
  for (int g = 0; g < ctx->ngw; ++g)
    for (int k = 0; k < 3; ++k) t[g + (size_t)ctx->ngw * k] = gk[k + 3 * (size_t)g];
  P.d_gk.ensure(sizeof(double) * (size_t)ctx->ldg * 3);
  DevBuf stage;
  stage.ensure(sizeof(double) * t.size());
  MC_HIP(hipMemcpy(stage.p, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_permute_real_in, dim3((ctx->ldg + 255) / 256, 3), dim3(256), 0, ctx->stream,
                     stage.as<double>(), ctx->ngw, P.d_gk.as<double>(), ctx->ldg, ctx->d_perm.as<int>(), ctx->ngw, 3);
  MC_HIP(hipGetLastError());
  MC_HIP(hipStreamSynchronize(ctx->stream));
  P.gk_ready = true;
}

void rnlfor_dev(mc_ctx* ctx, const double* f_host, int n, const double* fnl_host, const double* dfnl_host,
                double* fion_host) {
  auto& P = ctx->proj;
  MC_CHECK(P.ready, MC_ERR_ARG, "rnlfor: projectors not set (mc_set_projectors)");
  if (n <= 0 || P.nat == 0) return;
  const size_t nf = (size_t)P.nat * P.nhxs * n;
  if (fnl_host) {
    P.d_fnl.ensure(sizeof(double) * nf);
    MC_HIP(hipMemcpyAsync(P.d_fnl.p, fnl_host, sizeof(double) * nf, hipMemcpyHostToDevice, ctx->stream));
    P.fnl_nstate = n;
  }
  if (dfnl_host) {
    P.d_dfnl.ensure(sizeof(double) * nf * 3);
    MC_HIP(hipMemcpyAsync(P.d_dfnl.p, dfnl_host, sizeof(double) * nf * 3, hipMemcpyHostToDevice, ctx->stream));
    P.dfnl_nstate = n;
  }
  MC_CHECK(P.fnl_nstate == n, MC_ERR_ARG, "rnlfor: no fnl for this number of states (call rnlsm1 first)");
  MC_CHECK(P.dfnl_nstate == n, MC_ERR_ARG, "rnlfor: no dfnl for this number of states (call rnlsm2 first)");
  std::vector<int> sp_of(P.nat);
  for (int is = 0, isa = 0; is < P.nsp; ++is)
    for (int ia = 0; ia < P.na[is]; ++ia) sp_of[isa++] = is;
  // small staging area: sp_of (nat ints), ngh (nsp ints), f (n), out (3 nat)
  const size_t off_ngh = sizeof(int) * P.nat, off_f = (off_ngh + sizeof(int) * P.nsp + 15) / 16 * 16;
  const size_t off_out = off_f + sizeof(double) * n, total = off_out + sizeof(double) * 3 * P.nat;
  ctx->w_f.ensure(total);
  char* base = (char*)ctx->w_f.p;
  MC_HIP(hipMemcpyAsync(base, sp_of.data(), sizeof(int) * P.nat, hipMemcpyHostToDevice, ctx->stream));
  MC_HIP(hipMemcpyAsync(base + off_ngh, P.ngh.data(), sizeof(int) * P.nsp, hipMemcpyHostToDevice, ctx->stream));
  MC_HIP(hipMemcpyAsync(base + off_f, f_host, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_rnlfor, dim3(P.nat, 3), dim3(256), 0, ctx->stream, P.d_fnl.as<double>(),
                     P.d_dfnl.as<double>(), (const double*)(base + off_f), n, P.nat, P.nhxs, (const int*)base,
                     (const int*)(base + off_ngh), P.d_h.as<double>(), (double*)(base + off_out));
  MC_HIP(hipGetLastError());
  std::vector<double> out(3 * (size_t)P.nat);
  dl_add(ctx, out.data(), base + off_out, sizeof(double) * out.size());
  dl_finish(ctx);
  for (size_t e = 0; e < out.size(); ++e) fion_host[e] += out[e];
}

void rnlrh_dev(mc_ctx* ctx, const double* f_host, int n, double* enl) {
  auto& P = ctx->proj;
  MC_CHECK(P.ready, MC_ERR_ARG, "rnlrh: projectors not set (mc_set_projectors)");
  *enl = 0.0;
  if (n <= 0 || P.nat == 0) return;
  MC_CHECK(P.fnl_nstate == n, MC_ERR_ARG, "rnlrh: no fnl for this number of states (call rnlsm1 first)");
  std::vector<int> sp_of(P.nat);
  for (int is = 0, isa = 0; is < P.nsp; ++is)
    for (int ia = 0; ia < P.na[is]; ++ia) sp_of[isa++] = is;
  // small staging area: sp_of (nat ints), ngh (nsp ints), f (n), out (nat)
  const size_t off_ngh = sizeof(int) * P.nat, off_f = (off_ngh + sizeof(int) * P.nsp + 15) / 16 * 16;
  const size_t off_out = off_f + sizeof(double) * n, total = off_out + sizeof(double) * P.nat;
  ctx->w_f.ensure(total);
  char* base = (char*)ctx->w_f.p;
  MC_HIP(hipMemcpyAsync(base, sp_of.data(), sizeof(int) * P.nat, hipMemcpyHostToDevice, ctx->stream));
  MC_HIP(hipMemcpyAsync(base + off_ngh, P.ngh.data(), sizeof(int) * P.nsp, hipMemcpyHostToDevice, ctx->stream));
  MC_HIP(hipMemcpyAsync(base + off_f, f_host, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_rnlrh, dim3(P.nat), dim3(256), 0, ctx->stream, P.d_fnl.as<double>(),
                     (const double*)(base + off_f), n, P.nat, P.nhxs, (const int*)base, (const int*)(base + off_ngh),
                     P.d_h.as<double>(), (double*)(base + off_out));
  MC_HIP(hipGetLastError());
  std::vector<double> out((size_t)P.nat);
  dl_add(ctx, out.data(), base + off_out, sizeof(double) * out.size());
  dl_finish(ctx);
  double e = 0.0;
  for (double v : out) e += v;   // atoms in isa order
  *enl = e;
}

void fnonloc_dev(mc_ctx* ctx, double2* c2, const double* f_host, int n, const double* fnl_host) {
  auto& P = ctx->proj;
  MC_CHECK(P.ready, MC_ERR_ARG, "fnonloc: projectors not set (mc_set_projectors)");
  MC_CHECK(P.eigr_nat == P.nat, MC_ERR_ARG, "fnonloc: structure factors not set (mc_set_eigr)");
  if (n <= 0) return;
  if (fnl_host) {
    P.d_fnl.ensure(sizeof(double) * (size_t)P.nat * P.nhxs * n);
    MC_HIP(hipMemcpyAsync(P.d_fnl.p, fnl_host, sizeof(double) * (size_t)P.nat * P.nhxs * n, hipMemcpyHostToDevice,
                          ctx->stream));
    P.fnl_nstate = n;
  }
  MC_CHECK(P.fnl_nstate == n, MC_ERR_ARG, "fnonloc: no fnl for this number of states (call rnlsm1 first)");
  ctx->w_f2.ensure(sizeof(double) * n);
  MC_HIP(hipMemcpyAsync(ctx->w_f2.p, f_host, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
  int namax = 0;
  for (int is = 0; is < P.nsp; ++is) namax = std::max(namax, P.na[is]);
  ctx->w_dd.ensure(sizeof(double) * (size_t)std::max(namax, 1) * n);
  int isa0 = 0;
  for (int is = 0; is < P.nsp; ++is) {
    const int na = P.na[is];
    for (int iv = 0; iv < P.ngh[is] && na > 0; ++iv) {
      hipLaunchKernelGGL(k_fnl_dd, dim3((na + 63) / 64, n), dim3(64), 0, ctx->stream, P.d_fnl.as<double>(), P.nat,
                         P.nhxs, isa0, na, P.ngh[is], P.d_h.as<double>() + (size_t)P.nhxs * P.nhxs * is, iv, n,
                         ctx->w_dd.as<double>());
      MC_HIP(hipGetLastError());
      FnlEpi e;
      e.twnl = P.d_twnl.as<double>() + (size_t)(is * P.nhxs + iv) * ctx->ldg;
      e.f = ctx->w_f2.as<double>();
      e.lmod4 = ((P.lang[iv + (size_t)P.nhxs * is] % 4) + 4) % 4;
      gemm_nn(ctx, (const double*)(P.d_eigr.as<double2>() + (size_t)isa0 * ctx->ldg), wfn_ld(ctx),
              ctx->w_dd.as<double>(), na, wfn_rows(ctx), na, n, 1.0, 1.0, (double*)c2, wfn_ld(ctx), false, &e);
    }
    isa0 += na;
  }
}

}  // namespace mc
