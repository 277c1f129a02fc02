/*
 * ORACLE -- test infrastructure only.
 *
 * A plain-C FP64 CPU restatement of CPMD's plane-wave inner loop, used as the
 * checker of the GPU library.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; no product path calls it.
 *
 * It follows the reference's own algorithm and data conventions:
 *   - wavefunctions c(ig, i) on CPMD's half sphere, G order given by inyh;
 *   - the Gamma pair pack psi(nzh(G)) = c1 + i c2, psi(indz(G)) = conj(c1) + i conj(c2),
 *     G=0 set once (src/state_utils.mod.F90:171-189), with CPMD's *centred*
 *     grid index inyh (nzh/indz of src/rggen_utils.mod.F90:164-173);
 *   - inverse transform e^{+i}, unscaled; forward e^{-i} with 1/(n1 n2 n3)
 *     (src/fftmain_utils.mod.F90:86-155, src/mltfft_utils.mod.F90:513-517);
 *   - real-space arrays f(ix + kr1*(iy + kr2*iz)).
 * The transforms are an own mixed-radix recursive DFT, not CPMD's Goedecker
 * FFT (whose licence forbids copying); they are pinned against numpy.fft and a
 * brute-force DFT by the tests.
 */
#define _USE_MATH_DEFINES
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef double complex cplx;

/* ------------------------------------------------------------------------- */
/* 1D transform: recursive mixed radix (smallest prime factor first)          */
/* ------------------------------------------------------------------------- */
static int smallest_factor(int n) {
  for (int r = 2; r * r <= n; ++r)
    if (n % r == 0) return r;
  return n;
}

/* out[k] = sum_j in[j*stride] w^(j k), w = exp(sign 2 pi i / n) = tw[tws] (tw: table of
 * the top-level length N, tws = N/n).  `work` holds n entries of scratch. */
static void dft_rec(int n, const cplx* in, int stride, cplx* out, cplx* work, const cplx* tw, int tws) {
  if (n == 1) { out[0] = in[0]; return; }
  int r = smallest_factor(n);
  int m = n / r;
  if (r == n) {
    for (int k = 0; k < n; ++k) {
      cplx s = 0;
      for (int j = 0; j < n; ++j) s += in[(size_t)j * stride] * tw[(size_t)(((long long)j * k) % n) * tws];
      out[k] = s;
    }
    return;
  }
  /* sub-transforms of the r decimated sequences into work[q*m + k] */
  for (int q = 0; q < r; ++q)
    dft_rec(m, in + (size_t)q * stride, stride * r, work + (size_t)q * m, out, tw, tws * r);
  for (int s = 0; s < r; ++s)
    for (int k = 0; k < m; ++k) {
      int kk = k + m * s;
      cplx acc = 0;
      for (int q = 0; q < r; ++q) acc += work[(size_t)q * m + k] * tw[(size_t)(((long long)q * kk) % n) * tws];
      out[kk] = acc;
    }
}

static cplx* make_table(int n, int sign) {
  cplx* tw = malloc(sizeof(cplx) * n);
  for (int t = 0; t < n; ++t) {
    double a = sign * 2.0 * M_PI * (double)t / (double)n;
    tw[t] = cos(a) + I * sin(a);
  }
  return tw;
}

/* in-place transform of one strided line */
static void fft_line(int n, cplx* x, size_t stride, const cplx* tw, cplx* a, cplx* b, cplx* c) {
  for (int j = 0; j < n; ++j) a[j] = x[j * stride];
  dft_rec(n, a, 1, b, c, tw, 1);
  for (int j = 0; j < n; ++j) x[j * stride] = b[j];
}

/* 3D transform of f(n1,n2,n3) stored with leading dims (k1,k2): sign +1 inverse, -1 forward; unscaled */
void or_fft3d(int n1, int n2, int n3, int k1, int k2, double* data, int sign) {
  cplx* f = (cplx*)data;
  int nmax = n1 > n2 ? n1 : n2;
  nmax = nmax > n3 ? nmax : n3;
  cplx* a = malloc(sizeof(cplx) * nmax);
  cplx* b = malloc(sizeof(cplx) * nmax);
  cplx* c = malloc(sizeof(cplx) * nmax);
  cplx* t1 = make_table(n1, sign);
  cplx* t2 = make_table(n2, sign);
  cplx* t3 = make_table(n3, sign);
  for (int z = 0; z < n3; ++z)
    for (int y = 0; y < n2; ++y) fft_line(n1, f + (size_t)k1 * (y + (size_t)k2 * z), 1, t1, a, b, c);
  for (int z = 0; z < n3; ++z)
    for (int x = 0; x < n1; ++x) fft_line(n2, f + x + (size_t)k1 * k2 * z, (size_t)k1, t2, a, b, c);
  for (int y = 0; y < n2; ++y)
    for (int x = 0; x < n1; ++x) fft_line(n3, f + x + (size_t)k1 * y, (size_t)k1 * k2, t3, a, b, c);
  free(a); free(b); free(c); free(t1); free(t2); free(t3);
}

void or_fft1d(int n, double* data, int sign) {
  cplx* a = malloc(sizeof(cplx) * n);
  cplx* b = malloc(sizeof(cplx) * n);
  cplx* c = malloc(sizeof(cplx) * n);
  cplx* t = make_table(n, sign);
  fft_line(n, (cplx*)data, 1, t, a, b, c);
  free(a); free(b); free(c); free(t);
}

/* ------------------------------------------------------------------------- */
/* layout: CPMD's centred grid index of +G and -G (src/rggen_utils.mod.F90:164-173) */
/* ------------------------------------------------------------------------- */
typedef struct {
  int n1, n2, n3, kr1, kr2, kr3, ngw, geq0;
  const int* inyh; /* (3,ngw) 1-based */
  const double* hg;
  double tpiba2, omega;
} or_layout;

static void grid_index(const or_layout* L, int ig, size_t* nzh, size_t* indz) {
  int nh1 = L->n1 / 2 + 1, nh2 = L->n2 / 2 + 1, nh3 = L->n3 / 2 + 1;
  int i1 = L->inyh[3 * ig], i2 = L->inyh[3 * ig + 1], i3 = L->inyh[3 * ig + 2];
  /* 0-based versions of nzh = i1 + (i2-1) kr1 + (i3-1) kr1 kr2 and of indz for -G */
  *nzh = (size_t)(i1 - 1) + (size_t)L->kr1 * ((size_t)(i2 - 1) + (size_t)L->kr2 * (size_t)(i3 - 1));
  int j1 = 2 * nh1 - i1, j2 = 2 * nh2 - i2, j3 = 2 * nh3 - i3;
  *indz = (size_t)(j1 - 1) + (size_t)L->kr1 * ((size_t)(j2 - 1) + (size_t)L->kr2 * (size_t)(j3 - 1));
}

/* set_psi_2_states_g / set_psi_1_state_g, src/state_utils.mod.F90:133-189 */
static void pack_pair(const or_layout* L, const cplx* c1, const cplx* c2, cplx* psi, size_t nnr1) {
  memset(psi, 0, sizeof(cplx) * nnr1);
  for (int ig = 0; ig < L->ngw; ++ig) {
    size_t a, b;
    grid_index(L, ig, &a, &b);
    cplx v1 = c1[ig], v2 = c2 ? c2[ig] : 0;
    psi[a] = v1 + I * v2;
    psi[b] = conj(v1) + I * conj(v2);
  }
  if (L->geq0) {
    size_t a, b;
    grid_index(L, 0, &a, &b);
    psi[a] = c1[0] + I * (c2 ? c2[0] : 0);
  }
}

/* dotp, src/dotp_utils.mod.F90:58-101 */
static double dotp(const or_layout* L, const cplx* a, const cplx* b) {
  double s = 0.0;
  int start = 0;
  if (L->geq0) { s = creal(a[0]) * creal(b[0]); start = 1; }
  double t = 0.0;
  for (int ig = start; ig < L->ngw; ++ig) t += creal(a[ig]) * creal(b[ig]) + cimag(a[ig]) * cimag(b[ig]);
  return s + 2.0 * t;
}

/* rhoofr, src/rhoofr_utils.mod.F90:122-644 (closed shell, serial):
 * kin_energy (src/kin_energy_utils.mod.F90:28-110), pair loop (:305-409),
 * build_density_sum (src/density_utils.mod.F90:61-80), charge sums (:615-623). */
void or_rhoofr(const int* dims /* n1 n2 n3 kr1 kr2 kr3 ngw geq0 */, const int* inyh, const double* hg,
               double tpiba2, double omega, const double* c0d, int ldc, const double* f, int nstate,
               double* rhoe, double* ekin, double* csumg, double* csumr) {
  or_layout L = {dims[0], dims[1], dims[2], dims[3], dims[4], dims[5], dims[6], dims[7], inyh, hg, tpiba2, omega};
  const cplx* c0 = (const cplx*)c0d;
  size_t nnr1 = (size_t)L.kr1 * L.kr2 * L.kr3;
  double rsum = 0.0, xkin = 0.0;
  for (int i = 0; i < nstate; ++i) {
    if (f[i] == 0.0) continue;
    const cplx* c = c0 + (size_t)i * ldc;
    rsum += f[i] * dotp(&L, c, c);
    double sk = 0.0;
    for (int ig = 0; ig < L.ngw; ++ig) sk += hg[ig] * creal(conj(c[ig]) * c[ig]);
    xkin += f[i] * sk;
  }
  *ekin = xkin * tpiba2;
  memset(rhoe, 0, sizeof(double) * nnr1);
  cplx* psi = malloc(sizeof(cplx) * nnr1);
  for (int is1 = 0; is1 < nstate; is1 += 2) {
    int is2 = is1 + 1;
    int has2 = is2 < nstate;
    int tfcal = (f[is1] != 0.0) || (has2 && f[is2] != 0.0);
    if (!tfcal) continue;
    pack_pair(&L, c0 + (size_t)is1 * ldc, has2 ? c0 + (size_t)is2 * ldc : NULL, psi, nnr1);
    or_fft3d(L.n1, L.n2, L.n3, L.kr1, L.kr2, (double*)psi, +1);
    double coef3 = f[is1] / omega;
    double coef4 = has2 ? f[is2] / omega : 0.0;
    for (size_t l = 0; l < nnr1; ++l)
      rhoe[l] = rhoe[l] + coef3 * creal(psi[l]) * creal(psi[l]) + coef4 * cimag(psi[l]) * cimag(psi[l]);
  }
  free(psi);
  double rs = 0.0;
  for (size_t l = 0; l < nnr1; ++l) rs += rhoe[l];
  *csumg = rsum;
  *csumr = rs * omega / ((double)L.n1 * L.n2 * L.n3);
}

/* vpsi, src/vpsi_utils.mod.F90:120-732 (Gamma, closed shell, no preconditioned
 * kinetic term): pack (:400-435), invfftn (:443), psi *= V (:487-493),
 * fwfftn (:552), decode (:647-672), add_wfn (:717). */
void or_vpsi(const int* dims, const int* inyh, const double* hg, double tpiba2, double omega,
             const double* c0d, int ldc, double* c2d, const double* f, const double* vpot, int nstate) {
  or_layout L = {dims[0], dims[1], dims[2], dims[3], dims[4], dims[5], dims[6], dims[7], inyh, hg, tpiba2, omega};
  const cplx* c0 = (const cplx*)c0d;
  cplx* c2 = (cplx*)c2d;
  size_t nnr1 = (size_t)L.kr1 * L.kr2 * L.kr3;
  double scale = 1.0 / ((double)L.n1 * L.n2 * L.n3);
  cplx* psi = malloc(sizeof(cplx) * nnr1);
  for (int is1 = 0; is1 < nstate; is1 += 2) {
    int is2 = is1 + 1;
    int has2 = is2 < nstate;
    pack_pair(&L, c0 + (size_t)is1 * ldc, has2 ? c0 + (size_t)is2 * ldc : NULL, psi, nnr1);
    or_fft3d(L.n1, L.n2, L.n3, L.kr1, L.kr2, (double*)psi, +1);
    for (int z = 0; z < L.n3; ++z)
      for (int y = 0; y < L.n2; ++y)
        for (int x = 0; x < L.n1; ++x) {
          size_t l = x + (size_t)L.kr1 * (y + (size_t)L.kr2 * z);
          psi[l] = vpot[l] * psi[l];
        }
    or_fft3d(L.n1, L.n2, L.n3, L.kr1, L.kr2, (double*)psi, -1);
    double fi = f[is1] * 0.5;
    if (fi == 0.0) fi = 1.0;
    double fip1 = has2 ? f[is2] * 0.5 : 0.0;
    if (fip1 == 0.0) fip1 = 1.0;
    for (int ig = 0; ig < L.ngw; ++ig) {
      size_t a, b;
      grid_index(&L, ig, &a, &b);
      cplx psin = psi[a] * scale, psii = psi[b] * scale;
      cplx fp = psin + psii, fm = psin - psii;
      double g2 = tpiba2 * hg[ig];
      cplx t1 = -fi * (g2 * c0[(size_t)is1 * ldc + ig] + (creal(fp) + I * cimag(fm)));
      c2[(size_t)is1 * ldc + ig] += t1;
      if (has2) {
        cplx t2 = -fip1 * (g2 * c0[(size_t)is2 * ldc + ig] + (cimag(fp) - I * creal(fm)));
        c2[(size_t)is2 * ldc + ig] += t2;
      }
    }
  }
  free(psi);
}

/* ------------------------------------------------------------------------- */
/* Dense contractions (plain loops; column-major, CPMD G order, G=0 row 0)    */
/* ------------------------------------------------------------------------- */

/* ovlap, src/ovlap_utils.mod.F90:57-190: a = 2 Re(c1^H c2) - Re c1(0) Re c2(0) */
void or_ovlap(int ngw, int geq0, const double* c1d, const double* c2d, int ldc, int n, double* a) {
  const cplx* c1 = (const cplx*)c1d;
  const cplx* c2 = (const cplx*)c2d;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      const cplx* x = c1 + (size_t)i * ldc;
      const cplx* y = c2 + (size_t)j * ldc;
      for (int g = 0; g < ngw; ++g) s += creal(x[g]) * creal(y[g]) + cimag(x[g]) * cimag(y[g]);
      s *= 2.0;
      if (geq0) s -= creal(x[0]) * creal(y[0]);
      a[i + (size_t)n * j] = s;
    }
}

/* upper Cholesky S = U^T U in place (LAPACK dpotf2 'U' order); returns info */
int or_potrf_upper(double* s, int n) {
  for (int j = 0; j < n; ++j) {
    double d = s[j + (size_t)n * j];
    for (int k = 0; k < j; ++k) d -= s[k + (size_t)n * j] * s[k + (size_t)n * j];
    if (!(d > 0.0)) return j + 1;
    d = sqrt(d);
    s[j + (size_t)n * j] = d;
    for (int i = j + 1; i < n; ++i) {
      double t = s[j + (size_t)n * i];
      for (int k = 0; k < j; ++k) t -= s[k + (size_t)n * j] * s[k + (size_t)n * i];
      s[j + (size_t)n * i] = t / d;
    }
  }
  for (int j = 0; j < n; ++j)
    for (int i = j + 1; i < n; ++i) s[i + (size_t)n * j] = 0.0;
  return 0;
}

/* inverse of an upper triangular matrix (dtrti2 'U','N' semantics) */
void or_trtri_upper(const double* u, int n, double* x) {
  memset(x, 0, sizeof(double) * n * n);
  for (int j = 0; j < n; ++j) {
    x[j + (size_t)n * j] = 1.0 / u[j + (size_t)n * j];
    for (int i = j - 1; i >= 0; --i) {
      double s = 0.0;
      for (int k = i + 1; k <= j; ++k) s += u[i + (size_t)n * k] * x[k + (size_t)n * j];
      x[i + (size_t)n * j] = -s / u[i + (size_t)n * i];
    }
  }
}

/* c (real view 2*ngw x n, ld 2*ldc) <- c * m (m n x n) */
static void right_mul(int ngw, double* cd, int ldc, const double* m, int n) {
  double* row = malloc(sizeof(double) * n);
  for (int r = 0; r < 2 * ngw; ++r) {
    for (int j = 0; j < n; ++j) {
      double s = 0.0;
      for (int k = 0; k < n; ++k) s += cd[r + (size_t)2 * ldc * k] * m[k + (size_t)n * j];
      row[j] = s;
    }
    for (int j = 0; j < n; ++j) cd[r + (size_t)2 * ldc * j] = row[j];
  }
  free(row);
}

/* rgs, src/rgs_utils.mod.F90:31-93 with uinv :95-162; smat receives U^-1 */
int or_rgs(int ngw, int geq0, double* cp, int ldc, int n, double* smat) {
  double* s = malloc(sizeof(double) * n * n);
  or_ovlap(ngw, geq0, cp, cp, ldc, n, s);
  int info = or_potrf_upper(s, n);
  if (info) { free(s); return info; }
  or_trtri_upper(s, n, smat);
  right_mul(ngw, cp, ldc, smat, n);
  if (geq0)
    for (int i = 0; i < n; ++i) cp[(size_t)2 * ldc * i + 1] = 0.0;
  free(s);
  return 0;
}

/* symmetric eigen-decomposition by cyclic Jacobi rotations: a (n x n) is
 * destroyed, w eigenvalues, z eigenvectors (columns) */
void or_jacobi_eigen(double* a, int n, double* w, double* z) {
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) z[i + (size_t)n * j] = (i == j);
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < j; ++i) off += a[i + (size_t)n * j] * a[i + (size_t)n * j];
    if (off < 1e-30) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = a[p + (size_t)n * q];
        if (fabs(apq) < 1e-300) continue;
        double app = a[p + (size_t)n * p], aqq = a[q + (size_t)n * q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          double akp = a[k + (size_t)n * p], akq = a[k + (size_t)n * q];
          a[k + (size_t)n * p] = c * akp - s * akq;
          a[k + (size_t)n * q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          double apk = a[p + (size_t)n * k], aqk = a[q + (size_t)n * k];
          a[p + (size_t)n * k] = c * apk - s * aqk;
          a[q + (size_t)n * k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double zkp = z[k + (size_t)n * p], zkq = z[k + (size_t)n * q];
          z[k + (size_t)n * p] = c * zkp - s * zkq;
          z[k + (size_t)n * q] = s * zkp + c * zkq;
        }
      }
  }
  for (int i = 0; i < n; ++i) w[i] = a[i + (size_t)n * i];
}

/* lowdin, src/lowdin_utils.mod.F90:27-192 (closed shell):
 * A(i,j) = sum_k z(i,k) w_k^-1/2 z(j,k), C <- C A, Im c(G=0) = 0 */
void or_lowdin(int ngw, int geq0, double* cp, int ldc, int n, double* amat) {
  double* s = malloc(sizeof(double) * n * n);
  double* z = malloc(sizeof(double) * n * n);
  double* w = malloc(sizeof(double) * n);
  or_ovlap(ngw, geq0, cp, cp, ldc, n, s);
  or_jacobi_eigen(s, n, w, z);
  for (int k = 0; k < n; ++k) w[k] = 1.0 / sqrt(w[k]);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      double t = 0.0;
      for (int k = 0; k < n; ++k) t += w[k] * z[j + (size_t)n * k] * z[i + (size_t)n * k];
      amat[i + (size_t)n * j] = t;
    }
  right_mul(ngw, cp, ldc, amat, n);
  if (geq0)
    for (int i = 0; i < n; ++i) cp[(size_t)2 * ldc * i + 1] = 0.0;
  free(s); free(z); free(w);
}

/* rotate, src/rotate_utils.mod.F90:63-100: c2 = a c1 gam + b c2 */
void or_rotate(int ngw, double a, const double* c1, double b, double* c2, int ldc, const double* gam, int n) {
  double* row = malloc(sizeof(double) * n);
  for (int r = 0; r < 2 * ngw; ++r) {
    for (int j = 0; j < n; ++j) {
      double s = 0.0;
      for (int k = 0; k < n; ++k) s += c1[r + (size_t)2 * ldc * k] * gam[k + (size_t)n * j];
      row[j] = a * s + b * c2[r + (size_t)2 * ldc * j];
    }
    for (int j = 0; j < n; ++j) c2[r + (size_t)2 * ldc * j] = row[j];
  }
  free(row);
}

/* Lagrange projection, src/forces_driver.mod.F90:344-370 (iproj = 2) */
void or_lagrange(int ngw, int geq0, const double* c0, double* c2, int ldc, const double* f, int n) {
  double* gam = malloc(sizeof(double) * n * n);
  or_ovlap(ngw, geq0, c2, c0, ldc, n, gam);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) gam[i + (size_t)n * j] *= (f[i] != 0.0 ? 1.0 : 0.0) * (f[j] != 0.0 ? 1.0 : 0.0);
  or_rotate(ngw, -1.0, c0, 1.0, c2, ldc, gam, n);
  if (geq0)
    for (int i = 0; i < n; ++i) c2[(size_t)2 * ldc * i + 1] = 0.0;
  free(gam);
}

/* hesele, src/hesele_utils.mod.F90:32-142: vpp(ig) over ngw G vectors in CPMD
 * order.  prec: vp = vploc, then per species, projector pair (iv, jv) with
 * hmat(iv,jv,is) != 0 (Goedecker: equal l and m; wsg: diagonal)
 * vp += (hmat na) twnl(ig,iv) twnl(ig,jv) (:76-104), vp = ftpi hg + vp,
 * vp = max(vp, hthrs), vpp = 1/vp (:109-131); otherwise vpp = svar2. */
void or_hesele(int ngw, int prec, double svar2, double vploc, double hthrs, double tpiba2, const double* hg, int nsp,
               const int* na, int nhxs, const int* ngh, const double* hmat, const double* twnl, int ldtw,
               double* vpp) {
  for (int ig = 0; ig < ngw; ++ig) {
    if (!prec) {
      vpp[ig] = svar2;
      continue;
    }
    double v = vploc;
    for (int is = 0; is < nsp; ++is)
      for (int iv = 0; iv < ngh[is]; ++iv)
        for (int jv = 0; jv < ngh[is]; ++jv) {
          const double h = hmat[iv + (size_t)nhxs * (jv + (size_t)nhxs * is)];
          if (h == 0.0) continue;
          const double dd = h * na[is];
          v += dd * twnl[ig + (size_t)ldtw * (iv + (size_t)nhxs * is)] * twnl[ig + (size_t)ldtw * (jv + (size_t)nhxs * is)];
        }
    v = 0.5 * tpiba2 * hg[ig] + v;
    if (v < hthrs) v = hthrs;
    vpp[ig] = 1.0 / v;
  }
}

/* dotp of src/dotp_utils.mod.F90:58-101 on raw columns */
static double dotp_raw(int ngw, int geq0, const cplx* a, const cplx* b) {
  or_layout L;
  memset(&L, 0, sizeof L);
  L.ngw = ngw;
  L.geq0 = geq0;
  return dotp(&L, a, b);
}

/* csize, src/csize_utils.mod.F90:100-131 (closed shell, Gamma): over the first
 * nocc states (nocc = number with f > 1e-5), izamax (first largest
 * |Re| + |Im| in column-major order), gemax = its modulus;
 * cnorm = sqrt(sum_i dotp(c2_i, c2_i) / (nocc ngw)). */
void or_csize(int ngw, int geq0, const double* c2d, int ldc, const double* f, int n, double* gemax, double* cnorm) {
  const cplx* c2 = (const cplx*)c2d;
  int nocc = 0;
  for (int i = 0; i < n; ++i)
    if (f[i] > 1e-5) ++nocc;
  double best = -1.0, s = 0.0;
  *gemax = 0.0;
  for (int i = 0; i < nocc; ++i) {
    for (int ig = 0; ig < ngw; ++ig) {
      const cplx z = c2[ig + (size_t)ldc * i];
      const double k = fabs(creal(z)) + fabs(cimag(z));
      if (k > best) {
        best = k;
        *gemax = cabs(z);
      }
    }
    s += dotp_raw(ngw, geq0, c2 + (size_t)ldc * i, c2 + (size_t)ldc * i);
  }
  *cnorm = nocc ? sqrt(s / ((double)nocc * ngw)) : 0.0;
}

/* updwf steepest descent, src/updwf_utils.mod.F90:117-128:
 * c0(ig,i) = c0(ig,i) + vpp(ig) pf1 c2(ig,i) */
void or_sd_update(int ngw, double* c0d, const double* c2d, int ldc, int n, const double* vpp, double pf1) {
  cplx* c0 = (cplx*)c0d;
  const cplx* c2 = (const cplx*)c2d;
  for (int i = 0; i < n; ++i)
    for (int ig = 0; ig < ngw; ++ig) c0[ig + (size_t)ldc * i] += vpp[ig] * pf1 * c2[ig + (size_t)ldc * i];
}

/* rortho, src/rortog_utils.mod.F90:315-498; returns iterations, -1 if > maxit */
int or_rortho(int ngw, int geq0, const double* c0, double* cm, int ldc, double* gam, int tnon, int n, double epsog,
              int maxit) {
  size_t nn = (size_t)n * n;
  double *rho = malloc(sizeof(double) * nn), *sig = malloc(sizeof(double) * nn), *del = malloc(sizeof(double) * nn);
  double *gamn = malloc(sizeof(double) * nn), *scr = malloc(sizeof(double) * nn), *scr2 = malloc(sizeof(double) * nn);
#define M(a, i, j) a[(i) + (size_t)n * (j)]
  or_ovlap(ngw, geq0, c0, cm, ldc, n, rho);
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) { double v = 0.5 * (M(rho, i, j) + M(rho, j, i)); M(rho, i, j) = v; M(rho, j, i) = v; }
  for (size_t e = 0; e < nn; ++e) rho[e] = -rho[e];
  for (int i = 0; i < n; ++i) M(rho, i, i) += 1.0;
  or_ovlap(ngw, geq0, cm, cm, ldc, n, sig);
  for (size_t e = 0; e < nn; ++e) sig[e] = -sig[e];
  for (int i = 0; i < n; ++i) M(sig, i, i) += 1.0;
  if (tnon) {
    or_ovlap(ngw, geq0, c0, c0, ldc, n, del);
    for (int i = 0; i < n; ++i) M(del, i, i) -= 1.0;
    for (size_t e = 0; e < nn; ++e) del[e] *= -0.5;
  }
  for (size_t e = 0; e < nn; ++e) gam[e] = 0.5 * sig[e];
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int k = 0; k < n; ++k) s += M(rho, k, i) * M(rho, k, j);
      scr[i + (size_t)n * j] = 0.5 * s + 0.5 * M(sig, i, j);
    }
  memcpy(sig, scr, sizeof(double) * nn);
  int iter = 0;
  double difgam;
  do {
    ++iter;
    memcpy(gamn, sig, sizeof(double) * nn);
    for (size_t e = 0; e < nn; ++e) scr[e] = rho[e] - gam[e];
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int k = 0; k < n; ++k) s += M(scr, k, i) * M(scr, k, j);
        M(gamn, i, j) += -0.5 * s;
      }
    if (tnon) {
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          double s = 0.0;
          for (int k = 0; k < n; ++k) s += M(del, i, k) * M(gam, k, j);
          M(scr2, i, j) = s;
        }
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          double s = 0.0;
          for (int k = 0; k < n; ++k) s += M(gam, k, i) * M(scr2, k, j);
          M(gamn, i, j) += s;
        }
    }
    difgam = 0.0;
    for (size_t e = 0; e < nn; ++e) { double d = fabs(gamn[e] - gam[e]); if (d > difgam) difgam = d; }
    memcpy(gam, gamn, sizeof(double) * nn);
  } while (difgam > epsog * 0.5 && iter <= maxit);
  int ret = iter > maxit ? -1 : iter;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) { double v = 0.5 * (M(gam, i, j) + M(gam, j, i)); M(gam, i, j) = v; M(gam, j, i) = v; }
  or_rotate(ngw, 1.0, c0, 1.0, cm, ldc, gam, n);
#undef M
  free(rho); free(sig); free(del); free(gamn); free(scr); free(scr2);
  return ret;
}

/* (-i)^l exactly, as CPMD's integer power (0,-1)**l */
static cplx mi_pow(int l) {
  switch (((l % 4) + 4) % 4) {
    case 0: return 1.0;
    case 1: return -I;
    case 2: return -1.0;
    default: return I;
  }
}

/* rnlsm1, src/rnlsm1_utils.mod.F90:42-193 (Gamma, imagp = 1):
 * fnl(isa,iv,i) = 2 sum_G Re(conj(scr) c), scr = (-i)^l twnl eigr, G=0 halved */
void or_rnlsm1(int ngw, int geq0, int nsp, const int* na, int nhxs, const int* ngh, const int* lang,
               const double* twnl, int ldtw, const double* eigrd, int lde, const double* c0d, int ldc, int n,
               double* fnl) {
  const cplx* eigr = (const cplx*)eigrd;
  const cplx* c0 = (const cplx*)c0d;
  int nat = 0;
  for (int is = 0; is < nsp; ++is) nat += na[is];
  memset(fnl, 0, sizeof(double) * nat * nhxs * n);
  int isa0 = 0;
  for (int is = 0; is < nsp; ++is) {
    for (int iv = 0; iv < ngh[is]; ++iv) {
      cplx ci = mi_pow(lang[iv + nhxs * is]);
      for (int ia = 0; ia < na[is]; ++ia)
        for (int i = 0; i < n; ++i) {
          double s = 0.0;
          for (int g = 0; g < ngw; ++g) {
            cplx scr = twnl[g + (size_t)ldtw * (iv + nhxs * is)] * ci * eigr[g + (size_t)lde * (isa0 + ia)];
            if (geq0 && g == 0) scr *= 0.5;
            s += creal(scr) * creal(c0[g + (size_t)ldc * i]) + cimag(scr) * cimag(c0[g + (size_t)ldc * i]);
          }
          fnl[(isa0 + ia) + (size_t)nat * (iv + (size_t)nhxs * i)] = 2.0 * s;
        }
    }
    isa0 += na[is];
  }
}

/* fnonloc, src/fnonloc_utils.mod.F90:209-303 (Gamma, generic h(iv,jv)) */
void or_fnonloc(int ngw, int nsp, const int* na, int nhxs, const int* ngh, const int* lang, const double* hmat,
                const double* twnl, int ldtw, const double* eigrd, int lde, const double* fnl, const double* f,
                double* c2d, int ldc, int n) {
  const cplx* eigr = (const cplx*)eigrd;
  cplx* c2 = (cplx*)c2d;
  int nat = 0;
  for (int is = 0; is < nsp; ++is) nat += na[is];
  int isa0 = 0;
  for (int is = 0; is < nsp; ++is) {
    for (int iv = 0; iv < ngh[is]; ++iv) {
      for (int i = 0; i < n; ++i) {
        double ffi = f[i] < 1e-5 ? 1.0 : f[i];
        cplx ctm = -ffi * mi_pow(lang[iv + nhxs * is]);
        for (int g = 0; g < ngw; ++g) {
          cplx aux = 0;
          for (int ia = 0; ia < na[is]; ++ia) {
            double dd = 0.0;
            for (int jv = 0; jv < ngh[is]; ++jv)
              dd += fnl[(isa0 + ia) + (size_t)nat * (jv + (size_t)nhxs * i)] *
                    hmat[iv + nhxs * (jv + (size_t)nhxs * is)];
            aux += eigr[g + (size_t)lde * (isa0 + ia)] * dd;
          }
          c2[g + (size_t)ldc * i] += twnl[g + (size_t)ldtw * (iv + nhxs * is)] * ctm * aux;
        }
      }
    }
    isa0 += na[is];
  }
}

/* rnlsm2, src/rnlsm2_utils.mod.F90:321-482 (Gamma, imagp = 1):
 * dfnl(isa,iv,k,i) = 2 tpiba sum_G Re(conj(eiscr) c), eiscr = gk(k,G) twnl (-i)^(l+1) eigr,
 * G=0 halved (:441); gk (3 x ldgk) in units of 2 pi / alat */
void or_rnlsm2(int ngw, int geq0, double tpiba, const double* gk, int nsp, const int* na, int nhxs, const int* ngh,
               const int* lang, const double* twnl, int ldtw, const double* eigrd, int lde, const double* c0d, int ldc,
               int n, double* dfnl) {
  const cplx* eigr = (const cplx*)eigrd;
  const cplx* c0 = (const cplx*)c0d;
  int nat = 0;
  for (int is = 0; is < nsp; ++is) nat += na[is];
  memset(dfnl, 0, sizeof(double) * (size_t)nat * nhxs * 3 * n);
  for (int k = 0; k < 3; ++k) {
    int isa0 = 0;
    for (int is = 0; is < nsp; ++is) {
      for (int iv = 0; iv < ngh[is]; ++iv) {
        cplx ci = mi_pow(lang[iv + nhxs * is] + 1);
        for (int ia = 0; ia < na[is]; ++ia)
          for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int g = 0; g < ngw; ++g) {
              cplx e = gk[k + 3 * (size_t)g] * twnl[g + (size_t)ldtw * (iv + nhxs * is)] * ci *
                       eigr[g + (size_t)lde * (isa0 + ia)];
              if (geq0 && g == 0) e *= 0.5;
              s += creal(e) * creal(c0[g + (size_t)ldc * i]) + cimag(e) * cimag(c0[g + (size_t)ldc * i]);
            }
            dfnl[(isa0 + ia) + (size_t)nat * (iv + (size_t)nhxs * (k + 3 * (size_t)i))] = 2.0 * tpiba * s;
          }
      }
      isa0 += na[is];
    }
  }
}

/* rnlfor, src/rnlfor_utils.mod.F90:201-280 (Goedecker branch; BHS/TM :281+ is
 * the diagonal case of h): fion(k,isa) -= sum_i 2 f_i sum_{iv,jv} h(iv,jv)
 * dfnl(isa,iv,k,i) fnl(isa,jv,i), states with |f_i| <= 1e-12 skipped */
void or_rnlfor(int nsp, const int* na, int nhxs, const int* ngh, const double* hmat, const double* fnl,
               const double* dfnl, const double* f, int n, double* fion) {
  int nat = 0;
  for (int is = 0; is < nsp; ++is) nat += na[is];
  int isa0 = 0;
  for (int is = 0; is < nsp; ++is) {
    for (int iv = 0; iv < ngh[is]; ++iv)
      for (int jv = 0; jv < ngh[is]; ++jv) {
        const double h = hmat[iv + nhxs * (jv + (size_t)nhxs * is)];
        if (h == 0.0) continue;
        for (int i = 0; i < n; ++i) {
          if (fabs(f[i]) <= 1e-12) continue;
          const double tt = 2.0 * f[i] * h;
          for (int ia = 0; ia < na[is]; ++ia) {
            const int isa = isa0 + ia;
            const double fj = fnl[isa + (size_t)nat * (jv + (size_t)nhxs * i)];
            for (int k = 0; k < 3; ++k)
              fion[k + 3 * (size_t)isa] -= tt * dfnl[isa + (size_t)nat * (iv + (size_t)nhxs * (k + 3 * (size_t)i))] * fj;
          }
        }
      }
    isa0 += na[is];
  }
}

/* E_NL = sum_i f_i sum_isa sum_{iv,jv} h(iv,jv) fnl(isa,iv,i) fnl(isa,jv,i)
 * (the energy whose position derivative rnlfor returns) */
double or_enl(int nsp, const int* na, int nhxs, const int* ngh, const double* hmat, const double* fnl,
              const double* f, int n) {
  int nat = 0;
  for (int is = 0; is < nsp; ++is) nat += na[is];
  double e = 0.0;
  int isa0 = 0;
  for (int is = 0; is < nsp; ++is) {
    for (int iv = 0; iv < ngh[is]; ++iv)
      for (int jv = 0; jv < ngh[is]; ++jv) {
        const double h = hmat[iv + nhxs * (jv + (size_t)nhxs * is)];
        for (int i = 0; i < n; ++i)
          for (int ia = 0; ia < na[is]; ++ia) {
            const int isa = isa0 + ia;
            e += f[i] * h * fnl[isa + (size_t)nat * (iv + (size_t)nhxs * i)] *
                 fnl[isa + (size_t)nat * (jv + (size_t)nhxs * i)];
          }
      }
    isa0 += na[is];
  }
  return e;
}

/* ------------------------------------------------------------------------- */
/* local potential, Hartree, LDA and Ewald terms (vofrhoa / vofrhob / rpiiint) */
/* ------------------------------------------------------------------------- */

/* e^{-i G.R} for centred index triple (1-based inyh) and fractional position */
static cplx phase(const int* nr, const int* iny, const double* fr) {
  double t = 0.0;
  for (int k = 0; k < 3; ++k) t += (double)(iny[k] - (nr[k] / 2 + 1)) * fr[k];
  return cexp(-2.0 * M_PI * I * t);
}

/* eicalc, src/eicalc_utils.mod.F90:33-173: eivps(G) = sum_isa vps(is,G) e^{-iG.R},
 * eirop(G) = sum_isa rhops(is,G) e^{-iG.R}; vps, rhops (nsp, nhg), frac (3, nat) */
void or_eicalc(int nhg, const int* nr, const int* inyh, int nsp, const int* na, const double* vps,
               const double* rhops, const double* frac, double* eivpsd, double* eiropd) {
  cplx* eivps = (cplx*)eivpsd;
  cplx* eirop = (cplx*)eiropd;
  for (int g = 0; g < nhg; ++g) {
    cplx a = 0, b = 0;
    int isa = 0;
    for (int is = 0; is < nsp; ++is)
      for (int ia = 0; ia < na[is]; ++ia, ++isa) {
        cplx e = phase(nr, inyh + 3 * (size_t)g, frac + 3 * (size_t)isa);
        a += vps[is + (size_t)nsp * g] * e;
        b += rhops[is + (size_t)nsp * g] * e;
      }
    eivps[g] = a;
    eirop[g] = b;
  }
}

/* ppener, src/ppener_utils.mod.F90:23-108 (real parts of eh, ei, ee, eps);
 * vtemp(G) = scg rho_tot + eivps (G != 0), scg(0) rho_tot (G = 0) */
void or_ppener(int nhg, int geq0, const double* rhogd, const double* eivpsd, const double* eiropd, const double* scg,
               double* eh, double* ei, double* ee, double* eps, double* vploc, double* vtempd) {
  const cplx* rhog = (const cplx*)rhogd;
  const cplx* eivps = (const cplx*)eivpsd;
  const cplx* eirop = (const cplx*)eiropd;
  cplx* vtemp = (cplx*)vtempd;
  double sh = 0, si = 0, se = 0, sp = 0;
  int g1 = 0;
  *vploc = 0.0;
  if (geq0) {
    cplx vp = eivps[0], rp = eirop[0], rhet = rhog[0], rt = rhet + rp;
    *vploc = creal(vp);
    sp = creal(0.5 * vp * conj(rhet));
    sh = 0.5 * scg[0] * creal(rt) * creal(rt);
    si = creal(0.5 * scg[0] * rp * rp);
    se = creal(0.5 * scg[0] * rhet * rhet);
    vtemp[0] = scg[0] * rt;
    g1 = 1;
  }
  for (int g = g1; g < nhg; ++g) {
    cplx vp = eivps[g], rp = eirop[g], rhet = rhog[g], rt = rhet + rp;
    cplx vcg = scg[g] * rt;
    vtemp[g] = vcg + vp;
    sh += creal(vcg * conj(rt));
    si += scg[g] * creal(rp * conj(rp));
    se += scg[g] * creal(rhet * conj(rhet));
    sp += creal(conj(rhet) * vp);
  }
  *eh = sh;
  *ei = si;
  *ee = se;
  *eps = sp;
}

/* potfor, src/potfor_utils.mod.F90:38-201: fion(k,isa) += 2 Omega tpiba
 * Re(e^{-iG.R} (rhops scg conj(rho_tot) + vps conj(rho_e)) i gk_k), G != 0 */
void or_potfor(int nhg, int geq0, double omega, double tpiba, const double* gk, const int* nr, const int* inyh,
               int nsp, const int* na, const double* vps, const double* rhops, const double* scg, const double* frac,
               const double* rhogd, const double* eiropd, double* fion) {
  const cplx* rhog = (const cplx*)rhogd;
  const cplx* eirop = (const cplx*)eiropd;
  const double omtp = 2.0 * omega * tpiba;
  int isa = 0;
  for (int is = 0; is < nsp; ++is)
    for (int ia = 0; ia < na[is]; ++ia, ++isa) {
      double f[3] = {0, 0, 0};
      for (int g = geq0 ? 1 : 0; g < nhg; ++g) {
        cplx rhet = rhog[g], rt = rhet + eirop[g];
        cplx t = phase(nr, inyh + 3 * (size_t)g, frac + 3 * (size_t)isa) *
                 (rhops[is + (size_t)nsp * g] * scg[g] * conj(rt) + vps[is + (size_t)nsp * g] * conj(rhet));
        for (int k = 0; k < 3; ++k) f[k] += creal(t * (I * gk[k + 3 * (size_t)g]));
      }
      for (int k = 0; k < 3; ++k) fion[k + 3 * (size_t)isa] += omtp * f[k];
    }
}

/* LDA exchange-correlation per electron and potential, Goedecker-Teter-Hutter
 * Pade form (CPMD's FUNCTIONAL LDA = mfxcc_is_pade, src/nfunc_utils.mod.F90:251-419;
 * PRB 54, 1703 (1996)): eps(rs) = -(a0 + a1 rs + a2 rs^2 + a3 rs^3) /
 * (b1 rs + b2 rs^2 + b3 rs^3 + b4 rs^4), V = eps - rs/3 d eps/d rs, rho floored at 1e-24 */
void or_lda_pade(double rho, double* eps, double* vxc) {
  static const double a0 = 0.4581652932831429, a1 = 2.217058676663745, a2 = 0.7405551735357053,
                      a3 = 0.01968227878617998;
  static const double b1 = 1.0, b2 = 4.504130959426697, b3 = 1.110667363742916, b4 = 0.02359291751427506;
  const double r = rho > 1e-24 ? rho : 1e-24;
  const double rs = cbrt(3.0 / (4.0 * M_PI * r));
  const double num = a0 + rs * (a1 + rs * (a2 + rs * a3));
  const double den = rs * (b1 + rs * (b2 + rs * (b3 + rs * b4)));
  const double dnum = a1 + rs * (2.0 * a2 + rs * 3.0 * a3);
  const double dden = b1 + rs * (2.0 * b2 + rs * (3.0 * b3 + rs * 4.0 * b4));
  const double e = -num / den;
  const double de = -(dnum * den - num * dden) / (den * den);
  *eps = e;
  *vxc = e - rs / 3.0 * de;
}

/* xcener over the valid mesh (nr1 x nr2 x nr3 of the kr1 x kr2 x kr3 array):
 * v += V_xc(rho); sxc = sum eps rho_floored; vsum = sum V_xc max(rho, 0) */
void or_xcener(const int* nr, const int* kr, const double* rho, double* v, double* sxc, double* vsum) {
  double s = 0.0, w = 0.0;
  for (int z = 0; z < nr[2]; ++z)
    for (int y = 0; y < nr[1]; ++y)
      for (int x = 0; x < nr[0]; ++x) {
        const size_t e = x + (size_t)kr[0] * (y + (size_t)kr[1] * z);
        const double r = rho[e] > 1e-24 ? rho[e] : 1e-24;
        double ex, vx;
        or_lda_pade(r, &ex, &vx);
        v[e] += vx;
        s += ex * r;
        w += vx * (rho[e] > 0.0 ? rho[e] : 0.0);
      }
  *sxc = s;
  *vsum = w;
}

/* Ewald real-space sum esr and self energy eself (rpiiint, src/rpiiint_utils.mod.F90:29-200;
 * eself src/rinforce_utils.mod.F90:239-249); ht(3,3) column-major with ht(i,:) = a_i;
 * fion (3, nat) accumulates; pairs further than 20 rckj skipped */
void or_ewald(int nsp, const int* na, const double* zv, const double* raggio, const double* tau, const double* ht,
              int iesr, double* esr, double* eself, double* fion) {
  int nat = 0;
  for (int is = 0; is < nsp; ++is) nat += na[is];
  int* sp = malloc(sizeof(int) * (nat > 0 ? nat : 1));
  for (int is = 0, isa = 0; is < nsp; ++is)
    for (int ia = 0; ia < na[is]; ++ia) sp[isa++] = is;
  /* inverse cell for minimum-image differences */
  double a[3][3], inv[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a[i][j] = ht[i + 3 * j];   /* a[i] = lattice vector i */
  const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                     a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
  inv[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / det;
  inv[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
  inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
  inv[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / det;
  inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
  inv[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
  inv[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / det;
  inv[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
  inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
  double e = 0.0;
  for (int l = 0; l < nat; ++l)
    for (int m = l; m < nat; ++m) {
      const int k = sp[l], j = sp[m];
      const double zv2 = zv[k] * zv[j];
      if (fabs(zv2) < 1e-10) continue;
      const double rckj = sqrt(raggio[k] * raggio[k] + raggio[j] * raggio[j]);
      double d[3] = {0, 0, 0};
      const int tzero = (l == m);
      if (!tzero) {
        double dd[3], fr[3];
        for (int q = 0; q < 3; ++q) dd[q] = tau[q + 3 * l] - tau[q + 3 * m];
        for (int i = 0; i < 3; ++i) fr[i] = dd[0] * inv[0][i] + dd[1] * inv[1][i] + dd[2] * inv[2][i];
        for (int i = 0; i < 3; ++i) fr[i] -= nearbyint(fr[i]);
        for (int q = 0; q < 3; ++q) d[q] = fr[0] * a[0][q] + fr[1] * a[1][q] + fr[2] * a[2][q];
      }
      for (int ix = -iesr; ix <= iesr; ++ix)
        for (int iy = -iesr; iy <= iesr; ++iy)
          for (int iz = -iesr; iz <= iesr; ++iz) {
            if (tzero && ix == 0 && iy == 0 && iz == 0) continue;
            double r[3];
            for (int q = 0; q < 3; ++q) r[q] = d[q] + ix * a[0][q] + iy * a[1][q] + iz * a[2][q];
            const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            const double rl = sqrt(r2);
            const double arg = rl / rckj;
            if (arg > 20.0) continue;
            const double w = tzero ? 0.5 : 1.0;
            const double add = zv2 * erfc(arg) / rl;
            e += add * w;
            if (fion) {
              const double pre = 2.0 * zv2 / sqrt(M_PI) * exp(-arg * arg) / rckj;
              const double rep = w * (add + pre) / r2;
              for (int q = 0; q < 3; ++q) {
                fion[q + 3 * l] += rep * r[q];
                fion[q + 3 * m] -= rep * r[q];
              }
            }
          }
    }
  double s = 0.0;
  for (int is = 0; is < nsp; ++is) s += na[is] * zv[is] * zv[is] / raggio[is];
  *esr = e;
  *eself = s / sqrt(2.0 * M_PI);
  free(sp);
}

/* vofrho for a closed-shell Gamma density (vofrhoa + vofrhob, no core
 * correction, no gradient correction), dims = (nr1,nr2,nr3,kr1,kr2,kr3):
 *   rho (kr1,kr2,kr3) -> rho(G) (forward FFT, 1/N), eicalc, ppener, potfor (if
 *   fion), v(r) = Re invfft(vtemp on +-G) + V_xc(rho).
 * out[0..6] = ehep, ei*Omega, ee*Omega, epseu, vploc, exc, vxc (src/vofrhoa:96-122, vofrhob:459-461) */
void or_vofrho(const int* dims, int nhg, const int* inyh, const double* hg, const double* gk, double tpiba2,
               double omega, int geq0, int nsp, const int* na, const double* vps, const double* rhops,
               const double* frac, const double* rho, double* v, double* out, double* fion, double* rhog_out) {
  const int nr[3] = {dims[0], dims[1], dims[2]};
  const int kr[3] = {dims[3], dims[4], dims[5]};
  const size_t nnr = (size_t)kr[0] * kr[1] * kr[2];
  const double nn = (double)nr[0] * nr[1] * nr[2];
  cplx* w = calloc(nnr, sizeof(cplx));
  for (size_t e = 0; e < nnr; ++e) w[e] = rho[e];
  or_fft3d(nr[0], nr[1], nr[2], kr[0], kr[1], (double*)w, -1);
  cplx* rhog = malloc(sizeof(cplx) * nhg);
  cplx* eivps = malloc(sizeof(cplx) * nhg);
  cplx* eirop = malloc(sizeof(cplx) * nhg);
  cplx* vtemp = malloc(sizeof(cplx) * nhg);
  double* scg = malloc(sizeof(double) * nhg);
  for (int g = 0; g < nhg; ++g) {
    int idx[3];
    for (int k = 0; k < 3; ++k) {
      int m = inyh[k + 3 * g] - (nr[k] / 2 + 1);
      idx[k] = ((m % nr[k]) + nr[k]) % nr[k];
    }
    rhog[g] = w[idx[0] + (size_t)kr[0] * (idx[1] + (size_t)kr[1] * idx[2])] / nn;
    scg[g] = (geq0 && g == 0) ? 0.0 : 4.0 * M_PI / (tpiba2 * hg[g]);
  }
  or_eicalc(nhg, nr, inyh, nsp, na, vps, rhops, frac, (double*)eivps, (double*)eirop);
  double eh, ei, ee, eps, vploc;
  or_ppener(nhg, geq0, (double*)rhog, (double*)eivps, (double*)eirop, scg, &eh, &ei, &ee, &eps, &vploc,
            (double*)vtemp);
  if (fion)
    or_potfor(nhg, geq0, omega, sqrt(tpiba2), gk, nr, inyh, nsp, na, vps, rhops, scg, frac, (double*)rhog,
              (double*)eirop, fion);
  memset(w, 0, sizeof(cplx) * nnr);
  for (int g = 0; g < nhg; ++g) {
    int ip[3], im[3];
    for (int k = 0; k < 3; ++k) {
      int m = inyh[k + 3 * g] - (nr[k] / 2 + 1);
      ip[k] = ((m % nr[k]) + nr[k]) % nr[k];
      im[k] = ((-m % nr[k]) + nr[k]) % nr[k];
    }
    w[im[0] + (size_t)kr[0] * (im[1] + (size_t)kr[1] * im[2])] = conj(vtemp[g]);
    w[ip[0] + (size_t)kr[0] * (ip[1] + (size_t)kr[1] * ip[2])] = vtemp[g];
  }
  or_fft3d(nr[0], nr[1], nr[2], kr[0], kr[1], (double*)w, +1);
  for (size_t e = 0; e < nnr; ++e) v[e] = creal(w[e]);
  double sxc, vsum;
  or_xcener(nr, kr, rho, v, &sxc, &vsum);
  out[0] = eh * omega;
  out[1] = ei * omega;
  out[2] = ee * omega;
  out[3] = 2.0 * eps * omega;
  out[4] = vploc;
  out[5] = sxc * omega / nn;
  out[6] = vsum * omega / nn;
  if (rhog_out) memcpy(rhog_out, rhog, sizeof(cplx) * nhg);
  free(w); free(rhog); free(eivps); free(eirop); free(vtemp); free(scg);
}
