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
