/*
 * ORACLE -- the CPU baseline of bench.py and test infrastructure only.
 *
 * CPMD's CPU algorithm for rhoofr and vpsi over Gamma state pairs, with the
 * stick-sparse 3D transforms of fftnew (src/fftmain_utils.mod.F90:71-157,
 * sparse = .TRUE.) done by FFTW3-interface 1D transforms (the library
 * CPMD's __HAS_FFT_FFTW3 build calls, src/mltfft_utils.mod.F90:492-578; here
 * the FFTW3 interface of the image's Intel MKL), one state pair per OpenMP
 * thread:
 *   pack  : c1 + i c2 at +G, conj(c1) + i conj(c2) at -G into the sticks
 *           (src/state_utils.mod.F90:171-189);
 *   z pass: the sticks S x n3; y pass: the x-frequency columns Zw x n2 x n3;
 *   x pass: the full mesh n1 x n2 x n3 (inverse FFTW_BACKWARD, unscaled;
 *           forward FFTW_FORWARD scaled by 1/(n1 n2 n3) on the last pass);
 *   rhoofr: rho += f1/omega Re(psi)^2 + f2/omega Im(psi)^2
 *           (src/rhoofr_utils.mod.F90:305-409);
 *   vpsi  : psi *= V, forward, decode fp = psi(+G) + psi(-G), fm = psi(+G) -
 *           psi(-G), c2 += -f/2 (tpiba2 hg c + (Re fp, Im fm)) for the first
 *           state and (Im fp, -Re fm) for the second, f/2 = 0 -> 1
 *           (src/vpsi_utils.mod.F90:627-672).
 * Layout as the oracle: wavefunctions (ldc, nstate) complex in CPMD G order,
 * real-space arrays v(ix + kr1 (iy + kr2 iz)).  Checked against oracle.c in
 * tests/test_oracle.py.
 */
#include <complex.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* the FFTW3 interface (its public declarations) */
typedef double fftw_complex[2];
typedef struct fftw_plan_s* fftw_plan;
fftw_plan fftw_plan_many_dft(int rank, const int* n, int howmany, fftw_complex* in, const int* inembed, int istride,
                             int idist, fftw_complex* out, const int* onembed, int ostride, int odist, int sign,
                             unsigned flags);
void fftw_execute_dft(const fftw_plan p, fftw_complex* in, fftw_complex* out);
void fftw_destroy_plan(fftw_plan p);
#define FFTW_FORWARD (-1)
#define FFTW_BACKWARD (+1)
#define FFTW_ESTIMATE (1U << 6)

typedef double complex cplx;

typedef struct {
  int n1, n2, n3, kr1, kr2, kr3, ngw;
  int nstick, ncol;
  int *sp, *zp, *sm, *zm;     /* (ngw) stick and z index of +G and -G */
  int *scol, *sy;             /* (nstick) column and y of each stick */
  int *xcol;                  /* (ncol) x index of each column */
  const double* hg;
  double tpiba2, omega;
} cs_layout;

typedef struct {
  cplx *A, *B, *R;
  fftw_plan zi, zf, yi, yf, xi, xf;
} cs_work;

static int cmp_int(const void* a, const void* b) {
  const int x = *(const int*)a, y = *(const int*)b;
  return (x > y) - (x < y);
}

static int find_sorted(const int* v, int n, int key) {
  int lo = 0, hi = n - 1;
  while (lo <= hi) {
    const int mid = (lo + hi) / 2;
    if (v[mid] == key) return mid;
    if (v[mid] < key) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

/* the stick layout from the centred indices inyh(3, ngw) (1-based, nr/2 + 1 + m) */
static int layout_build(cs_layout* L, const int* dims, const int* inyh, const double* hg, double tpiba2,
                        double omega) {
  memset(L, 0, sizeof(*L));
  L->n1 = dims[0]; L->n2 = dims[1]; L->n3 = dims[2];
  L->kr1 = dims[3]; L->kr2 = dims[4]; L->kr3 = dims[5]; L->ngw = dims[6];
  L->hg = hg; L->tpiba2 = tpiba2; L->omega = omega;
  const int n1 = L->n1, n2 = L->n2, n3 = L->n3, ngw = L->ngw;
  int* keys = malloc(sizeof(int) * 2 * (size_t)ngw);
  L->zp = malloc(sizeof(int) * ngw); L->zm = malloc(sizeof(int) * ngw);
  int* kp = malloc(sizeof(int) * ngw);
  int* km = malloc(sizeof(int) * ngw);
  for (int g = 0; g < ngw; ++g) {
    const int m1 = inyh[3 * g] - (n1 / 2 + 1), m2 = inyh[3 * g + 1] - (n2 / 2 + 1), m3 = inyh[3 * g + 2] - (n3 / 2 + 1);
    const int xp = ((m1 % n1) + n1) % n1, yp = ((m2 % n2) + n2) % n2;
    const int xm = ((-m1 % n1) + n1) % n1, ym = ((-m2 % n2) + n2) % n2;
    kp[g] = xp * n2 + yp;
    km[g] = xm * n2 + ym;
    L->zp[g] = ((m3 % n3) + n3) % n3;
    L->zm[g] = ((-m3 % n3) + n3) % n3;
    keys[2 * g] = kp[g];
    keys[2 * g + 1] = km[g];
  }
  qsort(keys, 2 * (size_t)ngw, sizeof(int), cmp_int);
  int ns = 0;
  for (int i = 0; i < 2 * ngw; ++i)
    if (i == 0 || keys[i] != keys[i - 1]) keys[ns++] = keys[i];
  L->nstick = ns;
  L->sp = malloc(sizeof(int) * ngw); L->sm = malloc(sizeof(int) * ngw);
  for (int g = 0; g < ngw; ++g) {
    L->sp[g] = find_sorted(keys, ns, kp[g]);
    L->sm[g] = find_sorted(keys, ns, km[g]);
  }
  /* columns: the distinct x of the sticks (keys are sorted by x first) */
  L->scol = malloc(sizeof(int) * ns); L->sy = malloc(sizeof(int) * ns);
  L->xcol = malloc(sizeof(int) * ns);
  int nc = 0;
  for (int s = 0; s < ns; ++s) {
    const int x = keys[s] / n2;
    if (nc == 0 || L->xcol[nc - 1] != x) L->xcol[nc++] = x;
    L->scol[s] = nc - 1;
    L->sy[s] = keys[s] % n2;
  }
  L->ncol = nc;
  free(keys); free(kp); free(km);
  return 0;
}

static void layout_free(cs_layout* L) {
  free(L->sp); free(L->zp); free(L->sm); free(L->zm); free(L->scol); free(L->sy); free(L->xcol);
}

static void work_init(const cs_layout* L, cs_work* w) {
  const int n1 = L->n1, n2 = L->n2, n3 = L->n3;
  w->A = malloc(sizeof(cplx) * (size_t)L->nstick * n3);
  w->B = malloc(sizeof(cplx) * (size_t)L->ncol * n2 * n3);
  w->R = malloc(sizeof(cplx) * (size_t)n1 * n2 * n3);
  int n;
  n = n3;   /* z: S contiguous sticks */
  w->zi = fftw_plan_many_dft(1, &n, L->nstick, (fftw_complex*)w->A, NULL, 1, n3, (fftw_complex*)w->A, NULL, 1, n3,
                             FFTW_BACKWARD, FFTW_ESTIMATE);
  w->zf = fftw_plan_many_dft(1, &n, L->nstick, (fftw_complex*)w->A, NULL, 1, n3, (fftw_complex*)w->A, NULL, 1, n3,
                             FFTW_FORWARD, FFTW_ESTIMATE);
  n = n2;   /* y: one column [y][z], stride n3 */
  w->yi = fftw_plan_many_dft(1, &n, n3, (fftw_complex*)w->B, NULL, n3, 1, (fftw_complex*)w->B, NULL, n3, 1,
                             FFTW_BACKWARD, FFTW_ESTIMATE);
  w->yf = fftw_plan_many_dft(1, &n, n3, (fftw_complex*)w->B, NULL, n3, 1, (fftw_complex*)w->B, NULL, n3, 1,
                             FFTW_FORWARD, FFTW_ESTIMATE);
  n = n1;   /* x: the mesh [x][y][z], stride n2 n3 */
  w->xi = fftw_plan_many_dft(1, &n, n2 * n3, (fftw_complex*)w->R, NULL, n2 * n3, 1, (fftw_complex*)w->R, NULL,
                             n2 * n3, 1, FFTW_BACKWARD, FFTW_ESTIMATE);
  w->xf = fftw_plan_many_dft(1, &n, n2 * n3, (fftw_complex*)w->R, NULL, n2 * n3, 1, (fftw_complex*)w->R, NULL,
                             n2 * n3, 1, FFTW_FORWARD, FFTW_ESTIMATE);
}

static void work_free(cs_work* w) {
  fftw_destroy_plan(w->zi); fftw_destroy_plan(w->zf); fftw_destroy_plan(w->yi);
  fftw_destroy_plan(w->yf); fftw_destroy_plan(w->xi); fftw_destroy_plan(w->xf);
  free(w->A); free(w->B); free(w->R);
}

/* psi(x, y, z) of the pair (a, b) (b may be NULL) in w->R */
static void to_real(const cs_layout* L, cs_work* w, const cplx* a, const cplx* b) {
  const int n2 = L->n2, n3 = L->n3;
  const size_t col = (size_t)n2 * n3;
  memset(w->A, 0, sizeof(cplx) * (size_t)L->nstick * n3);
  for (int g = 0; g < L->ngw; ++g) {
    const cplx ca = a[g], cb = b ? b[g] : 0.0;
    w->A[(size_t)L->sp[g] * n3 + L->zp[g]] = ca + I * cb;
    if (g > 0) w->A[(size_t)L->sm[g] * n3 + L->zm[g]] = conj(ca) + I * conj(cb);
  }
  fftw_execute_dft(w->zi, (fftw_complex*)w->A, (fftw_complex*)w->A);
  memset(w->B, 0, sizeof(cplx) * (size_t)L->ncol * col);
  for (int s = 0; s < L->nstick; ++s)
    memcpy(w->B + (size_t)L->scol[s] * col + (size_t)L->sy[s] * n3, w->A + (size_t)s * n3, sizeof(cplx) * n3);
  for (int c = 0; c < L->ncol; ++c)
    fftw_execute_dft(w->yi, (fftw_complex*)(w->B + (size_t)c * col), (fftw_complex*)(w->B + (size_t)c * col));
  memset(w->R, 0, sizeof(cplx) * (size_t)L->n1 * col);
  for (int c = 0; c < L->ncol; ++c)
    memcpy(w->R + (size_t)L->xcol[c] * col, w->B + (size_t)c * col, sizeof(cplx) * col);
  fftw_execute_dft(w->xi, (fftw_complex*)w->R, (fftw_complex*)w->R);
}

/* forward transform of w->R; w->A holds the sticks, scaled by 1/(n1 n2 n3) */
static void to_sticks(const cs_layout* L, cs_work* w) {
  const int n2 = L->n2, n3 = L->n3;
  const size_t col = (size_t)n2 * n3;
  fftw_execute_dft(w->xf, (fftw_complex*)w->R, (fftw_complex*)w->R);
  for (int c = 0; c < L->ncol; ++c)
    memcpy(w->B + (size_t)c * col, w->R + (size_t)L->xcol[c] * col, sizeof(cplx) * col);
  for (int c = 0; c < L->ncol; ++c)
    fftw_execute_dft(w->yf, (fftw_complex*)(w->B + (size_t)c * col), (fftw_complex*)(w->B + (size_t)c * col));
  for (int s = 0; s < L->nstick; ++s)
    memcpy(w->A + (size_t)s * n3, w->B + (size_t)L->scol[s] * col + (size_t)L->sy[s] * n3, sizeof(cplx) * n3);
  fftw_execute_dft(w->zf, (fftw_complex*)w->A, (fftw_complex*)w->A);
}

/* rhoofr over the state pairs [0, npairs): rho (nnr1, CPMD layout) = sum f/omega psi^2 */
int cs_rhoofr(const int* dims, const int* inyh, const double* hg, double tpiba2, double omega, const double* c0d,
              int ldc, const double* f, int nstate, int npairs, double* rho) {
  cs_layout L;
  layout_build(&L, dims, inyh, hg, tpiba2, omega);
  const cplx* c0 = (const cplx*)c0d;
  const int n1 = L.n1, n2 = L.n2, n3 = L.n3;
  const size_t nnn = (size_t)n1 * n2 * n3;
  const int nt = omp_get_max_threads();
  cs_work* ws = malloc(sizeof(cs_work) * nt);
  double* acc = calloc((size_t)nt * nnn, sizeof(double));
  for (int t = 0; t < nt; ++t) work_init(&L, &ws[t]);
#pragma omp parallel for schedule(dynamic, 1)
  for (int p = 0; p < npairs; ++p) {
    const int t = omp_get_thread_num();
    const int is1 = 2 * p, is2 = is1 + 1;
    const int has2 = is2 < nstate;
    const double fa = f[is1] / omega, fb = has2 ? f[is2] / omega : 0.0;
    if (fa == 0.0 && fb == 0.0) continue;
    to_real(&L, &ws[t], c0 + (size_t)is1 * ldc, has2 ? c0 + (size_t)is2 * ldc : NULL);
    double* r = acc + (size_t)t * nnn;
    const cplx* R = ws[t].R;
    for (size_t e = 0; e < nnn; ++e) {
      const double x = creal(R[e]), y = cimag(R[e]);
      r[e] += fa * x * x + fb * y * y;
    }
  }
  /* sum the thread accumulators into CPMD's layout */
  memset(rho, 0, sizeof(double) * (size_t)L.kr1 * L.kr2 * L.kr3);
#pragma omp parallel for
  for (int x = 0; x < n1; ++x)
    for (int y = 0; y < n2; ++y)
      for (int z = 0; z < n3; ++z) {
        const size_t e = ((size_t)x * n2 + y) * n3 + z;
        double s = 0.0;
        for (int t = 0; t < nt; ++t) s += acc[(size_t)t * nnn + e];
        rho[x + (size_t)L.kr1 * (y + (size_t)L.kr2 * z)] = s;
      }
  for (int t = 0; t < nt; ++t) work_free(&ws[t]);
  free(ws); free(acc);
  layout_free(&L);
  return 0;
}

/* vpsi over the state pairs [0, npairs): c2 += -f/2 (tpiba2 hg c0 + P_G[2 V psi]) */
int cs_vpsi(const int* dims, const int* inyh, const double* hg, double tpiba2, double omega, const double* c0d,
            int ldc, double* c2d, const double* f, const double* vpot, int nstate, int npairs) {
  cs_layout L;
  layout_build(&L, dims, inyh, hg, tpiba2, omega);
  const cplx* c0 = (const cplx*)c0d;
  cplx* c2 = (cplx*)c2d;
  const int n1 = L.n1, n2 = L.n2, n3 = L.n3;
  const size_t nnn = (size_t)n1 * n2 * n3;
  double* vt = malloc(sizeof(double) * nnn);   /* V in the [x][y][z] order of the mesh buffer */
#pragma omp parallel for
  for (int x = 0; x < n1; ++x)
    for (int y = 0; y < n2; ++y)
      for (int z = 0; z < n3; ++z)
        vt[((size_t)x * n2 + y) * n3 + z] = vpot[x + (size_t)L.kr1 * (y + (size_t)L.kr2 * z)];
  const int nt = omp_get_max_threads();
  cs_work* ws = malloc(sizeof(cs_work) * nt);
  for (int t = 0; t < nt; ++t) work_init(&L, &ws[t]);
  const double scale = 1.0 / (double)nnn;
#pragma omp parallel for schedule(dynamic, 1)
  for (int p = 0; p < npairs; ++p) {
    const int t = omp_get_thread_num();
    cs_work* w = &ws[t];
    const int is1 = 2 * p, is2 = is1 + 1;
    const int has2 = is2 < nstate;
    const cplx* a = c0 + (size_t)is1 * ldc;
    const cplx* b = has2 ? c0 + (size_t)is2 * ldc : NULL;
    to_real(&L, w, a, b);
    for (size_t e = 0; e < nnn; ++e) w->R[e] *= vt[e];
    to_sticks(&L, w);
    double fi = f[is1] * 0.5;
    if (fi == 0.0) fi = 1.0;
    double fip1 = has2 ? f[is2] * 0.5 : 0.0;
    if (fip1 == 0.0) fip1 = 1.0;
    for (int g = 0; g < L.ngw; ++g) {
      const cplx psin = w->A[(size_t)L.sp[g] * n3 + L.zp[g]] * scale;
      const cplx psii = w->A[(size_t)L.sm[g] * n3 + L.zm[g]] * scale;
      const cplx fp = psin + psii, fm = psin - psii;
      const double g2 = tpiba2 * hg[g];
      c2[(size_t)is1 * ldc + g] += -fi * (g2 * a[g] + (creal(fp) + I * cimag(fm)));
      if (has2) c2[(size_t)is2 * ldc + g] += -fip1 * (g2 * b[g] + (cimag(fp) - I * creal(fm)));
    }
  }
  for (int t = 0; t < nt; ++t) work_free(&ws[t]);
  free(ws); free(vt);
  layout_free(&L);
  return 0;
}

int cs_threads(void) { return omp_get_max_threads(); }
