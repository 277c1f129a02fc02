// Device-side 1D transforms for gfx950: mixed-radix (2,3,4,5) Stockham
// autosort passes over a batch of lines held in LDS, executed cooperatively
// by one workgroup of kThreads threads.
//
// Sign convention (CPMD, src/mltfft_utils.mod.F90:513-517): the inverse
// transform (G -> r) is e^{+i 2 pi k x / n}, unscaled; the forward transform
// (r -> G) is e^{-i 2 pi k x / n}; the 1/(n1 n2 n3) scale is applied by the
// caller after the last forward pass.
//
// Stage s with radix R and p = product of the previous radices maps butterfly
// i in [0, n/R) of each line to inputs x[i + q n/R] (q < R), twiddles
// w^{q k n/(p R)} with k = i mod p, and outputs y[(i-k) R + k + s p].  Every
// thread keeps its butterflies in registers between the read and the write of
// a stage, so a stage is done in place with two barriers.
#pragma once

#include <hip/hip_runtime.h>

#include "mc_internal.h"

namespace mc {

constexpr int kThreads = 256;   // threads per workgroup of every transform kernel
constexpr int kMaxBfly = 4;     // butterflies per thread and stage (register budget)

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// multiply by conj(w) when DIR > 0 (inverse), by w when DIR < 0 (forward);
// the table holds w = exp(-2 pi i t / n).
template <int DIR>
__device__ __forceinline__ double2 twmul(double2 a, double2 w) {
  if (DIR < 0) return make_double2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
  return make_double2(a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y);
}
// multiply by DIR*i
template <int DIR>
__device__ __forceinline__ double2 mul_di(double2 a) {
  if (DIR < 0) return make_double2(a.y, -a.x);   // -i * a
  return make_double2(-a.y, a.x);                // +i * a
}

template <int R, int DIR>
__device__ __forceinline__ void dft(double2* v);

template <>
__device__ __forceinline__ void dft<2, -1>(double2* v) {
  double2 a = v[0], b = v[1];
  v[0] = cadd(a, b); v[1] = csub(a, b);
}
template <>
__device__ __forceinline__ void dft<2, 1>(double2* v) {
  double2 a = v[0], b = v[1];
  v[0] = cadd(a, b); v[1] = csub(a, b);
}

template <int DIR>
__device__ __forceinline__ void dft4(double2* v) {
  double2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
  double2 t2 = cadd(v[1], v[3]), t3 = mul_di<DIR>(csub(v[1], v[3]));
  v[0] = cadd(t0, t2); v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3); v[3] = csub(t1, t3);
}
template <> __device__ __forceinline__ void dft<4, -1>(double2* v) { dft4<-1>(v); }
template <> __device__ __forceinline__ void dft<4, 1>(double2* v) { dft4<1>(v); }

template <int DIR>
__device__ __forceinline__ void dft3(double2* v) {
  const double s3 = 0.86602540378443864676;   // sin(2 pi/3)
  double2 t1 = cadd(v[1], v[2]);
  double2 t2 = csub(v[1], v[2]);
  double2 y0 = cadd(v[0], t1);
  double2 m = make_double2(v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y);
  double2 r = mul_di<DIR>(make_double2(s3 * t2.x, s3 * t2.y));
  v[0] = y0; v[1] = cadd(m, r); v[2] = csub(m, r);
}
template <> __device__ __forceinline__ void dft<3, -1>(double2* v) { dft3<-1>(v); }
template <> __device__ __forceinline__ void dft<3, 1>(double2* v) { dft3<1>(v); }

template <int DIR>
__device__ __forceinline__ void dft5(double2* v) {
  const double c1 = 0.30901699437494742410;    // cos(2 pi/5)
  const double c2 = -0.80901699437494742410;   // cos(4 pi/5)
  const double s1 = 0.95105651629515357212;    // sin(2 pi/5)
  const double s2 = 0.58778525229247312917;    // sin(4 pi/5)
  double2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  double2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  double2 y0 = cadd(v[0], cadd(t1, t2));
  double2 a1 = make_double2(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
  double2 a2 = make_double2(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
  double2 b1 = mul_di<DIR>(make_double2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
  double2 b2 = mul_di<DIR>(make_double2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
  v[0] = y0;
  v[1] = cadd(a1, b1); v[4] = csub(a1, b1);
  v[2] = cadd(a2, b2); v[3] = csub(a2, b2);
}
template <> __device__ __forceinline__ void dft<5, -1>(double2* v) { dft5<-1>(v); }
template <> __device__ __forceinline__ void dft<5, 1>(double2* v) { dft5<1>(v); }

// One Stockham stage on `nlines` lines of length n with line stride `lds`
// (complex elements) in LDS.  tw: LDS copy of the n-point table.
template <int R, int DIR>
__device__ __forceinline__ void fft_stage(double2* buf, int nlines, int lds, int n, int p,
                                          const double2* tw) {
  const int nb = n / R;
  const int total = nlines * nb;
  const int tstride = n / (p * R);
  double2 v[kMaxBfly][R];
  int dst[kMaxBfly];
#pragma unroll
  for (int t = 0; t < kMaxBfly; ++t) {
    const int b = threadIdx.x + t * kThreads;
    dst[t] = -1;
    if (b < total) {
      const int line = b / nb;
      const int i = b - line * nb;
      const int k = i % p;
      double2* x = buf + line * lds;
#pragma unroll
      for (int q = 0; q < R; ++q) v[t][q] = x[i + q * nb];
      if (p > 1) {
#pragma unroll
        for (int q = 1; q < R; ++q) v[t][q] = twmul<DIR>(v[t][q], tw[q * k * tstride]);
      }
      dft<R, DIR>(v[t]);
      dst[t] = line * lds + (i - k) * R + k;
    }
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < kMaxBfly; ++t) {
    if (dst[t] >= 0) {
#pragma unroll
      for (int q = 0; q < R; ++q) buf[dst[t] + q * p] = v[t][q];
    }
  }
  __syncthreads();
}

// Full transform of nlines lines; requires nlines * n / min_radix <= kMaxBfly * kThreads.
template <int DIR>
__device__ void fft_lines(double2* buf, int nlines, int lds, const FftPlan& pl, const double2* tw) {
  int p = 1;
  for (int s = 0; s < pl.nf; ++s) {
    const int R = pl.rad[s];
    if (R == 4) fft_stage<4, DIR>(buf, nlines, lds, pl.n, p, tw);
    else if (R == 2) fft_stage<2, DIR>(buf, nlines, lds, pl.n, p, tw);
    else if (R == 3) fft_stage<3, DIR>(buf, nlines, lds, pl.n, p, tw);
    else fft_stage<5, DIR>(buf, nlines, lds, pl.n, p, tw);
    p *= R;
  }
}

// Largest number of lines of length n one workgroup may transform at once.
inline int max_lines(const FftPlan& pl) {
  int rmin = 5;
  for (int s = 0; s < pl.nf; ++s) rmin = pl.rad[s] < rmin ? pl.rad[s] : rmin;
  return (kMaxBfly * kThreads * rmin) / pl.n;
}

}  // namespace mc
