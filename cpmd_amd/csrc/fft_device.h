// Device-side 1D transforms for gfx950, templated on the arithmetic type
// (double for the FP64 path, float for the mixed-precision path).
//
// Sign convention (CPMD, src/mltfft_utils.mod.F90:513-517): the inverse
// transform (G -> r) is e^{+i 2 pi k x / n}, unscaled; the forward transform
// (r -> G) is e^{-i 2 pi k x / n}; the 1/(n1 n2 n3) scale is applied by the
// caller after the last forward pass.
//
// Two engines, both cooperative over one workgroup of kThreads threads on a
// batch of lines held in LDS (line stride `lds`):
//  * fft_fast<T, N, DIR>: lengths N = A*B known at compile time (A, B <= 16).
//    Four-step: each thread does one length-A DFT in registers over x[B n1 + n2]
//    (stride B), multiplies by w_N^{n2 k1}, the lines are transposed in LDS and
//    each thread does one length-B DFT in registers.  Two LDS round trips per
//    transform.  Register DFTs of composite length recurse at compile time into
//    radix 2/3/4/5/7 butterflies with exact constant twiddles (twiddle_consts.h).
//  * fft_lines<T, DIR>: any 2^a 3^b 5^c 7^d length from a runtime plan: mixed-radix
//    Stockham stages; stage s with radix R and p = product of previous radices
//    maps butterfly i of a line to inputs x[i + q n/R], twiddles w^{q k n/(pR)}
//    (k = i mod p) and outputs y[(i-k) R + k + s p], in place with two barriers.
#pragma once

#include <hip/hip_runtime.h>

#include "mc_internal.h"
#include "twiddle_consts.h"

namespace mc {

constexpr int kThreads = 256;   // threads per workgroup of every transform kernel
constexpr int kMaxBfly = 4;     // butterflies per thread and stage (generic engine)

template <typename T> struct V2;
template <> struct V2<double> { using type = double2; };
template <> struct V2<float> { using type = float2; };
template <typename T> using v2 = typename V2<T>::type;

template <typename V>
__device__ __forceinline__ V cadd(V a, V b) { V r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename V>
__device__ __forceinline__ V csub(V a, V b) { V r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
template <typename V>
__device__ __forceinline__ V cscale(V a, decltype(a.x) s) { V r; r.x = a.x * s; r.y = a.y * s; return r; }
template <typename V>
__device__ __forceinline__ V cmk(decltype(V().x) x, decltype(V().x) y) { V r; r.x = x; r.y = y; return r; }

// a * w when DIR < 0 (forward), a * conj(w) when DIR > 0 (inverse); w = exp(-2 pi i t/n)
template <int DIR, typename V>
__device__ __forceinline__ V twmul(V a, V w) {
  if (DIR < 0) return cmk<V>(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
  return cmk<V>(a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y);
}
// a * DIR * i
template <int DIR, typename V>
__device__ __forceinline__ V mul_di(V a) {
  if (DIR < 0) return cmk<V>(a.y, -a.x);
  return cmk<V>(-a.y, a.x);
}

// ---------------------------------------------------------------------------
// radix butterflies in registers
// ---------------------------------------------------------------------------
template <int DIR, typename V>
__device__ __forceinline__ void bfly2(V* v) {
  V a = v[0], b = v[1];
  v[0] = cadd(a, b); v[1] = csub(a, b);
}
template <int DIR, typename V>
__device__ __forceinline__ void bfly4(V* v) {
  V t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
  V t2 = cadd(v[1], v[3]), t3 = mul_di<DIR>(csub(v[1], v[3]));
  v[0] = cadd(t0, t2); v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3); v[3] = csub(t1, t3);
}
template <int DIR, typename V>
__device__ __forceinline__ void bfly3(V* v) {
  using T = decltype(V().x);
  const T s3 = (T)0.86602540378443864676;   // sin(2 pi/3)
  V t1 = cadd(v[1], v[2]);
  V t2 = csub(v[1], v[2]);
  V y0 = cadd(v[0], t1);
  V m = cmk<V>(v[0].x - (T)0.5 * t1.x, v[0].y - (T)0.5 * t1.y);
  V r = mul_di<DIR>(cscale(t2, s3));
  v[0] = y0; v[1] = cadd(m, r); v[2] = csub(m, r);
}
template <int DIR, typename V>
__device__ __forceinline__ void bfly5(V* v) {
  using T = decltype(V().x);
  const T c1 = (T)0.30901699437494742410;    // cos(2 pi/5)
  const T c2 = (T)-0.80901699437494742410;   // cos(4 pi/5)
  const T s1 = (T)0.95105651629515357212;    // sin(2 pi/5)
  const T s2 = (T)0.58778525229247312917;    // sin(4 pi/5)
  V t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  V t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  V y0 = cadd(v[0], cadd(t1, t2));
  V a1 = cmk<V>(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
  V a2 = cmk<V>(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
  V b1 = mul_di<DIR>(cmk<V>(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
  V b2 = mul_di<DIR>(cmk<V>(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
  v[0] = y0;
  v[1] = cadd(a1, b1); v[4] = csub(a1, b1);
  v[2] = cadd(a2, b2); v[3] = csub(a2, b2);
}

// 7-point DFT: t_k = v_k + v_{7-k}, u_k = v_k - v_{7-k} (k = 1..3);
// y_m = v_0 + sum_k cos(2 pi k m/7) t_k + DIR i sum_k sin(2 pi k m/7) u_k and
// y_{7-m} with the sine term negated (m = 1..3)
template <int DIR, typename V>
__device__ __forceinline__ void bfly7(V* v) {
  using T = decltype(V().x);
  V t[3], u[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    t[k] = cadd(v[k + 1], v[6 - k]);
    u[k] = csub(v[k + 1], v[6 - k]);
  }
  const V x0 = v[0];
  v[0] = cadd(x0, cadd(t[0], cadd(t[1], t[2])));
#pragma unroll
  for (int m = 1; m <= 3; ++m) {
    V a = x0, b = cmk<V>((T)0, (T)0);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int j = ((k + 1) * m) % 7;
      const T c = (T)TwC<7>::c[j], s = (T)TwC<7>::s[j];
      a = cmk<V>(a.x + c * t[k].x, a.y + c * t[k].y);
      b = cmk<V>(b.x + s * u[k].x, b.y + s * u[k].y);
    }
    b = mul_di<DIR>(b);
    v[m] = cadd(a, b);
    v[7 - m] = csub(a, b);
  }
}

template <int R>
struct Split {   // R = P * Q for the register recursion
  static constexpr int P = (R % 4 == 0 && R != 4) ? 4
                           : (R % 3 == 0 && R != 3) ? 3
                           : (R % 2 == 0 && R != 2) ? 2
                           : (R % 5 == 0 && R != 5) ? 5
                                                    : 7;
  static constexpr int Q = R / P;
};

// v[j] *= W_R^e (DIR < 0) or conj (DIR > 0), e compile-time
template <int R, int DIR, typename V>
__device__ __forceinline__ V twc(V a, int e) {
  using T = decltype(V().x);
  const T c = (T)TwC<R>::c[e], s = (T)TwC<R>::s[e];
  // W = c - i s
  if (DIR < 0) return cmk<V>(a.x * c + a.y * s, a.y * c - a.x * s);
  return cmk<V>(a.x * c - a.y * s, a.y * c + a.x * s);
}

// in-register DFT of length R: v[k] <- sum_n v[n] w^{DIR n k}
template <int R, int DIR, typename V>
__device__ __forceinline__ void dftn(V* v) {
  if constexpr (R == 1) {
    return;
  } else if constexpr (R == 2) {
    bfly2<DIR>(v);
  } else if constexpr (R == 3) {
    bfly3<DIR>(v);
  } else if constexpr (R == 4) {
    bfly4<DIR>(v);
  } else if constexpr (R == 5) {
    bfly5<DIR>(v);
  } else if constexpr (R == 7) {
    bfly7<DIR>(v);
  } else {
    constexpr int P = Split<R>::P, Q = Split<R>::Q;
    V y[R];
#pragma unroll
    for (int n2 = 0; n2 < Q; ++n2) {
      V u[P];
#pragma unroll
      for (int n1 = 0; n1 < P; ++n1) u[n1] = v[Q * n1 + n2];
      dftn<P, DIR>(u);
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) y[n2 * P + k1] = ((n2 * k1) % R == 0) ? u[k1] : twc<R, DIR>(u[k1], (n2 * k1) % R);
    }
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) {
      V u[Q];
#pragma unroll
      for (int n2 = 0; n2 < Q; ++n2) u[n2] = y[n2 * P + k1];
      dftn<Q, DIR>(u);
#pragma unroll
      for (int k2 = 0; k2 < Q; ++k2) v[k1 + P * k2] = u[k2];
    }
  }
}

// 16-point register DFT for the band of the wavefunction sphere (frequencies
// |m| < N/4 of an N = 16 B transform sit at k = 0..3 and 12..15 of the
// four-step's 16-point DFTs).  IN: the inputs 4..11 are zero and not read
// (the first radix-4 stage reduces to a +- d, a -+ DIR i d); OUT: only the
// outputs 0..3 and 12..15 are formed (4..11 keep what they held).  On the
// entries it forms it does dftn<16>'s arithmetic, so the values agree with it
// (up to the sign of zero results).
template <int DIR, bool IN, bool OUT, typename V>
__device__ __forceinline__ void dft16_band(V* v) {
  V y[16];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    V u[4];
    if constexpr (IN) {
      const V a = v[n2], d = v[12 + n2];
      const V md = mul_di<DIR>(d);
      u[0] = cadd(a, d);
      u[1] = csub(a, md);
      u[2] = csub(a, d);
      u[3] = cadd(a, md);
    } else {
#pragma unroll
      for (int n1 = 0; n1 < 4; ++n1) u[n1] = v[4 * n1 + n2];
      bfly4<DIR>(u);
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) y[n2 * 4 + k1] = ((n2 * k1) % 16 == 0) ? u[k1] : twc<16, DIR>(u[k1], (n2 * k1) % 16);
  }
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    V u[4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) u[n2] = y[n2 * 4 + k1];
    if constexpr (OUT) {
      const V t0 = cadd(u[0], u[2]), t1 = csub(u[0], u[2]);
      const V t2 = cadd(u[1], u[3]), t3 = mul_di<DIR>(csub(u[1], u[3]));
      v[k1] = cadd(t0, t2);
      v[k1 + 12] = csub(t1, t3);
    } else {
      bfly4<DIR>(u);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = u[k2];
    }
  }
}

// ---------------------------------------------------------------------------
// compile-time four-step engine
// ---------------------------------------------------------------------------
template <int N> struct FastShape { static constexpr int A = 0, B = 0; };
#define MC_FAST_SHAPE(n, a, b) \
  template <> struct FastShape<n> { static constexpr int A = a, B = b; };
MC_FAST_SHAPE(16, 4, 4)
MC_FAST_SHAPE(18, 6, 3)
MC_FAST_SHAPE(24, 6, 4)
MC_FAST_SHAPE(54, 9, 6)
MC_FAST_SHAPE(128, 16, 8)
MC_FAST_SHAPE(192, 16, 12)
MC_FAST_SHAPE(224, 16, 14)
MC_FAST_SHAPE(240, 16, 15)
#undef MC_FAST_SHAPE
// the lengths with a compile-time engine (host dispatch lists the same)
#define MC_FAST_SIZES(X) X(16) X(18) X(24) X(54) X(128) X(192) X(224) X(240)

// In-place transform of `nlines` lines (line stride lds) with tw = LDS copy of
// the N-point table; nlines * A <= kThreads.  Begins and ends with a barrier
// point (callers synchronise before, this returns after a __syncthreads).
template <typename T, int N, int DIR>
__device__ __forceinline__ void fft_fast(v2<T>* buf, int nlines, int lds, const v2<T>* tw) {
  using V = v2<T>;
  constexpr int A = FastShape<N>::A, B = FastShape<N>::B;
  static_assert(A * B == N && A >= 1 && B >= 1, "unsupported fast length");
  V v[A];
  const int t = threadIdx.x;
  // step 1: thread (line, n2) transforms x[B n1 + n2], n1 < A
  const bool act1 = t < nlines * B;
  const int l1 = act1 ? t / B : 0, n2 = act1 ? t - l1 * B : 0;
  V* x1 = buf + l1 * lds;
  if (act1) {
#pragma unroll
    for (int n1 = 0; n1 < A; ++n1) v[n1] = x1[B * n1 + n2];
    dftn<A, DIR>(v);
#pragma unroll
    for (int k1 = 1; k1 < A; ++k1) v[k1] = twmul<DIR>(v[k1], tw[n2 * k1]);
  }
  __syncthreads();
  if (act1) {
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) x1[k1 * B + n2] = v[k1];
  }
  __syncthreads();
  // step 2: thread (line, k1) transforms y[n2][k1], n2 < B
  const bool act2 = t < nlines * A;
  const int l2 = act2 ? t / A : 0, k1 = act2 ? t - l2 * A : 0;
  V* x2 = buf + l2 * lds;
  V u[B];
  if (act2) {
#pragma unroll
    for (int m = 0; m < B; ++m) u[m] = x2[k1 * B + m];
    dftn<B, DIR>(u);
  }
  __syncthreads();
  if (act2) {
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) x2[k1 + A * k2] = u[k2];
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// runtime-plan Stockham engine
// ---------------------------------------------------------------------------
template <int R, int DIR, typename V>
__device__ __forceinline__ void fft_stage(V* buf, int nlines, int lds, int n, int p, const V* tw) {
  const int nb = n / R;
  const int total = nlines * nb;
  const int tstride = n / (p * R);
  V v[kMaxBfly][R];
  int dst[kMaxBfly];
#pragma unroll
  for (int t = 0; t < kMaxBfly; ++t) {
    const int b = threadIdx.x + t * kThreads;
    dst[t] = -1;
    if (b < total) {
      const int line = b / nb;
      const int i = b - line * nb;
      const int k = i % p;
      V* x = buf + line * lds;
#pragma unroll
      for (int q = 0; q < R; ++q) v[t][q] = x[i + q * nb];
      if (p > 1) {
#pragma unroll
        for (int q = 1; q < R; ++q) v[t][q] = twmul<DIR>(v[t][q], tw[q * k * tstride]);
      }
      dftn<R, DIR>(v[t]);
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

template <typename T, int DIR>
__device__ void fft_lines(v2<T>* buf, int nlines, int lds, const FftPlan& pl, const v2<T>* tw) {
  int p = 1;
  for (int s = 0; s < pl.nf; ++s) {
    const int R = pl.rad[s];
    if (R == 4) fft_stage<4, DIR>(buf, nlines, lds, pl.n, p, tw);
    else if (R == 2) fft_stage<2, DIR>(buf, nlines, lds, pl.n, p, tw);
    else if (R == 3) fft_stage<3, DIR>(buf, nlines, lds, pl.n, p, tw);
    else if (R == 5) fft_stage<5, DIR>(buf, nlines, lds, pl.n, p, tw);
    else fft_stage<7, DIR>(buf, nlines, lds, pl.n, p, tw);
    p *= R;
  }
}

// one entry for both engines: N > 0 compile-time length, N == 0 runtime plan
template <typename T, int N, int DIR>
__device__ __forceinline__ void fft_any(v2<T>* buf, int nlines, int lds, const FftPlan& pl, const v2<T>* tw) {
  if constexpr (N > 0) fft_fast<T, N, DIR>(buf, nlines, lds, tw);
  else fft_lines<T, DIR>(buf, nlines, lds, pl, tw);
}

// Largest number of lines of length n one workgroup may transform at once (generic engine).
inline int max_lines(const FftPlan& pl) {
  int rmin = 5;
  for (int s = 0; s < pl.nf; ++s) rmin = pl.rad[s] < rmin ? pl.rad[s] : rmin;
  return (kMaxBfly * kThreads * rmin) / pl.n;
}

// Host side: the fast-engine factor A of length n (0 when n has no compile-time engine).
inline int fast_a(int n) {
  switch (n) {
#define MC_CASE(x) case x: return FastShape<x>::A;
    MC_FAST_SIZES(MC_CASE)
#undef MC_CASE
    default: return 0;
  }
}

}  // namespace mc
