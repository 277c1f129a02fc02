// rhoofr and vpsi on gfx950: the stick-sparse 3D transforms of Gamma state
// pairs, fused with the wavefunction pack, the |psi|^2 accumulation, the
// V(r) multiply and the Gamma decode + kinetic term.
//
// Per batch of P state pairs the data path through HBM is
//   zinv : C0 runs (stick order)  -> sticks  A[pair][slot][z]       (pack + z^-1)
//   yinv : A                       -> columns B[pair][z][col][y]     (y^-1)
//   xrho : B (all P pairs)         -> rho(r)  (x^-1, |psi|^2, rho read/written once per batch)
//   xvpsi: B (all P pairs), V(r)   -> B       (x^-1, *V, x^+1; V read once per batch)
//   yfwd : B                       -> A       (y^+1)
//   zfwd : A, C0, C2               -> C2      (z^+1, 1/(n1 n2 n3), decode, kinetic, c2 +=)
// Every kernel is templated on the arithmetic type T of the transforms and of
// the A/B workspaces (double: FP64 path; float: mixed-precision path; C0, C2,
// V and rho stay FP64 in HBM) and on the compile-time length N of its axis
// (0 = runtime plan).
// Reference: vpsi src/vpsi_utils.mod.F90:376-675 (one pair per iteration through
// invfftn/fwfftn, src/fftmain_utils.mod.F90:71-157) and rhoofr
// src/rhoofr_utils.mod.F90:305-409.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "fft_device.h"
#include "fft_launch.h"
#include "mc_internal.h"

namespace mc {

struct WParams {
  int n1, n2, n3, kr1, kr2, kr3;
  int ngw, ldg, n_own, n_slots, ncols, geq0;
  double tpiba2, omega;
  FftPlan p1, p2, p3;
  const int* ginfo;
  const double* hg;
  const int* own_start;
  const int* col_start;
  const int* col_slot;
  const int* col_gy;
  const int* col_gx;
  const double2* tw1;
  const double2* tw2;
  const double2* tw3;
};

constexpr int kMaxElem = 16;  // real-space elements per thread kept in registers (x pass)

// ---------------------------------------------------------------------------
// zinv: Gamma pack of a state pair (c1 + i c2 at +G, conj(c1) + i conj(c2) at
// -G, G=0 once; src/state_utils.mod.F90:171-189) into `lo` owner sticks and
// their mirrors, inverse z transform, write to A.  Optionally accumulates the
// per-state norms sum_G w_G |c|^2 (dotp, src/dotp_utils.mod.F90:58-101) and
// sum_G hg |c|^2 (kin_energy, src/kin_energy_utils.mod.F90:95-103) in FP64.
// ---------------------------------------------------------------------------
template <typename T, int N>
__global__ __launch_bounds__(kThreads) void k_zinv(WParams w, const double2* __restrict__ c0, int nstate,
                                                   int pair0, v2<T>* __restrict__ A, int lo,
                                                   double* __restrict__ red) {
  using V = v2<T>;
  const int n3 = N > 0 ? N : w.n3;
  V* tw = lds_base<T>();
  V* buf = tw + n3;
  const int pb = blockIdx.y;
  const int is1 = 2 * (pair0 + pb), is2 = is1 + 1;
  const bool has2 = is2 < nstate;
  const int o0 = blockIdx.x * lo;
  const int o1 = min(o0 + lo, w.n_own);
  const int nl = 2 * (o1 - o0);
  load_tw<T>(tw, w.tw3, n3);
  for (int e = threadIdx.x; e < nl * n3; e += kThreads) buf[e] = cmk<V>((T)0, (T)0);
  __syncthreads();
  const int j0 = w.own_start[o0], j1 = w.own_start[o1];
  const double2* ca = c0 + (size_t)is1 * w.ldg;
  const double2* cb = c0 + (size_t)is2 * w.ldg;
  double dot1 = 0.0, kin1 = 0.0, dot2 = 0.0, kin2 = 0.0;
  for (int j = j0 + threadIdx.x; j < j1; j += kThreads) {
    const int gi = w.ginfo[j];
    const int o = gi >> 12, l = gi & 4095;
    const int t = o - o0;
    const double2 a = ca[j];
    const double2 b = has2 ? cb[j] : make_double2(0.0, 0.0);
    buf[2 * t * n3 + l] = cmk<V>((T)(a.x - b.y), (T)(a.y + b.x));
    const bool g0 = (j == 0) && w.geq0;
    if (!g0) {
      const int ml = (o == 0) ? 2 * t : 2 * t + 1;
      buf[ml * n3 + (l == 0 ? 0 : n3 - l)] = cmk<V>((T)(a.x + b.y), (T)(b.x - a.y));
    }
    if (red) {
      const double na = a.x * a.x + a.y * a.y, nb = b.x * b.x + b.y * b.y;
      const double h = w.hg[j];
      dot1 += g0 ? a.x * a.x : 2.0 * na;
      dot2 += g0 ? b.x * b.x : 2.0 * nb;
      kin1 += h * na;
      kin2 += h * nb;
    }
  }
  __syncthreads();
  fft_any<T, N, 1>(buf, nl, n3, w.p3, tw);
  for (int e = threadIdx.x; e < nl * n3; e += kThreads) {
    const int line = e / n3, z = e - line * n3;
    const int o = o0 + (line >> 1);
    const bool mir = line & 1;
    if (mir && o == 0) continue;
    const int slot = mir ? w.n_own + o - 1 : o;
    A[((size_t)pb * w.n_slots + slot) * n3 + z] = buf[e];
  }
  if (red) {
    __shared__ double part[4][kThreads / 64];
    double v[4] = {warp_sum(dot1), warp_sum(kin1), warp_sum(dot2), warp_sum(kin2)};
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
    if (ln == 0)
      for (int k = 0; k < 4; ++k) part[k][wv] = v[k];
    __syncthreads();
    if (threadIdx.x < 4) {
      double s = 0.0;
      for (int q = 0; q < kThreads / 64; ++q) s += part[threadIdx.x][q];
      red[((size_t)pb * gridDim.x + blockIdx.x) * 4 + threadIdx.x] = s;
    }
  }
}

// ---------------------------------------------------------------------------
// y pass over one x-frequency column and a tile of z planes.
// DIR=+1: A (sticks of the column) -> B; DIR=-1: B -> A.
// ---------------------------------------------------------------------------
template <typename T, int N, int DIR>
__global__ __launch_bounds__(kThreads) void k_ypass(WParams w, v2<T>* __restrict__ A, v2<T>* __restrict__ B,
                                                    int tz) {
  using V = v2<T>;
  const int n2 = N > 0 ? N : w.n2;
  const int n3 = w.n3;
  V* tw = lds_base<T>();
  V* buf = tw + n2;
  const int c = blockIdx.x;
  const int z0 = blockIdx.y * tz;
  const int pb = blockIdx.z;
  const int nz = min(tz, n3 - z0);
  const int k0 = w.col_start[c], ncol = w.col_start[c + 1] - k0;
  load_tw<T>(tw, w.tw2, n2);
  const size_t abase = (size_t)pb * w.n_slots * n3;
  if (DIR > 0) {
    for (int e = threadIdx.x; e < nz * n2; e += kThreads) buf[e] = cmk<V>((T)0, (T)0);
    __syncthreads();
    for (int e = threadIdx.x; e < ncol * nz; e += kThreads) {
      const int k = e / nz, z = e - k * nz;
      const int slot = w.col_slot[k0 + k], gy = w.col_gy[k0 + k];
      buf[z * n2 + gy] = A[abase + (size_t)slot * n3 + z0 + z];
    }
  } else {
    for (int e = threadIdx.x; e < nz * n2; e += kThreads) {
      const int z = e / n2, y = e - z * n2;
      buf[e] = B[(((size_t)pb * n3 + z0 + z) * w.ncols + c) * n2 + y];
    }
  }
  __syncthreads();
  fft_any<T, N, DIR>(buf, nz, n2, w.p2, tw);
  if (DIR > 0) {
    for (int e = threadIdx.x; e < nz * n2; e += kThreads) {
      const int z = e / n2, y = e - z * n2;
      B[(((size_t)pb * n3 + z0 + z) * w.ncols + c) * n2 + y] = buf[e];
    }
  } else {
    for (int e = threadIdx.x; e < ncol * nz; e += kThreads) {
      const int k = e / nz, z = e - k * nz;
      const int slot = w.col_slot[k0 + k], gy = w.col_gy[k0 + k];
      A[abase + (size_t)slot * n3 + z0 + z] = buf[z * n2 + gy];
    }
  }
}

// The x passes gather, per pair, the ncols columns of one (z, y-tile) into ny
// zero-padded x lines.  On the FP32 path each thread owns up to kMaxCol column
// elements: they are fetched into registers one pair ahead (the loads of pair
// pb+1 are in flight during the transforms of pair pb) and scattered into LDS.
constexpr int kMaxCol = 8;

template <typename T>
__device__ __forceinline__ void fetch_cols(const WParams& w, const v2<T>* __restrict__ B, int pb, int z, int y0,
                                           int ny, v2<T> (&r)[kMaxCol]) {
  const size_t bbase = ((size_t)pb * w.n3 + z) * w.ncols;
  const int m = w.ncols * ny;
#pragma unroll
  for (int k = 0; k < kMaxCol; ++k) {
    const int e = threadIdx.x + k * kThreads;
    if (e < m) {
      const int c = e / ny, y = e - c * ny;
      r[k] = B[(bbase + c) * w.n2 + y0 + y];
    }
  }
}

// LDS position of each owned column element: y * n1 + x(column)
__device__ __forceinline__ void col_slots(const WParams& w, int ny, int n1, int (&dst)[kMaxCol]) {
  const int m = w.ncols * ny;
#pragma unroll
  for (int k = 0; k < kMaxCol; ++k) {
    const int e = threadIdx.x + k * kThreads;
    dst[k] = -1;
    if (e < m) {
      const int c = e / ny, y = e - c * ny;
      dst[k] = y * n1 + w.col_gx[c];
    }
  }
}

// direct gather (no register prefetch): the FP64 path, where the prefetch
// registers would cost occupancy
template <typename T>
__device__ __forceinline__ void load_xlines(const WParams& w, const v2<T>* __restrict__ B, int pb, int z, int y0,
                                            int ny, int n1, v2<T>* buf) {
  for (int e = threadIdx.x; e < ny * n1; e += kThreads) buf[e] = cmk<v2<T>>((T)0, (T)0);
  __syncthreads();
  const size_t bbase = ((size_t)pb * w.n3 + z) * w.ncols;
  for (int e = threadIdx.x; e < w.ncols * ny; e += kThreads) {
    const int c = e / ny, y = e - c * ny;
    buf[y * n1 + w.col_gx[c]] = B[(bbase + c) * w.n2 + y0 + y];
  }
  __syncthreads();
}

template <typename T>
__device__ __forceinline__ void put_cols(int ne, const int (&dst)[kMaxCol], const v2<T> (&r)[kMaxCol], v2<T>* buf) {
  for (int e = threadIdx.x; e < ne; e += kThreads) buf[e] = cmk<v2<T>>((T)0, (T)0);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kMaxCol; ++k)
    if (dst[k] >= 0) buf[dst[k]] = r[k];
  __syncthreads();
}

// ---------------------------------------------------------------------------
// xrho: inverse x pass of every pair of the batch for one (z, y-tile) and
// rho(r) += f1/omega Re(psi)^2 + f2/omega Im(psi)^2 (build_density_sum,
// src/density_utils.mod.F90:61-80), accumulated in FP64; pairs with
// f1 = f2 = 0 are skipped.
// ---------------------------------------------------------------------------
template <typename T, int N>
__global__ __launch_bounds__(kThreads) void k_xrho(WParams w, const v2<T>* __restrict__ B, int npairs, int pair0,
                                                   int nstate, const double* __restrict__ f,
                                                   double* __restrict__ rho, int ty) {
  using V = v2<T>;
  const int n1 = N > 0 ? N : w.n1;
  V* tw = lds_base<T>();
  V* buf = tw + n1;
  const int y0 = blockIdx.x * ty, z = blockIdx.y;
  const int ny = min(ty, w.n2 - y0);
  const int ne = ny * n1;
  load_tw<T>(tw, w.tw1, n1);
  double acc[kMaxElem];
#pragma unroll
  for (int k = 0; k < kMaxElem; ++k) acc[k] = 0.0;
  const double rom = 1.0 / w.omega;
  constexpr bool kPre = sizeof(T) == 4;   // register prefetch of the next pair (FP32 path)
  int dst[kMaxCol];
  V r[kMaxCol];
  if (kPre) {
    col_slots(w, ny, n1, dst);
    if (npairs > 0) fetch_cols<T>(w, B, 0, z, y0, ny, r);
  }
  for (int pb = 0; pb < npairs; ++pb) {
    const int is1 = 2 * (pair0 + pb), is2 = is1 + 1;
    const double fa = f[is1];
    const double fb = is2 < nstate ? f[is2] : 0.0;
    const bool skip = (fa == 0.0 && fb == 0.0);
    if (kPre) {
      if (!skip) put_cols<T>(ne, dst, r, buf);
      if (pb + 1 < npairs) fetch_cols<T>(w, B, pb + 1, z, y0, ny, r);
    } else if (!skip) {
      load_xlines<T>(w, B, pb, z, y0, ny, n1, buf);
    }
    if (skip) continue;
    const double c3 = fa * rom, c4 = fb * rom;
    fft_any<T, N, 1>(buf, ny, n1, w.p1, tw);
#pragma unroll
    for (int k = 0; k < kMaxElem; ++k) {
      const int e = threadIdx.x + k * kThreads;
      if (e < ne) {
        const V v = buf[e];
        const double vx = (double)v.x, vy = (double)v.y;
        acc[k] = acc[k] + c3 * vx * vx + c4 * vy * vy;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < kMaxElem; ++k) {
    const int e = threadIdx.x + k * kThreads;
    if (e < ne) {
      const int y = e / n1, x = e - y * n1;
      const size_t idx = (size_t)x + (size_t)w.kr1 * ((size_t)(y0 + y) + (size_t)w.kr2 * z);
      rho[idx] += acc[k];
    }
  }
}

// ---------------------------------------------------------------------------
// xvpsi: inverse x pass, psi(r) *= V(r), forward x pass, for every pair of the
// batch (vpsi:485-493 between invfftn and fwfftn); V is read once per batch.
// ---------------------------------------------------------------------------
template <typename T, int N>
__global__ __launch_bounds__(kThreads) void k_xvpsi(WParams w, v2<T>* __restrict__ B, int npairs,
                                                    const double* __restrict__ vpot, int ty) {
  using V = v2<T>;
  const int n1 = N > 0 ? N : w.n1;
  V* tw = lds_base<T>();
  V* buf = tw + n1;
  const int y0 = blockIdx.x * ty, z = blockIdx.y;
  const int ny = min(ty, w.n2 - y0);
  const int ne = ny * n1;
  load_tw<T>(tw, w.tw1, n1);
  T vr[kMaxElem];
#pragma unroll
  for (int k = 0; k < kMaxElem; ++k) {
    const int e = threadIdx.x + k * kThreads;
    vr[k] = (T)0;
    if (e < ne) {
      const int y = e / n1, x = e - y * n1;
      vr[k] = (T)vpot[(size_t)x + (size_t)w.kr1 * ((size_t)(y0 + y) + (size_t)w.kr2 * z)];
    }
  }
  constexpr bool kPre = sizeof(T) == 4;   // register prefetch of the next pair (FP32 path)
  int dst[kMaxCol];
  V r[kMaxCol];
  if (kPre) {
    col_slots(w, ny, n1, dst);
    if (npairs > 0) fetch_cols<T>(w, B, 0, z, y0, ny, r);
  }
  const int m = w.ncols * ny;
  for (int pb = 0; pb < npairs; ++pb) {
    if (kPre) {
      put_cols<T>(ne, dst, r, buf);
      if (pb + 1 < npairs) fetch_cols<T>(w, B, pb + 1, z, y0, ny, r);
    } else {
      load_xlines<T>(w, B, pb, z, y0, ny, n1, buf);
    }
    fft_any<T, N, 1>(buf, ny, n1, w.p1, tw);
#pragma unroll
    for (int k = 0; k < kMaxElem; ++k) {
      const int e = threadIdx.x + k * kThreads;
      if (e < ne) buf[e] = cscale(buf[e], vr[k]);
    }
    __syncthreads();
    fft_any<T, N, -1>(buf, ny, n1, w.p1, tw);
    const size_t bbase = ((size_t)pb * w.n3 + z) * w.ncols;
    if (kPre) {
#pragma unroll
      for (int k = 0; k < kMaxCol; ++k) {
        const int e = threadIdx.x + k * kThreads;
        if (e < m) {
          const int c = e / ny, y = e - c * ny;
          B[(bbase + c) * w.n2 + y0 + y] = buf[dst[k]];
        }
      }
    } else {
      for (int e = threadIdx.x; e < m; e += kThreads) {
        const int c = e / ny, y = e - c * ny;
        B[(bbase + c) * w.n2 + y0 + y] = buf[y * n1 + w.col_gx[c]];
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Fused x pass for the compile-time lengths N = A*B (FastShape).  One
// workgroup holds L consecutive y lines of one z plane; lane t works on line
// l = t % L (so 16 lanes read or write 16 consecutive y, one 128-byte run of
// a column of B) and on group g = t / L of the line:
//   phase 1 (g = n2 < B):  x[B n1 + n2], n1 < A, straight from B into
//            registers, DFT_A, twiddle w^{n2 k1}, to LDS buffer 1
//   phase 2 (g = k1 < A):  DFT_B of buffer 1 row k1 -> psi(x = k1 + A k2)
//            in registers (|psi|^2 accumulation or the V(r) multiply here)
//   phase 3 (g = k1):      forward DFT_B over k2, twiddle w^{k1 k1'}, to LDS
//            buffer 2 at A k1' + k1
//   phase 4 (g = k1' < B): DFT_A of buffer 2 row k1' -> psi(x = k1' + B k2'),
//            straight from registers to the sphere columns of B.
// Each phase is the four-step split of fft_fast (the forward one with A and B
// exchanged, so that phase 2's registers feed phase 3 directly).  With two
// buffers a pair takes two barriers (phase 1 -> 2, phase 3 -> 4): the next
// pair's phase 1 writes buffer 1 after every lane has passed the 3 -> 4
// barrier, hence after its phase 2 reads.  NB = 1 shares one buffer and adds
// the two barriers that sharing needs (the FP64 path, whose buffers are twice
// as large, and any length whose two buffers would pass 64 KiB).  LDS element
// (row group q, row r of the group, line l) sits at row index * L + l + 16 q:
// the 16-element shift per group puts the two 32-lane halves of a wave, which
// hold consecutive groups in phases 2 and 4, on disjoint 8-byte bank groups.
// ---------------------------------------------------------------------------
template <int N>
struct XTile {
  static constexpr int A = FastShape<N>::A, B = FastShape<N>::B;
  static constexpr int G = A > B ? A : B;
  static constexpr int L = kThreads / G;        // lines per workgroup
  static constexpr int kBuf = N * L + 16 * G;   // elements of one LDS buffer
};

// LDS bytes of the fused x kernels: tw, NB buffers, x -> column map
template <typename T, int N>
constexpr size_t xlds(int nb) {
  return sizeof(v2<T>) * (size_t)(N + nb * XTile<N>::kBuf) + sizeof(int) * N;
}
template <typename T, int N>
constexpr int xbufs() { return (sizeof(T) == 4 && xlds<T, N>(2) <= 65536) ? 2 : 1; }
template <typename T, int N>
constexpr size_t xfused_lds() { return xlds<T, N>(xbufs<T, N>()); }
// lengths whose fused kernels fit the 64 KiB of LDS of one workgroup
template <typename T, int N>
constexpr bool xfused_ok() {
  if constexpr (N > 0) return xfused_lds<T, N>() <= 65536;
  else return false;
}

template <typename T, int N>
struct XFused {
  using V = v2<T>;
  static constexpr int A = XTile<N>::A, B = XTile<N>::B, L = XTile<N>::L;
  static constexpr int NB = xbufs<T, N>();
  V* tw;
  V* buf[2];
  int* xcol;
  int l, g, ny;

  __device__ __forceinline__ void init(const WParams& w, int y0) {
    tw = lds_base<T>();
    buf[0] = tw + N;
    buf[1] = NB == 2 ? buf[0] + XTile<N>::kBuf : buf[0];
    xcol = reinterpret_cast<int*>(buf[0] + NB * XTile<N>::kBuf);
    l = threadIdx.x % L;
    g = threadIdx.x / L;
    ny = min(L, w.n2 - y0);
    load_tw<T>(tw, w.tw1, N);
    for (int x = threadIdx.x; x < N; x += kThreads) xcol[x] = -1;
    __syncthreads();
    for (int c = threadIdx.x; c < w.ncols; c += kThreads) xcol[w.col_gx[c]] = c;
  }
  // Byte offsets, within one pair's (z) slab of B, of the elements
  // x = B k + g of this lane's line (the positions of phase 1's inputs and
  // phase 4's outputs alike), kNoCol off the sphere columns: fixed for the
  // whole kernel.  32-bit unsigned bytes, so that the accesses take the
  // scalar-base + vector-offset form (a slab is far below 4 GiB).
  static constexpr unsigned kNoCol = 0xffffffffu;
  __device__ __forceinline__ void offsets(const WParams& w, int y0, unsigned (&off)[A]) const {
#pragma unroll
    for (int k = 0; k < A; ++k) {
      const int c = (g < B && l < ny) ? xcol[B * k + g] : -1;
      off[k] = c >= 0 ? (unsigned)((c * w.n2 + y0 + l) * sizeof(V)) : kNoCol;
    }
  }
  // phase 1 loads of one pair (slab base Bs): column elements, zero elsewhere
  __device__ __forceinline__ void fetch(const V* __restrict__ Bs, const unsigned (&off)[A], V (&r)[A]) const {
#pragma unroll
    for (int k = 0; k < A; ++k) r[k] = off[k] != kNoCol ? at(Bs, off[k]) : cmk<V>((T)0, (T)0);
  }
  __device__ __forceinline__ static const V& at(const V* p, unsigned b) {
    return *reinterpret_cast<const V*>(reinterpret_cast<const char*>(p) + b);
  }
  __device__ __forceinline__ static V& at(V* p, unsigned b) {
    return *reinterpret_cast<V*>(reinterpret_cast<char*>(p) + b);
  }
  template <int DIR>
  __device__ __forceinline__ void phase1(V (&v)[A], V* bf) const {
    if (g < B) {
      dftn<A, DIR>(v);
#pragma unroll
      for (int k1 = 1; k1 < A; ++k1) v[k1] = twmul<DIR>(v[k1], tw[g * k1]);
#pragma unroll
      for (int k1 = 0; k1 < A; ++k1) bf[(B * k1 + g) * L + l + 16 * k1] = v[k1];
    }
  }
  // u[k2] = psi(x = g + A k2), lanes g < A
  template <int DIR>
  __device__ __forceinline__ void phase2(V (&u)[B], const V* bf) const {
    if (g < A) {
#pragma unroll
      for (int m = 0; m < B; ++m) u[m] = bf[(B * g + m) * L + l + 16 * g];
      dftn<B, DIR>(u);
    }
  }
  template <int DIR>
  __device__ __forceinline__ void phase3(V (&u)[B], V* bf) const {
    if (g < A) {
      dftn<B, DIR>(u);
#pragma unroll
      for (int k1 = 1; k1 < B; ++k1) u[k1] = twmul<DIR>(u[k1], tw[g * k1]);
#pragma unroll
      for (int k1 = 0; k1 < B; ++k1) bf[(A * k1 + g) * L + l + 16 * k1] = u[k1];
    }
  }
  // v[k2] = psi(x = g + B k2), lanes g < B
  template <int DIR>
  __device__ __forceinline__ void phase4(V (&v)[A], const V* bf) const {
    if (g < B) {
#pragma unroll
      for (int m = 0; m < A; ++m) v[m] = bf[(A * g + m) * L + l + 16 * g];
      dftn<A, DIR>(v);
    }
  }
};

// xrho, fused form (same contract as k_xrho).  With two buffers the pairs
// alternate between them, one barrier per pair.
template <typename T, int N>
__global__ __launch_bounds__(kThreads) void k_xrho_f(WParams w, const v2<T>* __restrict__ Bg, int npairs, int pair0,
                                                     int nstate, const double* __restrict__ f,
                                                     double* __restrict__ rho) {
  using V = v2<T>;
  using X = XFused<T, N>;
  constexpr int A = X::A, B = X::B, L = X::L;
  X s;
  const int y0 = blockIdx.x * L, z = blockIdx.y;
  s.init(w, y0);
  __syncthreads();
  auto occupied = [&](int p) {
    const int i1 = 2 * (pair0 + p), i2 = i1 + 1;
    return f[i1] != 0.0 || (i2 < nstate && f[i2] != 0.0);
  };
  const double rom = 1.0 / w.omega;
  double acc[B];
#pragma unroll
  for (int k = 0; k < B; ++k) acc[k] = 0.0;
  unsigned off[A];
  s.offsets(w, y0, off);
  const size_t slab = (size_t)w.ncols * w.n2;   // elements of one (pair, z) slab of B
  auto base = [&](int p) { return Bg + ((size_t)p * w.n3 + z) * slab; };
  constexpr bool kPre = sizeof(T) == 4;
  V r[A];
  int pb = 0;
  while (pb < npairs && !occupied(pb)) ++pb;
  if (kPre && pb < npairs) s.fetch(base(pb), off, r);
  int par = 0;
  while (pb < npairs) {   // uniform over the workgroup
    int nx = pb + 1;
    while (nx < npairs && !occupied(nx)) ++nx;
    const int is1 = 2 * (pair0 + pb), is2 = is1 + 1;
    const double c3 = f[is1] * rom, c4 = (is2 < nstate ? f[is2] : 0.0) * rom;
    V v[A];
    if (kPre) {
#pragma unroll
      for (int k = 0; k < A; ++k) v[k] = r[k];
      if (nx < npairs) s.fetch(base(nx), off, r);
    } else {
      s.fetch(base(pb), off, v);
    }
    V* bf = par ? s.buf[1] : s.buf[0];
    s.template phase1<1>(v, bf);
    __syncthreads();
    V u[B];
    s.template phase2<1>(u, bf);
    if (s.g < A) {
#pragma unroll
      for (int k = 0; k < B; ++k) {
        const double vx = (double)u[k].x, vy = (double)u[k].y;
        acc[k] = acc[k] + c3 * vx * vx + c4 * vy * vy;
      }
    }
    if (X::NB == 1) __syncthreads();   // the next pair rewrites the only buffer
    par ^= (X::NB - 1);
    pb = nx;
  }
  // transpose the accumulators through LDS (rows of N + 1 doubles per line)
  // so that the read-modify-write of rho runs along x
  __syncthreads();
  double* st = reinterpret_cast<double*>(s.buf[0]);
  constexpr int LD = N + 1;
  static_assert(sizeof(double) * L * LD <= sizeof(V) * X::NB * XTile<N>::kBuf, "rho staging exceeds the buffers");
  if (s.g < A) {
#pragma unroll
    for (int k = 0; k < B; ++k) st[s.l * LD + s.g + A * k] = acc[k];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < s.ny * N; e += kThreads) {
    const int y = e / N, x = e - y * N;
    const size_t idx = (size_t)x + (size_t)w.kr1 * ((size_t)(y0 + y) + (size_t)w.kr2 * z);
    rho[idx] += st[y * LD + x];
  }
}

// xvpsi, fused form (same contract as k_xvpsi)
template <typename T, int N>
__global__ __launch_bounds__(kThreads) void k_xvpsi_f(WParams w, v2<T>* __restrict__ Bg, int npairs,
                                                      const double* __restrict__ vpot) {
  using V = v2<T>;
  using X = XFused<T, N>;
  constexpr int A = X::A, B = X::B, L = X::L;
  X s;
  const int y0 = blockIdx.x * L, z = blockIdx.y;
  s.init(w, y0);
  // V(r) of this tile, staged x-major (row stride L + 1) through buffer 1 so
  // that lane (l, g) picks up V(x = g + A k2, y0 + l) for its phase-2 values
  T* vst = reinterpret_cast<T*>(s.buf[1]);
  for (int e = threadIdx.x; e < L * N; e += kThreads) {
    const int y = e / N, x = e - y * N;
    vst[x * (L + 1) + y] =
        y < s.ny ? (T)vpot[(size_t)x + (size_t)w.kr1 * ((size_t)(y0 + y) + (size_t)w.kr2 * z)] : (T)0;
  }
  __syncthreads();
  T vr[B];
#pragma unroll
  for (int k = 0; k < B; ++k) vr[k] = s.g < A ? vst[(s.g + A * k) * (L + 1) + s.l] : (T)0;
  if (X::NB == 1) __syncthreads();
  unsigned off[A];
  s.offsets(w, y0, off);
  const size_t slab = (size_t)w.ncols * w.n2;   // elements of one (pair, z) slab of B
  constexpr bool kPre = sizeof(T) == 4;
  V r[A];
  if (kPre && npairs > 0) s.fetch(Bg + (size_t)z * slab, off, r);
  for (int pb = 0; pb < npairs; ++pb) {
    V* Bs = Bg + ((size_t)pb * w.n3 + z) * slab;
    V v[A];
    if (kPre) {
#pragma unroll
      for (int k = 0; k < A; ++k) v[k] = r[k];
      if (pb + 1 < npairs) s.fetch(Bs + (size_t)w.n3 * slab, off, r);
    } else {
      s.fetch(Bs, off, v);
    }
    s.template phase1<1>(v, s.buf[0]);
    __syncthreads();
    V u[B];
    s.template phase2<1>(u, s.buf[0]);
#pragma unroll
    for (int k = 0; k < B; ++k) u[k] = cscale(u[k], vr[k]);
    if (X::NB == 1) __syncthreads();
    s.template phase3<-1>(u, s.buf[1]);
    __syncthreads();
    s.template phase4<-1>(v, s.buf[1]);
#pragma unroll
    for (int k = 0; k < A; ++k)
      if (off[k] != X::kNoCol) X::at(Bs, off[k]) = v[k];
    if (X::NB == 1) __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// zfwd: forward z pass of owner sticks + mirrors, scale 1/(n1 n2 n3), Gamma
// decode fp = psi(G) + psi(-G), fm = psi(G) - psi(-G) and (in FP64)
//   c2(:,is1) += -fi   (tpiba2 hg c0(:,is1) + (Re fp, Im fm))
//   c2(:,is2) += -fip1 (tpiba2 hg c0(:,is2) + (Im fp, -Re fm))
// with fi = f/2 (f = 0 -> 1)  (vpsi_utils.mod.F90:627-672, add_wfn :717).
// ---------------------------------------------------------------------------
template <typename T, int N>
__global__ __launch_bounds__(kThreads) void k_zfwd(WParams w, const v2<T>* __restrict__ A,
                                                   const double2* __restrict__ c0, double2* __restrict__ c2,
                                                   const double* __restrict__ f, int nstate, int pair0, int lo,
                                                   double scale) {
  using V = v2<T>;
  const int n3 = N > 0 ? N : w.n3;
  V* tw = lds_base<T>();
  V* buf = tw + n3;
  const int pb = blockIdx.y;
  const int is1 = 2 * (pair0 + pb), is2 = is1 + 1;
  const bool has2 = is2 < nstate;
  const int o0 = blockIdx.x * lo;
  const int o1 = min(o0 + lo, w.n_own);
  const int nl = 2 * (o1 - o0);
  load_tw<T>(tw, w.tw3, n3);
  for (int e = threadIdx.x; e < nl * n3; e += kThreads) {
    const int line = e / n3, z = e - line * n3;
    const int o = o0 + (line >> 1);
    const bool mir = line & 1;
    V v = cmk<V>((T)0, (T)0);
    if (!(mir && o == 0)) {
      const int slot = mir ? w.n_own + o - 1 : o;
      v = A[((size_t)pb * w.n_slots + slot) * n3 + z];
    }
    buf[e] = v;
  }
  __syncthreads();
  fft_any<T, N, -1>(buf, nl, n3, w.p3, tw);
  double fi = f[is1] * 0.5;
  if (fi == 0.0) fi = 1.0;
  double fip1 = has2 ? f[is2] * 0.5 : 0.0;
  if (fip1 == 0.0) fip1 = 1.0;
  const int j0 = w.own_start[o0], j1 = w.own_start[o1];
  const double2* ca = c0 + (size_t)is1 * w.ldg;
  const double2* cb = c0 + (size_t)is2 * w.ldg;
  double2* da = c2 + (size_t)is1 * w.ldg;
  double2* db = c2 + (size_t)is2 * w.ldg;
  for (int j = j0 + threadIdx.x; j < j1; j += kThreads) {
    const int gi = w.ginfo[j];
    const int o = gi >> 12, l = gi & 4095;
    const int t = o - o0;
    const V a = buf[2 * t * n3 + l];
    const int ml = (o == 0) ? 2 * t : 2 * t + 1;
    const V b = buf[ml * n3 + (l == 0 ? 0 : n3 - l)];
    const double2 psin = make_double2((double)a.x * scale, (double)a.y * scale);
    const double2 psii = make_double2((double)b.x * scale, (double)b.y * scale);
    const double2 fp = cadd(psin, psii);
    const double2 fm = csub(psin, psii);
    const double g2 = w.tpiba2 * w.hg[j];
    const double2 x1 = ca[j];
    double2 r1 = da[j];
    r1.x += -fi * (g2 * x1.x + fp.x);
    r1.y += -fi * (g2 * x1.y + fm.y);
    da[j] = r1;
    if (has2) {
      const double2 x2 = cb[j];
      double2 r2 = db[j];
      r2.x += -fip1 * (g2 * x2.x + fp.y);
      r2.y += -fip1 * (g2 * x2.y - fm.x);
      db[j] = r2;
    }
  }
}

__global__ void k_scale(double* __restrict__ x, size_t n, double v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    x[i] *= v;
}

__global__ void k_fill(double* __restrict__ x, size_t n, double v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    x[i] = v;
}

// sum of x[0..n) into out[blockIdx.x]; fixed order for reproducibility
__global__ __launch_bounds__(kThreads) void k_sum_partial(const double* __restrict__ x, size_t n,
                                                          double* __restrict__ out) {
  double s = 0.0;
  for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads)
    s += x[i];
  s = warp_sum(s);
  __shared__ double part[kThreads / 64];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < kThreads / 64; ++q) t += part[q];
    out[blockIdx.x] = t;
  }
}

// per-pair sums of the zinv partials: red[(pair*gx + b)*4 + k] -> out[pair*4 + k]
__global__ void k_reduce_pairs(const double* __restrict__ red, int gx, double* __restrict__ out) {
  const int pair = blockIdx.x;
  if (threadIdx.x < 4) {
    double s = 0.0;
    for (int b = 0; b < gx; ++b) s += red[((size_t)pair * gx + b) * 4 + threadIdx.x];
    out[(size_t)pair * 4 + threadIdx.x] = s;
  }
}

// ---------------------------------------------------------------------------
// host drivers
// ---------------------------------------------------------------------------
static WParams make_params(mc_ctx* ctx) {
  WParams w;
  w.n1 = ctx->n1; w.n2 = ctx->n2; w.n3 = ctx->n3;
  w.kr1 = ctx->kr1; w.kr2 = ctx->kr2; w.kr3 = ctx->kr3;
  w.ngw = ctx->ngw; w.ldg = ctx->ldg; w.n_own = ctx->n_own; w.n_slots = ctx->n_slots;
  w.ncols = ctx->ncols; w.geq0 = ctx->geq0;
  w.tpiba2 = ctx->tpiba2; w.omega = ctx->omega;
  w.p1 = ctx->plan1; w.p2 = ctx->plan2; w.p3 = ctx->plan3;
  w.ginfo = ctx->d_ginfo.as<int>();
  w.hg = ctx->d_hg.as<double>();
  w.own_start = ctx->d_own_start.as<int>();
  w.col_start = ctx->d_col_start.as<int>();
  w.col_slot = ctx->d_col_slot.as<int>();
  w.col_gy = ctx->d_col_gy.as<int>();
  w.col_gx = ctx->d_col_gx.as<int>();
  w.tw1 = ctx->d_tw1.as<double2>();
  w.tw2 = ctx->d_tw2.as<double2>();
  w.tw3 = ctx->d_tw3.as<double2>();
  return w;
}

struct Shapes {
  int lo;   // owner sticks per zinv/zfwd workgroup
  int tz;   // z planes per y-pass workgroup
  int ty;   // x lines per x-pass workgroup
  int gx;   // zinv/zfwd workgroups per pair
};

template <typename T>
static Shapes shapes(mc_ctx* ctx) {
  Shapes s;
  MC_CHECK(lines_for<T>(ctx->plan3) >= 2, MC_ERR_UNSUPPORTED, "nr3 too large for the z pass");
  s.lo = lines_for<T>(ctx->plan3) / 2;
  s.tz = lines_for<T>(ctx->plan2);
  s.ty = std::min(lines_for<T>(ctx->plan1), (kMaxElem * kThreads) / ctx->n1);
  s.ty = std::min(s.ty, (kMaxCol * kThreads) / std::max(ctx->ncols, 1));
  MC_CHECK(s.ty >= 1, MC_ERR_UNSUPPORTED, "nr1 too large for the x pass");
  s.gx = (ctx->n_own + s.lo - 1) / s.lo;
  return s;
}

template <typename T>
static void ensure_workspace(mc_ctx* ctx, int P) {
  ctx->w_a.ensure(sizeof(v2<T>) * (size_t)P * ctx->n_slots * ctx->n3);
  ctx->w_b.ensure(sizeof(v2<T>) * (size_t)P * ctx->n3 * ctx->ncols * ctx->n2);
}

static void upload_f(mc_ctx* ctx, const double* f_host, int nstate) {
  ctx->w_f.ensure(sizeof(double) * std::max(nstate, 1));
  if (nstate)
    MC_HIP(hipMemcpyAsync(ctx->w_f.p, f_host, sizeof(double) * nstate, hipMemcpyHostToDevice, ctx->stream));
}

template <typename T>
static void rhoofr_impl(mc_ctx* ctx, const double2* c0, const double* f_host, int nstate, double* rho,
                        double* ekin, double* csumg, double* csumr) {
  using V = v2<T>;
  const size_t nnr1 = (size_t)ctx->kr1 * ctx->kr2 * ctx->kr3;
  const WParams w = make_params(ctx);
  const Shapes s = shapes<T>(ctx);
  const int npairs = (nstate + 1) / 2;
  const int P = std::max(1, std::min(ctx->batch_pairs_rho, npairs));
  ensure_workspace<T>(ctx, P);
  upload_f(ctx, f_host, nstate);
  const int nblk_sum = 1024;
  ctx->w_red.ensure(sizeof(double) * ((size_t)std::max(npairs, 1) * s.gx * 4 + (size_t)npairs * 4 + nblk_sum));
  double* red = ctx->w_red.as<double>();
  double* pair_sums = red + (size_t)std::max(npairs, 1) * s.gx * 4;
  double* rsum_part = pair_sums + (size_t)npairs * 4;
  LAUNCH(MC_KID_MISC, k_fill, dim3(1024), dim3(256), 0, rho, nnr1, 0.0);
  const size_t shz = sizeof(V) * (size_t)ctx->n3 * (1 + 2 * s.lo);
  const size_t shy = sizeof(V) * (size_t)ctx->n2 * (1 + s.tz);
  const size_t shx = sizeof(V) * (size_t)ctx->n1 * (1 + s.ty);
  V* A = ctx->w_a.as<V>();
  V* B = ctx->w_b.as<V>();
  for (int pair0 = 0; pair0 < npairs; pair0 += P) {
    const int np = std::min(P, npairs - pair0);
    with_length(ctx->n3, [&](auto NN) {
      LAUNCH(MC_KID_ZINV, (k_zinv<T, decltype(NN)::value>), dim3(s.gx, np), dim3(kThreads), shz, w, c0, nstate,
             pair0, A, s.lo, red + (size_t)pair0 * s.gx * 4);
    });
    with_length(ctx->n2, [&](auto NN) {
      LAUNCH(MC_KID_YINV, (k_ypass<T, decltype(NN)::value, 1>), dim3(ctx->ncols, (ctx->n3 + s.tz - 1) / s.tz, np),
             dim3(kThreads), shy, w, A, B, s.tz);
    });
    with_length(ctx->n1, [&](auto NN) {
      constexpr int N1 = decltype(NN)::value;
      if constexpr (xfused_ok<T, N1>()) {
        LAUNCH(MC_KID_XRHO, (k_xrho_f<T, N1>), dim3((ctx->n2 + XTile<N1>::L - 1) / XTile<N1>::L, ctx->n3),
               dim3(kThreads), (xfused_lds<T, N1>()), w, B, np, pair0, nstate, ctx->w_f.as<double>(), rho);
      } else {
        LAUNCH(MC_KID_XRHO, (k_xrho<T, N1>), dim3((ctx->n2 + s.ty - 1) / s.ty, ctx->n3), dim3(kThreads), shx, w, B,
               np, pair0, nstate, ctx->w_f.as<double>(), rho, s.ty);
      }
    });
  }
  if (npairs > 0) LAUNCH(MC_KID_MISC, k_reduce_pairs, dim3(npairs), dim3(64), 0, red, s.gx, pair_sums);
  LAUNCH(MC_KID_MISC, k_sum_partial, dim3(nblk_sum), dim3(kThreads), 0, rho, nnr1, rsum_part);
  std::vector<double> hs((size_t)npairs * 4 + nblk_sum);
  MC_HIP(hipMemcpyAsync(hs.data(), pair_sums, sizeof(double) * hs.size(), hipMemcpyDeviceToHost, ctx->stream));
  MC_HIP(hipStreamSynchronize(ctx->stream));
  double sg = 0.0, sk = 0.0, sr = 0.0;
  for (int p = 0; p < npairs; ++p) {
    const int is1 = 2 * p, is2 = is1 + 1;
    if (f_host[is1] != 0.0) { sg += f_host[is1] * hs[4 * p + 0]; sk += f_host[is1] * hs[4 * p + 1]; }
    if (is2 < nstate && f_host[is2] != 0.0) { sg += f_host[is2] * hs[4 * p + 2]; sk += f_host[is2] * hs[4 * p + 3]; }
  }
  for (int b = 0; b < nblk_sum; ++b) sr += hs[(size_t)npairs * 4 + b];
  double cr = sr * ctx->omega / ((double)ctx->n1 * ctx->n2 * ctx->n3);
  if (std::is_same<T, float>::value && cr > 0.0 && sg > 0.0) {
    // FP32 transforms lose ~1e-7 of the charge (a systematic rounding of the
    // norm, csumr - csumg ~ -1.6e-4 at 1024 electrons); rescale rho so that
    // its integral is the FP64 G-space charge csumg (the invariant CPMD's
    // csumr/csumg check enforces, rhoofr:625-635)
    const double scale = sg / cr;
    LAUNCH(MC_KID_MISC, k_scale, dim3(1024), dim3(256), 0, rho, nnr1, scale);
    cr *= scale;
  }
  if (ekin) *ekin = sk * ctx->tpiba2;
  if (csumg) *csumg = sg;
  if (csumr) *csumr = cr;
}

template <typename T>
static void vpsi_impl(mc_ctx* ctx, const double2* c0, double2* c2, const double* f_host, const double* vpot,
                      int nstate) {
  using V = v2<T>;
  const WParams w = make_params(ctx);
  const Shapes s = shapes<T>(ctx);
  const int npairs = (nstate + 1) / 2;
  if (npairs == 0) return;
  const int P = std::max(1, std::min(ctx->batch_pairs_vpsi, npairs));
  ensure_workspace<T>(ctx, P);
  upload_f(ctx, f_host, nstate);
  const size_t shz = sizeof(V) * (size_t)ctx->n3 * (1 + 2 * s.lo);
  const size_t shy = sizeof(V) * (size_t)ctx->n2 * (1 + s.tz);
  const size_t shx = sizeof(V) * (size_t)ctx->n1 * (1 + s.ty);
  const double scale = 1.0 / ((double)ctx->n1 * ctx->n2 * ctx->n3);
  V* A = ctx->w_a.as<V>();
  V* B = ctx->w_b.as<V>();
  for (int pair0 = 0; pair0 < npairs; pair0 += P) {
    const int np = std::min(P, npairs - pair0);
    with_length(ctx->n3, [&](auto NN) {
      LAUNCH(MC_KID_ZINV, (k_zinv<T, decltype(NN)::value>), dim3(s.gx, np), dim3(kThreads), shz, w, c0, nstate,
             pair0, A, s.lo, (double*)nullptr);
    });
    with_length(ctx->n2, [&](auto NN) {
      LAUNCH(MC_KID_YINV, (k_ypass<T, decltype(NN)::value, 1>), dim3(ctx->ncols, (ctx->n3 + s.tz - 1) / s.tz, np),
             dim3(kThreads), shy, w, A, B, s.tz);
    });
    with_length(ctx->n1, [&](auto NN) {
      constexpr int N1 = decltype(NN)::value;
      if constexpr (xfused_ok<T, N1>()) {
        LAUNCH(MC_KID_XVPSI, (k_xvpsi_f<T, N1>), dim3((ctx->n2 + XTile<N1>::L - 1) / XTile<N1>::L, ctx->n3),
               dim3(kThreads), (xfused_lds<T, N1>()), w, B, np, vpot);
      } else {
        LAUNCH(MC_KID_XVPSI, (k_xvpsi<T, N1>), dim3((ctx->n2 + s.ty - 1) / s.ty, ctx->n3), dim3(kThreads), shx, w,
               B, np, vpot, s.ty);
      }
    });
    with_length(ctx->n2, [&](auto NN) {
      LAUNCH(MC_KID_YFWD, (k_ypass<T, decltype(NN)::value, -1>), dim3(ctx->ncols, (ctx->n3 + s.tz - 1) / s.tz, np),
             dim3(kThreads), shy, w, A, B, s.tz);
    });
    with_length(ctx->n3, [&](auto NN) {
      LAUNCH(MC_KID_ZFWD, (k_zfwd<T, decltype(NN)::value>), dim3(s.gx, np), dim3(kThreads), shz, w, A, c0, c2,
             ctx->w_f.as<double>(), nstate, pair0, s.lo, scale);
    });
  }
}

void rhoofr_dev(mc_ctx* ctx, const double2* c0, const double* f_host, int nstate, double* rho, double* ekin,
                double* csumg, double* csumr) {
  if (ctx->fft_fp32)
    rhoofr_impl<float>(ctx, c0, f_host, nstate, rho, ekin, csumg, csumr);
  else
    rhoofr_impl<double>(ctx, c0, f_host, nstate, rho, ekin, csumg, csumr);
}

void vpsi_dev(mc_ctx* ctx, const double2* c0, double2* c2, const double* f_host, const double* vpot, int nstate) {
  if (ctx->fft_fp32)
    vpsi_impl<float>(ctx, c0, c2, f_host, vpot, nstate);
  else
    vpsi_impl<double>(ctx, c0, c2, f_host, vpot, nstate);
}

}  // namespace mc
