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
#include <mutex>
#include <optional>
#include <string>
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

// element (pair pb, column c, y, z) of the column workspace B[pair][col][y][z]
__device__ __forceinline__ size_t bidx(const WParams& w, int pb, int c, int y, int z) {
  return (((size_t)pb * w.ncols + c) * w.n2 + y) * w.n3 + z;
}

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
      buf[e] = B[bidx(w, pb, c, y, z0 + z)];
    }
  }
  __syncthreads();
  fft_any<T, N, DIR>(buf, nz, n2, w.p2, tw);
  if (DIR > 0) {
    for (int e = threadIdx.x; e < nz * n2; e += kThreads) {
      const int z = e / n2, y = e - z * n2;
      B[bidx(w, pb, c, y, z0 + z)] = buf[e];
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
  const int m = w.ncols * ny;
#pragma unroll
  for (int k = 0; k < kMaxCol; ++k) {
    const int e = threadIdx.x + k * kThreads;
    if (e < m) {
      const int c = e / ny, y = e - c * ny;
      r[k] = B[bidx(w, pb, c, y0 + y, z)];
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
  for (int e = threadIdx.x; e < w.ncols * ny; e += kThreads) {
    const int c = e / ny, y = e - c * ny;
    buf[y * n1 + w.col_gx[c]] = B[bidx(w, pb, c, y0 + y, z)];
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
    if (kPre) {
#pragma unroll
      for (int k = 0; k < kMaxCol; ++k) {
        const int e = threadIdx.x + k * kThreads;
        if (e < m) {
          const int c = e / ny, y = e - c * ny;
          B[bidx(w, pb, c, y0 + y, z)] = buf[dst[k]];
        }
      }
    } else {
      for (int e = threadIdx.x; e < m; e += kThreads) {
        const int c = e / ny, y = e - c * ny;
        B[bidx(w, pb, c, y0 + y, z)] = buf[y * n1 + w.col_gx[c]];
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Fused y and x passes for the compile-time lengths N = A*B (FastShape).
// A workgroup holds L consecutive z lines (z0 .. z0+L-1) of one (column, y)
// or (x column, y row); lane t works on line l = t % L and on group g = t / L
// of it.  With B laid out [pair][col][y][z], 16 lanes with consecutive l
// touch one 128-byte run of B, so every global access is coalesced straight
// from registers.  The phases are the four-step split of fft_fast:
//   phase 1 (g = n2 < B):  x[B n1 + n2], n1 < A, in registers, DFT_A,
//            twiddle w^{n2 k1}, to an LDS buffer
//   phase 2 (g = k1 < A):  DFT_B of row group k1 -> X(k1 + A k2) in registers
// and for xvpsi, whose registers after phase 2 feed the forward transform
// directly (the split with A and B exchanged):
//   phase 3 (g = k1):      DFT_B over k2, twiddle w^{k1 k1'}, to buffer 2
//   phase 4 (g = k1' < B): DFT_A of row group k1' -> Y(k1' + B k2').
// A buffer row (one index of the transform, L lines) is 16 elements for
// L = 16; rows are stored with bit 0 flipped in odd row groups (groups of
// even size), so that the two 32-lane halves of a wave, which hold
// consecutive groups or consecutive rows, land on opposite halves of the
// 64-bank LDS.  Two buffers alternate between pairs, one barrier per
// transform; NB = 1 (FP64 lengths whose two buffers pass 64 KiB, and the FP64
// path at large N) shares one buffer and adds the barriers that needs.
// ---------------------------------------------------------------------------
// register budget of the fused kernels: FP64 is held to 2 waves per SIMD
// (256 VGPRs); FP32 keeps its register prefetch unconstrained
#define F4_WAVES __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 1 : 2, 8)))

template <int N>
struct F4Tile {
  static constexpr int A = FastShape<N>::A, B = FastShape<N>::B;
  static constexpr int G = A > B ? A : B;
  static constexpr int L = kThreads / G;   // lines per workgroup
  static constexpr int kBuf = N * L;       // elements of one LDS buffer
};

// LDS bytes of the fused kernels: the two twiddle tables, nb buffers, one
// dynamic LDS the fused engine may take per workgroup (one buffer above 64 KB)
constexpr size_t kF4MaxLds = 98304;
// int map of length N
template <typename T, int N>
constexpr size_t f4_lds(int nb) {
  return sizeof(v2<T>) * (size_t)(2 * N + nb * F4Tile<N>::kBuf) + sizeof(int) * N;
}
template <typename T, int N>
constexpr int f4_bufs() { return f4_lds<T, N>(2) <= 65536 ? 2 : 1; }
template <typename T, int N>
constexpr size_t f4_lds() { return f4_lds<T, N>(f4_bufs<T, N>()); }
// lengths with a fused kernel that fits the 64 KiB of LDS of one workgroup
template <typename T, int N>
constexpr bool f4_ok() {
  if constexpr (N > 0) return f4_lds<T, N>() <= kF4MaxLds;
  else return false;
}
// Launches of the fused engine above 64 KB of dynamic LDS (a single FP64
// buffer at N = 240 takes 70 KB; gfx950 has 160 KB per CU) raise the
// kernel's dynamic LDS limit first.
template <typename K>
static void f4_allow_lds(K* kern, size_t bytes) {
  if (bytes > 65536)
    MC_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes));
}

// Buffer addressing.  Logical row r of row group q (groups of P rows) is
// stored at physical row r ^ (q & 1) when P is even (r otherwise), L lines
// per row.  Every access of a phase is then one of two per-lane base
// registers plus a compile-time offset:
//   write row P k + g (k compile-time):  P k L + wb[k & 1],  wb = {g L + l, (g ^ 1) L + l}
//   read row P g + m (m compile-time):   m L + rb[m & 1],    rb = {(P g + s) L + l, (P g - s) L + l}, s = g & 1
struct F4Bases {
  int wb[2], rb[2];
};
template <int P, int L>
__device__ __forceinline__ F4Bases f4_bases(int g, int l) {
  F4Bases b;
  const int s = (P % 2 == 0) ? (g & 1) : 0;
  b.wb[0] = g * L + l;
  b.wb[1] = ((P % 2 == 0) ? (g ^ 1) : g) * L + l;
  b.rb[0] = (P * g + s) * L + l;
  b.rb[1] = (P * g - s) * L + l;
  return b;
}

template <typename T, int N>
struct F4 {
  using V = v2<T>;
  static constexpr int A = F4Tile<N>::A, B = F4Tile<N>::B, L = F4Tile<N>::L;
  static constexpr int NB = f4_bufs<T, N>();
  static constexpr unsigned kNone = 0xffffffffu;
  V* t1;      // t1[k1 B + n2] = w^{n2 k1}, k1 < A, n2 < B (phase 1)
  V* t3;      // t3[k1 A + n2] = w^{n2 k1}, k1 < B, n2 < A (phase 3)
  V* buf[2];
  int* map;   // N ints after the buffers
  int l, g;
  F4Bases b12, b34;   // buffer bases of phases 1/2 (groups of B rows) and 3/4 (groups of A rows)

  __device__ __forceinline__ void init(const double2* table) {
    t1 = lds_base<T>();
    t3 = t1 + N;
    buf[0] = t3 + N;
    buf[1] = NB == 2 ? buf[0] + F4Tile<N>::kBuf : buf[0];
    map = reinterpret_cast<int*>(buf[0] + NB * F4Tile<N>::kBuf);
    l = threadIdx.x % L;
    g = threadIdx.x / L;
    b12 = f4_bases<B, L>(g, l);
    b34 = f4_bases<A, L>(g, l);
    for (int e = threadIdx.x; e < N; e += kThreads) {
      const int k1 = e / B, n2 = e - k1 * B;
      const double2 w1 = table[n2 * k1];
      t1[e] = cmk<V>((T)w1.x, (T)w1.y);
      const int j1 = e / A, m2 = e - j1 * A;
      const double2 w3 = table[m2 * j1];
      t3[e] = cmk<V>((T)w3.x, (T)w3.y);
    }
  }
  // 32-bit byte offsets from a uniform base: scalar-base + vector-offset accesses
  __device__ __forceinline__ static V ld(const V* p, unsigned b) {
    return *reinterpret_cast<const V*>(reinterpret_cast<const char*>(p) + b);
  }
  __device__ __forceinline__ static void st(V* p, unsigned b, V v) {
    *reinterpret_cast<V*>(reinterpret_cast<char*>(p) + b) = v;
  }
  template <int K>
  __device__ __forceinline__ static void fetch(const V* __restrict__ p, const unsigned (&off)[K], V (&r)[K]) {
#pragma unroll
    for (int k = 0; k < K; ++k) r[k] = off[k] != kNone ? ld(p, off[k]) : cmk<V>((T)0, (T)0);
  }
  template <int K>
  __device__ __forceinline__ static void store(V* __restrict__ p, const unsigned (&off)[K], const V (&r)[K]) {
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (off[k] != kNone) st(p, off[k], r[k]);
  }
  // PR (A = 16): the inputs of the band structure of dft16_band only (the
  // inverse passes of a wavefunction sphere with |m| < N/4)
  template <int DIR, bool PR = false>
  __device__ __forceinline__ void phase1(V (&v)[A], V* bf) const {
    if (g < B) {
      if constexpr (PR && A == 16) dft16_band<DIR, true, false>(v);
      else dftn<A, DIR>(v);
      const V* tp = t1 + g;
#pragma unroll
      for (int k1 = 1; k1 < A; ++k1) v[k1] = twmul<DIR>(v[k1], tp[k1 * B]);
      V* w0 = bf + b12.wb[0];
      V* w1 = bf + b12.wb[1];
#pragma unroll
      for (int k1 = 0; k1 < A; ++k1) ((k1 & 1) ? w1 : w0)[B * k1 * L] = v[k1];
    }
  }
  // u[k2] = X(g + A k2), lanes g < A
  template <int DIR>
  __device__ __forceinline__ void phase2(V (&u)[B], const V* bf) const {
    if (g < A) {
      const V* r0 = bf + b12.rb[0];
      const V* r1 = bf + b12.rb[1];
#pragma unroll
      for (int m = 0; m < B; ++m) u[m] = ((m & 1) ? r1 : r0)[m * L];
      dftn<B, DIR>(u);
    }
  }
  template <int DIR>
  __device__ __forceinline__ void phase3(V (&u)[B], V* bf) const {
    if (g < A) {
      dftn<B, DIR>(u);
      const V* tp = t3 + g;
#pragma unroll
      for (int k1 = 1; k1 < B; ++k1) u[k1] = twmul<DIR>(u[k1], tp[k1 * A]);
      V* w0 = bf + b34.wb[0];
      V* w1 = bf + b34.wb[1];
#pragma unroll
      for (int k1 = 0; k1 < B; ++k1) ((k1 & 1) ? w1 : w0)[A * k1 * L] = u[k1];
    }
  }
  // v[k2] = Y(g + B k2), lanes g < B; PR (A = 16): only k2 = 0..3, 12..15
  // (the outputs inside the band of a wavefunction sphere with |m| < N/4)
  template <int DIR, bool PR = false>
  __device__ __forceinline__ void phase4(V (&v)[A], const V* bf) const {
    if (g < B) {
      const V* r0 = bf + b34.rb[0];
      const V* r1 = bf + b34.rb[1];
#pragma unroll
      for (int m = 0; m < A; ++m) v[m] = ((m & 1) ? r1 : r0)[m * L];
      if constexpr (PR && A == 16) dft16_band<DIR, false, true>(v);
      else dftn<A, DIR>(v);
    }
  }
};

// ---------------------------------------------------------------------------
// y pass, fused form: workgroup (column c, z tile, pairs p0 .. p0+ppb-1).
// DIR=+1: sticks of column c in A -> B (dense y); DIR=-1: B -> sticks in A.
// ---------------------------------------------------------------------------
template <typename T, int N, int DIR, bool PR = false>
__global__ __launch_bounds__(kThreads) F4_WAVES void k_ypass_f(WParams w, v2<T>* __restrict__ Ag, v2<T>* __restrict__ Bg,
                                                      int npairs, int ppb) {
  using V = v2<T>;
  using X = F4<T, N>;
  constexpr int A = X::A, B = X::B, L = X::L;
  X s;
  s.init(w.tw2);
  const int c = blockIdx.x, z0 = blockIdx.y * L;
  const int p0 = blockIdx.z * ppb, p1 = min(npairs, p0 + ppb);
  const bool zin = z0 + s.l < w.n3;
  for (int y = threadIdx.x; y < N; y += kThreads) s.map[y] = -1;
  __syncthreads();
  const int k0 = w.col_start[c], k1 = w.col_start[c + 1];
  for (int k = k0 + threadIdx.x; k < k1; k += kThreads) s.map[w.col_gy[k]] = w.col_slot[k];
  __syncthreads();
  const size_t aslab = (size_t)w.n_slots * w.n3;   // A elements per pair
  const size_t bslab = (size_t)w.ncols * N * w.n3;  // B elements per pair
  const unsigned zoff = (unsigned)(z0 + s.l);
  // sparse (A) offsets: DIR>0 phase-1 inputs y = B k + g; DIR<0 phase-2 outputs y = g + A k
  constexpr int KA = DIR > 0 ? A : B;
  unsigned offa[KA];
#pragma unroll
  for (int k = 0; k < KA; ++k) {
    const bool act = zin && (DIR > 0 ? s.g < B : s.g < A);
    const int y = DIR > 0 ? B * k + s.g : s.g + A * k;
    const int slot = act ? s.map[y] : -1;
    offa[k] = slot >= 0 ? (unsigned)(((size_t)slot * w.n3 + zoff) * sizeof(V)) : X::kNone;
  }
  // dense (B) side, affine in k: DIR>0 phase-2 outputs y = g + A k; DIR<0
  // phase-1 inputs y = B k + g
  const bool actb = zin && (DIR > 0 ? s.g < A : s.g < B);
  const unsigned offb0 = (unsigned)((((size_t)c * N + s.g) * w.n3 + zoff) * sizeof(V));
  const unsigned dstep = (unsigned)((DIR > 0 ? A : B) * w.n3 * sizeof(V));
  constexpr bool kPre = sizeof(T) == 4;   // register prefetch of the next pair (FP32)
  V r[A];
  auto load = [&](int p, V (&dst)[A]) {
    if constexpr (DIR > 0) {
      X::fetch(Ag + (size_t)p * aslab, offa, dst);
    } else {
      const V* bp = Bg + (size_t)p * bslab;
#pragma unroll
      for (int k = 0; k < A; ++k) dst[k] = actb ? X::ld(bp, offb0 + k * dstep) : cmk<V>((T)0, (T)0);
    }
  };
  if (kPre && p0 < p1) load(p0, r);
  int par = 0;
  for (int p = p0; p < p1; ++p) {
    if (!kPre) load(p, r);
    V* bf = par ? s.buf[1] : s.buf[0];
    s.template phase1<DIR, PR>(r, bf);
    if (kPre && p + 1 < p1) load(p + 1, r);   // in flight through the rest of this pair
    __syncthreads();
    V u[B];
    s.template phase2<DIR>(u, bf);
    if constexpr (DIR > 0) {
      V* bp = Bg + (size_t)p * bslab;
      if (actb) {
#pragma unroll
        for (int k = 0; k < B; ++k) X::st(bp, offb0 + k * dstep, u[k]);
      }
    } else {
      X::store(Ag + (size_t)p * aslab, offa, u);
    }
    if (X::NB == 1) __syncthreads();
    par ^= (X::NB - 1);
  }
}

// x -> column map of the x kernels and the byte offsets, within one pair's B,
// of lane (l, g)'s elements x = B k + g (phase 1 inputs, phase 4 outputs)
template <typename X>
__device__ __forceinline__ void x_offsets(const X& s, const WParams& w, int y, int z0, unsigned (&off)[X::A]) {
  using V = typename X::V;
  const bool act = s.g < X::B && z0 + s.l < w.n3;
#pragma unroll
  for (int k = 0; k < X::A; ++k) {
    const int c = act ? s.map[X::B * k + s.g] : -1;
    off[k] = c >= 0 ? (unsigned)((((size_t)c * w.n2 + y) * w.n3 + z0 + s.l) * sizeof(V)) : X::kNone;
  }
}
template <typename X>
__device__ __forceinline__ void x_map(const X& s, const WParams& w, int n1) {
  for (int x = threadIdx.x; x < n1; x += kThreads) s.map[x] = -1;
  __syncthreads();
  for (int c = threadIdx.x; c < w.ncols; c += kThreads) s.map[w.col_gx[c]] = c;
}

// xrho, fused form (same contract as k_xrho): workgroup (z tile, y)
template <typename T, int N, bool PR = false>
__global__ __launch_bounds__(kThreads) F4_WAVES void k_xrho_f(WParams w, const v2<T>* __restrict__ Bg, int npairs, int pair0,
                                                     int nstate, const double* __restrict__ f,
                                                     double* __restrict__ rho) {
  using V = v2<T>;
  using X = F4<T, N>;
  constexpr int A = X::A, B = X::B, L = X::L;
  X s;
  s.init(w.tw1);
  const int z0 = blockIdx.x * L, y = blockIdx.y;
  const int nz = min(L, w.n3 - z0);
  x_map(s, w, N);
  __syncthreads();
  auto occupied = [&](int p) {
    const int i1 = 2 * (pair0 + p), i2 = i1 + 1;
    return f[i1] != 0.0 || (i2 < nstate && f[i2] != 0.0);
  };
  unsigned off[A];
  x_offsets(s, w, y, z0, off);
  const size_t bslab = (size_t)w.ncols * w.n2 * w.n3;
  const double rom = 1.0 / w.omega;
  double acc[B];
#pragma unroll
  for (int k = 0; k < B; ++k) acc[k] = 0.0;
  constexpr bool kPre = sizeof(T) == 4;
  V r[A];
  int pb = 0;
  while (pb < npairs && !occupied(pb)) ++pb;
  if (kPre && pb < npairs) X::fetch(Bg + (size_t)pb * bslab, off, r);
  int par = 0;
  while (pb < npairs) {   // uniform over the workgroup
    int nx = pb + 1;
    while (nx < npairs && !occupied(nx)) ++nx;
    const int is1 = 2 * (pair0 + pb), is2 = is1 + 1;
    const double c3 = f[is1] * rom, c4 = (is2 < nstate ? f[is2] : 0.0) * rom;
    if (!kPre) X::fetch(Bg + (size_t)pb * bslab, off, r);
    V* bf = par ? s.buf[1] : s.buf[0];
    s.template phase1<1, PR>(r, bf);
    if (kPre && nx < npairs) X::fetch(Bg + (size_t)nx * bslab, off, r);
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
  // transpose the accumulators through LDS (one row per z line) so that the
  // read-modify-write of rho runs along x
  __syncthreads();
  double* stg = reinterpret_cast<double*>(s.buf[0]);
  constexpr int LD = (sizeof(double) * L * (N + 1) <= sizeof(V) * X::NB * F4Tile<N>::kBuf) ? N + 1 : N;
  static_assert(sizeof(double) * L * LD <= sizeof(V) * X::NB * F4Tile<N>::kBuf, "rho staging exceeds the buffers");
  if (s.g < A) {
#pragma unroll
    for (int k = 0; k < B; ++k) stg[s.l * LD + s.g + A * k] = acc[k];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < nz * N; e += kThreads) {
    const int zl = e / N, x = e - zl * N;
    const size_t idx = (size_t)x + (size_t)w.kr1 * ((size_t)y + (size_t)w.kr2 * (z0 + zl));
    rho[idx] += stg[zl * LD + x];
  }
}

// xvpsi, fused form (same contract as k_xvpsi): workgroup (z tile, y).  On
// the FP32 path the next pair's loads are issued after phase kXPre.
constexpr int kXPre = 1;
template <typename T, int N, bool PR = false>
__global__ __launch_bounds__(kThreads) F4_WAVES void k_xvpsi_f(WParams w, v2<T>* __restrict__ Bg, int npairs,
                                                      const double* __restrict__ vpot) {
  using V = v2<T>;
  using X = F4<T, N>;
  constexpr int A = X::A, B = X::B, L = X::L;
  X s;
  s.init(w.tw1);
  const int z0 = blockIdx.x * L, y = blockIdx.y;
  const int nz = min(L, w.n3 - z0);
  x_map(s, w, N);
  // V(r) of this tile, staged x-major (row stride L + 1) through buffer 2 so
  // that lane (l, g) picks up V(x = g + A k2, z0 + l) for its phase-2 values
  T* vst = reinterpret_cast<T*>(s.buf[1]);
  static_assert(N * (L + 1) <= 2 * F4Tile<N>::kBuf, "V staging exceeds a buffer");
  for (int e = threadIdx.x; e < L * N; e += kThreads) {
    const int zl = e / N, x = e - zl * N;
    vst[x * (L + 1) + zl] =
        zl < nz ? (T)vpot[(size_t)x + (size_t)w.kr1 * ((size_t)y + (size_t)w.kr2 * (z0 + zl))] : (T)0;
  }
  __syncthreads();
  T vr[B];
#pragma unroll
  for (int k = 0; k < B; ++k) vr[k] = s.g < A ? vst[(s.g + A * k) * (L + 1) + s.l] : (T)0;
  if (X::NB == 1) __syncthreads();
  unsigned off[A];
  x_offsets(s, w, y, z0, off);
  const size_t bslab = (size_t)w.ncols * w.n2 * w.n3;
  constexpr bool kPre = sizeof(T) == 4;
  V r[A];
  if (kPre && npairs > 0) X::fetch(Bg, off, r);
  for (int pb = 0; pb < npairs; ++pb) {
    V* Bs = Bg + (size_t)pb * bslab;
    V v[A];
    if constexpr (kPre) {
      s.template phase1<1, PR>(r, s.buf[0]);
      if (kXPre == 1 && pb + 1 < npairs) X::fetch(Bs + bslab, off, r);   // in flight through phases 2-4
    } else {
      X::fetch(Bs, off, v);
      s.template phase1<1, PR>(v, s.buf[0]);
    }
    __syncthreads();
    V u[B];
    s.template phase2<1>(u, s.buf[0]);
#pragma unroll
    for (int k = 0; k < B; ++k) u[k] = cscale(u[k], vr[k]);
    if (X::NB == 1) __syncthreads();
    s.template phase3<-1>(u, s.buf[1]);
    if (kPre && kXPre == 3 && pb + 1 < npairs) X::fetch(Bs + bslab, off, r);   // in flight through phase 4
    __syncthreads();
    s.template phase4<-1, PR>(v, s.buf[1]);
    if constexpr (PR && A == 16) {   // outputs 4..11 lie outside the sphere's band (no column): not formed
#pragma unroll
      for (int k = 0; k < A; ++k)
        if ((k < 4 || k >= 12) && off[k] != X::kNone) X::st(Bs, off[k], v[k]);
    } else {
      X::store(Bs, off, v);
    }
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
                                                   double scale, int acc) {
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
    double2 r1 = acc ? da[j] : make_double2(0.0, 0.0);
    r1.x += -fi * (g2 * x1.x + fp.x);
    r1.y += -fi * (g2 * x1.y + fm.y);
    da[j] = r1;
    if (has2) {
      const double2 x2 = cb[j];
      double2 r2 = acc ? db[j] : make_double2(0.0, 0.0);
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
  int lo;   // owner sticks per zinv workgroup
  int lf;   // owner sticks per zfwd workgroup
  int tz;   // z planes per y-pass workgroup
  int ty;   // x lines per x-pass workgroup
  int gx;   // zinv workgroups per pair
  int gf;   // zfwd workgroups per pair
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
  // FP32 zfwd runs half as many sticks per workgroup as zinv.  Measured per
  // 128-pair launch at 192^3: zfwd 0.28 ms at 4 sticks against 0.31 ms at 8,
  // zinv 0.52 ms at 4 against 0.46 ms at 8
  s.lf = sizeof(T) == 4 ? std::max(1, s.lo / 2) : s.lo;
  s.gf = (ctx->n_own + s.lf - 1) / s.lf;
  return s;
}

// y pass of np pairs: the fused kernel (workgroups over columns, z tiles and
// runs of pairs, >= ~65536 workgroups) for compile-time lengths, else k_ypass
template <typename T, int DIR>
static void ypass(mc_ctx* ctx, const WParams& w, const Shapes& s, v2<T>* A, v2<T>* B, int np, size_t shy) {
  const int kid = DIR > 0 ? MC_KID_YINV : MC_KID_YFWD;
  with_length(ctx->n2, [&](auto NN) {
    constexpr int N2 = decltype(NN)::value;
    if constexpr (f4_ok<T, N2>()) {
      constexpr int L = F4Tile<N2>::L;
      const int zt = (ctx->n3 + L - 1) / L;
      // runs of pairs per (column, z tile): aimed at >= 65536 workgroups, i.e. a
      // few pairs per workgroup (at 192^3: 1.78 + 1.85 ms of y passes per CP
      // step against 1.97 + 2.06 ms aimed at 2048, 2.11 + 1.99 ms at one pair)
      const int runs = std::max(1, std::min(np, (65536 + ctx->ncols * zt - 1) / (ctx->ncols * zt)));
      const int ppb = (np + runs - 1) / runs;
      auto go = [&](auto PRC) {
        constexpr bool PR = decltype(PRC)::value;
        f4_allow_lds(&k_ypass_f<T, N2, DIR, PR>, f4_lds<T, N2>());
        LAUNCH(kid, (k_ypass_f<T, N2, DIR, PR>), dim3(ctx->ncols, zt, (np + ppb - 1) / ppb), dim3(kThreads),
               (f4_lds<T, N2>()), w, A, B, np, ppb);
      };
      if constexpr (DIR > 0 && FastShape<N2>::A == 16) {
        if (ctx->prune_y) go(std::true_type{});
        else go(std::false_type{});
      } else {
        go(std::false_type{});
      }
    } else {
      LAUNCH(kid, (k_ypass<T, N2, DIR>), dim3(ctx->ncols, (ctx->n3 + s.tz - 1) / s.tz, np), dim3(kThreads), shy, w,
             A, B, s.tz);
    }
  });
}

// Pairs per transform batch: at most P, and under a device memory limit as
// many as fit beside everything else the library holds and a reserve for the
// buffers the rest of a step allocates after the transforms (vofrho's dense
// mesh, contraction partials: max(256 MiB, limit / 16)); the workspaces are
// shrunk when they hold more; MC_ERR_ARG when not even one pair fits.  A later
// allocation that still does not fit frees them while they are idle
// (ws_reclaim), and the next transform call sizes them again.
template <typename T>
static int fit_pairs(mc_ctx* ctx, int P, size_t extra) {
  const size_t lim = mem_limit();
  if (!lim) return P;
  const size_t per = sizeof(v2<T>) * ((size_t)ctx->n_slots * ctx->n3 + (size_t)ctx->n3 * ctx->ncols * ctx->n2);
  const size_t held = ctx->w_a.bytes + ctx->w_b.bytes;
  const size_t reserve = std::max((size_t)256 << 20, lim / 16);
  const size_t others = mem_current() - std::min(held, mem_current()) + extra + reserve;
  const size_t avail = lim > others ? lim - others : 0;
  const long long fit = (long long)(avail / per);
  if (fit < 1)
    throw Error{MC_ERR_ARG, "device memory limit of " + std::to_string(lim) + " bytes leaves " +
                                std::to_string(avail) + " bytes for the transform workspaces; one state pair needs " +
                                std::to_string(per)};
  if (fit < P) P = (int)fit;
  if (held > (size_t)P * per) {   // shrink to what the batch needs
    ctx->w_a.release();
    ctx->w_b.release();
  }
  return P;
}

template <typename T>
static void ensure_workspace(mc_ctx* ctx, int P, bool need_b = true) {
  ctx->w_a.ensure(sizeof(v2<T>) * (size_t)P * ctx->n_slots * ctx->n3);
  if (need_b) ctx->w_b.ensure(sizeof(v2<T>) * (size_t)P * ctx->n3 * ctx->ncols * ctx->n2);
}

// Column reuse: the columns of all npairs pairs in w_keep, when reuse is on,
// no device memory limit is set and the allocation succeeds (else nullptr:
// the caller works in batches as usual)
template <typename T>
static v2<T>* keep_columns(mc_ctx* ctx, int npairs) {
  if (!ctx->col_reuse || mem_limit() != 0 || npairs <= 0) return nullptr;
  const size_t bytes = sizeof(v2<T>) * (size_t)npairs * ctx->n3 * ctx->ncols * ctx->n2;
  try {
    ctx->w_keep.ensure(bytes);
  } catch (const Error&) {
    (void)hipGetLastError();   // a failed hipMalloc leaves its error behind
    ctx->w_keep.release();
    return nullptr;
  }
  return ctx->w_keep.as<v2<T>>();
}

static void upload_f(mc_ctx* ctx, const double* f_host, int nstate) {
  ctx->w_f.ensure(sizeof(double) * std::max(nstate, 1));
  if (nstate)
    MC_HIP(hipMemcpyAsync(ctx->w_f.p, f_host, sizeof(double) * nstate, hipMemcpyHostToDevice, ctx->stream));
}

template <typename T>
static void rhoofr_impl(mc_ctx* ctx, const double2* c0, const double* f_host, int nstate, double* rho,
                        double* ekin, double* csumg, double* csumr, bool keep) {
  using V = v2<T>;
  const size_t nnr1 = (size_t)ctx->kr1 * ctx->kr2 * ctx->kr3;
  const WParams w = make_params(ctx);
  const Shapes s = shapes<T>(ctx);
  const int npairs = (nstate + 1) / 2;
  const int nblk_sum = 1024;
  WsUse use(ctx->w_a, ctx->w_b);
  ctx->keep_valid = false;
  V* K = keep ? keep_columns<T>(ctx, npairs) : nullptr;
  std::optional<WsUse> use_keep;
  if (K)
    use_keep.emplace(ctx->w_keep, ctx->w_keep);
  else if (mem_limit() != 0)
    ctx->w_keep.release();   // stale columns: room for the batch workspaces
  const size_t red_bytes =
      sizeof(double) * ((size_t)std::max(npairs, 1) * s.gx * 4 + (size_t)npairs * 4 + nblk_sum) +
      sizeof(double) * std::max(nstate, 1);
  const int P = fit_pairs<T>(ctx, std::max(1, std::min(ctx->batch_pairs_rho, npairs)),
                             std::max(red_bytes, ctx->w_red.bytes + ctx->w_f.bytes) - ctx->w_red.bytes - ctx->w_f.bytes);
  ensure_workspace<T>(ctx, P, K == nullptr);
  upload_f(ctx, f_host, nstate);
  ctx->w_red.ensure(sizeof(double) * ((size_t)std::max(npairs, 1) * s.gx * 4 + (size_t)npairs * 4 + nblk_sum));
  double* red = ctx->w_red.as<double>();
  double* pair_sums = red + (size_t)std::max(npairs, 1) * s.gx * 4;
  double* rsum_part = pair_sums + (size_t)npairs * 4;
  LAUNCH(MC_KID_MISC, k_fill, dim3(1024), dim3(256), 0, rho, nnr1, 0.0);
  const size_t shz = sizeof(V) * (size_t)ctx->n3 * (1 + 2 * s.lo);
  const size_t shy = sizeof(V) * (size_t)ctx->n2 * (1 + s.tz);
  const size_t shx = sizeof(V) * (size_t)ctx->n1 * (1 + s.ty);
  V* A = ctx->w_a.as<V>();
  const size_t bslab = (size_t)ctx->n3 * ctx->ncols * ctx->n2;
  for (int pair0 = 0; pair0 < npairs; pair0 += P) {
    const int np = std::min(P, npairs - pair0);
    V* B = K ? K + (size_t)pair0 * bslab : ctx->w_b.as<V>();
    with_length(ctx->n3, [&](auto NN) {
      LAUNCH(MC_KID_ZINV, (k_zinv<T, decltype(NN)::value>), dim3(s.gx, np), dim3(kThreads), shz, w, c0, nstate,
             pair0, A, s.lo, red + (size_t)pair0 * s.gx * 4);
    });
    ypass<T, 1>(ctx, w, s, A, B, np, shy);
    with_length(ctx->n1, [&](auto NN) {
      constexpr int N1 = decltype(NN)::value;
      if constexpr (f4_ok<T, N1>()) {
        auto go = [&](auto PRC) {
          constexpr bool PR = decltype(PRC)::value;
          f4_allow_lds(&k_xrho_f<T, N1, PR>, f4_lds<T, N1>());
          LAUNCH(MC_KID_XRHO, (k_xrho_f<T, N1, PR>), dim3((ctx->n3 + F4Tile<N1>::L - 1) / F4Tile<N1>::L, ctx->n2),
                 dim3(kThreads), (f4_lds<T, N1>()), w, B, np, pair0, nstate, ctx->w_f.as<double>(), rho);
        };
        if constexpr (FastShape<N1>::A == 16) {
          if (ctx->prune_x) go(std::true_type{});
          else go(std::false_type{});
        } else {
          go(std::false_type{});
        }
      } else {
        LAUNCH(MC_KID_XRHO, (k_xrho<T, N1>), dim3((ctx->n2 + s.ty - 1) / s.ty, ctx->n3), dim3(kThreads), shx, w, B,
               np, pair0, nstate, ctx->w_f.as<double>(), rho, s.ty);
      }
    });
  }
  if (K) {   // the columns of every pair, for the next vpsi on these wavefunctions
    ctx->keep_valid = true;
    ctx->keep_c0 = c0;
    ctx->keep_nstate = nstate;
    ctx->keep_tsize = (int)sizeof(T);
    ctx->keep_epoch = wfn_epoch();
  }
  if (npairs > 0) LAUNCH(MC_KID_MISC, k_reduce_pairs, dim3(npairs), dim3(64), 0, red, s.gx, pair_sums);
  LAUNCH(MC_KID_MISC, k_sum_partial, dim3(nblk_sum), dim3(kThreads), 0, rho, nnr1, rsum_part);
  std::vector<double> hs((size_t)npairs * 4 + nblk_sum);
  dl_add(ctx, hs.data(), pair_sums, sizeof(double) * hs.size());
  dl_finish(ctx);
  double sg = 0.0, sk = 0.0, sr = 0.0;
  for (int p = 0; p < npairs; ++p) {
    const int is1 = 2 * p, is2 = is1 + 1;
    if (f_host[is1] != 0.0) { sg += f_host[is1] * hs[4 * p + 0]; sk += f_host[is1] * hs[4 * p + 1]; }
    if (is2 < nstate && f_host[is2] != 0.0) { sg += f_host[is2] * hs[4 * p + 2]; sk += f_host[is2] * hs[4 * p + 3]; }
  }
  for (int b = 0; b < nblk_sum; ++b) sr += hs[(size_t)npairs * 4 + b];
  double cr = sr * ctx->omega / ((double)ctx->n1 * ctx->n2 * ctx->n3);
  ctx->last_csumg = sg;
  ctx->last_csumr_raw = cr;
  if (std::is_same<T, float>::value && cr > 0.0 && sg > 0.0) {
    // The FP32 engine contracts norms systematically: its DFT constants and
    // twiddles rounded to FP32 make each 192-point transform scale |x|^2 by
    // 1 - 5.9e-8 on average (the rounded sin(2 pi/3), cos(pi/8), ... lie
    // below their exact values), so the three inverse passes take 1.8e-7 of
    // the charge (csumr - csumg = -1.6e-4 at 1024 electrons).  rho is
    // rescaled so that its integral is the FP64 G-space charge csumg (the
    // invariant of CPMD's csumr / csumg check, rhoofr:625-635); the raw sum
    // stays available through mc_rhoofr_raw_charge.
    const double scale = sg / cr;
    LAUNCH(MC_KID_MISC, k_scale, dim3(1024), dim3(256), 0, rho, nnr1, scale);
    cr *= scale;
  }
  if (ekin) *ekin = sk * ctx->tpiba2;
  if (csumg) *csumg = sg;
  if (csumr) *csumr = cr;
}

// the pass sequence of vpsi over batches of at most P pairs, with the
// workspaces A, B (P pairs each) and the occupations f_dev given
template <typename T>
static void vpsi_run(mc_ctx* ctx, const double2* c0, double2* c2, const double* f_dev, const double* vpot,
                     int nstate, bool acc, int P, v2<T>* A, v2<T>* B, v2<T>* K = nullptr) {
  const WParams w = make_params(ctx);
  const Shapes s = shapes<T>(ctx);
  const int npairs = (nstate + 1) / 2;
  using V = v2<T>;
  const size_t shz = sizeof(V) * (size_t)ctx->n3 * (1 + 2 * s.lo);
  const size_t shf = sizeof(V) * (size_t)ctx->n3 * (1 + 2 * s.lf);
  const size_t shy = sizeof(V) * (size_t)ctx->n2 * (1 + s.tz);
  const size_t shx = sizeof(V) * (size_t)ctx->n1 * (1 + s.ty);
  const double scale = 1.0 / ((double)ctx->n1 * ctx->n2 * ctx->n3);
  const size_t bslab = (size_t)ctx->n3 * ctx->ncols * ctx->n2;
  for (int pair0 = 0; pair0 < npairs; pair0 += P) {
    const int np = std::min(P, npairs - pair0);
    if (K) {   // the inverse z and y passes of rhoofr, kept: start from its columns
      B = K + (size_t)pair0 * bslab;
    } else {
      with_length(ctx->n3, [&](auto NN) {
        LAUNCH(MC_KID_ZINV, (k_zinv<T, decltype(NN)::value>), dim3(s.gx, np), dim3(kThreads), shz, w, c0, nstate,
               pair0, A, s.lo, (double*)nullptr);
      });
      ypass<T, 1>(ctx, w, s, A, B, np, shy);
    }
    with_length(ctx->n1, [&](auto NN) {
      constexpr int N1 = decltype(NN)::value;
      if constexpr (f4_ok<T, N1>()) {
        auto go = [&](auto PRC) {
          constexpr bool PR = decltype(PRC)::value;
          f4_allow_lds(&k_xvpsi_f<T, N1, PR>, f4_lds<T, N1>());
          LAUNCH(MC_KID_XVPSI, (k_xvpsi_f<T, N1, PR>), dim3((ctx->n3 + F4Tile<N1>::L - 1) / F4Tile<N1>::L, ctx->n2),
                 dim3(kThreads), (f4_lds<T, N1>()), w, B, np, vpot);
        };
        if constexpr (FastShape<N1>::A == 16) {
          if (ctx->prune_x) go(std::true_type{});
          else go(std::false_type{});
        } else {
          go(std::false_type{});
        }
      } else {
        LAUNCH(MC_KID_XVPSI, (k_xvpsi<T, N1>), dim3((ctx->n2 + s.ty - 1) / s.ty, ctx->n3), dim3(kThreads), shx, w,
               B, np, vpot, s.ty);
      }
    });
    ypass<T, -1>(ctx, w, s, A, B, np, shy);
    with_length(ctx->n3, [&](auto NN) {
      LAUNCH(MC_KID_ZFWD, (k_zfwd<T, decltype(NN)::value>), dim3(s.gf, np), dim3(kThreads), shf, w, A, c0, c2,
             f_dev, nstate, pair0, s.lf, scale, acc ? 1 : 0);
    });
  }
}

template <typename T>
static void vpsi_impl(mc_ctx* ctx, const double2* c0, double2* c2, const double* f_host, const double* vpot,
                      int nstate, bool acc, bool reuse) {
  (void)shapes<T>(ctx);   // rejects unsupported lengths before any allocation
  const int npairs = (nstate + 1) / 2;
  if (npairs == 0) return;
  const size_t f_bytes = sizeof(double) * std::max(nstate, 1);
  WsUse use(ctx->w_a, ctx->w_b);
  // rhoofr's columns of these wavefunctions, if nothing wrote device
  // wavefunctions since (the epoch) and they are still held
  v2<T>* K = nullptr;
  if (reuse && ctx->col_reuse && ctx->keep_valid && ctx->keep_c0 == c0 && ctx->keep_nstate == nstate &&
      ctx->keep_tsize == (int)sizeof(T) && ctx->keep_epoch == wfn_epoch() && ctx->w_keep.p &&
      ctx->w_keep.bytes >= sizeof(v2<T>) * (size_t)npairs * ctx->n3 * ctx->ncols * ctx->n2)
    K = ctx->w_keep.as<v2<T>>();
  ctx->keep_valid = false;   // the x pass below overwrites them
  std::optional<WsUse> use_keep;
  if (K)
    use_keep.emplace(ctx->w_keep, ctx->w_keep);
  else if (mem_limit() != 0)
    ctx->w_keep.release();
  const int P = fit_pairs<T>(ctx, std::max(1, std::min(ctx->batch_pairs_vpsi, npairs)),
                             f_bytes > ctx->w_f.bytes ? f_bytes - ctx->w_f.bytes : 0);
  ensure_workspace<T>(ctx, P, K == nullptr);
  upload_f(ctx, f_host, nstate);
  vpsi_run<T>(ctx, c0, c2, ctx->w_f.as<double>(), vpot, nstate, acc, P, ctx->w_a.as<v2<T>>(),
              K ? nullptr : ctx->w_b.as<v2<T>>(), K);
}

// ---------------------------------------------------------------------------
// Allocation-free vpsi.  The occupations and the transform workspaces live in
// `scratch`, one fixed array per GPU that the code object carries, instead of
// context buffers that grow with hipMalloc.  In host code `scratch` names a
// host-side stand-in: only the address that hipGetSymbolAddress returns for
// the current device (kept per device by scratch_init) is passed to kernels
// and copies.  Every call on one GPU shares the array: calls that overlap in
// time on one GPU (other streams or contexts) race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[MC_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxScratchDevices] = {};

void* scratch_init(int device) {
  MC_CHECK(device >= 0 && device < kMaxScratchDevices, MC_ERR_UNSUPPORTED, "device ordinal beyond 63");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = -1;
    MC_HIP(hipGetDevice(&cur));
    MC_CHECK(cur == device, MC_ERR_INTERNAL, "scratch_init: device is not current");
    void* p = nullptr;
    MC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    MC_CHECK(p != nullptr, MC_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
    g_scratch_ptr[device] = p;
  }
  return g_scratch_ptr[device];
}

static size_t round256(size_t b) { return (b + 255) / 256 * 256; }

// the per-kernel timer creates events: it is held off for the call
struct TimerHold {
  mc_ctx* c;
  bool was;
  explicit TimerHold(mc_ctx* ctx) : c(ctx), was(ctx->timer.enabled) { c->timer.enabled = false; }
  ~TimerHold() { c->timer.enabled = was; }
};

template <typename T>
static void vpsi_noalloc_impl(mc_ctx* ctx, const double2* c0, double2* c2, const double* f_host,
                              const double* vpot, int nstate) {
  using V = v2<T>;
  (void)shapes<T>(ctx);
  const int npairs = (nstate + 1) / 2;
  if (npairs == 0) return;
  MC_CHECK(ctx->scratch != nullptr, MC_ERR_INTERNAL, "no scratch address for this device");
  const size_t fbytes = round256(sizeof(double) * (size_t)nstate);
  const size_t abytes = sizeof(V) * (size_t)ctx->n_slots * ctx->n3;           // sticks of one pair
  const size_t bbytes = sizeof(V) * (size_t)ctx->n3 * ctx->ncols * ctx->n2;   // columns of one pair
  const size_t cap = MC_NOALLOC_SCRATCH_BYTES;
  // P pairs take fbytes + round256(P abytes) + P bbytes <= cap
  long long P = 0;
  if (fbytes + 256 < cap) P = (long long)((cap - fbytes - 256) / (abytes + bbytes));
  P = std::min<long long>(P, std::min(ctx->batch_pairs_vpsi, npairs));
  if (P < 1)
    throw Error{MC_ERR_ARG, "mc_vpsi_dev_noalloc: the occupations and one state pair's transform workspaces need " +
                                std::to_string(fbytes + round256(abytes) + bbytes) + " bytes, beyond the " +
                                std::to_string(cap) + "-byte scratch"};
  unsigned char* base = static_cast<unsigned char*>(ctx->scratch);
  double* f_dev = reinterpret_cast<double*>(base);
  V* A = reinterpret_cast<V*>(base + fbytes);
  V* B = reinterpret_cast<V*>(base + fbytes + round256((size_t)P * abytes));
  MC_CHECK(fbytes + round256((size_t)P * abytes) + (size_t)P * bbytes <= cap, MC_ERR_INTERNAL,
           "scratch carving exceeds the array");
  TimerHold hold(ctx);
  MC_HIP(hipMemcpyAsync(f_dev, f_host, sizeof(double) * nstate, hipMemcpyHostToDevice, ctx->stream));
  vpsi_run<T>(ctx, c0, c2, f_dev, vpot, nstate, true, (int)P, A, B);
}

void vpsi_noalloc_dev(mc_ctx* ctx, const double2* c0, double2* c2, const double* f_host, const double* vpot,
                      int nstate) {
  if (ctx->fft_fp32)
    vpsi_noalloc_impl<float>(ctx, c0, c2, f_host, vpot, nstate);
  else
    vpsi_noalloc_impl<double>(ctx, c0, c2, f_host, vpot, nstate);
}

void rhoofr_dev(mc_ctx* ctx, const double2* c0, const double* f_host, int nstate, double* rho, double* ekin,
                double* csumg, double* csumr, bool keep) {
  if (ctx->fft_fp32)
    rhoofr_impl<float>(ctx, c0, f_host, nstate, rho, ekin, csumg, csumr, keep);
  else
    rhoofr_impl<double>(ctx, c0, f_host, nstate, rho, ekin, csumg, csumr, keep);
}

void vpsi_dev(mc_ctx* ctx, const double2* c0, double2* c2, const double* f_host, const double* vpot, int nstate,
              bool acc, bool reuse) {
  if (ctx->fft_fp32)
    vpsi_impl<float>(ctx, c0, c2, f_host, vpot, nstate, acc, reuse);
  else
    vpsi_impl<double>(ctx, c0, c2, f_host, vpot, nstate, acc, reuse);
}

// rows ngw .. ldg-1 of the first n states of c (ld ldg complex rows) <- 0
__global__ void k_zero_pad_rows(double2* __restrict__ c, int ngw, int ldg, int n) {
  const int np = ldg - ngw;
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (t >= (size_t)np * n) return;
  const int col = (int)(t / np), r = ngw + (int)(t - (size_t)col * np);
  c[(size_t)col * ldg + r] = make_double2(0.0, 0.0);
}

void forces_c2_dev(mc_ctx* ctx, const double2* c0, double2* c2, const double* f_host, const double* vpot,
                   int nstate) {
  if (nstate <= 0) return;
  if (ctx->ldg > ctx->ngw) {
    const size_t nt = (size_t)(ctx->ldg - ctx->ngw) * nstate;
    LAUNCH(MC_KID_MISC, k_zero_pad_rows, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c2, ctx->ngw, ctx->ldg,
           nstate);
  }
  vpsi_dev(ctx, c0, c2, f_host, vpot, nstate, false, true);
  if (ctx->proj.ready) fnonloc_dev(ctx, c2, f_host, nstate, nullptr);   // no projectors: local potential only
}

}  // namespace mc
