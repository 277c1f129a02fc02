// Split-bf16 A^T B contractions from pre-split operand planes (gfx950).
//
// The tall operands (wavefunctions, projectors: K = 2 ngw rows, a few hundred
// columns) are split once into NP bf16 parts, x = p0 + p1 (+ p2) to 16 (24)
// bits, and stored as "planes": blocks of 128 columns x 32 rows, the NP parts
// of a block contiguous (NP x 8 KiB), blocks of one 128-column group
// consecutive along K.  Inside a part, column c holds its 32 rows in 64 bytes
// as four 16-byte chunks (k = 8q .. 8q+7), chunk q stored at q ^ ((c >> 1) & 3):
// the MFMA operand reads (16 lanes = 16 columns, one chunk each) then hit 16
// distinct 16-byte bank groups.
//
// The contraction streams the blocks of a 128 x 128 tile through a three-stage
// LDS ring with global_load_lds_dwordx4 (each wave copies 1 KiB runs straight
// into LDS, no registers), so two K slices are in flight while the third is
// multiplied.  Each 32-row slice is summed by the MFMA from a zero
// accumulator (the MFMA aligns its products to the accumulator it is given
// and drops what falls below its last bit) and added to an FP32 accumulator
// on the VALU; a workgroup covers at most 64 slices and the chunk partials are
// reduced in FP64.  Products: NP = 3 sums p1 p1, p2 q0,
// p0 q2, p1 q0, p0 q1, p0 q0 (six MFMAs, error ~2^-16 of |x y| per term and
// unbiased); NP = 2 sums p1 q0, p0 q1, p0 q0.
//
// Diagonal of square overlaps: the products dropped by the six-term split
// (p1 q2, p2 q1, p2 q2) have random signs also on a diagonal sum_k a_k^2, so
// the diagonal carries no systematic bias (unlike a two-part split, whose
// dropped lo*lo is always positive there) and needs no separate FP64 sum.
//
// MFMA operand layout (16x16x32 bf16): lane l holds A[i = l%16][k = 8(l/16) ..
// 8(l/16)+7], B[k = 8(l/16) ..][j = l%16], and D[i = 4(l/16) + r][j = l%16] in
// accumulator register r.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mc_internal.h"

namespace mc {

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int kPT = 128;                  // columns per block / tile edge
constexpr int kPK = 32;                   // rows per block (one MFMA depth)
constexpr int kPart = kPT * kPK * 2;      // bytes of one part of one block
constexpr int kStages = 3;
constexpr int kXcdP = 8;
constexpr int kMaxSlices = 64;            // K chunk (slices of 32 rows) per workgroup: FP32 accumulation
constexpr int kTnWaves = 8;               // waves per k_tn_planes workgroup (two per SIMD)
constexpr int kTnWorkgroups = 512;        // k_tn_planes workgroups aimed at (two per CU)

// chunk q (16 bytes) of column c of a part sits at slot swz(c, q): columns are
// 64 bytes apart, so in a 128-byte bank row columns c and c + 2 share the
// slots; the XOR with (c >> 1) & 3 gives any 8 consecutive columns (the lanes
// of one LDS read phase) 8 distinct slots
__device__ __forceinline__ int swz(int c, int q) { return q ^ ((c >> 1) & 3); }

typedef __attribute__((address_space(3))) void lds_void;

// 16 bytes per lane from src straight into LDS at dst + 16 * lane (dst
// uniform over the wave); outside the kernel template so that the host pass
// does not instantiate the device-only builtin
__device__ __forceinline__ void lds_copy16(const void* src, lds_void* dst) {
  __builtin_amdgcn_global_load_lds(src, dst, 16, 0, 0);
}

__device__ __forceinline__ int xcd_tile_p(int total) {
  const int id = blockIdx.x;
  const int per = (total + kXcdP - 1) / kXcdP;
  return (id % kXcdP) * per + id / kXcdP;
}

// split 8 FP32 values into NP bf16 parts (round to nearest at each step)
template <int NP>
__device__ __forceinline__ void split_parts(float (&v)[8], bf8 (&p)[NP]) {
#pragma unroll
  for (int q = 0; q < NP; ++q)
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const __bf16 h = (__bf16)v[u];
      p[q][u] = h;
      v[u] -= (float)h;
    }
}

template <int NP>
__device__ __forceinline__ void store_parts(__bf16* out, size_t blk, int c, int q, const bf8 (&p)[NP]) {
#pragma unroll
  for (int r = 0; r < NP; ++r)
    *reinterpret_cast<bf8*>(out + (blk * NP + r) * (kPart / 2) + c * kPK + swz(c, q) * 8) = p[r];
}

}  // namespace

// X (K x n, column-major, leading dimension ldx doubles; ldx and the column
// starts 16-byte aligned) -> planes.  One thread per (block, column, chunk),
// chunk fastest: four lanes read one column's 256 contiguous bytes.
template <int NP>
__global__ __launch_bounds__(256) void k_split_planes(const double* __restrict__ X, int ldx, int K, int n, int nkb,
                                                      size_t nblk, __bf16* __restrict__ out) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  const int q = (int)(t & 3), c = (int)((t >> 2) & (kPT - 1));
  const size_t blk = t >> 9;
  if (blk >= nblk) return;
  const int cb = (int)(blk / nkb), kb = (int)(blk - (size_t)cb * nkb);
  const int col = cb * kPT + c, k0 = kb * kPK + 8 * q;
  float v[8];
  if (col < n && k0 + 8 <= K) {
    const double2* src = reinterpret_cast<const double2*>(X + (size_t)col * ldx + k0);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double2 d = src[u];
      v[2 * u] = (float)d.x;
      v[2 * u + 1] = (float)d.y;
    }
  } else {
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = (col < n && k0 + u < K) ? (float)X[(size_t)col * ldx + k0 + u] : 0.f;
  }
  bf8 p[NP];
  split_parts<NP>(v, p);
  store_parts<NP>(out, blk, c, q, p);
}

// velupa + posupa with the planes of both coefficient sets as a by-product:
// cp = c0 + dt (cm + dtm c2) (c2 may be null) over all ldx stored rows, and
// the bf16 planes of c0 and of cp for the nkb = ceil(K / 32) row blocks of the
// contractions (layout of k_split_planes; rows >= K and columns >= n zero).
// One wave per (512-row segment, column): lane l moves rows 2l + 128u (u < 4)
// as 16-byte runs, so every load and store of c0, c2, cm, cp covers 1 KiB of a
// column; the fp32 values pass through LDS so that lane l then holds rows
// 8l..8l+7, the plane chunk it splits and stores.  cm is read, not written: the
// caller's final rotation sets the velocity (cp_final - c0) / dt.
// KICK (velupa before rortv): cp = cm + dtm c2 instead, written over cm (cp ==
// cm: every element is read before the same thread writes it), with the
// planes of c0 and of the kicked cm.  C0P false: the planes of c0 are already
// held (kept from an earlier call) and are not written; with KICK, c0 is then
// not read at all.
template <int NP, bool KICK, bool C0P>
__global__ __launch_bounds__(256) void k_posupa_planes(const double* __restrict__ c0, const double* __restrict__ c2,
                                                       const double* cm, double* cp, int ldx, int K,
                                                       int n, int ncolp, int nseg, int nkb, double dtm, double dt,
                                                       __bf16* __restrict__ pl0, __bf16* __restrict__ plp) {
  __shared__ float4 lds[4][2][128];   // [wave][c0 | cp][512 rows]
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const size_t unit = (size_t)blockIdx.x * 4 + wv;
  const bool live = unit < (size_t)ncolp * nseg;
  // columns fastest: the waves of a workgroup write neighbouring columns of a
  // plane block, whose 64-byte column chunks share 128-byte lines
  const int seg = live ? (int)(unit / ncolp) : 0;
  const int col = live ? (int)(unit - (size_t)seg * ncolp) : 0;
  float2* s0 = reinterpret_cast<float2*>(lds[wv][0]);
  float2* sp = reinterpret_cast<float2*>(lds[wv][1]);
  const int r0 = seg * 512 + 2 * lane;
  const size_t o = (size_t)col * ldx + r0;
  const bool any = live && col < n;
  double2 x[4], w[4], f[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const bool ok = any && r0 + 128 * u < ldx;
    x[u] = ok && (C0P || !KICK) ? *reinterpret_cast<const double2*>(c0 + o + 128 * u) : make_double2(0.0, 0.0);
    w[u] = ok ? *reinterpret_cast<const double2*>(cm + o + 128 * u) : make_double2(0.0, 0.0);
    f[u] = ok && c2 ? *reinterpret_cast<const double2*>(c2 + o + 128 * u) : make_double2(0.0, 0.0);
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    w[u].x += dtm * f[u].x;
    w[u].y += dtm * f[u].y;
    f[u] = KICK ? w[u] : make_double2(x[u].x + dt * w[u].x, x[u].y + dt * w[u].y);   // f now holds cp
    const bool in = r0 + 128 * u < K;                                  // K even: whole complex rows
    if (C0P) s0[lane + 64 * u] = in ? make_float2((float)x[u].x, (float)x[u].y) : make_float2(0.f, 0.f);
    sp[lane + 64 * u] = in ? make_float2((float)f[u].x, (float)f[u].y) : make_float2(0.f, 0.f);
  }
  __syncthreads();
  const int kb = seg * 16 + (lane >> 2);
  if (live && kb < nkb) {
    const size_t blk = (size_t)(col / kPT) * nkb + kb;
    const int c = col % kPT, q = lane & 3;
    float v[8];
    bf8 p[NP];
    if constexpr (C0P) {
      const float4 a0 = lds[wv][0][2 * lane], a1 = lds[wv][0][2 * lane + 1];
      v[0] = a0.x; v[1] = a0.y; v[2] = a0.z; v[3] = a0.w; v[4] = a1.x; v[5] = a1.y; v[6] = a1.z; v[7] = a1.w;
      split_parts<NP>(v, p);
      store_parts<NP>(pl0, blk, c, q, p);
    }
    const float4 b0 = lds[wv][1][2 * lane], b1 = lds[wv][1][2 * lane + 1];
    v[0] = b0.x; v[1] = b0.y; v[2] = b0.z; v[3] = b0.w; v[4] = b1.x; v[5] = b1.y; v[6] = b1.z; v[7] = b1.w;
    split_parts<NP>(v, p);
    store_parts<NP>(plp, blk, c, q, p);
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    if (any && r0 + 128 * u < ldx) *reinterpret_cast<double2*>(cp + o + 128 * u) = f[u];
  }
}

// The projector operand of rnlsm1 / rnlsm2 (as ProjA in gemm.hip) -> planes:
// row 2g / 2g+1 = Re / Im of (-i)^l tw_g eigr(atom, g), tw_g = twnl_g [gk_g]
// (halved at G = 0).
template <int NP>
__global__ __launch_bounds__(256) void k_proj_planes(ProjA pa, int K, int m, int nkb, size_t nblk,
                                                     __bf16* __restrict__ out) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  const int q = (int)(t & 3), c = (int)((t >> 2) & (kPT - 1));
  const size_t blk = t >> 9;
  if (blk >= nblk) return;
  const int cb = (int)(blk / nkb), kb = (int)(blk - (size_t)cb * nkb);
  const int col = cb * kPT + c, k0 = kb * kPK + 8 * q;
  float v[8];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int g = (k0 >> 1) + u;
    double re = 0.0, im = 0.0;
    if (col < m && 2 * g + 1 < K) {
      const double2 e = pa.eigr[(size_t)col * pa.ldg + g];
      double tw = pa.twnl[g];
      if (pa.gkk) tw *= pa.gkk[g];
      if (pa.geq0 && g == 0) tw *= 0.5;
      switch (pa.lmod4) {   // (-i)^l * tw * e
        case 0: re = tw * e.x; im = tw * e.y; break;
        case 1: re = tw * e.y; im = -tw * e.x; break;
        case 2: re = -tw * e.x; im = -tw * e.y; break;
        default: re = -tw * e.y; im = tw * e.x; break;
      }
    }
    v[2 * u] = (float)re;
    v[2 * u + 1] = (float)im;
  }
  bf8 p[NP];
  split_parts<NP>(v, p);
  store_parts<NP>(out, blk, c, q, p);
}

// Stacked projector variants (one thread per column, chunk of the first
// variant's geometry; eigr, twnl read once for all nv variants): variant v of
// column c at column block v * ncbv + c / 128; lbase = l mod 4 of the group.
template <int NP>
__global__ __launch_bounds__(256) void k_proj_planes_stacked(ProjA pa, const double* __restrict__ gk, int lbase, int v0,
                                                             int nv, int K, int m, int nkb, int ncbv, size_t nblk,
                                                             __bf16* __restrict__ out) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  const int q = (int)(t & 3), c = (int)((t >> 2) & (kPT - 1));
  const size_t blk = t >> 9;
  if (blk >= nblk) return;
  const int cb = (int)(blk / nkb), kb = (int)(blk - (size_t)cb * nkb);
  const int col = cb * kPT + c, k0 = kb * kPK + 8 * q;
  double er[4], ei[4], tw[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int g = (k0 >> 1) + u;
    er[u] = ei[u] = tw[u] = 0.0;
    if (col < m && 2 * g + 1 < K) {
      double2 e;
      if (pa.ph1) {   // eigr from the phase tables, as k_eigr forms it
        const int isa = pa.isa0 + col;
        const double2 a = pa.ph1[(size_t)isa * (pa.n1 + 1) + pa.gmw[3 * g] + pa.n1 / 2];
        const double2 b = pa.ph2[(size_t)isa * (pa.n2 + 1) + pa.gmw[3 * g + 1] + pa.n2 / 2];
        const double2 c = pa.ph3[(size_t)isa * (pa.n3 + 1) + pa.gmw[3 * g + 2] + pa.n3 / 2];
        const double2 ab = make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
        e = make_double2(ab.x * c.x - ab.y * c.y, ab.x * c.y + ab.y * c.x);
      } else {
        e = pa.eigr[(size_t)col * pa.ldg + g];
      }
      er[u] = e.x;
      ei[u] = e.y;
      tw[u] = pa.twnl[g] * ((pa.geq0 && g == 0) ? 0.5 : 1.0);
    }
  }
  for (int vv = 0; vv < nv; ++vv) {
    const int v = v0 + vv;
    const int lm = (lbase + (v > 0 ? 1 : 0)) & 3;
    float x[8];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int g = (k0 >> 1) + u;
      double w = tw[u];
      if (v > 0 && w != 0.0) w *= gk[(size_t)(v - 1) * pa.ldg + g];
      double re, im;
      switch (lm) {   // (-i)^l * w * e
        case 0: re = w * er[u]; im = w * ei[u]; break;
        case 1: re = w * ei[u]; im = -w * er[u]; break;
        case 2: re = -w * er[u]; im = -w * ei[u]; break;
        default: re = -w * ei[u]; im = w * er[u]; break;
      }
      x[2 * u] = (float)re;
      x[2 * u + 1] = (float)im;
    }
    bf8 p[NP];
    split_parts<NP>(x, p);
    store_parts<NP>(out, ((size_t)vv * ncbv + cb) * nkb + kb, c, q, p);
  }
}

// ---------------------------------------------------------------------------
// A^T B partials from planes: P[s] (m x n, FP64) over the K blocks
// [s kbper, (s+1) kbper); launch ceil(total/8)*8 workgroups, total = nsplit *
// tiles, tiles = tm tn, or with upper (m == n) the tm (tm + 1) / 2 tiles on or
// above the diagonal.
// ---------------------------------------------------------------------------
template <int NP>
__global__ __launch_bounds__(64 * kTnWaves, 1) void k_tn_planes(const __bf16* __restrict__ Apl,
                                                                const __bf16* __restrict__ Bpl, int nkb, int m, int n,
                                                                int kbper, double* __restrict__ P, int upper, int tm,
                                                                int tn, int total) {
  constexpr int NW = kTnWaves;                     // waves: 2 x NW/2 blocks of 64 x (128 / (NW/2)) of the tile
  constexpr int NBC = 8 / NW * 2;                  // 16-column MFMA tiles per wave (4 with 4 waves, 2 with 8)
  constexpr int kStage = 2 * NP * kPart;           // bytes of one ring stage (A block, B block)
  constexpr int kRuns = 2 * NP * kPart / 1024;     // 1 KiB runs per stage
  static_assert(kRuns % NW == 0, "runs per wave");
  constexpr int kPer = kRuns / NW;                 // copies per wave and stage
  static_assert(kPer < 64, "vmcnt range");
  constexpr int kWait1 = (kPer & 15) | 0x70 | 0xF00 | ((kPer >> 4) << 14);
  constexpr int kWait0 = 0x70 | 0xF00;
  __shared__ __attribute__((aligned(16))) unsigned char ring[kStages * kStage];
  const int t = xcd_tile_p(total);
  if (t >= total) return;
  const int tiles = tile_count(upper, tm, tn);
  const int s = t / tiles;
  int ti, tj;
  tile_ij(t - s * tiles, upper, tm, tn, ti, tj);
  const int kb0 = s * kbper, kb1 = min(nkb, kb0 + kbper);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const unsigned char* Ab = reinterpret_cast<const unsigned char*>(Apl) + (size_t)ti * nkb * NP * kPart;
  const unsigned char* Bb = reinterpret_cast<const unsigned char*>(Bpl) + (size_t)tj * nkb * NP * kPart;
  // wave w copies runs w, w + NW, ... of a stage: run j < NP*8 from the A
  // block, the rest from the B block
  auto issue = [&](int kb, int stage) {
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int j = wave + NW * i;
      const bool isb = j >= NP * 8;
      const unsigned char* src =
          (isb ? Bb : Ab) + (size_t)kb * NP * kPart + (size_t)(isb ? j - NP * 8 : j) * 1024 + lane * 16;
      lds_copy16(src, (lds_void*)(ring + stage * kStage + j * 1024));
    }
  };
  const int wr = (wave / (NW / 2)) * 64, wc = (wave % (NW / 2)) * (16 * NBC);
  const int li = lane & 15, lq = lane >> 4;
  int offa[4], offb[NBC];   // byte offsets of this lane's chunk in a part
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int r = wr + 16 * a + li;
    offa[a] = r * (kPK * 2) + swz(r, lq) * 16;
  }
#pragma unroll
  for (int b = 0; b < NBC; ++b) {
    const int c = wc + 16 * b + li;
    offb[b] = c * (kPK * 2) + swz(c, lq) * 16;
  }
  // Accumulation: the six MFMAs of one slice and element start from zero,
  // small products first, so that none is added to a larger partial sum
  // (the MFMA aligns its products to the accumulator it is given and drops
  // what falls below its last bit, a same-signed loss on sums of squares);
  // the slice sums go to an FP32 accumulator on the VALU, which is added to
  // an FP64 one every kMaxSlices slices; the host reduces the workgroups'
  // FP64 partials in a fixed order.
  f4 acc32[4][NBC];
  double acc64[4][NBC][4];   // the chunk sums in FP64
  const f4 z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < NBC; ++b) {
      acc32[a][b] = z;
#pragma unroll
      for (int r = 0; r < 4; ++r) acc64[a][b][r] = 0.0;
    }
  const int nsl = kb1 - kb0;
  if (nsl > 0) issue(kb0, 0);
  if (nsl > 1) issue(kb0 + 1, 1);
#pragma unroll 1
  for (int i = 0; i < nsl; ++i) {
    // this wave's copies of slice i have landed once at most the next
    // slice's copies are outstanding (s_waitcnt vmcnt(N), gfx9 encoding:
    // vmcnt[3:0] | expcnt[6:4] | lgkmcnt[11:8] | vmcnt[5:4] << 14)
    if (i + 1 < nsl) __builtin_amdgcn_s_waitcnt(kWait1);
    else __builtin_amdgcn_s_waitcnt(kWait0);
    // One barrier per slice, without a memory fence (a fence would wait for
    // every outstanding copy and drain the ring; the "memory" clobber keeps
    // the compiler from moving LDS accesses across it): past it, every
    // wave's copies of slice i have landed and every wave is done with slice
    // i - 1, whose stage is refilled with slice i + 2.
    asm volatile("s_barrier" ::: "memory");
    if (i + 2 < nsl) issue(kb0 + i + 2, (i + 2) % kStages);
    const unsigned char* st = ring + (i % kStages) * kStage;
    bf8 ap[NP][4];
#pragma unroll
    for (int p = 0; p < NP; ++p)
#pragma unroll
      for (int a = 0; a < 4; ++a) ap[p][a] = *reinterpret_cast<const bf8*>(st + p * kPart + offa[a]);
#pragma unroll
    for (int b = 0; b < NBC; ++b) {
      bf8 bp[NP];
#pragma unroll
      for (int p = 0; p < NP; ++p) bp[p] = *reinterpret_cast<const bf8*>(st + (NP + p) * kPart + offb[b]);
      f4 u[4];
      // the four row blocks' chains in lockstep (independent consecutive MFMAs)
      if (NP == 2) {
#pragma unroll
        for (int a = 0; a < 4; ++a) u[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[1][a], bp[0], z, 0, 0, 0);
#pragma unroll
        for (int a = 0; a < 4; ++a) u[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[0][a], bp[1], u[a], 0, 0, 0);
      } else {
#pragma unroll
        for (int a = 0; a < 4; ++a) u[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[1][a], bp[1], z, 0, 0, 0);
#pragma unroll
        for (int a = 0; a < 4; ++a)
          u[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[NP - 1][a], bp[0], u[a], 0, 0, 0);
#pragma unroll
        for (int a = 0; a < 4; ++a)
          u[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[0][a], bp[NP - 1], u[a], 0, 0, 0);
#pragma unroll
        for (int a = 0; a < 4; ++a) u[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[1][a], bp[0], u[a], 0, 0, 0);
#pragma unroll
        for (int a = 0; a < 4; ++a) u[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[0][a], bp[1], u[a], 0, 0, 0);
      }
#pragma unroll
      for (int a = 0; a < 4; ++a) u[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[0][a], bp[0], u[a], 0, 0, 0);
#pragma unroll
      for (int a = 0; a < 4; ++a) acc32[a][b] += u[a];
    }
    if ((i + 1) % kMaxSlices == 0 || i + 1 == nsl) {   // a chunk of <= kMaxSlices slices ends
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < NBC; ++b) {
#pragma unroll
          for (int r = 0; r < 4; ++r) acc64[a][b][r] += (double)acc32[a][b][r];
          acc32[a][b] = z;
        }
    }
  }
  double* Ps = P + (size_t)s * m * n;
  const int i0 = ti * kPT, j0 = tj * kPT;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < NBC; ++b) {
      const int j = j0 + wc + 16 * b + li;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + wr + 16 * a + 4 * lq + r;
        if (i < m && j < n) Ps[(size_t)j * m + i] = acc64[a][b][r];
      }
    }
}

// ---------------------------------------------------------------------------
// D = alpha A G + beta D with the small G (k x n) as two-part bf16 planes
// (split once per product) and A (M x k, column-major, the tall wavefunction
// or projector operand) split per workgroup: 128 x 128 tiles of D, 4 waves of
// 64 x 64, K slices of 32.  A slices are fetched as FP64 into registers one
// slice ahead, split into hi/lo and staged in LDS as [row][k] (stride 40);
// G blocks are copied into a two-stage LDS ring with global_load_lds.  Each
// slice product hi*hi + hi*lo + lo*hi starts from a zero accumulator and is
// added to the FP32 tile sum on the VALU.  Epilogue as k_x3_nn (plain
// alpha/beta, or the fnonloc twnl * ctm factors).  upper_g: G upper
// triangular (the k loop stops after the diagonal block).
// ---------------------------------------------------------------------------
// PL: the epilogue also writes the planes of D_new (FnlEpi::pl); SQ: the
// per-block column sums of |D_new|^2 (FnlEpi::colsq).  Compile-time switches,
// so that the other products carry none of their code.
template <bool PL, bool SQ>
__global__ __launch_bounds__(256, 3) void k_nn_planes(const double* __restrict__ A, int lda,
                                                      const __bf16* __restrict__ Gpl, int gnkb, int M, int k, int n,
                                                      double alpha, double beta, double* __restrict__ D, int ldd,
                                                      int upper_g, FnlEpi epi, int tn, int total) {
  constexpr int kLdA = 40;
  constexpr int kGStage = 2 * kPart;   // two parts of one G block
  __shared__ __attribute__((aligned(16))) __bf16 As[2 * kPT * kLdA];
  __shared__ __attribute__((aligned(16))) unsigned char Gs[2 * kGStage];
  const int t = xcd_tile_p(total);
  if (t >= total) return;
  const int rt = t / tn, ct = t - rt * tn;
  const int i0 = rt * kPT, j0 = ct * kPT;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;
  const int li = lane & 15, lq = lane >> 4;
  const int kmax = upper_g ? min(k, j0 + kPT) : k;
  const int nsl = (kmax + kPK - 1) / kPK;
  // A slice (128 rows x 32 k): wave w loads k = 8w .. 8w+7 of rows i0+2l, i0+2l+1
  const int ar = 2 * lane, ak = 8 * wave;
  double2 ra[8];
  auto fetch_a = [&](int kb) {
    const int row = i0 + ar;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int kk = kb * kPK + ak + u;
      ra[u] = (kk < kmax && row < M) ? *reinterpret_cast<const double2*>(A + (size_t)kk * lda + row)
                                     : double2{0.0, 0.0};
    }
  };
  const unsigned char* Gb = reinterpret_cast<const unsigned char*>(Gpl) + (size_t)ct * gnkb * kGStage;
  auto issue_g = [&](int kb, int stage) {
#pragma unroll
    for (int i = 0; i < kGStage / 1024 / 4; ++i) {
      const int j = wave + 4 * i;
      lds_copy16(Gb + (size_t)kb * kGStage + j * 1024 + lane * 16, (lds_void*)(Gs + stage * kGStage + j * 1024));
    }
  };
  int offb[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int c = wc + 16 * b + li;
    offb[b] = c * (kPK * 2) + swz(c, lq) * 16;
  }
  const int offa = li * kLdA + 8 * lq;
  f4 acc[4][4];
  const f4 z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = z;
  if (nsl > 0) {
    fetch_a(0);
    issue_g(0, 0);
  }
  for (int s = 0; s < nsl; ++s) {
    {   // A slice -> hi / lo rows ar, ar + 1, k = ak .. ak + 7
      float v0[8], v1[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        v0[u] = (float)ra[u].x;
        v1[u] = (float)ra[u].y;
      }
      bf8 p0[2], p1[2];
      split_parts<2>(v0, p0);
      split_parts<2>(v1, p1);
      *reinterpret_cast<bf8*>(As + ar * kLdA + ak) = p0[0];
      *reinterpret_cast<bf8*>(As + kPT * kLdA + ar * kLdA + ak) = p0[1];
      *reinterpret_cast<bf8*>(As + (ar + 1) * kLdA + ak) = p1[0];
      *reinterpret_cast<bf8*>(As + kPT * kLdA + (ar + 1) * kLdA + ak) = p1[1];
    }
    __syncthreads();   // A slice staged and G block s landed (the fence waits for the copies)
    if (s + 1 < nsl) {
      fetch_a(s + 1);
      issue_g(s + 1, (s + 1) & 1);
    }
    const unsigned char* gs = Gs + (s & 1) * kGStage;
    bf8 ah[4], al[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      ah[a] = *reinterpret_cast<const bf8*>(As + (wr + 16 * a) * kLdA + offa);
      al[a] = *reinterpret_cast<const bf8*>(As + kPT * kLdA + (wr + 16 * a) * kLdA + offa);
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const bf8 bh = *reinterpret_cast<const bf8*>(gs + offb[b]);
      const bf8 bl = *reinterpret_cast<const bf8*>(gs + kPart + offb[b]);
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        f4 u = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh, z, 0, 0, 0);
        u = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl, u, 0, 0, 0);
        u = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh, u, 0, 0, 0);
        acc[a][b] += u;
      }
    }
    // every wave is done reading A slice s and G stage s & 1 (their reads
    // completed before the MFMAs); no fence: the next slice's copies stay in flight
    asm volatile("s_barrier" ::: "memory");
  }
  // epilogue: lane holds rows 4 lq + 16 a + wr + i0 .. +3 of column j0 + wc + 16 b + li.
  // Each wave stages its 64 x 16 block of columns wc + 16 b .. through its own
  // 4 KiB of Gs (fp32, 4-row groups XOR-swizzled by column), then reads it back
  // as rows 2 pr, 2 pr + 1 of two columns per instruction: the global reads and
  // writes of D (and of A and vcm) are 256-byte column runs.  The last loop
  // iteration ended with a barrier and no copy into Gs is in flight.
  float* st = reinterpret_cast<float*>(Gs) + wave * (64 * 16);
  const int hl = lane >> 5, pr = lane & 31;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
      *reinterpret_cast<f4*>(st + li * 64 + (((4 * a + lq) ^ li) * 4)) = acc[a][b];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // four column pairs at a time: their loads (D, and c0 for the velocity)
    // first, then the arithmetic and the stores
#pragma unroll
    for (int h = 0; h < 8; h += 4) {
      double2 dv[4], av[4];
      float2 xs[4];
#pragma unroll
      for (int it = 0; it < 4; ++it) {
        const int c = 2 * (h + it) + hl, row = 2 * pr;
        xs[it] = *reinterpret_cast<const float2*>(st + c * 64 + (((row >> 2) ^ c) * 4) + (row & 3));
        const int j = j0 + wc + 16 * b + c, i = i0 + wr + row;   // i even, M even: rows i, i + 1
        dv[it] = av[it] = make_double2(0.0, 0.0);
        if (j < n && i < M) {
          const double* d = D + (size_t)j * ldd + i;
          if (epi.twnl || beta != 0.0) dv[it] = make_double2(d[0], d[1]);
          if (!epi.twnl && epi.vcm) {
            const double* a = A + (size_t)j * lda + i;
            av[it] = make_double2(a[0], a[1]);
          }
        }
      }
#pragma unroll
      for (int it = 0; it < 4; ++it) {
        const int c = 2 * (h + it) + hl, row = 2 * pr;
        const int j = j0 + wc + 16 * b + c, i = i0 + wr + row;
        // FP32 copy of D_new for the planes, over this lane's own staging slot
        // (read above by this lane only); zero outside D
        float2* sy = reinterpret_cast<float2*>(st + c * 64 + (((row >> 2) ^ c) * 4) + (row & 3));
        // |D_new|^2 of this lane's complex row, over the staging words of its
        // column (the column's values were all read above, in one instruction)
        double* sq = reinterpret_cast<double*>(st) + c * 32 + pr;
        if (j >= n || i >= M) {
          if constexpr (PL) *sy = make_float2(0.f, 0.f);
          if constexpr (SQ) *sq = 0.0;
          continue;
        }
        double* d = D + (size_t)j * ldd + i;
        const double x0 = (double)xs[it].x, x1 = (double)xs[it].y;
        if (epi.twnl) {
          double f = epi.f[j];
          if (f < 1e-5) f = 1.0;
          const double tw = epi.twnl[i >> 1];
          double yr, yi;
          switch (epi.lmod4) {   // ctm = -f (-i)^l ; y = tw * ctm * x
            case 0: yr = -f * tw * x0; yi = -f * tw * x1; break;
            case 1: yr = -f * tw * x1; yi = f * tw * x0; break;
            case 2: yr = f * tw * x0; yi = f * tw * x1; break;
            default: yr = f * tw * x1; yi = -f * tw * x0; break;
          }
          d[0] = dv[it].x + yr;
          d[1] = dv[it].y + yi;
        } else {
          const double y0 = (beta == 0.0 ? 0.0 : beta * dv[it].x) + alpha * x0;
          const double y1 = (beta == 0.0 ? 0.0 : beta * dv[it].y) + alpha * x1;
          d[0] = y0;
          d[1] = y1;
          if (epi.vcm) {
            double* v = epi.vcm + (size_t)j * ldd + i;
            v[0] = epi.vodt * (y0 - av[it].x);
            v[1] = epi.vodt * (y1 - av[it].y);
          }
          if constexpr (PL) *sy = i < epi.pl_rows ? make_float2((float)y0, (float)y1) : make_float2(0.f, 0.f);
          if constexpr (SQ) *sq = (epi.sq_g0 && i == 0) ? 0.0 : y0 * y0 + y1 * y1;
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if constexpr (PL) {
      // planes of D_new for the wave's 64 rows x 16 columns: lane (c, r) splits
      // rows 8 r8 .. 8 r8 + 7 of column c, r8 = r and r + 4 (the chunks of
      // split_planes: rows >= M and columns >= n zero, as staged)
      const int c = lane & 15, j = j0 + wc + 16 * b + c;
#pragma unroll
      for (int rr = 0; rr < 2; ++rr) {
        const int r8 = (lane >> 4) + 4 * rr;
        const int kb = (i0 + wr + 8 * r8) >> 5;
        if (kb >= epi.pl_nkb) continue;
        const float4 lo = *reinterpret_cast<const float4*>(st + c * 64 + (((2 * r8) ^ c) * 4));
        const float4 hi = *reinterpret_cast<const float4*>(st + c * 64 + (((2 * r8 + 1) ^ c) * 4));
        float v[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
        const size_t blk = (size_t)(j / kPT) * epi.pl_nkb + kb;
        const int q = ((wr + 8 * r8) >> 3) & 3;
        if (epi.pl_np == 3) {
          bf8 p[3];
          split_parts<3>(v, p);
          store_parts<3>(epi.pl, blk, j % kPT, q, p);
        } else {
          bf8 p[2];
          split_parts<2>(v, p);
          store_parts<2>(epi.pl, blk, j % kPT, q, p);
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");   // every lane's reads before the next b's staging
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if constexpr (SQ) {
      // lane c < 16 sums column c's 32 complex rows of this wave in order
      if (lane < 16) {
        const double* q = reinterpret_cast<const double*>(st) + lane * 32;
        double t = 0.0;
        for (int r = 0; r < 32; ++r) t += q[r];
        const int j = j0 + wc + 16 * b + lane;
        if (j < n) epi.colsq[(size_t)j * epi.sq_nrb + (i0 + wr) / kColsqRows] = t;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
}

// out[j] = 2 sum_r P[j * nrb + r] + (g0 ? D(0, j)^2 : 0): one workgroup per
// column, thread t summing r = t, t + 256, ... in order, then a fixed tree
__global__ __launch_bounds__(256) void k_colsq_finish(const double* __restrict__ P, int nrb,
                                                      const double* __restrict__ D, int ldd, int g0,
                                                      double* __restrict__ out) {
  const int j = blockIdx.x;
  const double* p = P + (size_t)j * nrb;
  double s = 0.0;
  for (int r = threadIdx.x; r < nrb; r += 256) s += p[r];
  __shared__ double part[256];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double d0 = g0 ? D[(size_t)j * ldd] : 0.0;
    out[j] = 2.0 * part[0] + d0 * d0;
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
size_t planes_bytes(int K, int n, int np) {
  const size_t nkb = (K + kPK - 1) / kPK, ncb = (n + kPT - 1) / kPT;
  return nkb * ncb * (size_t)np * kPart;
}

void split_planes(mc_ctx* ctx, const double* X, int ldx, int K, int n, int np, DevBuf& out) {
  MC_CHECK(ldx % 2 == 0 && (reinterpret_cast<uintptr_t>(X) & 15) == 0, MC_ERR_INTERNAL,
           "split_planes: operand not 16-byte aligned");
  const int nkb = (K + kPK - 1) / kPK;
  const size_t nblk = (size_t)nkb * ((n + kPT - 1) / kPT);
  out.ensure(planes_bytes(K, n, np));
  const dim3 grid((unsigned)(nblk * 2));   // 512 threads per block
  timer_begin(ctx, MC_KID_GEMM);
  if (np == 3)
    hipLaunchKernelGGL(k_split_planes<3>, grid, dim3(256), 0, ctx->stream, X, ldx, K, n, nkb, nblk, out.as<__bf16>());
  else
    hipLaunchKernelGGL(k_split_planes<2>, grid, dim3(256), 0, ctx->stream, X, ldx, K, n, nkb, nblk, out.as<__bf16>());
  MC_HIP(hipGetLastError());
  timer_end(ctx, MC_KID_GEMM);
}

static void update_planes(mc_ctx* ctx, bool kick, const double* c0, const double* c2, const double* cm, double* cp,
                          int ldx, int K, int n, double dtm, double dt) {
  auto& C = ctx->planes;
  MC_CHECK(C.scoped, MC_ERR_INTERNAL, "posupa_planes outside a plane scope");
  MC_CHECK(ldx % kPK == 0 && K % 2 == 0 && K <= ldx && (reinterpret_cast<uintptr_t>(c0) & 15) == 0 &&
               (reinterpret_cast<uintptr_t>(cm) & 15) == 0 && (reinterpret_cast<uintptr_t>(cp) & 15) == 0 &&
               (!c2 || (reinterpret_cast<uintptr_t>(c2) & 15) == 0),
           MC_ERR_INTERNAL, "posupa_planes: operands not 16-byte aligned");
  const int np = ctx->x3_parts == 2 ? 2 : 3;
  const int nkb = (K + kPK - 1) / kPK, ncolp = (n + kPT - 1) / kPT * kPT, nseg = (ldx + 511) / 512;
  // the planes of c0 kept from an earlier call (this call's only epoch count
  // since; c0 is neither cp nor cm) stay in slot 0 and are not rewritten
  const bool have0 = kept_planes(ctx, c0, K, n, np, 1) && (const void*)c0 != cp && (const void*)c0 != cm;
  if (!have0) {
    C.keep_src = nullptr;
    C.buf[0].ensure(planes_bytes(K, n, np));
  }
  C.buf[1].ensure(planes_bytes(K, n, np));
  const dim3 grid((unsigned)(((size_t)ncolp * nseg + 3) / 4));
  timer_begin(ctx, MC_KID_MISC);
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, grid, dim3(256), 0, ctx->stream, c0, c2, cm, cp, ldx, K, n, ncolp, nseg, nkb,
                       c2 ? dtm : 0.0, dt, C.buf[0].as<__bf16>(), C.buf[1].as<__bf16>());
  };
  if (np == 3) {
    if (kick) have0 ? go(k_posupa_planes<3, true, false>) : go(k_posupa_planes<3, true, true>);
    else have0 ? go(k_posupa_planes<3, false, false>) : go(k_posupa_planes<3, false, true>);
  } else {
    if (kick) have0 ? go(k_posupa_planes<2, true, false>) : go(k_posupa_planes<2, true, true>);
    else have0 ? go(k_posupa_planes<2, false, false>) : go(k_posupa_planes<2, false, true>);
  }
  MC_HIP(hipGetLastError());
  timer_end(ctx, MC_KID_MISC);
  const void* src[2] = {c0, cp};
  for (int i = 0; i < 2; ++i) {
    C.src[i] = src[i];
    C.np[i] = np;
    C.K[i] = K;
    C.n[i] = n;
  }
  C.next = 0;
}

void posupa_planes(mc_ctx* ctx, const double* c0, const double* c2, const double* cm, double* cp, int ldx, int K,
                   int n, double dtm, double dt) {
  update_planes(ctx, false, c0, c2, cm, cp, ldx, K, n, dtm, dt);
}

void kick_planes(mc_ctx* ctx, const double* c0, const double* c2, double* cm, int ldx, int K, int n, double dtm) {
  update_planes(ctx, true, c0, c2, cm, cm, ldx, K, n, dtm, 0.0);
}

bool kept_planes(const mc_ctx* ctx, const void* X, int K, int n, int np, int lag) {
  const auto& C = ctx->planes;
  return ctx->col_reuse && X && C.keep_src == X && C.keep_np == np && C.keep_K == K && C.keep_n == n &&
         C.buf[0].p && C.buf[0].bytes >= planes_bytes(K, n, np) && C.keep_epoch + (unsigned long long)lag == wfn_epoch();
}

void keep_new_planes(mc_ctx* ctx, const void* X, int K, int n) {
  auto& C = ctx->planes;
  C.src[0] = nullptr;   // slot 0 no longer holds the scope's first operand
  C.keep_src = ctx->col_reuse ? X : nullptr;
  C.keep_np = ctx->x3_parts == 2 ? 2 : 3;
  C.keep_K = K;
  C.keep_n = n;
  C.keep_epoch = wfn_epoch();
}

void keep_c0_planes(mc_ctx* ctx, const void* c0, int K, int n) {
  auto& C = ctx->planes;
  const int np = ctx->x3_parts == 2 ? 2 : 3;
  if (!ctx->col_reuse || C.src[0] != c0 || C.np[0] != np || C.K[0] != K || C.n[0] != n) {
    C.keep_src = nullptr;
    return;
  }
  C.keep_src = c0;
  C.keep_np = np;
  C.keep_K = K;
  C.keep_n = n;
  C.keep_epoch = wfn_epoch();
}

void proj_planes(mc_ctx* ctx, const ProjA& pa, int K, int m, int np, DevBuf& out) {
  const int nkb = (K + kPK - 1) / kPK;
  const size_t nblk = (size_t)nkb * ((m + kPT - 1) / kPT);
  out.ensure(planes_bytes(K, m, np));
  const dim3 grid((unsigned)(nblk * 2));
  timer_begin(ctx, MC_KID_GEMM);
  if (np == 3)
    hipLaunchKernelGGL(k_proj_planes<3>, grid, dim3(256), 0, ctx->stream, pa, K, m, nkb, nblk, out.as<__bf16>());
  else
    hipLaunchKernelGGL(k_proj_planes<2>, grid, dim3(256), 0, ctx->stream, pa, K, m, nkb, nblk, out.as<__bf16>());
  MC_HIP(hipGetLastError());
  timer_end(ctx, MC_KID_GEMM);
}

void proj_planes_stacked(mc_ctx* ctx, const ProjA& base, const double* gk, int v0, int nv, int K, int na, int np,
                         DevBuf& out) {
  const int nkb = (K + kPK - 1) / kPK, ncbv = (na + kPT - 1) / kPT;
  const size_t nblk = (size_t)nkb * ncbv;
  out.ensure(nblk * nv * (size_t)np * kPart);
  const dim3 grid((unsigned)(nblk * 2));
  timer_begin(ctx, MC_KID_GEMM);
  if (np == 3)
    hipLaunchKernelGGL(k_proj_planes_stacked<3>, grid, dim3(256), 0, ctx->stream, base, gk, base.lmod4, v0, nv, K, na,
                       nkb, ncbv, nblk, out.as<__bf16>());
  else
    hipLaunchKernelGGL(k_proj_planes_stacked<2>, grid, dim3(256), 0, ctx->stream, base, gk, base.lmod4, v0, nv, K, na,
                       nkb, ncbv, nblk, out.as<__bf16>());
  MC_HIP(hipGetLastError());
  timer_end(ctx, MC_KID_GEMM);
}

void gemm_nn_planes(mc_ctx* ctx, const double* A, int lda, const double* G, int ldgm, int M, int k, int n,
                    double alpha, double beta, double* D, int ldd, bool upper_g, const FnlEpi& epi_in) {
  FnlEpi epi = epi_in;
  MC_CHECK(!epi.colsq || (!epi.twnl && kPT % kColsqRows == 0), MC_ERR_INTERNAL,
           "column sums with the fnonloc epilogue");
  if (epi.pl) {
    MC_CHECK(!epi.twnl, MC_ERR_INTERNAL, "planes of D_new with the fnonloc epilogue");
    MC_CHECK(epi.pl_rows > 0 && epi.pl_rows <= M && epi.pl_rows % 2 == 0, MC_ERR_INTERNAL,
             "planes of D_new: bad row count");
    epi.pl_nkb = (epi.pl_rows + kPK - 1) / kPK;   // the plane layout of pl_rows rows (split_planes' K)
    epi.pl_np = ctx->x3_parts == 2 ? 2 : 3;
  }
  split_planes(ctx, G, ldgm, k, n, 2, ctx->w_plg);
  const int gnkb = (k + kPK - 1) / kPK;
  const int tm = (M + kPT - 1) / kPT, tn = (n + kPT - 1) / kPT;
  MC_CHECK(!epi.colsq || epi.sq_nrb >= tm * (kPT / kColsqRows), MC_ERR_INTERNAL, "column-sum blocks too few");
  const int total = tm * tn;
  const int launch = (total + kXcdP - 1) / kXcdP * kXcdP;
  timer_begin(ctx, MC_KID_GEMM);
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(launch), dim3(256), 0, ctx->stream, A, lda, ctx->w_plg.as<__bf16>(), gnkb, M, k, n,
                       alpha, beta, D, ldd, upper_g ? 1 : 0, epi, tn, total);
  };
  MC_CHECK(!(epi.pl && epi.colsq), MC_ERR_INTERNAL, "planes and column sums in one product");
  if (epi.pl)
    go(k_nn_planes<true, false>);
  else if (epi.colsq)
    go(k_nn_planes<false, true>);
  else
    go(k_nn_planes<false, false>);
  MC_HIP(hipGetLastError());
  timer_end(ctx, MC_KID_GEMM);
}

void colsq_finish(mc_ctx* ctx, const double* colsq, int nrb, int n, const double* D, int ldd, int g0, double* out) {
  if (n <= 0) return;
  timer_begin(ctx, MC_KID_MISC);
  hipLaunchKernelGGL(k_colsq_finish, dim3(n), dim3(256), 0, ctx->stream, colsq, nrb, D, ldd, g0, out);
  MC_HIP(hipGetLastError());
  timer_end(ctx, MC_KID_MISC);
}

int tn_planes_partials(mc_ctx* ctx, const DevBuf& Apl, const DevBuf& Bpl, int np, int K, int m, int n,
                       bool symmetric) {
  const int nkb = (K + kPK - 1) / kPK;
  const int tm = (m + kPT - 1) / kPT, tn = (n + kPT - 1) / kPT;
  if (symmetric && m != n) throw Error{MC_ERR_INTERNAL, "symmetric A^T B with m != n"};
  const int tiles = tile_count(symmetric, tm, tn);
  // one workgroup per CU (the ring takes 96-144 KiB of LDS): about two rounds
  // of 256 workgroups, each over a contiguous K range that it sums in FP32
  // chunks of at most kMaxSlices slices added in FP64 in registers, writing
  // one FP64 partial tile (the fewer partials, the less traffic here and in
  // the reduction)
  int nsplit = std::max(1, std::min(nkb, kTnWorkgroups / tiles));
  const int kbper = (nkb + nsplit - 1) / nsplit;
  nsplit = (nkb + kbper - 1) / kbper;
  ctx->w_gemm.ensure(sizeof(double) * (size_t)nsplit * m * n);
  const int total = nsplit * tiles;
  const int launch = (total + kXcdP - 1) / kXcdP * kXcdP;
  timer_begin(ctx, MC_KID_TN_PLANES);
  if (np == 3)
    hipLaunchKernelGGL(k_tn_planes<3>, dim3(launch), dim3(64 * kTnWaves), 0, ctx->stream, Apl.as<__bf16>(), Bpl.as<__bf16>(),
                       nkb, m, n, kbper, ctx->w_gemm.as<double>(), symmetric ? 1 : 0, tm, tn, total);
  else
    hipLaunchKernelGGL(k_tn_planes<2>, dim3(launch), dim3(64 * kTnWaves), 0, ctx->stream, Apl.as<__bf16>(), Bpl.as<__bf16>(),
                       nkb, m, n, kbper, ctx->w_gemm.as<double>(), symmetric ? 1 : 0, tm, tn, total);
  MC_HIP(hipGetLastError());
  timer_end(ctx, MC_KID_TN_PLANES);
  return nsplit;
}

// out = alpha A^T B [- A(0,:)^T B(0,:)] from planes of A (K x m) and B (K x n);
// A64 / B64 are the FP64 operands (row 0 for the G = 0 correction);
// symmetric: upper tiles only, mirrored
void gemm_tn_planes(mc_ctx* ctx, const DevBuf& Apl, const DevBuf& Bpl, int np, int K, int m, int n, double alpha,
                    const double* A64, int lda, const double* B64, int ldb, bool g0corr, bool symmetric, double* out,
                    int ldo) {
  if (m <= 0 || n <= 0) return;
  const int nsplit = tn_planes_partials(ctx, Apl, Bpl, np, K, m, n, symmetric);
  reduce_tn(ctx, nsplit, m, n, alpha, A64, lda, B64, ldb, g0corr, symmetric, out, ldo);
}

}  // namespace mc
