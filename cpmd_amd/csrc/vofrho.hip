// The density -> potential step on gfx950 (vofrhoa + vofrhob, closed shell,
// LDA) and the Ewald real-space sum, all FP64:
//
//   rho(r) --x,y,z forward passes--> W = N rho(G)        (dense mesh, fwfftn)
//   per G (thread per G, loop over atoms with per-atom phase tables):
//       eivps, eirop (eicalc), vtemp and the ppener energy sums
//   W <- 0; W(+-G) = vtemp, conj(vtemp)
//   per G x atom block: potfor partial forces, reduced in a fixed order
//   W --z,y,x inverse passes--> v(r) = Re W + V_xc(rho(r)), E_xc sums (fused
//   into the last x pass)
//
// References: vofrhoa src/vofrhoa_utils.mod.F90:41-144 (fwfftn :92, potfor
// :101, ppener :103-122), eicalc src/eicalc_utils.mod.F90:33-173, ppener
// src/ppener_utils.mod.F90:23-108, potfor src/potfor_utils.mod.F90:38-201,
// vofrhob src/vofrhob_utils.mod.F90:80-461 (v(+-G) :150-154, invfftn :155,
// xcener :220, exc/vxc scaling :459-461), LDA Pade (CPMD's FUNCTIONAL LDA)
// src/nfunc_utils.mod.F90:251-419, rpiiint src/rpiiint_utils.mod.F90:29-200,
// eself src/rinforce_utils.mod.F90:239-249.
//
// e^{-iG.R} is the product of three per-atom tables e^{-2 pi i m f_k} over the
// centred index m of each axis (the ei1/ei2/ei3 factorisation of eicalc and
// potfor); the tables are built on the device from the fractional positions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "fft_launch.h"
#include "mc_internal.h"

namespace mc {

namespace {

constexpr int kRedBlocks = 1024;     // workgroups of the per-G and x passes

struct DParams {
  int n1, n2, n3, kr1, kr2, kr3;
  FftPlan p1, p2, p3;
  const double2* tw1;
  const double2* tw2;
  const double2* tw3;
};

DParams make_dparams(mc_ctx* ctx) {
  DParams d;
  d.n1 = ctx->n1; d.n2 = ctx->n2; d.n3 = ctx->n3;
  d.kr1 = ctx->kr1; d.kr2 = ctx->kr2; d.kr3 = ctx->kr3;
  d.p1 = ctx->plan1; d.p2 = ctx->plan2; d.p3 = ctx->plan3;
  d.tw1 = ctx->d_tw1.as<double2>();
  d.tw2 = ctx->d_tw2.as<double2>();
  d.tw3 = ctx->d_tw3.as<double2>();
  return d;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// block sum of NV values per thread (kThreads threads); thread q < nvalid stores
// the sum of value q to dst[q]
template <int NV>
__device__ __forceinline__ void block_sum_store(const double (&v)[NV], double* dst, int nvalid) {
  __shared__ double part[NV][kThreads / 64];
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const double s = warp_sum(v[q]);
    if (ln == 0) part[q][wv] = s;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < nvalid; q += kThreads) {
    double s = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) s += part[q][w];
    dst[q] = s;
  }
  __syncthreads();
}

}  // namespace

// ---------------------------------------------------------------------------
// dense transforms of W (n1 x n2 x n3 complex, x fastest, unpadded)
// ---------------------------------------------------------------------------

// forward x pass from the real density rho (kr1 x kr2 x kr3): ty lines per block
template <int N>
__global__ __launch_bounds__(kThreads) void k_dx_fwd_real(DParams d, const double* __restrict__ rho,
                                                          double2* __restrict__ W, int ty) {
  const int n1 = N > 0 ? N : d.n1;
  double2* tw = lds_base<double>();
  double2* buf = tw + n1;
  const int y0 = blockIdx.x * ty, z = blockIdx.y;
  const int ny = min(ty, d.n2 - y0);
  load_tw<double>(tw, d.tw1, n1);
  for (int e = threadIdx.x; e < ny * n1; e += kThreads) {
    const int y = e / n1, x = e - y * n1;
    buf[e] = make_double2(rho[(size_t)x + (size_t)d.kr1 * ((size_t)(y0 + y) + (size_t)d.kr2 * z)], 0.0);
  }
  __syncthreads();
  fft_any<double, N, -1>(buf, ny, n1, d.p1, tw);
  for (int e = threadIdx.x; e < ny * n1; e += kThreads) {
    const int y = e / n1, x = e - y * n1;
    W[(size_t)x + (size_t)n1 * ((size_t)(y0 + y) + (size_t)d.n2 * z)] = buf[e];
  }
}

// y (AXIS 2) or z (AXIS 3) pass in place: tl consecutive x lines per block
template <int N, int DIR, int AXIS>
__global__ __launch_bounds__(kThreads) void k_dyz(DParams d, double2* __restrict__ W, int tl) {
  const int n = N > 0 ? N : (AXIS == 2 ? d.n2 : d.n3);
  const FftPlan& pl = AXIS == 2 ? d.p2 : d.p3;
  double2* tw = lds_base<double>();
  double2* buf = tw + n;
  const int ld = n + 1;   // line stride in LDS (odd: no bank aliasing between lines)
  const int x0 = blockIdx.x * tl;
  const int nl = min(tl, d.n1 - x0);
  const int other = blockIdx.y;   // z for the y pass, y for the z pass
  const size_t stride = AXIS == 2 ? (size_t)d.n1 : (size_t)d.n1 * d.n2;
  const size_t base = AXIS == 2 ? (size_t)x0 + (size_t)d.n1 * d.n2 * other : (size_t)x0 + (size_t)d.n1 * other;
  load_tw<double>(tw, AXIS == 2 ? d.tw2 : d.tw3, n);
  for (int e = threadIdx.x; e < nl * n; e += kThreads) {
    const int p = e / nl, j = e - p * nl;
    buf[j * ld + p] = W[base + j + stride * p];
  }
  __syncthreads();
  fft_any<double, N, DIR>(buf, nl, ld, pl, tw);
  for (int e = threadIdx.x; e < nl * n; e += kThreads) {
    const int p = e / nl, j = e - p * nl;
    W[base + j + stride * p] = buf[j * ld + p];
  }
}

// inverse x pass to real space fused with the LDA: v(r) = Re W(r) + V_xc(rho(r))
// on the valid mesh (padding of v set to 0); per-block partial sums of
// eps_xc rho and V_xc max(rho, 0).  Pade form of Goedecker, Teter, Hutter
// (PRB 54, 1703 (1996)) with rho floored at 1e-24.
__device__ __forceinline__ void lda_pade(double rho, double& eps, double& vxc) {
  const double a0 = 0.4581652932831429, a1 = 2.217058676663745, a2 = 0.7405551735357053,
               a3 = 0.01968227878617998;
  const double b1 = 1.0, b2 = 4.504130959426697, b3 = 1.110667363742916, b4 = 0.02359291751427506;
  const double rs = cbrt(0.75 / (M_PI * rho));
  const double num = a0 + rs * (a1 + rs * (a2 + rs * a3));
  const double den = rs * (b1 + rs * (b2 + rs * (b3 + rs * b4)));
  const double dnum = a1 + rs * (2.0 * a2 + rs * 3.0 * a3);
  const double dden = b1 + rs * (2.0 * b2 + rs * (3.0 * b3 + rs * 4.0 * b4));
  eps = -num / den;
  const double de = -(dnum * den - num * dden) / (den * den);
  vxc = eps - rs / 3.0 * de;
}

template <int N>
__global__ __launch_bounds__(kThreads) void k_dx_inv_xc(DParams d, const double2* __restrict__ W,
                                                        const double* __restrict__ rho, double* __restrict__ v,
                                                        int ty, double* __restrict__ part) {
  const int n1 = N > 0 ? N : d.n1;
  double2* tw = lds_base<double>();
  double2* buf = tw + n1;
  const int y0 = blockIdx.x * ty, z = blockIdx.y;
  const int ny = min(ty, d.n2 - y0);
  load_tw<double>(tw, d.tw1, n1);
  for (int e = threadIdx.x; e < ny * n1; e += kThreads) {
    const int y = e / n1, x = e - y * n1;
    buf[e] = W[(size_t)x + (size_t)n1 * ((size_t)(y0 + y) + (size_t)d.n2 * z)];
  }
  __syncthreads();
  fft_any<double, N, 1>(buf, ny, n1, d.p1, tw);
  double acc[2] = {0.0, 0.0};
  for (int e = threadIdx.x; e < ny * d.kr1; e += kThreads) {
    const int y = e / d.kr1, x = e - y * d.kr1;
    const size_t idx = (size_t)x + (size_t)d.kr1 * ((size_t)(y0 + y) + (size_t)d.kr2 * z);
    if (x < n1) {
      const double r0 = rho[idx];
      const double r = r0 > 1e-24 ? r0 : 1e-24;
      double ex, vx;
      lda_pade(r, ex, vx);
      v[idx] = buf[y * n1 + x].x + vx;
      acc[0] += ex * r;
      acc[1] += vx * (r0 > 0.0 ? r0 : 0.0);
    } else {
      v[idx] = 0.0;
    }
  }
  block_sum_store<2>(acc, part + 2 * ((size_t)blockIdx.y * gridDim.x + blockIdx.x), 2);
}

// zero the padding planes y >= n2 and z >= n3 of v (the x passes write y < n2, z < n3)
__global__ void k_pad_zero(double* __restrict__ v, int n2, int n3, int kr1, int kr2, int kr3) {
  const size_t nn = (size_t)kr1 * kr2 * kr3;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nn; e += (size_t)gridDim.x * blockDim.x) {
    const size_t yz = e / kr1;
    const int y = (int)(yz % kr2), z = (int)(yz / kr2);
    if (y >= n2 || z >= n3) v[e] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// per-G work
// ---------------------------------------------------------------------------

// phase tables: ph_k[isa][m + n_k/2] = e^{-2 pi i m f_k(isa)}, m in [-n_k/2, n_k/2]
__global__ void k_phase_tables(const double* __restrict__ frac, int nat, int n1, int n2, int n3,
                               double2* __restrict__ ph1, double2* __restrict__ ph2, double2* __restrict__ ph3) {
  const int isa = blockIdx.y;
  const int k = blockIdx.z;
  const int n = k == 0 ? n1 : (k == 1 ? n2 : n3);
  double2* ph = k == 0 ? ph1 : (k == 1 ? ph2 : ph3);
  const double f = frac[3 * isa + k];
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t <= n; t += gridDim.x * blockDim.x) {
    const int m = t - n / 2;
    // reduce m f to [-1/2, 1/2] before the sin/cos of 2 pi x
    double x = (double)m * f;
    x -= nearbyint(x);
    double s, c;
    sincospi(2.0 * x, &s, &c);
    ph[(size_t)isa * (n + 1) + t] = make_double2(c, -s);
  }
}

struct GParams {
  int nhg, geq0, nsp, nat;
  int h1, h2, h3;            // n_k / 2
  const int* gm;             // (3, nhg) centred indices m
  const int* gidx;           // (2, nhg) linear W index of +G, -G
  const double* scg;         // (nhg)
  const double* vps;         // (nsp, nhg)
  const double* rhops;       // (nsp, nhg)
  const int* sp_start;       // (nsp + 1) first atom of each species
  const double2* ph1;
  const double2* ph2;
  const double2* ph3;
  const int* sidx;           // (nhg) stick of each G
  int nstick, L3;            // sticks; L3 = n3 + 1 positions m3 + h3 per stick
  const double* sf;          // (nsp) blocks of (nstick x 2 L3): Re | Im of the structure factor
};

// eicalc + ppener: rho_e(G) = W(+G)/N, eivps, eirop, vtemp; block partials of
// (eh, ei, ee, eps) over G != 0; the G = 0 terms are added on the host.  The
// species structure factors sum_{a in sp} e^{-iG.R_a} come from the stick GEMM.
__global__ __launch_bounds__(kThreads) void k_eicalc_ppener(GParams g, int n1, int n2, int n3,
                                                            const double2* __restrict__ W, double scale,
                                                            double2* __restrict__ rhog, double2* __restrict__ eirop,
                                                            double2* __restrict__ eivps, double2* __restrict__ vtemp,
                                                            double* __restrict__ part) {
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int ig = blockIdx.x * blockDim.x + threadIdx.x; ig < g.nhg; ig += gridDim.x * blockDim.x) {
    const double2 w = W[g.gidx[2 * ig]];
    const double2 re = make_double2(w.x * scale, w.y * scale);
    double2 ev = make_double2(0.0, 0.0), er = make_double2(0.0, 0.0);
    const size_t ps = (size_t)(g.gm[3 * ig + 2] + g.h3) * g.nstick + g.sidx[ig];
    for (int is = 0; is < g.nsp; ++is) {
      const double* S = g.sf + (size_t)is * g.nstick * 2 * g.L3;
      const double2 sf = make_double2(S[ps], S[ps + (size_t)g.L3 * g.nstick]);
      const double vp = g.vps[is + (size_t)g.nsp * ig], rp = g.rhops[is + (size_t)g.nsp * ig];
      ev.x += vp * sf.x; ev.y += vp * sf.y;
      er.x += rp * sf.x; er.y += rp * sf.y;
    }
    rhog[ig] = re;
    eirop[ig] = er;
    eivps[ig] = ev;
    const double2 rt = make_double2(re.x + er.x, re.y + er.y);
    const double s = g.scg[ig];
    if (g.geq0 && ig == 0) {
      vtemp[ig] = make_double2(s * rt.x, s * rt.y);
    } else {
      vtemp[ig] = make_double2(s * rt.x + ev.x, s * rt.y + ev.y);
      acc[0] += s * (rt.x * rt.x + rt.y * rt.y);
      acc[1] += s * (er.x * er.x + er.y * er.y);
      acc[2] += s * (re.x * re.x + re.y * re.y);
      acc[3] += re.x * ev.x + re.y * ev.y;          // Re(conj(rho_e) eivps)
    }
  }
  block_sum_store<4>(acc, part + 4 * (size_t)blockIdx.x, 4);
}

// W(+G) = vtemp, W(-G) = conj(vtemp) (G = 0: vtemp); W zeroed before
__global__ void k_scatter_vtemp(GParams g, const double2* __restrict__ vtemp, double2* __restrict__ W) {
  for (int ig = blockIdx.x * blockDim.x + threadIdx.x; ig < g.nhg; ig += gridDim.x * blockDim.x) {
    const double2 v = vtemp[ig];
    if (g.geq0 && ig == 0) {
      W[g.gidx[0]] = v;
    } else {
      W[g.gidx[2 * ig]] = v;
      W[g.gidx[2 * ig + 1]] = make_double2(v.x, -v.y);
    }
  }
}

// ---------------------------------------------------------------------------
// Structure factors and local-potential ion forces as FP64 GEMMs over the
// z-sticks (m1, m2) of the density set.  With e^{-iG.R_a} = p1_a(m1) p2_a(m2)
// p3_a(m3) (phase tables) and Q(s, a) = p1_a(m1_s) p2_a(m2_s):
//   S_sp(s, m3) = sum_{a in sp} Q(s, a) p3_a(m3)                        (eicalc)
//   T_k(s, a)   = sum_m3 p3_a(m3) X_sp(G) G_k,  G = (m1_s, m2_s, m3)     (potfor)
//   F_k(a)      = -sum_s Im(Q(s, a) T_k(s, a))
// Complex products run as real GEMMs: one operand holds (re, im) pairs along
// the contraction index, the other the real 2 x 2 expansion of p3.
// ---------------------------------------------------------------------------

// Q (nstick x 2 nat, ld nstick): columns 2a, 2a + 1 = Re, Im of Q(s, a)
__global__ void k_stick_q(const int* __restrict__ stk, int nstick, int n1, int n2, const double2* __restrict__ ph1,
                          const double2* __restrict__ ph2, double* __restrict__ Q) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x, a = blockIdx.y;
  if (s >= nstick) return;
  const double2 p = cmul(ph1[(size_t)a * (n1 + 1) + stk[2 * s] + n1 / 2], ph2[(size_t)a * (n2 + 1) + stk[2 * s + 1] + n2 / 2]);
  Q[(size_t)(2 * a) * nstick + s] = p.x;
  Q[(size_t)(2 * a + 1) * nstick + s] = p.y;
}

// Bs (2 nat x 2 L3, ld 2 nat), S = Q Bs: rows 2a, 2a + 1 of column t (Re S)
// hold (Re p3, -Im p3), of column L3 + t (Im S) (Im p3, Re p3).
// Bf (2 L3 x 2 nat, ld 2 L3), T = Y Bf: per species a block of columns
// [Re | Im] (atom a of a species with first atom a0 and na atoms: Re column
// 2 a0 + (a - a0), Im column 2 a0 + na + (a - a0)); rows 2t, 2t + 1 hold
// (Re p3, -Im p3) in the Re column and (Im p3, Re p3) in the Im column.
__global__ void k_stick_p3(const double2* __restrict__ ph3, int nat, int n3, const int* __restrict__ sp_of,
                           const int* __restrict__ sp_start, double* __restrict__ Bs, double* __restrict__ Bf) {
  const int L3 = n3 + 1;
  const int t = blockIdx.x * blockDim.x + threadIdx.x, a = blockIdx.y;
  if (t >= L3) return;
  const double2 p = ph3[(size_t)a * L3 + t];
  const size_t lds = 2 * (size_t)nat;
  Bs[(size_t)t * lds + 2 * a] = p.x;
  Bs[(size_t)t * lds + 2 * a + 1] = -p.y;
  Bs[(size_t)(L3 + t) * lds + 2 * a] = p.y;
  Bs[(size_t)(L3 + t) * lds + 2 * a + 1] = p.x;
  const int sp = sp_of[a], a0 = sp_start[sp], na = sp_start[sp + 1] - a0;
  const size_t ldf = 2 * (size_t)L3;
  double* re = Bf + (size_t)(2 * a0 + (a - a0)) * ldf;
  double* im = Bf + (size_t)(2 * a0 + na + (a - a0)) * ldf;
  re[2 * t] = p.x;
  re[2 * t + 1] = -p.y;
  im[2 * t] = p.y;
  im[2 * t + 1] = p.x;
}

// Y (nsp blocks of 3 nstick x 2 L3, ld 3 nstick, zeroed before): row
// k nstick + s, columns 2t, 2t + 1 = Re, Im of X_sp(G) G_k for G != 0 on stick
// s at t = m3 + h3, X_sp = rhops_sp scg conj(rho_tot) + vps_sp conj(rho_e)
__global__ void k_stick_y(GParams g, const double* __restrict__ gk, const double2* __restrict__ rhog,
                          const double2* __restrict__ eirop, double* __restrict__ Y) {
  const size_t ldy = 3 * (size_t)g.nstick, blk = ldy * 2 * g.L3;
  for (int ig = blockIdx.x * blockDim.x + threadIdx.x; ig < g.nhg; ig += gridDim.x * blockDim.x) {
    if (g.geq0 && ig == 0) continue;
    const int s = g.sidx[ig], t = g.gm[3 * ig + 2] + g.h3;
    const double2 re = rhog[ig], er = eirop[ig];
    const double sc = g.scg[ig];
    const double2 A = make_double2(sc * (re.x + er.x), -sc * (re.y + er.y));   // scg conj(rho_tot)
    const double2 B = make_double2(re.x, -re.y);                              // conj(rho_e)
    const double gv[3] = {gk[3 * ig], gk[3 * ig + 1], gk[3 * ig + 2]};
    for (int is = 0; is < g.nsp; ++is) {
      const double rp = g.rhops[is + (size_t)g.nsp * ig], vp = g.vps[is + (size_t)g.nsp * ig];
      const double2 X = make_double2(rp * A.x + vp * B.x, rp * A.y + vp * B.y);
      double* y = Y + is * blk + s;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        y[(size_t)(2 * t) * ldy + (size_t)k * g.nstick] = X.x * gv[k];
        y[(size_t)(2 * t + 1) * ldy + (size_t)k * g.nstick] = X.y * gv[k];
      }
    }
  }
}

// fout[3a + k] = -scale sum_s Im(Q(s, a) T_k(s, a)); T (3 nstick x 2 nat, ld
// 3 nstick) in the column order of Bf.  One workgroup per (a, k), fixed order.
__global__ __launch_bounds__(kThreads) void k_stick_force(const double* __restrict__ Q, const double* __restrict__ T,
                                                          int nstick, const int* __restrict__ sp_of,
                                                          const int* __restrict__ sp_start, double scale,
                                                          double* __restrict__ fout) {
  const int a = blockIdx.x, k = blockIdx.y;
  const int sp = sp_of[a], a0 = sp_start[sp], na = sp_start[sp + 1] - a0;
  const size_t ldt = 3 * (size_t)nstick;
  const double* qr = Q + (size_t)(2 * a) * nstick;
  const double* qi = qr + nstick;
  const double* tr = T + (size_t)(2 * a0 + (a - a0)) * ldt + (size_t)k * nstick;
  const double* ti = T + (size_t)(2 * a0 + na + (a - a0)) * ldt + (size_t)k * nstick;
  double acc[1] = {0.0};
  for (int s = threadIdx.x; s < nstick; s += kThreads) acc[0] += qr[s] * ti[s] + qi[s] * tr[s];
  acc[0] *= -scale;
  block_sum_store<1>(acc, fout + 3 * (size_t)a + k, 1);
}

// ---------------------------------------------------------------------------
// Ewald real-space sum (rpiiint): block per atom l, threads over partners m;
// e_l = sum_m sum_images w zv_l zv_m erfc(r / rckj) / r with w = 1/2 (each
// unordered pair is visited from both ends; the self images of l carry the
// reference's 1/2), F_l = sum_{m != l} sum_images (erfc term + Gaussian term) r / r^2
// ---------------------------------------------------------------------------
struct EParams {
  int nat, iesr;
  double a[9];      // a[3 i + q]: component q of lattice vector i
  double inv[9];    // inv[3 q + i]: fractional coordinate i of a Cartesian vector
  const double* tau;
  const int* sp_of;
  const double* zv;
  const double* raggio;
};

__global__ __launch_bounds__(kThreads) void k_ewald(EParams p, double* __restrict__ out) {
  const int l = blockIdx.x;
  const int kl = p.sp_of[l];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int m = threadIdx.x; m < p.nat; m += kThreads) {
    const int km = p.sp_of[m];
    const double zv2 = p.zv[kl] * p.zv[km];
    if (fabs(zv2) < 1e-10) continue;
    const double rckj = sqrt(p.raggio[kl] * p.raggio[kl] + p.raggio[km] * p.raggio[km]);
    const bool tzero = (m == l);
    double d[3] = {0.0, 0.0, 0.0};
    if (!tzero) {
      double dd[3], fr[3];
      for (int q = 0; q < 3; ++q) dd[q] = p.tau[3 * l + q] - p.tau[3 * m + q];
      for (int i = 0; i < 3; ++i) fr[i] = dd[0] * p.inv[i] + dd[1] * p.inv[3 + i] + dd[2] * p.inv[6 + i];
      for (int i = 0; i < 3; ++i) fr[i] -= nearbyint(fr[i]);
      for (int q = 0; q < 3; ++q) d[q] = fr[0] * p.a[q] + fr[1] * p.a[3 + q] + fr[2] * p.a[6 + q];
    }
    for (int ix = -p.iesr; ix <= p.iesr; ++ix)
      for (int iy = -p.iesr; iy <= p.iesr; ++iy)
        for (int iz = -p.iesr; iz <= p.iesr; ++iz) {
          if (tzero && ix == 0 && iy == 0 && iz == 0) continue;
          double r[3];
          for (int q = 0; q < 3; ++q) r[q] = d[q] + ix * p.a[q] + iy * p.a[3 + q] + iz * p.a[6 + q];
          const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
          const double rl = sqrt(r2);
          const double arg = rl / rckj;
          if (arg > 20.0) continue;
          const double add = zv2 * erfc(arg) / rl;
          acc[0] += 0.5 * add;
          if (!tzero) {
            const double rep = (add + 2.0 * zv2 / sqrt(M_PI) * exp(-arg * arg) / rckj) / r2;
            acc[1] += rep * r[0];
            acc[2] += rep * r[1];
            acc[3] += rep * r[2];
          }
        }
  }
  block_sum_store<4>(acc, out + 4 * (size_t)l, 4);
}

// projector structure factors on the wavefunction rows (device order):
// eigr(j, isa) = e^{-iG_j.R_isa} from the phase tables, 0 on padding rows
__global__ void k_eigr(const int* __restrict__ gmw, int ngw, int ldg, int n1, int n2, int n3,
                       const double2* __restrict__ ph1, const double2* __restrict__ ph2,
                       const double2* __restrict__ ph3, double2* __restrict__ eigr) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, isa = blockIdx.y;
  if (j >= ldg) return;
  double2 e = make_double2(0.0, 0.0);
  if (j < ngw) {
    const double2 a = ph1[(size_t)isa * (n1 + 1) + gmw[3 * j] + n1 / 2];
    const double2 b = ph2[(size_t)isa * (n2 + 1) + gmw[3 * j + 1] + n2 / 2];
    const double2 c = ph3[(size_t)isa * (n3 + 1) + gmw[3 * j + 2] + n3 / 2];
    e = cmul(cmul(a, b), c);
  }
  eigr[(size_t)isa * ldg + j] = e;
}

// ---------------------------------------------------------------------------
// host drivers
// ---------------------------------------------------------------------------
void set_density_gset(mc_ctx* ctx, int nhg, const int* inyh, const double* hg, const double* gk) {
  auto& D = ctx->dens;
  MC_CHECK(nhg >= 1, MC_ERR_ARG, "set_density_gset: nhg < 1");
  const int nr[3] = {ctx->n1, ctx->n2, ctx->n3};
  std::vector<int> gm((size_t)3 * nhg), gidx((size_t)2 * nhg);
  std::vector<double> scg(nhg);
  for (int ig = 0; ig < nhg; ++ig) {
    int ip[3], im[3];
    for (int k = 0; k < 3; ++k) {
      const int m = inyh[3 * (size_t)ig + k] - (nr[k] / 2 + 1);
      MC_CHECK(m >= -nr[k] / 2 && m <= nr[k] / 2, MC_ERR_ARG, "set_density_gset: G index outside the mesh");
      gm[3 * (size_t)ig + k] = m;
      ip[k] = ((m % nr[k]) + nr[k]) % nr[k];
      im[k] = ((-m % nr[k]) + nr[k]) % nr[k];
    }
    gidx[2 * (size_t)ig] = ip[0] + nr[0] * (ip[1] + nr[1] * ip[2]);
    gidx[2 * (size_t)ig + 1] = im[0] + nr[0] * (im[1] + nr[1] * im[2]);
    if (ig == 0 && ctx->geq0) {
      MC_CHECK(gm[0] == 0 && gm[1] == 0 && gm[2] == 0, MC_ERR_ARG, "set_density_gset: first G must be G = 0");
      scg[ig] = 0.0;
    } else {
      MC_CHECK(hg[ig] > 0.0, MC_ERR_ARG, "set_density_gset: G = 0 after the first position");
      scg[ig] = 4.0 * M_PI / (ctx->tpiba2 * hg[ig]);
    }
  }
  // z-sticks: the distinct (m1, m2) in increasing order and the stick of each G
  const long long w2 = nr[1] + 1;
  std::vector<long long> key(nhg);
  for (int ig = 0; ig < nhg; ++ig)
    key[ig] = (long long)(gm[3 * (size_t)ig] + nr[0] / 2) * w2 + (gm[3 * (size_t)ig + 1] + nr[1] / 2);
  std::vector<long long> uk(key);
  std::sort(uk.begin(), uk.end());
  uk.erase(std::unique(uk.begin(), uk.end()), uk.end());
  const int nstick = (int)uk.size();
  std::vector<int> sidx(nhg), stk(2 * (size_t)nstick);
  for (int ig = 0; ig < nhg; ++ig) sidx[ig] = (int)(std::lower_bound(uk.begin(), uk.end(), key[ig]) - uk.begin());
  for (int st = 0; st < nstick; ++st) {
    stk[2 * (size_t)st] = (int)(uk[st] / w2) - nr[0] / 2;
    stk[2 * (size_t)st + 1] = (int)(uk[st] % w2) - nr[1] / 2;
  }
  D.nstick = nstick;
  D.d_sidx.ensure(sizeof(int) * sidx.size());
  D.d_stk.ensure(sizeof(int) * stk.size());
  MC_HIP(hipMemcpy(D.d_sidx.p, sidx.data(), sizeof(int) * sidx.size(), hipMemcpyHostToDevice));
  MC_HIP(hipMemcpy(D.d_stk.p, stk.data(), sizeof(int) * stk.size(), hipMemcpyHostToDevice));
  D.nhg = nhg;
  D.d_gm.ensure(sizeof(int) * gm.size());
  D.d_gidx.ensure(sizeof(int) * gidx.size());
  D.d_scg.ensure(sizeof(double) * nhg);
  D.d_gk.ensure(sizeof(double) * 3 * (size_t)nhg);
  MC_HIP(hipMemcpy(D.d_gm.p, gm.data(), sizeof(int) * gm.size(), hipMemcpyHostToDevice));
  MC_HIP(hipMemcpy(D.d_gidx.p, gidx.data(), sizeof(int) * gidx.size(), hipMemcpyHostToDevice));
  MC_HIP(hipMemcpy(D.d_scg.p, scg.data(), sizeof(double) * nhg, hipMemcpyHostToDevice));
  MC_HIP(hipMemcpy(D.d_gk.p, gk, sizeof(double) * 3 * (size_t)nhg, hipMemcpyHostToDevice));
  D.ready = true;
  D.pp_ready = false;
}

void set_local_pp(mc_ctx* ctx, int nsp, const int* na, const double* zv, const double* raggio, const double* vps,
                  const double* rhops) {
  auto& D = ctx->dens;
  MC_CHECK(D.ready, MC_ERR_ARG, "set_local_pp: density G set not set (mc_set_density_gset)");
  MC_CHECK(nsp >= 1, MC_ERR_ARG, "set_local_pp: nsp < 1");
  D.nsp = nsp;
  D.na.assign(na, na + nsp);
  D.zv.assign(zv, zv + nsp);
  D.raggio.assign(raggio, raggio + nsp);
  D.nat = 0;
  std::vector<int> start(nsp + 1, 0);
  for (int is = 0; is < nsp; ++is) {
    MC_CHECK(na[is] >= 0 && raggio[is] > 0.0, MC_ERR_ARG, "set_local_pp: bad na / raggio");
    start[is + 1] = start[is] + na[is];
  }
  D.nat = start[nsp];
  std::vector<int> sp_of(std::max(D.nat, 1), 0);
  for (int is = 0; is < nsp; ++is)
    for (int isa = start[is]; isa < start[is + 1]; ++isa) sp_of[isa] = is;
  D.d_sp_start.ensure(sizeof(int) * (nsp + 1));
  D.d_sp_of.ensure(sizeof(int) * sp_of.size());
  D.d_zv.ensure(sizeof(double) * nsp);
  D.d_raggio.ensure(sizeof(double) * nsp);
  D.d_vps.ensure(sizeof(double) * (size_t)nsp * D.nhg);
  D.d_rhops.ensure(sizeof(double) * (size_t)nsp * D.nhg);
  MC_HIP(hipMemcpy(D.d_sp_start.p, start.data(), sizeof(int) * (nsp + 1), hipMemcpyHostToDevice));
  MC_HIP(hipMemcpy(D.d_sp_of.p, sp_of.data(), sizeof(int) * sp_of.size(), hipMemcpyHostToDevice));
  MC_HIP(hipMemcpy(D.d_zv.p, zv, sizeof(double) * nsp, hipMemcpyHostToDevice));
  MC_HIP(hipMemcpy(D.d_raggio.p, raggio, sizeof(double) * nsp, hipMemcpyHostToDevice));
  MC_HIP(hipMemcpy(D.d_vps.p, vps, sizeof(double) * (size_t)nsp * D.nhg, hipMemcpyHostToDevice));
  MC_HIP(hipMemcpy(D.d_rhops.p, rhops, sizeof(double) * (size_t)nsp * D.nhg, hipMemcpyHostToDevice));
  D.pp_ready = true;
  D.ions_ready = false;
}

void set_ions(mc_ctx* ctx, const double* tau, const double* ht) {
  auto& D = ctx->dens;
  MC_CHECK(D.pp_ready, MC_ERR_ARG, "set_ions: local pseudopotential not set (mc_set_local_pp)");
  const int nat = D.nat;
  double a[9], inv[9];
  for (int i = 0; i < 3; ++i)
    for (int q = 0; q < 3; ++q) a[3 * i + q] = ht[i + 3 * q];   // ht(i, q): component q of vector i
  const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                     a[2] * (a[3] * a[7] - a[4] * a[6]);
  MC_CHECK(std::fabs(det) > 0.0, MC_ERR_ARG, "set_ions: singular cell");
  // adjugate / det: A^-1 of A[i][q] = a[3i+q]
  inv[0] = (a[4] * a[8] - a[5] * a[7]) / det;
  inv[1] = (a[2] * a[7] - a[1] * a[8]) / det;
  inv[2] = (a[1] * a[5] - a[2] * a[4]) / det;
  inv[3] = (a[5] * a[6] - a[3] * a[8]) / det;
  inv[4] = (a[0] * a[8] - a[2] * a[6]) / det;
  inv[5] = (a[2] * a[3] - a[0] * a[5]) / det;
  inv[6] = (a[3] * a[7] - a[4] * a[6]) / det;
  inv[7] = (a[1] * a[6] - a[0] * a[7]) / det;
  inv[8] = (a[0] * a[4] - a[1] * a[3]) / det;
  // inv holds A^-1 row-major: inv[3q+i] = (A^-1)[q][i], and d = A^T f gives f_i = sum_q d_q inv[3q+i]
  const double* t = inv;
  std::vector<double> frac(3 * (size_t)std::max(nat, 1), 0.0);
  for (int isa = 0; isa < nat; ++isa)
    for (int i = 0; i < 3; ++i) {
      double f = 0.0;
      for (int q = 0; q < 3; ++q) f += tau[3 * (size_t)isa + q] * t[3 * q + i];
      frac[3 * (size_t)isa + i] = f;
    }
  for (int k = 0; k < 9; ++k) {
    D.a[k] = a[k];
    D.inv[k] = t[k];
  }
  D.tau.assign(tau, tau + 3 * (size_t)nat);
  // tau and the fractional coordinates in one asynchronous upload: d_tau
  // holds [tau | frac], d_frac points at its second half
  std::vector<double> both(6 * (size_t)std::max(nat, 1), 0.0);
  std::copy(tau, tau + 3 * (size_t)nat, both.begin());
  std::copy(frac.begin(), frac.end(), both.begin() + 3 * (size_t)std::max(nat, 1));
  D.d_tau.ensure(sizeof(double) * both.size());
  upload_async(ctx, D.d_tau.p, both.data(), sizeof(double) * both.size());
  const double* d_frac = D.d_tau.as<double>() + 3 * (size_t)std::max(nat, 1);
  D.d_ph1.ensure(sizeof(double2) * (size_t)std::max(nat, 1) * (ctx->n1 + 1));
  D.d_ph2.ensure(sizeof(double2) * (size_t)std::max(nat, 1) * (ctx->n2 + 1));
  D.d_ph3.ensure(sizeof(double2) * (size_t)std::max(nat, 1) * (ctx->n3 + 1));
  if (nat > 0) {
    hipLaunchKernelGGL(k_phase_tables, dim3(1, nat, 3), dim3(256), 0, ctx->stream, d_frac, nat,
                       ctx->n1, ctx->n2, ctx->n3, D.d_ph1.as<double2>(), D.d_ph2.as<double2>(),
                       D.d_ph3.as<double2>());
    MC_HIP(hipGetLastError());
  }
  D.ions_ready = true;
  ++D.ions_serial;
}

static GParams make_gparams(mc_ctx* ctx) {
  auto& D = ctx->dens;
  GParams g;
  g.nhg = D.nhg;
  g.geq0 = ctx->geq0;
  g.nsp = D.nsp;
  g.nat = D.nat;
  g.h1 = ctx->n1 / 2;
  g.h2 = ctx->n2 / 2;
  g.h3 = ctx->n3 / 2;
  g.gm = D.d_gm.as<int>();
  g.gidx = D.d_gidx.as<int>();
  g.scg = D.d_scg.as<double>();
  g.vps = D.d_vps.as<double>();
  g.rhops = D.d_rhops.as<double>();
  g.sp_start = D.d_sp_start.as<int>();
  g.ph1 = D.d_ph1.as<double2>();
  g.ph2 = D.d_ph2.as<double2>();
  g.ph3 = D.d_ph3.as<double2>();
  g.sidx = D.d_sidx.as<int>();
  g.nstick = D.nstick;
  g.L3 = ctx->n3 + 1;
  g.sf = nullptr;
  return g;
}

void vofrho_dev(mc_ctx* ctx, const double* rho, double* v, double* energies, double* fion_host) {
  auto& D = ctx->dens;
  MC_CHECK(D.ready && D.pp_ready && D.ions_ready, MC_ERR_ARG,
           "vofrho: density G set, local pseudopotential and ions must be set first");
  const DParams d = make_dparams(ctx);
  const GParams g = make_gparams(ctx);
  const size_t nw = (size_t)ctx->n1 * ctx->n2 * ctx->n3;
  const double nn = (double)nw;
  D.w_W.ensure(sizeof(double2) * nw);
  D.w_g.ensure(sizeof(double2) * 4 * (size_t)D.nhg);
  double2* W = D.w_W.as<double2>();
  double2* rhog = D.w_g.as<double2>();
  double2* eirop = rhog + D.nhg;
  double2* eivps = eirop + D.nhg;
  double2* vtemp = eivps + D.nhg;
  const int tx = std::min(lines_for<double>(ctx->plan1), 16);
  // y / z passes keep lines at stride n + 1 in LDS: fit table + lines in kLdsBudget
  auto fit = [](int l, int n) { return std::max(1, std::min(l, (kLdsBudget / 16 - n) / (n + 1))); };
  const int t2 = fit(lines_for<double>(ctx->plan2), ctx->n2), t3 = fit(lines_for<double>(ctx->plan3), ctx->n3);
  const size_t shx = sizeof(double2) * (size_t)ctx->n1 * (1 + tx);
  const size_t sh2 = sizeof(double2) * ((size_t)ctx->n2 + (size_t)(ctx->n2 + 1) * t2);
  const size_t sh3 = sizeof(double2) * ((size_t)ctx->n3 + (size_t)(ctx->n3 + 1) * t3);
  const int gx_blocks = (ctx->n2 + tx - 1) / tx;
  const int n_xblk = gx_blocks * ctx->n3;
  // partial-sum area: ppener (4 per block), xc (2 per x block), ion forces (3 per atom)
  const size_t npart = 4 * (size_t)kRedBlocks + 2 * (size_t)n_xblk + 3 * (size_t)std::max(D.nat, 1);
  D.w_part.ensure(sizeof(double) * npart);
  double* p_pp = D.w_part.as<double>();
  double* p_xc = p_pp + 4 * kRedBlocks;
  double* f_out = p_xc + 2 * (size_t)n_xblk;
  // stick operands: Q, the p3 expansions, the species structure factors
  const int ns = D.nstick, L3 = ctx->n3 + 1, nat = D.nat;
  const size_t sfblk = (size_t)ns * 2 * L3;
  D.w_q.ensure(sizeof(double) * (size_t)ns * 2 * std::max(nat, 1));
  D.w_bs.ensure(sizeof(double) * 4 * (size_t)std::max(nat, 1) * L3);
  D.w_bf.ensure(sizeof(double) * 4 * (size_t)std::max(nat, 1) * L3);
  D.w_sf.ensure(sizeof(double) * sfblk * D.nsp);
  double* Q = D.w_q.as<double>();
  double* Sf = D.w_sf.as<double>();
  if (nat > 0) {
    LAUNCH(MC_KID_VOFRHO, k_stick_q, dim3((ns + 255) / 256, nat), dim3(256), 0, D.d_stk.as<int>(), ns, ctx->n1,
           ctx->n2, g.ph1, g.ph2, Q);
    LAUNCH(MC_KID_VOFRHO, k_stick_p3, dim3((L3 + 255) / 256, nat), dim3(256), 0, g.ph3, nat, ctx->n3,
           D.d_sp_of.as<int>(), D.d_sp_start.as<int>(), D.w_bs.as<double>(), D.w_bf.as<double>());
  }
  {
    int a0 = 0;
    for (int is = 0; is < D.nsp; ++is) {
      const int na = D.na[is];
      if (na > 0)
        gemm_nn_f64(ctx, Q + (size_t)2 * a0 * ns, ns, D.w_bs.as<double>() + 2 * a0, 2 * nat, ns, 2 * na, 2 * L3,
                    1.0, 0.0, Sf + is * sfblk, ns, false, nullptr, MC_KID_VOFRHO);
      else
        LAUNCH(MC_KID_VOFRHO, k_fill, dim3(256), dim3(256), 0, Sf + is * sfblk, sfblk, 0.0);
      a0 += na;
    }
  }
  GParams gs = g;
  gs.sf = Sf;

  // forward transform of rho
  with_length(ctx->n1, [&](auto NN) {
    LAUNCH(MC_KID_VOFRHO, (k_dx_fwd_real<decltype(NN)::value>), dim3(gx_blocks, ctx->n3), dim3(kThreads), shx, d,
           rho, W, tx);
  });
  with_length(ctx->n2, [&](auto NN) {
    LAUNCH(MC_KID_VOFRHO, (k_dyz<decltype(NN)::value, -1, 2>), dim3((ctx->n1 + t2 - 1) / t2, ctx->n3),
           dim3(kThreads), sh2, d, W, t2);
  });
  with_length(ctx->n3, [&](auto NN) {
    LAUNCH(MC_KID_VOFRHO, (k_dyz<decltype(NN)::value, -1, 3>), dim3((ctx->n1 + t3 - 1) / t3, ctx->n2),
           dim3(kThreads), sh3, d, W, t3);
  });
  // per-G energies and potential
  LAUNCH(MC_KID_VOFRHO, k_eicalc_ppener, dim3(kRedBlocks), dim3(kThreads), 0, gs, ctx->n1, ctx->n2, ctx->n3, W,
         1.0 / nn, rhog, eirop, eivps, vtemp, p_pp);
  if (fion_host && nat > 0) {
    // potfor: Y = X G_k on the sticks, T = Y Bf per species, F = -Im sum_s Q T
    const size_t yblk = 3 * sfblk;
    D.w_y.ensure(sizeof(double) * yblk * D.nsp);
    D.w_t.ensure(sizeof(double) * 3 * (size_t)ns * 2 * nat);
    double* Y = D.w_y.as<double>();
    double* T = D.w_t.as<double>();
    LAUNCH(MC_KID_VOFRHO, k_fill, dim3(1024), dim3(256), 0, Y, yblk * D.nsp, 0.0);
    LAUNCH(MC_KID_VOFRHO, k_stick_y, dim3(kRedBlocks), dim3(256), 0, gs, D.d_gk.as<double>(), rhog, eirop, Y);
    int a0 = 0;
    for (int is = 0; is < D.nsp; ++is) {
      const int na = D.na[is];
      if (na > 0)
        gemm_nn_f64(ctx, Y + is * yblk, 3 * ns, D.w_bf.as<double>() + (size_t)2 * a0 * 2 * L3, 2 * L3, 3 * ns,
                    2 * L3, 2 * na, 1.0, 0.0, T + (size_t)2 * a0 * 3 * ns, 3 * ns, false, nullptr, MC_KID_VOFRHO);
      a0 += na;
    }
    LAUNCH(MC_KID_VOFRHO, k_stick_force, dim3(nat, 3), dim3(kThreads), 0, Q, T, ns, D.d_sp_of.as<int>(),
           D.d_sp_start.as<int>(), 2.0 * ctx->omega * std::sqrt(ctx->tpiba2), f_out);
  }
  LAUNCH(MC_KID_VOFRHO, k_fill, dim3(1024), dim3(256), 0, (double*)W, 2 * nw, 0.0);
  LAUNCH(MC_KID_VOFRHO, k_scatter_vtemp, dim3(kRedBlocks), dim3(kThreads), 0, g, vtemp, W);
  with_length(ctx->n3, [&](auto NN) {
    LAUNCH(MC_KID_VOFRHO, (k_dyz<decltype(NN)::value, 1, 3>), dim3((ctx->n1 + t3 - 1) / t3, ctx->n2),
           dim3(kThreads), sh3, d, W, t3);
  });
  with_length(ctx->n2, [&](auto NN) {
    LAUNCH(MC_KID_VOFRHO, (k_dyz<decltype(NN)::value, 1, 2>), dim3((ctx->n1 + t2 - 1) / t2, ctx->n3),
           dim3(kThreads), sh2, d, W, t2);
  });
  with_length(ctx->n1, [&](auto NN) {
    LAUNCH(MC_KID_VOFRHO, (k_dx_inv_xc<decltype(NN)::value>), dim3(gx_blocks, ctx->n3), dim3(kThreads), shx, d,
           W, rho, v, tx, p_xc);
  });
  if (ctx->kr2 > ctx->n2 || ctx->kr3 > ctx->n3)
    LAUNCH(MC_KID_VOFRHO, k_pad_zero, dim3(1024), dim3(256), 0, v, ctx->n2, ctx->n3, ctx->kr1, ctx->kr2, ctx->kr3);
  // host: fixed-order sums of the partials and the G = 0 terms
  std::vector<double> hp(4 * (size_t)kRedBlocks + 2 * (size_t)n_xblk);
  dl_add(ctx, hp.data(), p_pp, sizeof(double) * hp.size());
  std::vector<double> hf(3 * (size_t)D.nat);
  if (fion_host && D.nat > 0)
    dl_add(ctx, hf.data(), f_out, sizeof(double) * hf.size());
  double2 g0[3] = {};
  if (ctx->geq0) {
    dl_add(ctx, &g0[0], rhog, sizeof(double2));
    dl_add(ctx, &g0[1], eirop, sizeof(double2));
    dl_add(ctx, &g0[2], eivps, sizeof(double2));
  }
  dl_finish(ctx);
  double s[4] = {0, 0, 0, 0};
  for (int b = 0; b < kRedBlocks; ++b)
    for (int q = 0; q < 4; ++q) s[q] += hp[4 * (size_t)b + q];
  double sxc = 0.0, vsum = 0.0;
  for (int b = 0; b < n_xblk; ++b) {
    sxc += hp[4 * (size_t)kRedBlocks + 2 * (size_t)b];
    vsum += hp[4 * (size_t)kRedBlocks + 2 * (size_t)b + 1];
  }
  double vploc = 0.0;
  if (ctx->geq0) {
    // scg(G=0) = 0: only the local pseudopotential term survives at G = 0
    vploc = g0[2].x;
    s[3] += 0.5 * (g0[2].x * g0[0].x + g0[2].y * g0[0].y);
  }
  if (energies) {
    energies[0] = s[0] * ctx->omega;         // ehep
    energies[1] = s[1] * ctx->omega;         // ei Omega
    energies[2] = s[2] * ctx->omega;         // ee Omega
    energies[3] = 2.0 * s[3] * ctx->omega;   // epseu
    energies[4] = vploc;
    energies[5] = sxc * ctx->omega / nn;     // exc
    energies[6] = vsum * ctx->omega / nn;    // vxc
  }
  if (fion_host)
    for (size_t e = 0; e < hf.size(); ++e) fion_host[e] += hf[e];
}

void ewald(mc_ctx* ctx, int iesr, double* esr, double* eself, double* fion_host) {
  auto& D = ctx->dens;
  MC_CHECK(D.pp_ready && D.ions_ready, MC_ERR_ARG, "ewald: local pseudopotential and ions must be set first");
  MC_CHECK(iesr >= 0 && iesr <= 8, MC_ERR_ARG, "ewald: iesr outside [0, 8]");
  double es = 0.0;
  for (int is = 0; is < D.nsp; ++is) es += D.na[is] * D.zv[is] * D.zv[is] / D.raggio[is];
  if (eself) *eself = es / std::sqrt(2.0 * M_PI);
  const int nat = D.nat;
  if (nat == 0) {
    if (esr) *esr = 0.0;
    return;
  }
  EParams p;
  p.nat = nat;
  p.iesr = iesr;
  for (int k = 0; k < 9; ++k) {
    p.a[k] = D.a[k];
    p.inv[k] = D.inv[k];
  }
  p.tau = D.d_tau.as<double>();
  p.sp_of = D.d_sp_of.as<int>();
  p.zv = D.d_zv.as<double>();
  p.raggio = D.d_raggio.as<double>();
  D.w_part.ensure(sizeof(double) * 4 * (size_t)nat);
  LAUNCH(MC_KID_VOFRHO, k_ewald, dim3(nat), dim3(kThreads), 0, p, D.w_part.as<double>());
  std::vector<double> h(4 * (size_t)nat);
  dl_add(ctx, h.data(), D.w_part.p, sizeof(double) * h.size());
  dl_finish(ctx);
  double e = 0.0;
  for (int l = 0; l < nat; ++l) {
    e += h[4 * (size_t)l];
    if (fion_host)
      for (int q = 0; q < 3; ++q) fion_host[3 * (size_t)l + q] += h[4 * (size_t)l + 1 + q];
  }
  if (esr) *esr = e;
}

}  // namespace mc

namespace mc {

// phfac on the device: the projector structure factors of the current ions
// (replaces the host eigr upload after an ionic move)
void phfac(mc_ctx* ctx) {
  auto& D = ctx->dens;
  auto& P = ctx->proj;
  MC_CHECK(D.ions_ready, MC_ERR_ARG, "phfac: ions not set (mc_set_ions)");
  MC_CHECK(P.ready && P.nat == D.nat, MC_ERR_ARG, "phfac: projector atoms differ from the ions of mc_set_ions");
  P.d_eigr.ensure(sizeof(double2) * (size_t)ctx->ldg * std::max(D.nat, 1));
  P.eigr_nat = D.nat;
  P.eigr_serial = D.ions_serial;
  if (D.nat == 0) return;
  LAUNCH(MC_KID_VOFRHO, k_eigr, dim3((ctx->ldg + 255) / 256, D.nat), dim3(256), 0, ctx->d_gmw.as<int>(), ctx->ngw,
         ctx->ldg, ctx->n1, ctx->n2, ctx->n3, D.d_ph1.as<double2>(), D.d_ph2.as<double2>(), D.d_ph3.as<double2>(),
         P.d_eigr.as<double2>());
}

}  // namespace mc
