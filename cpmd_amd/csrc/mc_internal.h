// Internal declarations of the micpmd library (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/micpmd.h"

namespace mc {

// ---------------------------------------------------------------------------
// error handling: every ABI entry point returns a status and leaves a message
// in a thread-local buffer (mc_last_error()).
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);
void clear_error();

struct Error {
  int code;
  std::string msg;
};

#define MC_HIP(call)                                                           \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      throw ::mc::Error{MC_ERR_HIP, std::string(#call) + ": " +                \
                                        hipGetErrorString(e_) + " (" +         \
                                        __FILE__ + ":" +                       \
                                        std::to_string(__LINE__) + ")"};       \
  } while (0)

// Every entry point counts one wavefunction epoch, except those that never
// write device wavefunction memory (MC_API_BEGIN_RO); an intermediate kept
// from one call for a later one on the same wavefunctions is valid only while
// the count is unchanged.
unsigned long long wfn_epoch();
void wfn_epoch_bump();
// true inside an entry point declared with MC_API_BEGIN_RO (this thread)
bool in_ro_call();
void ro_call_begin();

#define MC_API_BEGIN      \
  try {                   \
    mc::clear_error();    \
    mc::wfn_epoch_bump();
#define MC_API_BEGIN_RO \
  try {                 \
    mc::clear_error();  \
    mc::ro_call_begin();
#define MC_API_END                                   \
  }                                                  \
  catch (const mc::Error& e) {                       \
    mc::set_error(e.msg);                            \
    return e.code;                                   \
  }                                                  \
  catch (const std::exception& e) {                  \
    mc::set_error(std::string("internal: ") + e.what()); \
    return MC_ERR_INTERNAL;                          \
  }                                                  \
  return MC_OK;

#define MC_CHECK(cond, code, msg)                                              \
  do {                                                                         \
    if (!(cond)) throw ::mc::Error{(code), std::string(msg)};                  \
  } while (0)

// ---------------------------------------------------------------------------
// 1D transform plan: n = prod(rad[0..nf)), executed as mixed-radix Stockham
// passes in LDS.  Twiddles exp(-2*pi*i*t/n), t in [0,n), live in device memory.
// ---------------------------------------------------------------------------
constexpr int kMaxStages = 12;
struct FftPlan {
  int n;
  int nf;
  int rad[kMaxStages];
};

FftPlan make_plan(int n);  // throws on unsupported lengths

// Device memory the library holds (all contexts and mc_dev_alloc blocks)
// and the limit of mc_set_mem_limit (0: none).
size_t mem_limit();
size_t mem_current();

// Device buffer owned by the context; grows on demand (counted against the
// device memory limit: MC_ERR_ARG when it would be exceeded).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int in_use = 0;   // workspaces: > 0 while a call uses them (never reclaimed then)
  void ensure(size_t nbytes);
  void release();
  ~DevBuf() { release(); }
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

// Reclaimable workspaces (the transform batch buffers of every context).  An
// allocation that would take the library beyond the device memory limit first
// frees the registered workspaces that no call is using, then tries again;
// the next transform call sizes them again to what is left.
void ws_register(DevBuf* b);
void ws_unregister(DevBuf* b);
// Marks two workspaces in use for a scope (before they are sized).
struct WsUse {
  DevBuf &a, &b;
  WsUse(DevBuf& x, DevBuf& y);
  ~WsUse();
};

// bf16 part planes of up to two tall operands (gemm_planes.hip).  Inside a
// PlaneScope (a library function whose wavefunction operands do not change
// while it runs) the planes of an operand are kept and reused by the next
// product with the same operand; outside one they are rebuilt per product.
struct PlaneCache {
  bool scoped = false;
  const void* src[2] = {nullptr, nullptr};
  int np[2] = {0, 0}, K[2] = {0, 0}, n[2] = {0, 0};
  DevBuf buf[2];
  int next = 0;
  // Kept planes (column reuse on): buf[0] holds the planes of the device
  // wavefunctions keep_src as of wavefunction epoch keep_epoch, across calls
  // (rnlsm's split of c0, reused by the next rortv and posupa); nullptr: none
  const void* keep_src = nullptr;
  int keep_np = 0, keep_K = 0, keep_n = 0;
  unsigned long long keep_epoch = 0;
};

// Per-kernel timing (HIP events around launches), enabled on request.
struct KernelTimer {
  bool enabled = false;
  unsigned long long mask = ~0ull;   // kernel ids timed while enabled (bit kid)
  std::vector<hipEvent_t> ev;  // pairs (start, stop)
  std::vector<int> kid;        // kernel id of each pair
  std::vector<hipEvent_t> pool;   // events of flushed pairs, reused
  double total_ms[MC_KID_COUNT] = {0};
  long long count[MC_KID_COUNT] = {0};
};

// Operand of rnlsm1 built on the fly: column ia of the real view of
// (-i)^l twnl(G) eigr(G, ia), G=0 halved (src/rnlsm1_utils.mod.F90:84-150).
struct ProjA {
  const double2* eigr;   // (ldg, nat) device layout
  const double* twnl;    // (ldg) of this projector
  int ldg;
  int lmod4;             // l mod 4 of (-i)^l
  int geq0;
  const double* gkk = nullptr;   // (ldg) one Cartesian component of G (rnlsm2), or nullptr
  // the per-axis phase tables eigr was built from (set when they still match
  // it): the projector plane builds form eigr(g, isa0 + col) from them instead
  // of reading the (ldg, nat) table; nullptr: read eigr
  const int* gmw = nullptr;      // int[3 ngw] centred Miller indices of the rows
  const double2 *ph1 = nullptr, *ph2 = nullptr, *ph3 = nullptr;
  int n1 = 0, n2 = 0, n3 = 0, isa0 = 0;
};

// Epilogue of fnonloc (src/fnonloc_utils.mod.F90:272-300):
// c2(g,i) += twnl(g) ctm_i auxc(g,i), ctm_i = -f_i (-i)^l, f_i < 1e-5 -> 1.
struct FnlEpi {
  const double* twnl;   // (ldg) of this projector; nullptr for a plain GEMM
  const double* f;      // (n) occupations on the device
  int lmod4;
  // plain GEMM only: also vcm(i, j) = vodt (D_new(i, j) - A(i, j)) with vcm
  // at the leading dimension of D (A and D of one layout); nullptr: off
  double* vcm = nullptr;
  double vodt = 0.0;
  // plain GEMM, split-bf16 path only: also the bf16 planes of the first
  // pl_rows rows of D_new (pl_np parts, the layout of split_planes for K =
  // pl_rows, rows beyond it zero); nullptr: off.  The caller sets pl_rows, the
  // host wrapper pl_nkb and pl_np.
  __bf16* pl = nullptr;
  int pl_rows = 0, pl_nkb = 0, pl_np = 0;
  // plain GEMM, split-bf16 path only: colsq[j * sq_nrb + r] = the sum over
  // the complex rows of 64-row block r of column j of |D_new|^2, the G = 0
  // row left out when sq_g0 (rekine's norms, finished by colsq_finish);
  // nullptr: off
  double* colsq = nullptr;
  int sq_g0 = 0, sq_nrb = 0;
};

// Nonlocal pseudopotential data of the current geometry.
struct Projectors {
  bool ready = false;
  int nsp = 0, nhxs = 0, nat = 0, eigr_nat = -1, fnl_nstate = -1;
  int eigr_serial = -1;   // the phase tables' version d_eigr was built from
  std::vector<int> na, ngh, lang;
  std::vector<double> h_h;
  DevBuf d_twnl;   // (ldg, nhxs, nsp)
  DevBuf d_h;      // (nhxs, nhxs, nsp)
  DevBuf d_eigr;   // (ldg, nat) complex
  DevBuf d_fnl;    // (nat, nhxs, nstate)
  DevBuf d_gk;     // (ldg, 3) G vectors in units of 2 pi / alat (rnlsm2)
  bool gk_ready = false;
  DevBuf d_dfnl;   // (nat, nhxs, 3, nstate)
  int dfnl_nstate = -1;
};

// Density G set, local pseudopotential and ions of vofrho / ewald (vofrho.hip).
struct DensitySet {
  bool ready = false, pp_ready = false, ions_ready = false;
  int ions_serial = 0;   // counts mc_set_ions calls (the phase tables' version)
  int nhg = 0, nsp = 0, nat = 0;
  std::vector<int> na;
  std::vector<double> zv, raggio, tau;
  double a[9] = {}, inv[9] = {};   // lattice vectors a[3i+q]; fractional f_i = sum_q d_q inv[3q+i]
  DevBuf d_gm, d_gidx, d_scg, d_gk;
  DevBuf d_sp_start, d_sp_of, d_zv, d_raggio, d_vps, d_rhops;
  DevBuf d_tau, d_ph1, d_ph2, d_ph3;   // d_tau: [tau (3, nat) | fractional coordinates (3, nat)]
  DevBuf w_W, w_g, w_part;
  // z-sticks of the density set: the distinct (m1, m2) (d_stk, sorted) and
  // the stick of each G (d_sidx); operands of the stick GEMMs of vofrho
  int nstick = 0;
  DevBuf d_stk, d_sidx;
  DevBuf w_q, w_bs, w_bf, w_sf, w_y, w_t;
};

}  // namespace mc

// The context: static layout of the plane-wave set on one GPU plus workspaces.
struct mc_ctx {
  mc_ctx() {
    mc::ws_register(&w_a);
    mc::ws_register(&w_b);
    mc::ws_register(&w_keep);
  }
  ~mc_ctx() {
    mc::ws_unregister(&w_a);
    mc::ws_unregister(&w_b);
    mc::ws_unregister(&w_keep);
  }
  mc_ctx(const mc_ctx&) = delete;
  mc_ctx& operator=(const mc_ctx&) = delete;

  // mesh and cell
  int n1 = 0, n2 = 0, n3 = 0;     // transform lengths
  int kr1 = 0, kr2 = 0, kr3 = 0;  // padded real-space leading dimensions
  int ngw = 0;                    // wavefunction plane waves (half sphere)
  int ldg = 0;                    // device leading dimension of wavefunctions (complex)
  double tpiba2 = 0.0, omega = 0.0;
  int geq0 = 1;
  int device = 0;
  hipStream_t stream = nullptr;   // borrowed or owned
  bool own_stream = false;
  // pinned staging buffer of small asynchronous uploads (mc::upload_async)
  void* h_stage = nullptr;
  size_t h_stage_bytes = 0;
  hipEvent_t h_stage_done = nullptr;   // recorded after the last copy out of h_stage
  // pinned staging buffer of device-to-host result copies (mc::dl_add / dl_finish)
  void* h_down = nullptr;
  size_t h_down_bytes = 0, h_down_used = 0;
  struct DlCopy {
    void* dst;
    size_t off, bytes;
  };
  std::vector<DlCopy> h_down_pending;

  // stick layout (see context.hip for the definitions)
  int n_own = 0;        // owner sticks (hold half-sphere G's); owner 0 is (0,0)
  int n_slots = 0;      // sticks = 2*n_own - 1
  int ncols = 0;        // distinct x-frequency columns among sticks
  std::vector<int> perm;          // stick-order position j -> CPMD index ig
  std::vector<int> h_own_start;   // n_own+1 run offsets into stick-ordered G
  std::vector<int> h_col_start;   // ncols+1 offsets into col_slot/col_gy
  int max_col_len = 0;

  mc::FftPlan plan1, plan2, plan3;

  // device tables
  mc::DevBuf d_perm;       // int[ngw]     stick-order -> CPMD order
  mc::DevBuf d_gmw;        // int[3 ngw]   centred indices (m1, m2, m3) of the device rows
  mc::DevBuf d_ginfo;      // int[ngw]     (owner << 12) | z index
  mc::DevBuf d_hg;         // double[ngw]  |G|^2 (tpiba2 units), stick order
  mc::DevBuf d_own_start;  // int[n_own+1]
  mc::DevBuf d_col_start;  // int[ncols+1]
  mc::DevBuf d_col_slot;   // int[n_slots] slot ids grouped by column
  mc::DevBuf d_col_gy;     // int[n_slots] y frequency index (0..n2-1)
  mc::DevBuf d_col_gx;     // int[ncols]   x frequency index (0..n1-1)
  mc::DevBuf d_tw1, d_tw2, d_tw3;  // twiddles, double2[n]

  // workspaces
  mc::DevBuf w_a;          // sticks:  [pair][slot][z]       complex
  mc::DevBuf w_b;          // columns: [pair][z][col][y]     complex
  // column reuse (mc_ctx_set_column_reuse): rhoofr_dev keeps the columns of
  // every pair after the inverse y pass in w_keep, and the next vpsi on the
  // same wavefunctions, precision and epoch starts from them
  bool col_reuse = false;
  mc::DevBuf w_keep;
  bool keep_valid = false;
  const void* keep_c0 = nullptr;
  int keep_nstate = 0, keep_tsize = 0;
  unsigned long long keep_epoch = 0;
  mc::DevBuf w_c0, w_c2;   // staging of host wavefunctions (device layout)
  mc::DevBuf w_v;          // staging of host potential / density
  mc::DevBuf w_red;        // reduction partials
  mc::DevBuf w_f;          // occupations on device

  mc::DevBuf w_gemm;       // split-K partials
  mc::DevBuf w_small[8];    // n x n matrices (7: single overlaps, rotations and rortv's gamy)
  mc::DevBuf w_red2, w_info, w_wfn, w_f2, w_dd;
  mc::DevBuf w_colsq;   // per-block column sums of a colsq epilogue (rekine fused into rortv)
  // preconditioner of the electronic gradient (mc_hesele), device rows
  mc::DevBuf d_vpp, w_hw;
  bool vpp_ready = false;
  double vpp0 = 0.0;   // vpp at G = 0 (odiis' damping step)
  bool vpp_prec = false;
  // ODIIS history (mc_odiis_dev): ring of mdiis coefficient / gradient sets
  struct Odiis {
    int istate = 0, ndiis = 0, nowv = 0, nsize = 0, nocc = 0, mdiis = 0, nstate = 0;
    double eold = 0.0, gamma = 0.0;
    double diism[20 * 20] = {};
    double grmax[20] = {}, gimax[20] = {}, gnorm[20] = {};   // grimax of each slot
    mc::DevBuf pme, gde, d_ff, w_part;
  } odiis;
  mc::Projectors proj;

  bool fft_fp32 = false;   // rhoofr/vpsi transforms and workspaces in FP32
  // the wavefunction sphere lies inside |m1| < n1/4 (resp. |m2| < n2/4): the
  // inverse x (y) passes and the forward x pass of the 16 x B engines skip
  // the frequencies outside that band (dft16_band); true whenever the mesh
  // is at least twice the sphere's diameter (dual >= 4)
  bool prune_x = false, prune_y = false;
  bool gemm_x3 = false;    // tall wavefunction contractions as split-bf16 MFMA products
  int x3_parts = 3;        // bf16 parts per operand of the split A^T B products (2 or 3)
  mc::PlaneCache planes;   // bf16 planes of the wavefunction operands of the A^T B products
  mc::DevBuf w_plp;        // bf16 planes of the projector operand
  mc::DevBuf w_plg;        // bf16 planes of the small right operand of A G products
  mc::DensitySet dens;     // vofrho / ewald data

  // pairs per transform batch: larger batches amortise the per-launch tails
  // (512 states on 192^3: 128 / 64 pairs, FP64 workspace 6.9 / 3.4 GB)
  int batch_pairs_vpsi = 64;
  int batch_pairs_rho = 128;

  // charge sums of the last rhoofr: G space and real space before any rescale
  double last_csumg = 0.0, last_csumr_raw = 0.0;

  // device address of this GPU's allocation-free scratch array (wfft.hip)
  void* scratch = nullptr;

  mc::KernelTimer timer;
};

namespace mc {
// launch helpers used by the drivers (timed when ctx->timer.enabled)
void timer_begin(mc_ctx* ctx, int kid);
void timer_end(mc_ctx* ctx, int kid);
void timer_flush(mc_ctx* ctx);

__global__ void k_zero(double* __restrict__ x, size_t n);
__global__ void k_fill(double* __restrict__ x, size_t n, double v);
__global__ void k_axpy(double alpha, const double* __restrict__ x, double* __restrict__ y, size_t n);

// Output tiles of an A^T B launch: the tm x tn tiles row-major, or with upper
// (a symmetric product, tm == tn) only the tm (tm + 1) / 2 tiles on or above
// the diagonal, row by row (row ti holds tm - ti of them).
__host__ __device__ inline int tile_count(bool upper, int tm, int tn) { return upper ? tm * (tm + 1) / 2 : tm * tn; }
__device__ __forceinline__ void tile_ij(int tile, bool upper, int tm, int tn, int& ti, int& tj) {
  if (upper) {
    ti = 0;
    while (tile >= tm - ti) {
      tile -= tm - ti;
      ++ti;
    }
    tj = ti + tile;
  } else {
    ti = tile / tn;
    tj = tile - ti * tn;
  }
}

// vofrho / ewald (vofrho.hip)
void set_density_gset(mc_ctx* ctx, int nhg, const int* inyh, const double* hg, const double* gk);
void set_local_pp(mc_ctx* ctx, int nsp, const int* na, const double* zv, const double* raggio, const double* vps,
                  const double* rhops);
void set_ions(mc_ctx* ctx, const double* tau, const double* ht);
void vofrho_dev(mc_ctx* ctx, const double* rho, double* v, double* energies, double* fion_host);
void ewald(mc_ctx* ctx, int iesr, double* esr, double* eself, double* fion_host);
void phfac(mc_ctx* ctx);

// contractions (gemm.hip)
void gemm_tn(mc_ctx* ctx, const double* A, int lda, const double* B, int ldb, int K, int m, int n, double alpha,
             bool g0corr, bool symmetric, double* out, int ldo, const ProjA* proj);
// the FP64 kernel of gemm_nn whatever the context precision, timed under kid
void gemm_nn_f64(mc_ctx* ctx, const double* A, int lda, const double* G, int ldgm, int M, int k, int n, double alpha,
                 double beta, double* D, int ldd, bool upper_g, const FnlEpi* epi, int kid);
void gemm_nn(mc_ctx* ctx, const double* A, int lda, const double* G, int ldgm, int M, int k, int n, double alpha,
             double beta, double* D, int ldd, bool upper_g, const FnlEpi* epi);
// out = alpha * sum of the nsplit partials in ctx->w_gemm (- G=0 term), mirrored if symmetric
void reduce_tn(mc_ctx* ctx, int nsplit, int m, int n, double alpha, const double* A, int lda, const double* B, int ldb,
               bool g0corr, bool symmetric, double* out, int ldo);
// rows [r0, r0 + m) of partials with ldp rows: out = alpha sum of the nsplit partials
void reduce_tn_rows(mc_ctx* ctx, int nsplit, int ldp, int r0, int m, int n, double alpha, double* out, int ldo);
// split-bf16 (3-term, FP32 accumulation) variants used in mixed precision (gemm_x3.hip)
constexpr int kX3Min = 16384;   // rows of K (tn) or M (nn) from which the split kernels are used
// bf16 part planes and the contraction that streams them (gemm_planes.hip)
size_t planes_bytes(int K, int n, int np);
void split_planes(mc_ctx* ctx, const double* X, int ldx, int K, int n, int np, DevBuf& out);
void proj_planes(mc_ctx* ctx, const ProjA& pa, int K, int m, int np, DevBuf& out);
// cp = c0 + dt (cm + dtm c2) (c2 may be null; cm is not written), writing the
// planes of c0 and cp into the plane cache of the current PlaneScope
void posupa_planes(mc_ctx* ctx, const double* c0, const double* c2, const double* cm, double* cp, int ldx, int K,
                   int n, double dtm, double dt);
// cm += dtm c2 in place, writing the planes of c0 and of the new cm into the
// plane cache of the current PlaneScope
void kick_planes(mc_ctx* ctx, const double* c0, const double* c2, double* cm, int ldx, int K, int n, double dtm);
// Kept planes of X in the cache's slot 0: lag 0 for a read-only call, 1 for a
// call that counted its own epoch and has written nothing X covers yet
bool kept_planes(const mc_ctx* ctx, const void* X, int K, int n, int np, int lag);
// marks slot 0 (the planes of c0 the last posupa_planes / kick_planes of
// this call left there) as kept, when column reuse is on
void keep_c0_planes(mc_ctx* ctx, const void* c0, int K, int n);
// marks slot 0 as the planes of X, written there by a final rotation's
// epilogue (FnlEpi::pl), when column reuse is on
void keep_new_planes(mc_ctx* ctx, const void* X, int K, int n);
// whether gemm_nn takes the split-bf16 plane form for these operands (the
// form that supports FnlEpi::pl and FnlEpi::colsq)
bool nn_plane_form(const mc_ctx* ctx, const double* A, int lda, const double* G, int ldgm, int M, int ldd);
// out[j] = 2 sum_r colsq[j * nrb + r] + (g0 ? D(0, j)^2 : 0) (a fixed-order
// sum): the Gamma-point norms <d_j|d_j> from the partials of a colsq epilogue
void colsq_finish(mc_ctx* ctx, const double* colsq, int nrb, int n, const double* D, int ldd, int g0, double* out);
// rows per colsq block (the nrb of a product with M rows: ceil(M / kColsqRows))
constexpr int kColsqRows = 64;
void cp_posupa_dev(mc_ctx* ctx, const double2* c0, const double2* c2, double2* cm, double2* cp, int nstate, double dtm,
                   double dt);
void cp_velocity_dev(mc_ctx* ctx, const double2* c0, const double2* cp, double2* cm, int nstate, double dt);
void posupa_dev(mc_ctx* ctx, const double2* c0, const double2* c2, double2* cm, double2* cp, int n, double dtm,
                double dt, double epsog, int maxit, double* gam_host, int* iters);
void gemm_tn_planes(mc_ctx* ctx, const DevBuf& Apl, const DevBuf& Bpl, int np, int K, int m, int n, double alpha,
                    const double* A64, int lda, const double* B64, int ldb, bool g0corr, bool symmetric, double* out,
                    int ldo);
// D = alpha A G + beta D (epilogue as gemm_nn) with G as bf16 planes (gemm_planes.hip)
void gemm_nn_planes(mc_ctx* ctx, const double* A, int lda, const double* G, int ldgm, int M, int k, int n,
                    double alpha, double beta, double* D, int ldd, bool upper_g, const FnlEpi& epi);
// partials of A^T B from planes into ctx->w_gemm (m x n each); returns their number
int tn_planes_partials(mc_ctx* ctx, const DevBuf& Apl, const DevBuf& Bpl, int np, int K, int m, int n, bool symmetric);
// planes of the projector variants of one (species, projector) group, stacked
// along the columns in blocks of ceil(na/128)*128: variant v = 0 (fnl:
// (-i)^l twnl eigr) and v = 1..3 (dfnl: (-i)^(l+1) twnl gk_v eigr), for
// v in [v0, v0 + nv)
void proj_planes_stacked(mc_ctx* ctx, const ProjA& base, const double* gk, int v0, int nv, int K, int na, int np,
                         DevBuf& out);
// keeps operand planes for reuse while alive (see PlaneCache)
struct PlaneScope {
  mc_ctx* c;
  explicit PlaneScope(mc_ctx* ctx);
  ~PlaneScope();
};
void proj_contract_x3(mc_ctx* ctx, const ProjA& base, const double* gk, int v0, int nv, const double* B, int ldb,
                      int K, int na, int n, double* const* outs, const int* ldos, const double* alphas);
// copy bytes from host memory to the device on the context stream without
// waiting for the stream: through the pinned staging buffer, which is reused
// once the previous copy out of it has completed
void upload_async(mc_ctx* ctx, void* dst, const void* src, size_t bytes);
// Device-to-host copies of results through the pinned download stage: dl_add
// queues a copy of bytes from src (device) on the context stream, dl_finish
// waits for the stream and moves every queued copy to its host destination
// (pageable destinations would make the runtime stage each copy itself).
void dl_add(mc_ctx* ctx, void* dst, const void* src, size_t bytes);
void dl_finish(mc_ctx* ctx);
void gemm_tn_x3(mc_ctx* ctx, const double* A, int lda, const double* B, int ldb, int K, int m, int n, double alpha,
                bool g0corr, bool symmetric, double* out, int ldo, const ProjA* proj);
void gemm_nn_x3(mc_ctx* ctx, const double* A, int lda, const double* G, int ldgm, int M, int k, int n, double alpha,
                double beta, double* D, int ldd, bool upper_g, const FnlEpi* epi);

// orthonormalisation and projectors (ortho.hip)
void ovlap_dev(mc_ctx* ctx, int n, double* a_dev, int lda, const double2* c1, const double2* c2);
void rgs_dev(mc_ctx* ctx, double2* c, int n, double* smat_host);
void lowdin_dev(mc_ctx* ctx, double2* c, int n, int* iters);
void rotate_dev(mc_ctx* ctx, double a, const double2* c1, double b, double2* c2, const double* gam_dev, int n);
void lagrange_dev(mc_ctx* ctx, const double2* c0, double2* c2, const double* f_host, int n);
void rortog_dev(mc_ctx* ctx, const double2* c0, double2* cm, double* gam_host, int tnon, int n, double epsog,
                int maxit, int* iters);
// gamy_host (n x n, may be null) is filled by a queued download: dl_finish
// c2 != nullptr: velupa's kick cm += dtm c2 first
// dotp_dev (device, n doubles; may be null): <cm_i|cm_i> of the result, from
// the rotation's epilogue when the product takes the split-bf16 plane form;
// returns whether dotp_dev was written
bool rortv_dev(mc_ctx* ctx, const double2* c0, double2* cm, double* gamy_host, int n, const double2* c2 = nullptr,
               double dtm = 0.0, double* dotp_dev = nullptr);
void set_projectors(mc_ctx* ctx, int nsp, const int* na, int nhxs, const int* ngh, const int* lang,
                    const double* hmat, const double* twnl, int ldtw);
void set_eigr(mc_ctx* ctx, const double* eigr, int lde, int nat);
void rnlsm1_dev(mc_ctx* ctx, const double2* c0, int n, double* fnl_host);
void set_gk(mc_ctx* ctx, const double* gk);
void rnlsm2_dev(mc_ctx* ctx, const double2* c0, int n, double* dfnl_host);
void rnlsm_dev(mc_ctx* ctx, const double2* c0, int n, bool tfor, double* fnl_host, double* dfnl_host);
void rnlfor_dev(mc_ctx* ctx, const double* f_host, int n, const double* fnl_host, const double* dfnl_host,
                double* fion_host);
void rnlrh_dev(mc_ctx* ctx, const double* f_host, int n, double* enl);
void fnonloc_dev(mc_ctx* ctx, double2* c2, const double* f_host, int n, const double* fnl_host);
void wfn_upload(mc_ctx* ctx, const double* host, int ldc, int nstate, double2* dev);
void wfn_download(mc_ctx* ctx, const double2* dev, int nstate, double* host, int ldc);

// driver entry points implemented in wfft.hip
// keep / reuse: the column reuse of mc_ctx_set_column_reuse applies (the
// device-resident entry points; host-pointer ones stage through buffers
// that the next call overwrites)
void rhoofr_dev(mc_ctx* ctx, const double2* c0, const double* f_host, int nstate,
                double* rho, double* ekin, double* csumg, double* csumr, bool keep = false);
void vpsi_dev(mc_ctx* ctx, const double2* c0, double2* c2, const double* f_host,
              const double* vpot, int nstate, bool acc = true, bool reuse = false);
void forces_c2_dev(mc_ctx* ctx, const double2* c0, double2* c2, const double* f_host,
                   const double* vpot, int nstate);
// address of the allocation-free scratch array on `device` (the current
// device), looked up once per device with hipGetSymbolAddress
void* scratch_init(int device);
// vpsi (accumulating) with the workspaces and occupations in the scratch array
void vpsi_noalloc_dev(mc_ctx* ctx, const double2* c0, double2* c2, const double* f_host, const double* vpot,
                      int nstate);
}  // namespace mc
