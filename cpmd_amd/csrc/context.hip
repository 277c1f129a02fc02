// Context creation, stick layout, wavefunction permutes and the small C-ABI
// helpers of the micpmd library.
//
// Stick layout (the library's own; CPMD's x-ray layout of
// src/fftprp_utils.mod.F90:209-285 plays the same role):
//  * a stick is a line along z at fixed (m1, m2) Miller indices;
//  * an owner stick holds half-sphere G vectors (CPMD's half sphere,
//    src/loadpa_utils.mod.F90:283-300: m1>0 | m1=0,m2>0 | m1=m2=0,m3>=0);
//    owner 0 is (0,0), the others are ordered by (m1, m2);
//  * the mirror of owner o>0 is (-m1,-m2); it receives conj(c) at -m3;
//  * slots: owner o -> slot o, mirror of owner o -> slot n_own + o - 1;
//  * wavefunction rows on the device are permuted into stick order: the G
//    vectors of owner o form the run [own_start[o], own_start[o+1]), sorted by
//    m3, so G=0 is row 0 of the device layout;
//  * a column is the set of sticks with the same m1 (x frequency); the y pass
//    works per column.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "fft_device.h"
#include "mc_internal.h"

namespace mc {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

static std::atomic<unsigned long long> g_wfn_epoch{0};
static thread_local bool g_ro_call = false;
unsigned long long wfn_epoch() { return g_wfn_epoch.load(std::memory_order_acquire); }
void wfn_epoch_bump() {
  g_ro_call = false;
  g_wfn_epoch.fetch_add(1, std::memory_order_acq_rel);
}
bool in_ro_call() { return g_ro_call; }
void ro_call_begin() { g_ro_call = true; }

// ---------------------------------------------------------------------------
// device memory accounting: every allocation of the library (context buffers
// and mc_dev_alloc) is counted; with a limit set, an allocation that would
// take the count beyond it fails with MC_ERR_ARG instead of being made
// ---------------------------------------------------------------------------
namespace {
struct MemTrack {
  std::mutex mu;
  size_t cur = 0, peak = 0, limit = 0;
  std::map<void*, size_t> user;   // mc_dev_alloc'ed blocks
  std::vector<DevBuf*> ws;        // reclaimable workspaces
};
MemTrack& mem() {
  static MemTrack* t = new MemTrack();   // never destroyed: buffers may be freed at exit
  return *t;
}
// count `add` bytes that replace `drop` counted bytes; throws beyond the limit
void mem_take(size_t add, size_t drop, const char* what) {
  auto& t = mem();
  std::lock_guard<std::mutex> lk(t.mu);
  const size_t after = t.cur - drop + add;
  if (t.limit && add > drop && after > t.limit)
    throw Error{MC_ERR_ARG, std::string(what) + ": device memory limit of " + std::to_string(t.limit) +
                                " bytes exceeded (" + std::to_string(t.cur - drop) + " in use, " +
                                std::to_string(add) + " requested)"};
  t.cur = after;
  t.peak = std::max(t.peak, t.cur);
}
void mem_give(size_t bytes) {
  auto& t = mem();
  std::lock_guard<std::mutex> lk(t.mu);
  t.cur -= std::min(bytes, t.cur);
}
// frees the registered workspaces no call is using (other than `keep`);
// returns the bytes freed.  They are detached under the lock, so a call that
// marks them in use afterwards finds them empty and sizes them again.  hipFree
// waits for the device, so kernels still queued on them finish first.
size_t ws_reclaim(const DevBuf* keep) {
  std::vector<std::pair<void*, size_t>> drop;
  {
    auto& t = mem();
    std::lock_guard<std::mutex> lk(t.mu);
    for (DevBuf* b : t.ws)
      if (b != keep && b->in_use == 0 && b->p) {
        drop.emplace_back(b->p, b->bytes);
        b->p = nullptr;
        b->bytes = 0;
      }
  }
  size_t freed = 0;
  for (const auto& d : drop) {
    (void)hipFree(d.first);
    mem_give(d.second);
    freed += d.second;
  }
  return freed;
}
// mem_take with one retry after freeing the idle workspaces
void take_or_reclaim(size_t add, const char* what, const DevBuf* self) {
  try {
    mem_take(add, 0, what);
  } catch (const Error&) {
    if (!ws_reclaim(self)) throw;
    mem_take(add, 0, what);
  }
}
}  // namespace

void ws_register(DevBuf* b) {
  std::lock_guard<std::mutex> lk(mem().mu);
  mem().ws.push_back(b);
}
void ws_unregister(DevBuf* b) {
  std::lock_guard<std::mutex> lk(mem().mu);
  auto& v = mem().ws;
  v.erase(std::remove(v.begin(), v.end(), b), v.end());
}
WsUse::WsUse(DevBuf& x, DevBuf& y) : a(x), b(y) {
  std::lock_guard<std::mutex> lk(mem().mu);
  ++a.in_use;
  ++b.in_use;
}
WsUse::~WsUse() {
  std::lock_guard<std::mutex> lk(mem().mu);
  --a.in_use;
  --b.in_use;
}

size_t mem_limit() {
  std::lock_guard<std::mutex> lk(mem().mu);
  return mem().limit;
}
size_t mem_current() {
  std::lock_guard<std::mutex> lk(mem().mu);
  return mem().cur;
}

void DevBuf::ensure(size_t nbytes) {
  if (nbytes <= bytes) return;
  release();
  if (nbytes == 0) return;
  take_or_reclaim(nbytes, "device buffer", this);
  const hipError_t e = hipMalloc(&p, nbytes);
  if (e != hipSuccess) {
    p = nullptr;
    mem_give(nbytes);
    MC_HIP(e);
  }
  bytes = nbytes;
}
void upload_async(mc_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return;
  if (ctx->h_stage_done) MC_HIP(hipEventSynchronize(ctx->h_stage_done));
  if (bytes > ctx->h_stage_bytes) {
    if (ctx->h_stage) MC_HIP(hipHostFree(ctx->h_stage));
    ctx->h_stage = nullptr;
    ctx->h_stage_bytes = 0;
    const size_t nb = std::max(bytes, (size_t)65536);
    MC_HIP(hipHostMalloc(&ctx->h_stage, nb, hipHostMallocDefault));
    ctx->h_stage_bytes = nb;
  }
  if (!ctx->h_stage_done) MC_HIP(hipEventCreateWithFlags(&ctx->h_stage_done, hipEventDisableTiming));
  std::memcpy(ctx->h_stage, src, bytes);
  MC_HIP(hipMemcpyAsync(dst, ctx->h_stage, bytes, hipMemcpyHostToDevice, ctx->stream));
  MC_HIP(hipEventRecord(ctx->h_stage_done, ctx->stream));
}

void dl_finish(mc_ctx* ctx) {
  MC_HIP(hipStreamSynchronize(ctx->stream));
  for (const auto& c : ctx->h_down_pending)
    std::memcpy(c.dst, static_cast<const char*>(ctx->h_down) + c.off, c.bytes);
  ctx->h_down_pending.clear();
  ctx->h_down_used = 0;
}

void dl_add(mc_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return;
  size_t off = (ctx->h_down_used + 255) & ~(size_t)255;
  if (off + bytes > ctx->h_down_bytes) {
    if (!ctx->h_down_pending.empty()) dl_finish(ctx);
    off = 0;
    if (bytes > ctx->h_down_bytes) {
      if (ctx->h_down) MC_HIP(hipHostFree(ctx->h_down));
      ctx->h_down = nullptr;
      ctx->h_down_bytes = 0;
      const size_t nb = std::max(bytes, (size_t)1 << 20);
      MC_HIP(hipHostMalloc(&ctx->h_down, nb, hipHostMallocDefault));
      ctx->h_down_bytes = nb;
    }
  }
  MC_HIP(hipMemcpyAsync(static_cast<char*>(ctx->h_down) + off, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  ctx->h_down_pending.push_back({dst, off, bytes});
  ctx->h_down_used = off + bytes;
}

void DevBuf::release() {
  if (p) {
    (void)hipFree(p);
    mem_give(bytes);
  }
  p = nullptr;
  bytes = 0;
}

FftPlan make_plan(int n) {
  FftPlan pl{};
  pl.n = n;
  pl.nf = 0;
  MC_CHECK(n >= 2 && n <= 2048, MC_ERR_UNSUPPORTED,
           "transform length " + std::to_string(n) + " outside [2, 2048]");
  int m = n;
  // radix-4 first, then one radix-2, then 3s, 5s and 7s (the 7s: lengths of
  // the FFTW table of src/fftchk_utils.mod.F90:73-97)
  while (m % 4 == 0) { pl.rad[pl.nf++] = 4; m /= 4; }
  while (m % 2 == 0) { pl.rad[pl.nf++] = 2; m /= 2; }
  while (m % 3 == 0) { pl.rad[pl.nf++] = 3; m /= 3; }
  while (m % 5 == 0) { pl.rad[pl.nf++] = 5; m /= 5; }
  while (m % 7 == 0) { pl.rad[pl.nf++] = 7; m /= 7; }
  MC_CHECK(m == 1 && pl.nf <= kMaxStages, MC_ERR_UNSUPPORTED,
           "transform length " + std::to_string(n) + " is not of the form 2^a 3^b 5^c 7^d");
  return pl;
}

static void upload_twiddles(DevBuf& buf, int n) {
  std::vector<double> tw(2 * (size_t)n);
  for (int t = 0; t < n; ++t) {
    // exp(-2 pi i t/n) evaluated in long double and rounded once to double
    long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)n;
    tw[2 * t] = (double)cosl(a);
    tw[2 * t + 1] = (double)sinl(a);
  }
  buf.ensure(sizeof(double) * 2 * n);
  MC_HIP(hipMemcpy(buf.p, tw.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice));
}

static inline int wrap(int m, int n) { int r = m % n; return r < 0 ? r + n : r; }

struct HostLayout {
  int n_own = 0, n_slots = 0, ncols = 0, max_col_len = 0;
  std::vector<int> perm, ginfo, own_start, col_start, col_slot, col_gy, col_gx;
  std::vector<int> gmw;   // (3, ngw) centred indices of the device rows
  std::vector<double> hg;
};

// Builds the stick layout from CPMD's inyh/hg; host only.
static HostLayout build_layout(const mc_layout* lay) {
  MC_CHECK(lay != nullptr, MC_ERR_ARG, "null layout");
  const int n1 = lay->nr1, n2 = lay->nr2, n3 = lay->nr3;
  MC_CHECK(n1 > 1 && n2 > 1 && n3 > 1, MC_ERR_ARG, "mesh dimensions must be > 1");
  MC_CHECK(n3 < 4096, MC_ERR_UNSUPPORTED, "nr3 must be < 4096");
  MC_CHECK(lay->kr1 >= n1 && lay->kr2 >= n2 && lay->kr3 >= n3, MC_ERR_ARG,
           "leading dimensions kr must be >= nr");
  MC_CHECK(lay->ngw > 0 && lay->inyh && lay->hg, MC_ERR_ARG, "empty G set");
  const int ngw = lay->ngw;
  const int nh1 = n1 / 2 + 1, nh2 = n2 / 2 + 1, nh3 = n3 / 2 + 1;
  HostLayout L;
  // owner sticks keyed by (m1, m2)
  std::map<std::pair<int, int>, int> own_of;
  std::vector<std::tuple<int, int, int>> mill(ngw);
  for (int ig = 0; ig < ngw; ++ig) {
    int m1 = lay->inyh[3 * ig] - nh1, m2 = lay->inyh[3 * ig + 1] - nh2, m3 = lay->inyh[3 * ig + 2] - nh3;
    MC_CHECK(2 * std::abs(m1) < n1 && 2 * std::abs(m2) < n2 && 2 * std::abs(m3) < n3, MC_ERR_ARG,
             "G vector " + std::to_string(ig + 1) + " outside the mesh (aliasing)");
    mill[ig] = {m1, m2, m3};
    own_of[{m1, m2}] = 0;
  }
  MC_CHECK(own_of.count({0, 0}) || !lay->geq0, MC_ERR_ARG, "geq0 set but no (0,0) stick");
  // CPMD half sphere: a stick and its mirror never both own G vectors.
  for (auto& kv : own_of) {
    auto k = kv.first;
    if (k.first == 0 && k.second == 0) continue;
    MC_CHECK(!own_of.count({-k.first, -k.second}), MC_ERR_UNSUPPORTED,
             "G set is not a CPMD half sphere: sticks (" + std::to_string(k.first) + "," +
                 std::to_string(k.second) + ") and its mirror both hold G vectors");
  }
  // order owners: (0,0) first, then by (m1, m2)
  std::vector<std::pair<int, int>> owners;
  if (own_of.count({0, 0})) owners.push_back({0, 0});
  for (auto& kv : own_of)
    if (!(kv.first.first == 0 && kv.first.second == 0)) owners.push_back(kv.first);
  std::sort(owners.begin() + (own_of.count({0, 0}) ? 1 : 0), owners.end());
  const bool has_center = own_of.count({0, 0}) > 0;
  MC_CHECK(has_center, MC_ERR_UNSUPPORTED, "G set without the (0,0) stick is not supported");
  for (size_t o = 0; o < owners.size(); ++o) own_of[owners[o]] = (int)o;
  L.n_own = (int)owners.size();
  MC_CHECK(L.n_own < (1 << 19), MC_ERR_UNSUPPORTED, "too many sticks");
  L.n_slots = 2 * L.n_own - 1;
  // stick-ordered G: sort by (owner, m3)
  std::vector<int> idx(ngw);
  for (int i = 0; i < ngw; ++i) idx[i] = i;
  std::vector<int> owner_of_g(ngw);
  for (int ig = 0; ig < ngw; ++ig) owner_of_g[ig] = own_of[{std::get<0>(mill[ig]), std::get<1>(mill[ig])}];
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
    if (owner_of_g[a] != owner_of_g[b]) return owner_of_g[a] < owner_of_g[b];
    return std::get<2>(mill[a]) < std::get<2>(mill[b]);
  });
  L.perm = idx;
  L.gmw.resize(3 * (size_t)ngw);
  for (int j = 0; j < ngw; ++j) {
    L.gmw[3 * (size_t)j] = std::get<0>(mill[idx[j]]);
    L.gmw[3 * (size_t)j + 1] = std::get<1>(mill[idx[j]]);
    L.gmw[3 * (size_t)j + 2] = std::get<2>(mill[idx[j]]);
  }
  L.ginfo.resize(ngw);
  L.hg.resize(ngw);
  L.own_start.assign(L.n_own + 1, 0);
  for (int j = 0; j < ngw; ++j) {
    const int ig = idx[j];
    const int o = owner_of_g[ig];
    L.ginfo[j] = (o << 12) | wrap(std::get<2>(mill[ig]), n3);
    L.hg[j] = lay->hg[ig];
    L.own_start[o + 1]++;
  }
  for (int o = 0; o < L.n_own; ++o) L.own_start[o + 1] += L.own_start[o];
  if (lay->geq0) {
    MC_CHECK(std::get<0>(mill[idx[0]]) == 0 && std::get<1>(mill[idx[0]]) == 0 &&
                 std::get<2>(mill[idx[0]]) == 0,
             MC_ERR_ARG, "geq0 set but G=0 missing");
    MC_CHECK(idx[0] == 0, MC_ERR_ARG, "geq0 set but G=0 is not the first G vector");
  }
  // columns over all slots
  std::map<int, std::vector<std::pair<int, int>>> cols;  // m1 -> (m2, slot)
  for (int o = 0; o < L.n_own; ++o) {
    auto k = owners[o];
    cols[k.first].push_back({k.second, o});
    if (o > 0) cols[-k.first].push_back({-k.second, L.n_own + o - 1});
  }
  L.ncols = (int)cols.size();
  L.col_start.push_back(0);
  for (auto& kv : cols) {
    auto v = kv.second;
    std::sort(v.begin(), v.end());
    for (auto& e : v) {
      L.col_slot.push_back(e.second);
      L.col_gy.push_back(wrap(e.first, n2));
    }
    L.col_start.push_back((int)L.col_slot.size());
    L.col_gx.push_back(wrap(kv.first, n1));
    L.max_col_len = std::max(L.max_col_len, (int)v.size());
  }
  return L;
}

template <class T>
static void upload_vec(DevBuf& b, const std::vector<T>& v) {
  b.ensure(sizeof(T) * std::max<size_t>(v.size(), 1));
  if (!v.empty()) MC_HIP(hipMemcpy(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
}

// ---------------------------------------------------------------------------
// permutes between CPMD row order (staging, pitch ngw) and the device layout
// ---------------------------------------------------------------------------
__global__ void k_permute_in(const double2* __restrict__ stage, double2* __restrict__ c,
                             const int* __restrict__ perm, int ngw, int ldg, int nstate) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (i >= nstate || j >= ldg) return;
  c[(size_t)i * ldg + j] = (j < ngw) ? stage[(size_t)i * ngw + perm[j]] : make_double2(0.0, 0.0);
}
__global__ void k_permute_out(const double2* __restrict__ c, double2* __restrict__ stage,
                              const int* __restrict__ perm, int ngw, int ldg, int nstate) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (i >= nstate || j >= ngw) return;
  stage[(size_t)i * ngw + perm[j]] = c[(size_t)i * ldg + j];
}

__global__ void k_zero(double* __restrict__ x, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    x[i] = 0.0;
}

// y += alpha x: 16-byte accesses, four per lane in flight, when both are
// 16-byte aligned
__global__ void k_axpy(double alpha, const double* __restrict__ x, double* __restrict__ y, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(y)) & 15) != 0) {
    for (; i < n; i += stride) y[i] += alpha * x[i];
    return;
  }
  const double2* __restrict__ x2 = reinterpret_cast<const double2*>(x);
  double2* __restrict__ y2 = reinterpret_cast<double2*>(y);
  const size_t n2 = n / 2;
  for (; i + 3 * stride < n2; i += 4 * stride) {
    double2 a[4], b[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      a[k] = x2[i + k * stride];
      b[k] = y2[i + k * stride];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) y2[i + k * stride] = make_double2(b[k].x + alpha * a[k].x, b[k].y + alpha * a[k].y);
  }
  for (; i < n2; i += stride) {
    const double2 a = x2[i], b = y2[i];
    y2[i] = make_double2(b.x + alpha * a.x, b.y + alpha * a.y);
  }
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) y[n - 1] += alpha * x[n - 1];
}

// y = alpha x + beta y (beta = 0: y is not read)
__global__ void k_axpby(double alpha, const double* __restrict__ x, double beta, double* __restrict__ y, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    y[i] = beta == 0.0 ? alpha * x[i] : alpha * x[i] + beta * y[i];
}

// velupa + posupa in one pass over double2 coefficients:
//   cm += dtm c2 (src/velupa_utils.mod.F90:55-58, dtm = nnlst dt/(2 emass)),
//   cp = c0 + dt cm (src/posupa_utils.mod.F90:69-73)
__global__ void k_cp_posupa(const double2* __restrict__ c0, const double2* __restrict__ c2, double2* __restrict__ cm,
                            double2* __restrict__ cp, double dtm, double dt, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double2 a = c0[i], f = c2 ? c2[i] : make_double2(0.0, 0.0);
    double2 v = cm[i];
    v.x += dtm * f.x;
    v.y += dtm * f.y;
    cm[i] = v;
    cp[i] = make_double2(a.x + dt * v.x, a.y + dt * v.y);
  }
}

// velocities after the orthonormality constraint: cm = (cp - c0) / dt, which
// equals CPMD's cm += c0 gamx / dt (posupa:150-158) for cp = c0 + dt cm + c0 gamx
__global__ void k_cp_velocity(const double2* __restrict__ c0, const double2* __restrict__ cp,
                              double2* __restrict__ cm, double odt, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double2 a = c0[i], b = cp[i];
    cm[i] = make_double2(odt * (b.x - a.x), odt * (b.y - a.y));
  }
}

void cp_posupa_dev(mc_ctx* ctx, const double2* c0, const double2* c2, double2* cm, double2* cp, int nstate, double dtm,
                   double dt) {
  const size_t n = (size_t)ctx->ldg * nstate;
  if (!n) return;
  timer_begin(ctx, MC_KID_MISC);
  hipLaunchKernelGGL(k_cp_posupa, dim3(2048), dim3(256), 0, ctx->stream, c0, c2, cm, cp, c2 ? dtm : 0.0, dt, n);
  MC_HIP(hipGetLastError());
  timer_end(ctx, MC_KID_MISC);
}

void cp_velocity_dev(mc_ctx* ctx, const double2* c0, const double2* cp, double2* cm, int nstate, double dt) {
  const size_t n = (size_t)ctx->ldg * nstate;
  if (!n) return;
  timer_begin(ctx, MC_KID_MISC);
  hipLaunchKernelGGL(k_cp_velocity, dim3(2048), dim3(256), 0, ctx->stream, c0, cp, cm, 1.0 / dt, n);
  MC_HIP(hipGetLastError());
  timer_end(ctx, MC_KID_MISC);
}

// out[i] = Re a0 Re b0 + 2 sum_{G != 0} Re(conj(a) b) for column i (dotp,
// src/dotp_utils.mod.F90:58-101); one workgroup per state, fixed order
__global__ __launch_bounds__(256) void k_dotp_cols(const double2* __restrict__ a, const double2* __restrict__ b,
                                                   int ngw, int ldg, int geq0, double* __restrict__ out) {
  const int i = blockIdx.x;
  const double2* x = a + (size_t)ldg * i;
  const double2* y = b + (size_t)ldg * i;
  // eight independent partial sums per lane: eight loads of each operand in
  // flight (fixed summation order, so the result is reproducible)
  const int g0 = geq0 ? 1 : 0;
  double p[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  int g = g0 + threadIdx.x;
  for (; g + 7 * 256 < ngw; g += 8 * 256) {
    double2 u[8], v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      u[k] = x[g + 256 * k];
      v[k] = y[g + 256 * k];
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) p[k] += u[k].x * v[k].x + u[k].y * v[k].y;
  }
  for (int k = 0; g < ngw; g += 256, ++k) p[k & 7] += x[g].x * y[g].x + x[g].y * y[g].y;
  double s = ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7]));
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  __shared__ double part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double t = part[0] + part[1] + part[2] + part[3];
    out[i] = 2.0 * t + (geq0 ? x[0].x * y[0].x : 0.0);
  }
}

void wfn_upload(mc_ctx* ctx, const double* host, int ldc, int nstate, double2* dev) {
  MC_CHECK(ldc >= ctx->ngw && nstate >= 0, MC_ERR_ARG, "wfn_upload: ldc < ngw or nstate < 0");
  if (nstate == 0) return;
  const size_t sb = sizeof(double2) * (size_t)ctx->ngw * nstate;
  ctx->w_c0.ensure(sb);
  MC_HIP(hipMemcpy2DAsync(ctx->w_c0.p, sizeof(double2) * ctx->ngw, host, sizeof(double2) * (size_t)ldc,
                          sizeof(double2) * ctx->ngw, nstate, hipMemcpyHostToDevice, ctx->stream));
  dim3 grid((ctx->ldg + 255) / 256, nstate);
  hipLaunchKernelGGL(k_permute_in, grid, dim3(256), 0, ctx->stream, ctx->w_c0.as<double2>(), dev,
                     ctx->d_perm.as<int>(), ctx->ngw, ctx->ldg, nstate);
  MC_HIP(hipGetLastError());
  // the staging buffer is reused by the next upload
  MC_HIP(hipStreamSynchronize(ctx->stream));
}

void wfn_download(mc_ctx* ctx, const double2* dev, int nstate, double* host, int ldc) {
  MC_CHECK(ldc >= ctx->ngw && nstate >= 0, MC_ERR_ARG, "wfn_download: ldc < ngw or nstate < 0");
  if (nstate == 0) return;
  const size_t sb = sizeof(double2) * (size_t)ctx->ngw * nstate;
  ctx->w_c2.ensure(sb);
  dim3 grid((ctx->ngw + 255) / 256, nstate);
  hipLaunchKernelGGL(k_permute_out, grid, dim3(256), 0, ctx->stream, dev, ctx->w_c2.as<double2>(),
                     ctx->d_perm.as<int>(), ctx->ngw, ctx->ldg, nstate);
  MC_HIP(hipGetLastError());
  MC_HIP(hipMemcpy2DAsync(host, sizeof(double2) * (size_t)ldc, ctx->w_c2.p, sizeof(double2) * ctx->ngw,
                          sizeof(double2) * ctx->ngw, nstate, hipMemcpyDeviceToHost, ctx->stream));
  MC_HIP(hipStreamSynchronize(ctx->stream));
}

// ---------------------------------------------------------------------------
// kernel timer
// ---------------------------------------------------------------------------
static bool timed(const mc_ctx* ctx, int kid) {
  return ctx->timer.enabled && ((ctx->timer.mask >> kid) & 1ull);
}
static hipEvent_t timer_event(KernelTimer& t) {
  hipEvent_t e;
  if (!t.pool.empty()) {
    e = t.pool.back();
    t.pool.pop_back();
  } else {
    MC_HIP(hipEventCreate(&e));
  }
  return e;
}
void timer_begin(mc_ctx* ctx, int kid) {
  if (!timed(ctx, kid)) return;
  auto& t = ctx->timer;
  const hipEvent_t a = timer_event(t), b = timer_event(t);
  t.ev.push_back(a);
  t.ev.push_back(b);
  t.kid.push_back(kid);
  MC_HIP(hipEventRecord(a, ctx->stream));
}
void timer_end(mc_ctx* ctx, int kid) {
  if (!timed(ctx, kid)) return;
  MC_HIP(hipEventRecord(ctx->timer.ev.back(), ctx->stream));
}
void timer_flush(mc_ctx* ctx) {
  auto& t = ctx->timer;
  if (t.ev.empty()) return;
  MC_HIP(hipStreamSynchronize(ctx->stream));
  for (size_t k = 0; k < t.kid.size(); ++k) {
    float ms = 0.f;
    MC_HIP(hipEventElapsedTime(&ms, t.ev[2 * k], t.ev[2 * k + 1]));
    t.total_ms[t.kid[k]] += ms;
    t.count[t.kid[k]] += 1;
    t.pool.push_back(t.ev[2 * k]);
    t.pool.push_back(t.ev[2 * k + 1]);
  }
  t.ev.clear();
  t.kid.clear();
}

// hesele (src/hesele_utils.mod.F90:32-142): the diagonal preconditioner of the
// electronic gradient on the device rows, with prec
//   vpp = 1 / max(0.5 tpiba2 hg + vploc + sum_is sum_iv,jv (h na) tw_iv tw_jv, hthrs)
// (hw = h na per species, zero outside the (l, m) blocks), else vpp = svar2;
// padding rows get 0 so that updates leave them zero.
__global__ void k_hesele(const double* __restrict__ hg, const double* __restrict__ twnl, int ldg, int ngw,
                         int nhxs, int nsp, const double* __restrict__ hw, double vploc, double ftpi, double hthrs,
                         int prec, double svar2, double* __restrict__ vpp) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ldg) return;
  if (j >= ngw) {
    vpp[j] = 0.0;
    return;
  }
  if (!prec) {
    vpp[j] = svar2;
    return;
  }
  double v = vploc;
  for (int is = 0; is < nsp; ++is)
    for (int iv = 0; iv < nhxs; ++iv) {
      const double ti = twnl[(size_t)ldg * (iv + nhxs * is) + j];
      for (int jv = 0; jv < nhxs; ++jv) {
        const double dd = hw[iv + nhxs * (jv + nhxs * is)];
        if (dd != 0.0) v += dd * ti * twnl[(size_t)ldg * (jv + nhxs * is) + j];
      }
    }
  v = ftpi * hg[j] + v;
  if (v < hthrs) v = hthrs;
  vpp[j] = 1.0 / v;
}

// csize (src/csize_utils.mod.F90:100-131) per state i: out[6i] = dotp(c2_i, c2_i)
// in a fixed order (lanes strided by the block, then per wave, then the waves in
// order); out[6i+1..3] = the izamax pick over the state:
// largest |Re| + |Im| (BLAS dcabs1), ties to the smallest CPMD index
// i ngw + ig, then that index and the modulus of the element; out[6i+4..5] =
// max |Re| and max |Im| (grimax, src/odiis_utils.mod.F90:408-450).
constexpr int kCsizeThreads = 1024;   // 16 waves per state: enough loads in flight per CU

__global__ __launch_bounds__(kCsizeThreads) void k_csize_cols(const double2* __restrict__ c2,
                                                              const int* __restrict__ perm, int ngw, int ldg,
                                                              int geq0, double* __restrict__ out) {
  const int i = blockIdx.x;
  const double2* x = c2 + (size_t)ldg * i;
  double s = 0.0, key = -1.0, idx = 0.0, mod = 0.0, mre = 0.0, mim = 0.0;
  for (int g = threadIdx.x; g < ngw; g += kCsizeThreads) {
    const double2 v = x[g];
    if (g >= geq0) s += v.x * v.x + v.y * v.y;
    mre = fmax(mre, fabs(v.x));
    mim = fmax(mim, fabs(v.y));
    const double k = fabs(v.x) + fabs(v.y);
    const double id = (double)i * ngw + perm[g];
    if (k > key || (k == key && id < idx)) {
      key = k;
      idx = id;
      mod = hypot(v.x, v.y);
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_down(s, off, 64);
    mre = fmax(mre, __shfl_down(mre, off, 64));
    mim = fmax(mim, __shfl_down(mim, off, 64));
    const double k = __shfl_down(key, off, 64), id = __shfl_down(idx, off, 64), m = __shfl_down(mod, off, 64);
    if (k > key || (k == key && id < idx)) {
      key = k;
      idx = id;
      mod = m;
    }
  }
  constexpr int nw = kCsizeThreads / 64;
  __shared__ double part[nw][6];
  if ((threadIdx.x & 63) == 0) {
    const int w = threadIdx.x >> 6;
    part[w][0] = s;
    part[w][1] = key;
    part[w][2] = idx;
    part[w][3] = mod;
    part[w][4] = mre;
    part[w][5] = mim;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0, mr = 0.0, mi = 0.0;
    int b = 0;
    for (int w = 0; w < nw; ++w) {
      t += part[w][0];
      mr = fmax(mr, part[w][4]);
      mi = fmax(mi, part[w][5]);
      if (part[w][1] > part[b][1] || (part[w][1] == part[b][1] && part[w][2] < part[b][2])) b = w;
    }
    out[6 * i] = 2.0 * t + (geq0 ? x[0].x * x[0].x : 0.0);
    out[6 * i + 1] = part[b][1];
    out[6 * i + 2] = part[b][2];
    out[6 * i + 3] = part[b][3];
    out[6 * i + 4] = mr;
    out[6 * i + 5] = mi;
  }
}

// Preconditioned steepest descent of updwf (src/updwf_utils.mod.F90:118-125):
// c0(g, i) += (vpp(g) pf1) c2(g, i), one complex element per thread.
__global__ void k_sd_update(double2* __restrict__ c0, const double2* __restrict__ c2, const double* __restrict__ vpp,
                            double pf1, int ldg) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ldg) return;
  const size_t o = (size_t)ldg * blockIdx.y + j;
  const double w = vpp[j] * pf1;
  const double2 g = c2[o];
  double2 a = c0[o];
  a.x += w * g.x;
  a.y += w * g.y;
  c0[o] = a;
}

// odiis (src/odiis_utils.mod.F90:492-740) kernels.  Ring slot w of pme/gde
// holds the occupied coefficients and the negated gradient of one iteration.
__global__ void k_odiis_store(const double2* __restrict__ c0, const double2* __restrict__ c2,
                              double2* __restrict__ pme, double2* __restrict__ gde, int ldg) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ldg) return;
  const size_t o = (size_t)ldg * blockIdx.y + j;
  const double2 g = c2[o];
  pme[o] = c0[o];
  gde[o] = make_double2(-g.x, -g.y);
}

// part[i nocc + k] = dotp(gde_i(:, k), 2 ff_k^2 v^2 gde_w(:, k)), v = vpp / (1 + vpp gamma);
// one workgroup per (state, slot), fixed order as k_dotp_cols
__global__ __launch_bounds__(256) void k_odiis_bdot(const double2* __restrict__ gde, size_t slot, int w,
                                                    const double* __restrict__ vpp, double gamma,
                                                    const double* __restrict__ ff, int ngw, int ldg, int nocc,
                                                    int geq0, double* __restrict__ part) {
  const int k = blockIdx.x, i = blockIdx.y;
  const double2* a = gde + slot * i + (size_t)ldg * k;
  const double2* b = gde + slot * w + (size_t)ldg * k;
  const double s2 = 2.0 * ff[k] * ff[k];
  double s = 0.0;
  for (int g = threadIdx.x + (geq0 ? 1 : 0); g < ngw; g += 256) {
    const double v = gamma != 0.0 ? vpp[g] / (1.0 + vpp[g] * gamma) : vpp[g];
    const double wgt = s2 * v * v;
    s += a[g].x * (b[g].x * wgt) + a[g].y * (b[g].y * wgt);
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  __shared__ double pt[4];
  if ((threadIdx.x & 63) == 0) pt[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 2.0 * (pt[0] + pt[1] + pt[2] + pt[3]);
    if (geq0) {
      const double v = gamma != 0.0 ? vpp[0] / (1.0 + vpp[0] * gamma) : vpp[0];
      t += a[0].x * (b[0].x * (s2 * v * v));
    }
    part[(size_t)i * nocc + k] = t;
  }
}

struct DiisCoef {
  double vc[20];
  int nv;
};

// c0(:, k < nocc) = sum_i vc_i pme_i - sum_i vc_i ff_k v gde_i;  c0(:, k >= nocc) += svar2 c2
__global__ void k_odiis_combine(double2* __restrict__ c0, const double2* __restrict__ c2,
                                const double2* __restrict__ pme, const double2* __restrict__ gde, size_t slot,
                                const double* __restrict__ vpp, double gamma, const double* __restrict__ ff,
                                DiisCoef dc, int ldg, int nocc, double svar2) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ldg) return;
  const int k = blockIdx.y;
  const size_t o = (size_t)ldg * k + j;
  if (k >= nocc) {
    const double2 g = c2[o];
    double2 a = c0[o];
    a.x += svar2 * g.x;
    a.y += svar2 * g.y;
    c0[o] = a;
    return;
  }
  const double v = gamma != 0.0 ? vpp[j] / (1.0 + vpp[j] * gamma) : vpp[j];
  double2 s = make_double2(0.0, 0.0);
  for (int i = 0; i < dc.nv; ++i) {
    const double2 p = pme[slot * i + o];
    s.x += dc.vc[i] * p.x;
    s.y += dc.vc[i] * p.y;
  }
  for (int i = 0; i < dc.nv; ++i) {
    const double2 g = gde[slot * i + o];
    const double cw = dc.vc[i] * ff[k] * v;
    s.x -= cw * g.x;
    s.y -= cw * g.y;
  }
  c0[o] = s;
}

// Minimum-norm least-squares solution of the symmetric system b x = v (n <= 21)
// through a cyclic Jacobi eigendecomposition, dropping eigenvalues below
// rcond max|lambda|: the result of LAPACK dgelss for a symmetric matrix.
static void sym_lstsq(std::vector<double> b, int n, std::vector<double>& v, double rcond) {
  std::vector<double> z((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) z[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int p = 0; p < n; ++p)
      for (int q = 0; q < n; ++q) {
        tot += b[(size_t)p * n + q] * b[(size_t)p * n + q];
        if (p != q) off += b[(size_t)p * n + q] * b[(size_t)p * n + q];
      }
    if (off <= 1e-30 * tot) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = b[(size_t)p * n + q];
        if (apq == 0.0) continue;
        const double app = b[(size_t)p * n + p], aqq = b[(size_t)q * n + q];
        const double th = 0.5 * (aqq - app) / apq;
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
        for (int r = 0; r < n; ++r) {   // columns p, q
          const double brp = b[(size_t)r * n + p], brq = b[(size_t)r * n + q];
          b[(size_t)r * n + p] = c * brp - sn * brq;
          b[(size_t)r * n + q] = sn * brp + c * brq;
        }
        for (int r = 0; r < n; ++r) {   // rows p, q
          const double bpr = b[(size_t)p * n + r], bqr = b[(size_t)q * n + r];
          b[(size_t)p * n + r] = c * bpr - sn * bqr;
          b[(size_t)q * n + r] = sn * bpr + c * bqr;
        }
        for (int r = 0; r < n; ++r) {
          const double zrp = z[(size_t)r * n + p], zrq = z[(size_t)r * n + q];
          z[(size_t)r * n + p] = c * zrp - sn * zrq;
          z[(size_t)r * n + q] = sn * zrp + c * zrq;
        }
      }
  }
  double lmax = 0.0;
  for (int i = 0; i < n; ++i) lmax = std::max(lmax, std::fabs(b[(size_t)i * n + i]));
  std::vector<double> x(n, 0.0);
  for (int e = 0; e < n; ++e) {
    const double l = b[(size_t)e * n + e];
    if (std::fabs(l) <= rcond * lmax || l == 0.0) continue;
    double d = 0.0;
    for (int r = 0; r < n; ++r) d += z[(size_t)r * n + e] * v[r];
    d /= l;
    for (int r = 0; r < n; ++r) x[r] += d * z[(size_t)r * n + e];
  }
  v = x;
}

}  // namespace mc

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* mc_last_error(void) { return mc::g_last_error.c_str(); }
int mc_version(void) { return 100; }

int mc_device_count(int* count) {
  MC_API_BEGIN
  MC_CHECK(count, MC_ERR_ARG, "null count");
  int n = 0;
  MC_HIP(hipGetDeviceCount(&n));
  *count = n;
  MC_API_END
}

int mc_layout_stats(const mc_layout* lay, int* n_sticks, int* n_columns, int* n_owner) {
  MC_API_BEGIN
  mc::HostLayout L = mc::build_layout(lay);
  if (n_sticks) *n_sticks = L.n_slots;
  if (n_columns) *n_columns = L.ncols;
  if (n_owner) *n_owner = L.n_own;
  MC_API_END
}

int mc_ctx_create(const mc_layout* lay, mc_ctx** out) {
  MC_API_BEGIN
  MC_CHECK(out, MC_ERR_ARG, "null output pointer");
  *out = nullptr;
  mc::HostLayout L = mc::build_layout(lay);
  auto* ctx = new mc_ctx();
  try {
    MC_HIP(hipGetDevice(&ctx->device));
    ctx->n1 = lay->nr1; ctx->n2 = lay->nr2; ctx->n3 = lay->nr3;
    ctx->kr1 = lay->kr1; ctx->kr2 = lay->kr2; ctx->kr3 = lay->kr3;
    ctx->ngw = lay->ngw;
    ctx->ldg = (lay->ngw + 63) / 64 * 64;
    ctx->tpiba2 = lay->tpiba2;
    ctx->omega = lay->omega;
    ctx->geq0 = lay->geq0 ? 1 : 0;
    ctx->plan1 = mc::make_plan(ctx->n1);
    ctx->plan2 = mc::make_plan(ctx->n2);
    ctx->plan3 = mc::make_plan(ctx->n3);
    ctx->n_own = L.n_own;
    ctx->n_slots = L.n_slots;
    ctx->ncols = L.ncols;
    ctx->max_col_len = L.max_col_len;
    ctx->perm = L.perm;
    ctx->h_own_start = L.own_start;
    ctx->h_col_start = L.col_start;
    MC_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    ctx->own_stream = true;
    mc::upload_vec(ctx->d_perm, L.perm);
    mc::upload_vec(ctx->d_gmw, L.gmw);
    mc::upload_vec(ctx->d_ginfo, L.ginfo);
    mc::upload_vec(ctx->d_hg, L.hg);
    mc::upload_vec(ctx->d_own_start, L.own_start);
    mc::upload_vec(ctx->d_col_start, L.col_start);
    mc::upload_vec(ctx->d_col_slot, L.col_slot);
    mc::upload_vec(ctx->d_col_gy, L.col_gy);
    mc::upload_vec(ctx->d_col_gx, L.col_gx);
    int mx = 0, my = 0;   // largest |m1| of a column, |m2| of a stick
    for (int gx : L.col_gx) mx = std::max(mx, std::abs(gx <= ctx->n1 / 2 ? gx : gx - ctx->n1));
    for (int gy : L.col_gy) my = std::max(my, std::abs(gy <= ctx->n2 / 2 ? gy : gy - ctx->n2));
    ctx->prune_x = 4 * mx < ctx->n1;
    ctx->prune_y = 4 * my < ctx->n2;
    mc::upload_twiddles(ctx->d_tw1, ctx->n1);
    mc::upload_twiddles(ctx->d_tw2, ctx->n2);
    mc::upload_twiddles(ctx->d_tw3, ctx->n3);
    ctx->scratch = mc::scratch_init(ctx->device);
  } catch (...) {
    if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    throw;
  }
  *out = ctx;
  MC_API_END
}

int mc_ctx_destroy(mc_ctx* ctx) {
  MC_API_BEGIN
  if (!ctx) return MC_OK;
  (void)hipStreamSynchronize(ctx->stream);
  for (auto e : ctx->timer.ev) (void)hipEventDestroy(e);
  for (auto e : ctx->timer.pool) (void)hipEventDestroy(e);
  if (ctx->h_stage_done) (void)hipEventDestroy(ctx->h_stage_done);
  if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
  if (ctx->h_down) (void)hipHostFree(ctx->h_down);
  if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  MC_API_END
}

int mc_ctx_set_stream(mc_ctx* ctx, void* hip_stream) {
  MC_API_BEGIN
  MC_CHECK(ctx, MC_ERR_ARG, "null context");
  MC_HIP(hipStreamSynchronize(ctx->stream));
  if (hip_stream) {
    if (ctx->own_stream) MC_HIP(hipStreamDestroy(ctx->stream));
    ctx->stream = (hipStream_t)hip_stream;
    ctx->own_stream = false;
  } else if (!ctx->own_stream) {
    MC_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    ctx->own_stream = true;
  }
  MC_API_END
}

int mc_ctx_synchronize(mc_ctx* ctx) {
  MC_API_BEGIN_RO
  MC_CHECK(ctx, MC_ERR_ARG, "null context");
  MC_HIP(hipStreamSynchronize(ctx->stream));
  MC_API_END
}

int mc_ctx_set_batch(mc_ctx* ctx, int pairs_rhoofr, int pairs_vpsi) {
  MC_API_BEGIN
  MC_CHECK(ctx, MC_ERR_ARG, "null context");
  MC_CHECK(pairs_rhoofr >= 1 && pairs_vpsi >= 1 && pairs_rhoofr <= 4096 && pairs_vpsi <= 4096,
           MC_ERR_ARG, "batch sizes must be in [1, 4096]");
  ctx->batch_pairs_rho = pairs_rhoofr;
  ctx->batch_pairs_vpsi = pairs_vpsi;
  MC_API_END
}

int mc_ctx_set_column_reuse(mc_ctx* ctx, int on) {
  MC_API_BEGIN
  MC_CHECK(ctx, MC_ERR_ARG, "null context");
  MC_HIP(hipStreamSynchronize(ctx->stream));
  ctx->col_reuse = on != 0;
  ctx->keep_valid = false;
  ctx->planes.keep_src = nullptr;
  if (!ctx->col_reuse) ctx->w_keep.release();
  MC_API_END
}

size_t mc_wfn_ld(const mc_ctx* ctx) { return ctx ? (size_t)ctx->ldg : 0; }

int mc_set_precision(mc_ctx* ctx, int fft_precision) {
  MC_API_BEGIN
  MC_CHECK(ctx, MC_ERR_ARG, "null context");
  MC_CHECK(fft_precision >= MC_PREC_FP64 && fft_precision <= MC_PREC_MIXED_BF16X3, MC_ERR_ARG,
           "unknown precision mode");
  MC_HIP(hipStreamSynchronize(ctx->stream));
  ctx->fft_fp32 = fft_precision != MC_PREC_FP64;
  ctx->gemm_x3 = fft_precision == MC_PREC_MIXED || fft_precision == MC_PREC_MIXED_BF16X3;
  ctx->x3_parts = fft_precision == MC_PREC_MIXED_BF16X3 ? 2 : 3;
  MC_API_END
}

int mc_dev_alloc(size_t bytes, void** ptr) {
  MC_API_BEGIN_RO   // a new buffer holds no wavefunctions yet (mc_dev_free counts an epoch)
  MC_CHECK(ptr, MC_ERR_ARG, "null pointer");
  *ptr = nullptr;
  if (bytes) {
    mc::take_or_reclaim(bytes, "mc_dev_alloc", nullptr);
    const hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
      *ptr = nullptr;
      mc::mem_give(bytes);
      MC_HIP(e);
    }
    std::lock_guard<std::mutex> lk(mc::mem().mu);
    mc::mem().user[*ptr] = bytes;
  }
  MC_API_END
}
int mc_dev_free(void* ptr) {
  MC_API_BEGIN
  if (ptr) {
    MC_HIP(hipFree(ptr));
    size_t b = 0;
    {
      std::lock_guard<std::mutex> lk(mc::mem().mu);
      auto it = mc::mem().user.find(ptr);
      if (it != mc::mem().user.end()) {
        b = it->second;
        mc::mem().user.erase(it);
      }
    }
    mc::mem_give(b);
  }
  MC_API_END
}

int mc_set_mem_limit(size_t bytes) {
  MC_API_BEGIN
  std::lock_guard<std::mutex> lk(mc::mem().mu);
  mc::mem().limit = bytes;
  MC_API_END
}
int mc_mem_stats(size_t* current, size_t* peak, size_t* limit) {
  MC_API_BEGIN_RO
  std::lock_guard<std::mutex> lk(mc::mem().mu);
  if (current) *current = mc::mem().cur;
  if (peak) *peak = mc::mem().peak;
  if (limit) *limit = mc::mem().limit;
  MC_API_END
}
int mc_mem_reset_peak(void) {
  MC_API_BEGIN
  std::lock_guard<std::mutex> lk(mc::mem().mu);
  mc::mem().peak = mc::mem().cur;
  MC_API_END
}
int mc_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  MC_API_BEGIN
  if (bytes) MC_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  MC_API_END
}
int mc_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  MC_API_BEGIN_RO
  if (bytes) MC_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  MC_API_END
}
int mc_memset_d(void* dst, int value, size_t bytes) {
  MC_API_BEGIN
  if (bytes) MC_HIP(hipMemset(dst, value, bytes));
  MC_API_END
}

int mc_wfn_zero(mc_ctx* ctx, void* c_dev, int nstate) {
  MC_API_BEGIN
  MC_CHECK(ctx && c_dev && nstate >= 0, MC_ERR_ARG, "bad argument");
  const size_t n = (size_t)ctx->ldg * nstate * 2;
  if (n) {
    hipLaunchKernelGGL(mc::k_zero, dim3(1024), dim3(256), 0, ctx->stream, (double*)c_dev, n);
    MC_HIP(hipGetLastError());
  }
  MC_API_END
}

int mc_wfn_axpy(mc_ctx* ctx, double alpha, const void* x_dev, void* y_dev, int nstate) {
  MC_API_BEGIN
  MC_CHECK(ctx && x_dev && y_dev && nstate >= 0, MC_ERR_ARG, "bad argument");
  const size_t n = (size_t)ctx->ldg * nstate * 2;
  if (n) {
    hipLaunchKernelGGL(mc::k_axpy, dim3(2048), dim3(256), 0, ctx->stream, alpha, (const double*)x_dev,
                       (double*)y_dev, n);
    MC_HIP(hipGetLastError());
  }
  MC_API_END
}


int mc_cp_posupa(mc_ctx* ctx, const void* c0_dev, const void* c2_dev, void* cm_dev, void* cp_dev, int nstate,
                 double dtm, double dt) {
  MC_API_BEGIN
  MC_CHECK(ctx && c0_dev && c2_dev && cm_dev && cp_dev && nstate >= 0, MC_ERR_ARG, "bad argument");
  MC_CHECK(cp_dev != c0_dev && cp_dev != c2_dev && cp_dev != cm_dev, MC_ERR_ARG, "cp must not alias c0, c2 or cm");
  mc::cp_posupa_dev(ctx, (const double2*)c0_dev, (const double2*)c2_dev, (double2*)cm_dev, (double2*)cp_dev, nstate,
                    dtm, dt);
  MC_API_END
}

int mc_cp_velocity(mc_ctx* ctx, const void* c0_dev, const void* cp_dev, void* cm_dev, int nstate, double dt) {
  MC_API_BEGIN
  MC_CHECK(ctx && c0_dev && cp_dev && cm_dev && nstate >= 0 && dt != 0.0, MC_ERR_ARG, "bad argument");
  mc::cp_velocity_dev(ctx, (const double2*)c0_dev, (const double2*)cp_dev, (double2*)cm_dev, nstate, dt);
  MC_API_END
}

int mc_wfn_axpby(mc_ctx* ctx, double alpha, const void* x_dev, double beta, void* y_dev, int nstate) {
  MC_API_BEGIN
  MC_CHECK(ctx && x_dev && y_dev && nstate >= 0, MC_ERR_ARG, "bad argument");
  const size_t n = (size_t)ctx->ldg * nstate * 2;
  if (n) {
    hipLaunchKernelGGL(mc::k_axpby, dim3(2048), dim3(256), 0, ctx->stream, alpha, (const double*)x_dev, beta,
                       (double*)y_dev, n);
    MC_HIP(hipGetLastError());
  }
  MC_API_END
}

int mc_hesele(mc_ctx* ctx, int prec, double svar2, double vploc, double hthrs) {
  MC_API_BEGIN
  MC_CHECK(ctx, MC_ERR_ARG, "null context");
  MC_CHECK(!prec || hthrs > 0.0, MC_ERR_ARG, "hesele: hthrs must be positive");
  const auto& P = ctx->proj;
  const int nsp = P.ready ? P.nsp : 0, nhxs = P.ready ? P.nhxs : 0;
  // h na per species, the dd of hesele's Goedecker / wsg branches
  std::vector<double> hw((size_t)nhxs * nhxs * nsp + 1, 0.0);
  for (int is = 0; is < nsp; ++is)
    for (int jv = 0; jv < P.ngh[is]; ++jv)
      for (int iv = 0; iv < P.ngh[is]; ++iv) {
        const size_t o = iv + (size_t)nhxs * (jv + (size_t)nhxs * is);
        hw[o] = P.h_h[o] * P.na[is];
      }
  ctx->w_hw.ensure(sizeof(double) * hw.size());
  MC_HIP(hipMemcpyAsync(ctx->w_hw.p, hw.data(), sizeof(double) * hw.size(), hipMemcpyHostToDevice, ctx->stream));
  ctx->d_vpp.ensure(sizeof(double) * ctx->ldg);
  hipLaunchKernelGGL(mc::k_hesele, dim3((ctx->ldg + 255) / 256), dim3(256), 0, ctx->stream, ctx->d_hg.as<double>(),
                     nsp ? P.d_twnl.as<double>() : nullptr, ctx->ldg, ctx->ngw, nhxs, nsp, ctx->w_hw.as<double>(),
                     vploc, 0.5 * ctx->tpiba2, hthrs, prec, svar2, ctx->d_vpp.as<double>());
  MC_HIP(hipGetLastError());
  // hw is a host temporary; vpp(G=0) for odiis
  MC_HIP(hipMemcpyAsync(&ctx->vpp0, ctx->d_vpp.p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  MC_HIP(hipStreamSynchronize(ctx->stream));
  ctx->vpp_ready = true;
  ctx->vpp_prec = prec != 0;
  MC_API_END
}

int mc_csize_dev(mc_ctx* ctx, const void* c2_dev, const double* f, int nstate, double* gemax, double* cnorm) {
  MC_API_BEGIN
  MC_CHECK(ctx && c2_dev && f && gemax && cnorm && nstate >= 0, MC_ERR_ARG, "bad argument");
  int nocc = 0;
  for (int i = 0; i < nstate; ++i)
    if (f[i] > 1e-5) ++nocc;
  *gemax = 0.0;
  *cnorm = 0.0;
  if (nocc == 0) return MC_OK;
  ctx->w_red2.ensure(sizeof(double) * 6 * nocc);
  hipLaunchKernelGGL(mc::k_csize_cols, dim3(nocc), dim3(mc::kCsizeThreads), 0, ctx->stream, (const double2*)c2_dev,
                     ctx->d_perm.as<int>(), ctx->ngw, ctx->ldg, ctx->geq0, ctx->w_red2.as<double>());
  MC_HIP(hipGetLastError());
  std::vector<double> r((size_t)6 * nocc);
  mc::dl_add(ctx, r.data(), ctx->w_red2.p, sizeof(double) * r.size());
  mc::dl_finish(ctx);
  double s = 0.0;
  int b = 0;
  for (int i = 0; i < nocc; ++i) {
    s += r[6 * i];
    // states in order: on equal keys the earlier state has the smaller index
    if (r[6 * i + 1] > r[6 * b + 1]) b = i;
  }
  *gemax = r[6 * b + 3];
  *cnorm = std::sqrt(s / ((double)nocc * ctx->ngw));
  MC_API_END
}

int mc_sd_update_dev(mc_ctx* ctx, void* c0_dev, const void* c2_dev, int nstate, double pf1) {
  MC_API_BEGIN
  MC_CHECK(ctx && c0_dev && c2_dev && nstate >= 0, MC_ERR_ARG, "bad argument");
  MC_CHECK(ctx->vpp_ready, MC_ERR_ARG, "sd_update: call mc_hesele first");
  MC_CHECK(nstate <= 65535, MC_ERR_ARG, "sd_update: nstate > 65535");
  if (nstate == 0) return MC_OK;
  mc::timer_begin(ctx, MC_KID_MISC);
  hipLaunchKernelGGL(mc::k_sd_update, dim3((ctx->ldg + 255) / 256, nstate), dim3(256), 0, ctx->stream,
                     (double2*)c0_dev, (const double2*)c2_dev, ctx->d_vpp.as<double>(), pf1, ctx->ldg);
  MC_HIP(hipGetLastError());
  mc::timer_end(ctx, MC_KID_MISC);
  MC_API_END
}

int mc_odiis_dev(mc_ctx* ctx, void* c0_dev, const void* c2_dev, const double* f, int nstate, double etot,
                 double svar2, int mdiis, int nreset, int* reinit) {
  MC_API_BEGIN
  MC_CHECK(ctx && c0_dev && c2_dev && f && reinit && nstate > 0, MC_ERR_ARG, "bad argument");
  MC_CHECK(mdiis >= 1 && mdiis <= 20, MC_ERR_ARG, "odiis: mdiis must be in [1, 20]");
  MC_CHECK(nstate <= 65535, MC_ERR_ARG, "odiis: nstate > 65535");
  MC_CHECK(ctx->vpp_ready, MC_ERR_ARG, "odiis: call mc_hesele first");
  auto& D = ctx->odiis;
  const size_t slot = (size_t)ctx->ldg * nstate;   // one ring slot (occupied columns used)
  if (*reinit || D.istate == 0 || D.mdiis != mdiis || D.nstate != nstate) {
    D.istate = 0;
    D.eold = 9999.0;
    D.gamma = 0.0;
    *reinit = 0;
    D.nocc = 0;
    for (int i = 0; i < nstate; ++i)
      if (f[i] > 1e-5) ++D.nocc;
    D.ndiis = 0;
    D.mdiis = mdiis;
    D.nstate = nstate;
    D.pme.ensure(sizeof(double2) * slot * mdiis);
    D.gde.ensure(sizeof(double2) * slot * mdiis);
    std::vector<double> ff(nstate, 1.0);
    for (int k = 0; k < nstate; ++k)
      if (ctx->vpp_prec && f[k] > 0.1) ff[k] = 1.0 / f[k];
    D.d_ff.ensure(sizeof(double) * nstate);
    // stream-ordered: kernels of an earlier call may still read d_ff
    mc::upload_async(ctx, D.d_ff.p, ff.data(), sizeof(double) * nstate);
  } else {
    // the damping of the low-G steps grows while the energy rises (:541-551)
    if (ctx->geq0) D.gamma = etot - D.eold > 0.0 ? D.gamma + 1.0 / ctx->vpp0 : 0.75 * D.gamma;
    // insufficient progress over the history: start again (:553-569)
    if (D.ndiis > nreset && nreset >= 3) {
      const int nowm = D.nowv == mdiis ? 1 : D.nowv + 1;
      const double g1 = std::sqrt(D.grmax[D.nowv - 1] * D.grmax[D.nowv - 1] + D.gimax[D.nowv - 1] * D.gimax[D.nowv - 1]);
      const double g2 = std::sqrt(D.grmax[nowm - 1] * D.grmax[nowm - 1] + D.gimax[nowm - 1] * D.gimax[nowm - 1]);
      const double raim = 1.0 - (double)mdiis / 100.0;
      if (g1 / g2 > raim && D.gnorm[D.nowv - 1] / D.gnorm[nowm - 1] > raim) {
        D.istate = 0;
        D.gamma = 0.0;
        *reinit = 1;   // returned to the caller, as odiis returns reinit (:584-593)
      }
    }
  }
  const int nocc = D.nocc;
  if (D.istate == 0) D.ndiis = D.nowv = D.nsize = 0;   // updis(.., 0)
  D.istate = 2;
  D.nowv = D.ndiis % mdiis + 1;                         // updis(.., 1)
  D.ndiis += 1;
  D.nsize = std::min(D.ndiis + 1, mdiis + 1);
  const int w = D.nowv - 1, nv = D.nsize - 1;
  if (nocc > 0) {
    hipLaunchKernelGGL(mc::k_odiis_store, dim3((ctx->ldg + 255) / 256, nocc), dim3(256), 0, ctx->stream,
                       (const double2*)c0_dev, (const double2*)c2_dev, D.pme.as<double2>() + slot * w,
                       D.gde.as<double2>() + slot * w, ctx->ldg);
    MC_HIP(hipGetLastError());
    D.w_part.ensure(sizeof(double) * ((size_t)nv * nocc + 6 * (size_t)nocc));
    // grimax of the new gradient (|-c2| = |c2|): max |Re|, max |Im|, gnorm
    hipLaunchKernelGGL(mc::k_csize_cols, dim3(nocc), dim3(mc::kCsizeThreads), 0, ctx->stream, (const double2*)c2_dev,
                       ctx->d_perm.as<int>(), ctx->ngw, ctx->ldg, ctx->geq0, D.w_part.as<double>() + (size_t)nv * nocc);
    MC_HIP(hipGetLastError());
    hipLaunchKernelGGL(mc::k_odiis_bdot, dim3(nocc, nv), dim3(256), 0, ctx->stream, D.gde.as<double2>(), slot, w,
                       ctx->d_vpp.as<double>(), D.gamma, D.d_ff.as<double>(), ctx->ngw, ctx->ldg, nocc,
                       ctx->geq0, D.w_part.as<double>());
    MC_HIP(hipGetLastError());
    std::vector<double> part((size_t)nv * nocc + 6 * (size_t)nocc);
    mc::dl_add(ctx, part.data(), D.w_part.p, sizeof(double) * part.size());
    mc::dl_finish(ctx);
    const double* gi = part.data() + (size_t)nv * nocc;
    double sn = 0.0, mr = 0.0, mi = 0.0;
    for (int k = 0; k < nocc; ++k) {
      sn += gi[6 * k];
      mr = std::max(mr, gi[6 * k + 4]);
      mi = std::max(mi, gi[6 * k + 5]);
    }
    D.grmax[w] = mr;
    D.gimax[w] = mi == 0.0 ? 2.220446049250313e-16 : mi;
    D.gnorm[w] = std::sqrt(sn / ((double)nocc * ctx->ngw));
    for (int i = 0; i < nv; ++i) {
      double t = 0.0;
      for (int k = 0; k < nocc; ++k) t += part[(size_t)i * nocc + k];
      D.diism[i + 20 * w] = t;
      D.diism[w + 20 * i] = t;
    }
  }
  // the DIIS equations with the Lagrange row (:612-631), solved like solve/dgelss
  const int ns = D.nsize;
  std::vector<double> bc((size_t)ns * ns, 0.0), vc(ns, 0.0);
  for (int i = 0; i < nv; ++i)
    for (int j = 0; j < nv; ++j) bc[(size_t)i * ns + j] = D.diism[i + 20 * j];
  for (int i = 0; i < nv; ++i) {
    bc[(size_t)i * ns + nv] = -1.0;
    bc[(size_t)nv * ns + i] = -1.0;
  }
  vc[nv] = -1.0;
  mc::sym_lstsq(bc, ns, vc, 0.2e-15);
  mc::DiisCoef dc{};
  dc.nv = nv;
  for (int i = 0; i < nv; ++i) dc.vc[i] = vc[i];
  mc::timer_begin(ctx, MC_KID_MISC);
  hipLaunchKernelGGL(mc::k_odiis_combine, dim3((ctx->ldg + 255) / 256, nstate), dim3(256), 0, ctx->stream,
                     (double2*)c0_dev, (const double2*)c2_dev, D.pme.as<double2>(), D.gde.as<double2>(), slot,
                     ctx->d_vpp.as<double>(), D.gamma, D.d_ff.as<double>(), dc, ctx->ldg, nocc, svar2);
  MC_HIP(hipGetLastError());
  mc::timer_end(ctx, MC_KID_MISC);
  D.eold = etot;
  MC_API_END
}

int mc_odiis_release(mc_ctx* ctx) {
  MC_API_BEGIN
  MC_CHECK(ctx, MC_ERR_ARG, "null context");
  MC_HIP(hipStreamSynchronize(ctx->stream));
  auto& D = ctx->odiis;
  D.pme.release();
  D.gde.release();
  D.w_part.release();
  D.istate = 0;
  D.mdiis = 0;
  MC_API_END
}

int mc_wfn_dotp(mc_ctx* ctx, const void* a_dev, const void* b_dev, int nstate, double* out) {
  MC_API_BEGIN
  MC_CHECK(ctx && a_dev && b_dev && out && nstate >= 0, MC_ERR_ARG, "bad argument");
  if (nstate == 0) return MC_OK;
  ctx->w_red2.ensure(sizeof(double) * nstate);
  hipLaunchKernelGGL(mc::k_dotp_cols, dim3(nstate), dim3(256), 0, ctx->stream, (const double2*)a_dev,
                     (const double2*)b_dev, ctx->ngw, ctx->ldg, ctx->geq0, ctx->w_red2.as<double>());
  MC_HIP(hipGetLastError());
  MC_HIP(hipMemcpyAsync(out, ctx->w_red2.p, sizeof(double) * nstate, hipMemcpyDeviceToHost, ctx->stream));
  MC_HIP(hipStreamSynchronize(ctx->stream));
  MC_API_END
}

}  // extern "C"

namespace mc {
// rekine (src/rekine_utils.mod.F90:34-49, no tmass): ekinc = emass sum_i
// dotp(cm_i, cm_i); the per-state sums are queued for download (dl_finish
// fills out[0..n) on the host)
static void rekine_queue(mc_ctx* ctx, const double2* cm, int n, double* out) {
  ctx->w_red2.ensure(sizeof(double) * n);
  hipLaunchKernelGGL(k_dotp_cols, dim3(n), dim3(256), 0, ctx->stream, cm, cm, ctx->ngw, ctx->ldg, ctx->geq0,
                     ctx->w_red2.as<double>());
  MC_HIP(hipGetLastError());
  dl_add(ctx, out, ctx->w_red2.p, sizeof(double) * n);
}
static double rekine_sum(const std::vector<double>& d, double emass) {
  double s = 0.0;
  for (double v : d) s += v;
  return s * emass;
}
}  // namespace mc

extern "C" {

int mc_rekine_dev(mc_ctx* ctx, const void* cm_dev, int nstate, double emass, double* ekinc) {
  MC_API_BEGIN
  MC_CHECK(ctx && cm_dev && ekinc && nstate >= 0, MC_ERR_ARG, "rekine: bad argument");
  *ekinc = 0.0;
  if (nstate == 0) return MC_OK;
  std::vector<double> d(nstate);
  mc::rekine_queue(ctx, (const double2*)cm_dev, nstate, d.data());
  mc::dl_finish(ctx);
  *ekinc = mc::rekine_sum(d, emass);
  MC_API_END
}

int mc_rortv_dev(mc_ctx* ctx, const void* c0_dev, void* cm_dev, double* gamy, int nstate) {
  MC_API_BEGIN
  MC_CHECK(ctx && c0_dev && cm_dev && nstate >= 0, MC_ERR_ARG, "rortv: bad argument");
  MC_CHECK(c0_dev != cm_dev, MC_ERR_ARG, "rortv: c0 and cm must not alias");
  mc::rortv_dev(ctx, (const double2*)c0_dev, (double2*)cm_dev, gamy, nstate);
  mc::dl_finish(ctx);
  MC_API_END
}

int mc_cp_velupa_rortv_dev(mc_ctx* ctx, const void* c0_dev, const void* c2_dev, void* cm_dev, int nstate,
                           double dtm, double emass, double* gamy, double* ekinc) {
  MC_API_BEGIN
  MC_CHECK(ctx && c0_dev && cm_dev && nstate >= 0, MC_ERR_ARG, "velupa/rortv: bad argument");
  MC_CHECK(c0_dev != cm_dev && c2_dev != cm_dev, MC_ERR_ARG, "velupa/rortv: cm must not alias c0 or c2");
  if (ekinc) *ekinc = 0.0;
  if (nstate == 0) return MC_OK;
  // velupa(c0, cm, c2, nstate, 1): cm += dt / (2 emass) c2, then rortv; in
  // mixed precision the rotation's epilogue gives rekine's norms
  std::vector<double> d;
  if (ekinc) {
    d.resize(nstate);
    ctx->w_red2.ensure(sizeof(double) * nstate);
  }
  const bool fused = mc::rortv_dev(ctx, (const double2*)c0_dev, (double2*)cm_dev, gamy, nstate,
                                   (const double2*)c2_dev, dtm, ekinc ? ctx->w_red2.as<double>() : nullptr);
  if (ekinc) {
    if (fused)
      mc::dl_add(ctx, d.data(), ctx->w_red2.p, sizeof(double) * nstate);
    else
      mc::rekine_queue(ctx, (const double2*)cm_dev, nstate, d.data());
  }
  mc::dl_finish(ctx);
  if (ekinc) *ekinc = mc::rekine_sum(d, emass);
  MC_API_END
}

int mc_wfn_upload(mc_ctx* ctx, const double* c_host, int ldc, int nstate, void* c_dev) {
  MC_API_BEGIN
  MC_CHECK(ctx && c_host && c_dev, MC_ERR_ARG, "null argument");
  mc::wfn_upload(ctx, c_host, ldc, nstate, (double2*)c_dev);
  MC_HIP(hipStreamSynchronize(ctx->stream));
  MC_API_END
}

int mc_wfn_download(mc_ctx* ctx, const void* c_dev, int nstate, double* c_host, int ldc) {
  MC_API_BEGIN_RO
  MC_CHECK(ctx && c_host && c_dev, MC_ERR_ARG, "null argument");
  mc::wfn_download(ctx, (const double2*)c_dev, nstate, c_host, ldc);
  MC_API_END
}

int mc_rhoofr_dev(mc_ctx* ctx, const void* c0_dev, const double* f, int nstate, void* rhoe_dev,
                  double* ekin, double* csumg, double* csumr) {
  MC_API_BEGIN_RO
  MC_CHECK(ctx && c0_dev && f && rhoe_dev, MC_ERR_ARG, "null argument");
  MC_CHECK(nstate >= 0, MC_ERR_ARG, "nstate < 0");
  mc::rhoofr_dev(ctx, (const double2*)c0_dev, f, nstate, (double*)rhoe_dev, ekin, csumg, csumr, true);
  MC_API_END
}

int mc_rhoofr(mc_ctx* ctx, const double* c0, int ldc, const double* f, int nstate, double* rhoe,
              double* ekin, double* csumg, double* csumr) {
  MC_API_BEGIN
  MC_CHECK(ctx && c0 && f && rhoe, MC_ERR_ARG, "null argument");
  MC_CHECK(nstate >= 0, MC_ERR_ARG, "nstate < 0");
  const size_t nnr1 = (size_t)ctx->kr1 * ctx->kr2 * ctx->kr3;
  mc::DevBuf cdev;
  cdev.ensure(sizeof(double2) * (size_t)ctx->ldg * std::max(nstate, 1));
  mc::wfn_upload(ctx, c0, ldc, nstate, cdev.as<double2>());
  ctx->w_v.ensure(sizeof(double) * nnr1);
  mc::rhoofr_dev(ctx, cdev.as<double2>(), f, nstate, ctx->w_v.as<double>(), ekin, csumg, csumr);
  MC_HIP(hipMemcpyAsync(rhoe, ctx->w_v.p, sizeof(double) * nnr1, hipMemcpyDeviceToHost, ctx->stream));
  MC_HIP(hipStreamSynchronize(ctx->stream));
  MC_API_END
}

int mc_vpsi_dev(mc_ctx* ctx, const void* c0_dev, void* c2_dev, const double* f, const void* vpot_dev,
                int nstate) {
  MC_API_BEGIN_RO
  MC_CHECK(ctx && c0_dev && c2_dev && f && vpot_dev, MC_ERR_ARG, "null argument");
  MC_CHECK(nstate >= 0, MC_ERR_ARG, "nstate < 0");
  mc::vpsi_dev(ctx, (const double2*)c0_dev, (double2*)c2_dev, f, (const double*)vpot_dev, nstate, true, true);
  MC_API_END
}

int mc_rhoofr_raw_charge(mc_ctx* ctx, double* csumg, double* csumr_raw) {
  MC_API_BEGIN_RO
  MC_CHECK(ctx, MC_ERR_ARG, "null context");
  if (csumg) *csumg = ctx->last_csumg;
  if (csumr_raw) *csumr_raw = ctx->last_csumr_raw;
  MC_API_END
}

int mc_vpsi_dev_noalloc(mc_ctx* ctx, const void* c0_dev, void* c2_dev, const double* f, const void* vpot_dev,
                        int nstate) {
  MC_API_BEGIN
  MC_CHECK(ctx && c0_dev && c2_dev && f && vpot_dev, MC_ERR_ARG, "null argument");
  MC_CHECK(nstate >= 0, MC_ERR_ARG, "nstate < 0");
  MC_HIP(hipSetDevice(ctx->device));
  mc::vpsi_noalloc_dev(ctx, (const double2*)c0_dev, (double2*)c2_dev, f, (const double*)vpot_dev, nstate);
  MC_API_END
}

int mc_forces_c2_dev(mc_ctx* ctx, const void* c0_dev, void* c2_dev, const double* f, const void* vpot_dev,
                     int nstate) {
  MC_API_BEGIN_RO
  MC_CHECK(ctx && c0_dev && c2_dev && f && vpot_dev, MC_ERR_ARG, "null argument");
  MC_CHECK(nstate >= 0, MC_ERR_ARG, "nstate < 0");
  mc::forces_c2_dev(ctx, (const double2*)c0_dev, (double2*)c2_dev, f, (const double*)vpot_dev, nstate);
  MC_API_END
}

int mc_vpsi(mc_ctx* ctx, const double* c0, int ldc, double* c2, const double* f, const double* vpot,
            int nstate) {
  MC_API_BEGIN
  MC_CHECK(ctx && c0 && c2 && f && vpot, MC_ERR_ARG, "null argument");
  MC_CHECK(nstate >= 0, MC_ERR_ARG, "nstate < 0");
  const size_t nnr1 = (size_t)ctx->kr1 * ctx->kr2 * ctx->kr3;
  mc::DevBuf c0d, c2d;
  c0d.ensure(sizeof(double2) * (size_t)ctx->ldg * std::max(nstate, 1));
  c2d.ensure(sizeof(double2) * (size_t)ctx->ldg * std::max(nstate, 1));
  mc::wfn_upload(ctx, c0, ldc, nstate, c0d.as<double2>());
  mc::wfn_upload(ctx, c2, ldc, nstate, c2d.as<double2>());
  ctx->w_v.ensure(sizeof(double) * nnr1);
  MC_HIP(hipMemcpyAsync(ctx->w_v.p, vpot, sizeof(double) * nnr1, hipMemcpyHostToDevice, ctx->stream));
  mc::vpsi_dev(ctx, c0d.as<double2>(), c2d.as<double2>(), f, ctx->w_v.as<double>(), nstate);
  mc::wfn_download(ctx, c2d.as<double2>(), nstate, c2, ldc);
  MC_API_END
}

int mc_timer_enable(mc_ctx* ctx, int on) {
  MC_API_BEGIN_RO
  MC_CHECK(ctx, MC_ERR_ARG, "null context");
  mc::timer_flush(ctx);
  ctx->timer.enabled = on != 0;
  MC_API_END
}
int mc_timer_select(mc_ctx* ctx, unsigned long long kid_mask) {
  MC_API_BEGIN_RO
  MC_CHECK(ctx, MC_ERR_ARG, "null context");
  mc::timer_flush(ctx);
  ctx->timer.mask = kid_mask;
  MC_API_END
}
int mc_timer_reset(mc_ctx* ctx) {
  MC_API_BEGIN_RO
  MC_CHECK(ctx, MC_ERR_ARG, "null context");
  mc::timer_flush(ctx);
  for (int k = 0; k < MC_KID_COUNT; ++k) { ctx->timer.total_ms[k] = 0; ctx->timer.count[k] = 0; }
  MC_API_END
}
int mc_timer_get(mc_ctx* ctx, int kid, double* total_ms, long long* count) {
  MC_API_BEGIN_RO
  MC_CHECK(ctx && kid >= 0 && kid < MC_KID_COUNT, MC_ERR_ARG, "bad kernel id");
  mc::timer_flush(ctx);
  if (total_ms) *total_ms = ctx->timer.total_ms[kid];
  if (count) *count = ctx->timer.count[kid];
  MC_API_END
}

}  // extern "C"
