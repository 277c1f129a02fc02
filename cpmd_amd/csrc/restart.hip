// CPMD RESTART files (format "STUTTGART VERSION 3.0", sequential
// unformatted): writer and reader of the sections that carry the state of a
// Gamma-point run -- cell (2), species (3), coordinates (4), ionic
// velocities (5), initial coordinates (6), cutoff and mesh (7), states and
// plane waves (8), the coefficients C0 (9) and their velocities CM (10) and
// the occupations (23).  Every other section of sections 2..99 is written as
// an empty marker and skipped when read.
//
// Record framing: each Fortran WRITE is one record, a 4-byte little-endian
// length before and after the payload (the sequential-access layout of
// gfortran and flang).  A section starts with the marker record
// (irecord int32, fpos int64) followed by irecord records (|irecord| for
// the wavefunction layout of section 9).  Host-only code: no kernels.
//
// Reference: zhwwf / wv30 src/wv30_utils.mod.F90:99-757 (sections 1-10 at
// :304-535, 23 at :854-871, 39-99 at :1220-1266), w_wfnio
// src/wfnio_utils.mod.F90:52-233, wr30wfn / wrwfn
// src/wr30wfn_utils.mod.F90:460-693, compress src/compress.F90:2-89, rv30
// src/rv30_utils.mod.F90:286-735, r_wfnio src/wfnio_utils.mod.F90:235-330.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mc_internal.h"

namespace mc {
namespace {

constexpr size_t kMaxRecord = 0x7fffffffu - 16;   // single-record payload limit of the 4-byte framing

class RecWriter {
 public:
  explicit RecWriter(const char* path) : f_(std::fopen(path, "wb")), path_(path) {
    MC_CHECK(f_, MC_ERR_ARG, std::string("restart: cannot open ") + path + " for writing");
  }
  ~RecWriter() {
    if (f_) std::fclose(f_);
  }
  // one Fortran record: the bytes collected in b
  void rec(const std::vector<unsigned char>& b) { raw(b.data(), b.size()); }
  void raw(const void* p, size_t n) {
    MC_CHECK(n <= kMaxRecord, MC_ERR_UNSUPPORTED, "restart: record beyond 2 GiB");
    const int32_t m = (int32_t)n;
    put(&m, 4);
    put(p, n);
    put(&m, 4);
  }
  void close() {
    FILE* f = f_;
    f_ = nullptr;
    MC_CHECK(f && std::fclose(f) == 0, MC_ERR_ARG, std::string("restart: error closing ") + path_);
  }

 private:
  void put(const void* p, size_t n) {
    MC_CHECK(std::fwrite(p, 1, n, f_) == n, MC_ERR_ARG, std::string("restart: write error on ") + path_);
  }
  FILE* f_;
  std::string path_;
};

// payload builder: items appended in the order of a Fortran I/O list
struct Rec {
  std::vector<unsigned char> b;
  Rec& i4(int32_t v) { return add(&v, 4); }
  Rec& i8(int64_t v) { return add(&v, 8); }
  Rec& r8(double v) { return add(&v, 8); }
  Rec& l4(bool v) { return i4(v ? 1 : 0); }   // LOGICAL(4): .TRUE. = 1
  Rec& add(const void* p, size_t n) {
    const auto* c = static_cast<const unsigned char*>(p);
    b.insert(b.end(), c, c + n);
    return *this;
  }
};

void marker(RecWriter& w, int32_t irecord, int64_t fpos = 0) { w.rec(Rec().i4(irecord).i8(fpos).b); }

// BLAS izamax over n complex values: the first index of the largest |Re| + |Im|
size_t izamax(const double* c, size_t n) {
  size_t best = 0;
  double bv = -1.0;
  for (size_t k = 0; k < n; ++k) {
    const double v = std::fabs(c[2 * k]) + std::fabs(c[2 * k + 1]);
    if (v > bv) { bv = v; best = k; }
  }
  return best;
}

// per-state records of wrwfn (icompr = 1: nx = 2 ngw reals, icmp = their 8-byte words
// as 2 nx int32): (nx), (icmp(1:2 nx))
void write_states(RecWriter& w, const double* c, int ldc, int ngw, int nstate) {
  for (int i = 0; i < nstate; ++i) {
    w.rec(Rec().i4(2 * ngw).b);
    w.raw(c + 2 * (size_t)ldc * i, sizeof(double) * 2 * (size_t)ngw);
  }
}

// section 9 through w_wfnio (tag 'C0', iwfntype 1, icompr 1): marker
// (-(2 nstate + 3), pos = -1), (iwfntype), (nstate, ngwks, icompr, scale),
// (tlsd, nsup, nsdown), the states; scale = |c| of the izamax element
void write_c0(RecWriter& w, const double* c, int ldc, int ngw, int nstate) {
  std::vector<double> packed;
  const double* src = c;
  if (ldc != ngw) {   // izamax runs over the contiguous (ngwk, nstate) array
    packed.resize(2 * (size_t)ngw * nstate);
    for (int i = 0; i < nstate; ++i)
      std::memcpy(&packed[2 * (size_t)ngw * i], c + 2 * (size_t)ldc * i, sizeof(double) * 2 * ngw);
    src = packed.data();
  }
  const size_t k = izamax(src, (size_t)ngw * nstate);
  const double scale = nstate > 0 && ngw > 0 ? std::hypot(src[2 * k], src[2 * k + 1]) : 0.0;
  marker(w, -(2 * nstate + 3), -1);
  w.rec(Rec().i4(1).b);
  w.rec(Rec().i4(nstate).i4(ngw).i4(1).r8(scale).b);
  w.rec(Rec().l4(false).i4(0).i4(0).b);
  write_states(w, c, ldc, ngw, nstate);
}

// section 10 through wr30wfn (tag 'NIL', icompr 1): marker (2 nstate + 1, 0),
// (nstate, ngwks, icompr, scale), the states; scale = |x| of the idamax element
// of the real view
void write_cm(RecWriter& w, const double* c, int ldc, int ngw, int nstate) {
  double scale = 0.0;
  for (int i = 0; i < nstate; ++i)
    for (size_t e = 0; e < 2 * (size_t)ngw; ++e) scale = std::max(scale, std::fabs(c[2 * (size_t)ldc * i + e]));
  marker(w, 2 * nstate + 1, 0);
  w.rec(Rec().i4(nstate).i4(ngw).i4(1).r8(scale).b);
  write_states(w, c, ldc, ngw, nstate);
}

void write_positions(RecWriter& w, const double* t, int nat) {
  marker(w, nat);
  for (int a = 0; a < nat; ++a) w.rec(Rec().r8(t[3 * a]).r8(t[3 * a + 1]).r8(t[3 * a + 2]).b);
}

class RecReader {
 public:
  explicit RecReader(const char* path) : f_(std::fopen(path, "rb")), path_(path) {
    MC_CHECK(f_, MC_ERR_ARG, std::string("restart: cannot open ") + path);
  }
  ~RecReader() {
    if (f_) std::fclose(f_);
  }
  // the next record's payload
  std::vector<unsigned char> rec() {
    const int32_t n = len();
    std::vector<unsigned char> b((size_t)n);
    MC_CHECK(n == 0 || std::fread(b.data(), 1, (size_t)n, f_) == (size_t)n, MC_ERR_ARG, "restart: truncated record");
    MC_CHECK(len() == n, MC_ERR_ARG, "restart: record length markers differ (not a sequential unformatted file)");
    return b;
  }
  // the next record's payload into dst (exactly n bytes expected)
  void rec_into(void* dst, size_t n) {
    const int32_t m = len();
    MC_CHECK((size_t)m == n, MC_ERR_ARG, "restart: unexpected record length");
    MC_CHECK(std::fread(dst, 1, n, f_) == n, MC_ERR_ARG, "restart: truncated record");
    MC_CHECK(len() == m, MC_ERR_ARG, "restart: record length markers differ");
  }
  void skip(long long nrec) {
    for (long long k = 0; k < nrec; ++k) {
      const int32_t n = len();
      MC_CHECK(std::fseek(f_, n, SEEK_CUR) == 0, MC_ERR_ARG, "restart: seek error");
      MC_CHECK(len() == n, MC_ERR_ARG, "restart: record length markers differ");
    }
  }

 private:
  int32_t len() {
    int32_t n = 0;
    MC_CHECK(std::fread(&n, 1, 4, f_) == 4, MC_ERR_ARG, std::string("restart: end of file in ") + path_);
    MC_CHECK(n >= 0, MC_ERR_UNSUPPORTED, "restart: continued (subdivided) records are not supported");
    return n;
  }
  FILE* f_;
  std::string path_;
};

template <class T>
T get(const std::vector<unsigned char>& b, size_t off) {
  MC_CHECK(off + sizeof(T) <= b.size(), MC_ERR_ARG, "restart: record shorter than its I/O list");
  T v;
  std::memcpy(&v, b.data() + off, sizeof(T));
  return v;
}

int32_t read_marker(RecReader& r) { return get<int32_t>(r.rec(), 0); }

// wavefunction section (9 or 10) after its marker irecord: fills c (ldc x
// nstate, may be null) and returns the number of states in the file
int read_wfn_section(RecReader& r, int32_t irecord, double* c, int ldc, int nstate_max, int* ngw_out) {
  if (irecord == 0) return 0;
  if (irecord < 0) {   // w_wfnio layout: (iwfntype) first
    const int32_t t = get<int32_t>(r.rec(), 0);
    MC_CHECK(t == 1, MC_ERR_UNSUPPORTED, "restart: wavefunction type " + std::to_string(t) + " (only 1)");
  }
  const auto hb = r.rec();
  const int32_t n0 = get<int32_t>(hb, 0), ngw0 = get<int32_t>(hb, 4), icompr = get<int32_t>(hb, 8);
  MC_CHECK(icompr == 1, MC_ERR_UNSUPPORTED,
           "restart: compressed coefficients (WCOMPB " + std::to_string(icompr) + "); only uncompressed records");
  if (irecord < 0) r.skip(1);   // (tlsd, nsup, nsdown)
  if (ngw_out) *ngw_out = ngw0;
  for (int i = 0; i < n0; ++i) {
    const int32_t nx = get<int32_t>(r.rec(), 0);
    MC_CHECK(nx == 2 * ngw0, MC_ERR_ARG, "restart: state record of unexpected size");
    if (c && i < nstate_max) {
      MC_CHECK(ngw0 <= ldc, MC_ERR_ARG, "restart: ldc smaller than the file's plane-wave count");
      r.rec_into(c + 2 * (size_t)ldc * i, sizeof(double) * 2 * (size_t)ngw0);
    } else {
      r.skip(1);
    }
  }
  return n0;
}

struct ReadOut {
  int* na = nullptr;
  double *tau0 = nullptr, *velp = nullptr, *taui = nullptr, *f = nullptr, *c0 = nullptr, *cm = nullptr;
  int ldc = 0;
};

// reads the file; with `out` null only the header and the section flags
void read_restart(const char* path, mc_restart_header* h, const ReadOut* out) {
  RecReader r(path);
  const auto s1 = r.rec();
  const std::string str(s1.begin(), s1.end());
  MC_CHECK(str.find("STUTTGART VERSION 3.0") != std::string::npos, MC_ERR_ARG,
           std::string("restart: ") + path + " is not a CPMD RESTART file (header)");
  MC_CHECK(str.find("STREAM") == std::string::npos, MC_ERR_UNSUPPORTED, "restart: STREAM files are not supported");
  mc_restart_header H{};
  int32_t irec = 0;
  // 2: cell
  irec = read_marker(r);
  if (irec == 2) {
    const auto a = r.rec();
    H.ibrav = get<int32_t>(a, 0);
    H.indpg = get<int32_t>(a, 4);
    const auto b = r.rec();
    for (int k = 0; k < 6; ++k) H.celldm[k] = get<double>(b, 8 * k);
  } else {
    r.skip(std::abs(irec));
  }
  // 3: species
  irec = read_marker(r);
  std::vector<int32_t> na;
  if (irec == 2) {
    H.nsp = get<int32_t>(r.rec(), 0);
    const auto b = r.rec();
    na.resize(H.nsp);
    for (int s = 0; s < H.nsp; ++s) na[s] = get<int32_t>(b, 4 * s);
    for (int v : na) H.nat += v;
  } else {
    r.skip(std::abs(irec));
  }
  // 4, 5, 6: coordinates, velocities, initial coordinates
  double* dst[3] = {out ? out->tau0 : nullptr, out ? out->velp : nullptr, out ? out->taui : nullptr};
  const int flag[3] = {0, MC_RST_VELP, MC_RST_TAUI};
  for (int s = 0; s < 3; ++s) {
    irec = read_marker(r);
    if (irec > 0) {
      MC_CHECK(H.nat == 0 || irec == H.nat, MC_ERR_ARG, "restart: coordinate count differs from the species");
      H.sections |= flag[s];
      for (int a = 0; a < irec; ++a) {
        const auto b = r.rec();
        if (dst[s])
          for (int q = 0; q < 3; ++q) dst[s][3 * a + q] = get<double>(b, 8 * q);
      }
    } else {
      r.skip(std::abs(irec));
    }
  }
  // 7: cutoff, electrons, mesh
  irec = read_marker(r);
  if (irec == 1) {
    const auto b = r.rec();
    H.ecut = get<double>(b, 0);
    H.cdual = get<double>(b, 8);
    H.dual = get<int32_t>(b, 16) != 0;
    H.nel = get<int32_t>(b, 20);
    H.nr1s = get<int32_t>(b, 24);
    H.nr2s = get<int32_t>(b, 28);
    H.nr3s = get<int32_t>(b, 32);
  } else {
    r.skip(std::abs(irec));
  }
  // 8: states, k points, plane waves
  irec = read_marker(r);
  if (irec == 1) {
    const auto b = r.rec();
    H.nstate = get<int32_t>(b, 0);
    MC_CHECK(get<int32_t>(b, 4) == 0, MC_ERR_UNSUPPORTED, "restart: k-point file (only Gamma)");
    H.ngwks = get<int32_t>(b, 8);
    H.ngwls = get<int32_t>(b, 12);
    H.nhgs = get<int32_t>(b, 16);
    H.nhgls = get<int32_t>(b, 20);
  } else {
    r.skip(std::abs(irec));
  }
  // 9, 10: C0, CM
  for (int s = 0; s < 2; ++s) {
    irec = read_marker(r);
    double* c = out ? (s == 0 ? out->c0 : out->cm) : nullptr;
    int ngw0 = 0;
    const int n0 = read_wfn_section(r, irec, c, out ? out->ldc : 0, H.nstate, &ngw0);
    if (irec != 0) {
      H.sections |= s == 0 ? MC_RST_C0 : MC_RST_CM;
      MC_CHECK(n0 == H.nstate && ngw0 == H.ngwks, MC_ERR_ARG,
               "restart: wavefunction section sizes differ from section 8");
    }
  }
  // 11..22 skipped, 23 occupations
  for (int sec = 11; sec <= 22; ++sec) r.skip(std::abs(read_marker(r)));
  irec = read_marker(r);
  if (irec >= 1) {
    const auto b = r.rec();
    MC_CHECK(b.size() >= sizeof(double) * (size_t)H.nstate, MC_ERR_ARG, "restart: occupation record too short");
    if (out && out->f)
      for (int i = 0; i < H.nstate; ++i) out->f[i] = get<double>(b, 8 * (size_t)i);
    H.sections |= MC_RST_OCC;
    r.skip(irec - 1);
  }
  if (out && out->na)
    for (int s = 0; s < H.nsp; ++s) out->na[s] = na[s];
  *h = H;
}

void write_restart(const char* path, const mc_restart_header* h, const int* na, const double* tau0,
                   const double* velp, const double* taui, const double* f, const double* c0, const double* cm,
                   int ldc) {
  MC_CHECK(path && h && na && tau0, MC_ERR_ARG, "restart: null argument");
  MC_CHECK(h->nsp >= 1 && h->nstate >= 0 && h->ngwks >= 1, MC_ERR_ARG, "restart: bad sizes");
  int nat = 0;
  for (int s = 0; s < h->nsp; ++s) {
    MC_CHECK(na[s] >= 0, MC_ERR_ARG, "restart: negative atom count");
    nat += na[s];
  }
  MC_CHECK(h->nat == 0 || h->nat == nat, MC_ERR_ARG, "restart: nat differs from sum(na)");
  const int sec = h->sections;
  MC_CHECK(!(sec & MC_RST_VELP) || velp, MC_ERR_ARG, "restart: velocities flagged but null");
  MC_CHECK(!(sec & MC_RST_TAUI) || taui, MC_ERR_ARG, "restart: initial coordinates flagged but null");
  MC_CHECK(!(sec & MC_RST_C0) || c0, MC_ERR_ARG, "restart: C0 flagged but null");
  MC_CHECK(!(sec & MC_RST_CM) || cm, MC_ERR_ARG, "restart: CM flagged but null");
  MC_CHECK(!(sec & MC_RST_OCC) || f, MC_ERR_ARG, "restart: occupations flagged but null");
  MC_CHECK(!(sec & (MC_RST_C0 | MC_RST_CM)) || ldc >= h->ngwks, MC_ERR_ARG, "restart: ldc < ngwks");
  RecWriter w(path);
  // 1: header, CHARACTER(len=80)
  std::string str = " STUTTGART VERSION 3.0";
  str.resize(80, ' ');
  w.raw(str.data(), str.size());
  // 2: ibrav, indpg; celldm(1:6)
  marker(w, 2);
  w.rec(Rec().i4(h->ibrav).i4(h->indpg).b);
  {
    Rec r;
    for (int k = 0; k < 6; ++k) r.r8(h->celldm[k]);
    w.rec(r.b);
  }
  // 3: nsp; na(1:nsp)
  marker(w, 2);
  w.rec(Rec().i4(h->nsp).b);
  {
    Rec r;
    for (int s = 0; s < h->nsp; ++s) r.i4(na[s]);
    w.rec(r.b);
  }
  // 4, 5, 6
  write_positions(w, tau0, nat);
  if (sec & MC_RST_VELP) write_positions(w, velp, nat); else marker(w, 0);
  if (sec & MC_RST_TAUI) write_positions(w, taui, nat); else marker(w, 0);
  // 7: ecut, cdual, dual, nel, nr1s, nr2s, nr3s
  marker(w, 1);
  w.rec(Rec().r8(h->ecut).r8(h->cdual).l4(h->dual != 0).i4(h->nel).i4(h->nr1s).i4(h->nr2s).i4(h->nr3s).b);
  // 8: n, nkpts0 (0: Gamma), ngwks, ngwls, nhgs, nhgls
  marker(w, 1);
  w.rec(Rec().i4(h->nstate).i4(0).i4(h->ngwks).i4(h->ngwls).i4(h->nhgs).i4(h->nhgls).b);
  // 9, 10
  if (sec & MC_RST_C0) write_c0(w, c0, ldc, h->ngwks, h->nstate); else marker(w, 0, -1);
  if (sec & MC_RST_CM) write_cm(w, cm, ldc, h->ngwks, h->nstate); else marker(w, 0);
  // 11..22 empty
  for (int s = 11; s <= 22; ++s) marker(w, 0);
  // 23: occupations of the one (Gamma) k point
  if (sec & MC_RST_OCC) {
    marker(w, 1);
    Rec r;
    for (int i = 0; i < h->nstate; ++i) r.r8(f[i]);
    w.rec(r.b);
  } else {
    marker(w, 0);
  }
  // 24..99 empty
  for (int s = 24; s <= 99; ++s) marker(w, 0);
  w.close();
}

}  // namespace
}  // namespace mc

extern "C" {

int mc_restart_write(const char* path, const mc_restart_header* h, const int* na, const double* tau0,
                     const double* velp, const double* taui, const double* f, const double* c0, const double* cm,
                     int ldc) {
  MC_API_BEGIN
  mc::write_restart(path, h, na, tau0, velp, taui, f, c0, cm, ldc);
  MC_API_END
}

int mc_restart_read_header(const char* path, mc_restart_header* h) {
  MC_API_BEGIN
  MC_CHECK(path && h, MC_ERR_ARG, "restart: null argument");
  mc::read_restart(path, h, nullptr);
  MC_API_END
}

int mc_restart_read(const char* path, mc_restart_header* h, int* na, double* tau0, double* velp, double* taui,
                    double* f, double* c0, double* cm, int ldc) {
  MC_API_BEGIN
  MC_CHECK(path && h, MC_ERR_ARG, "restart: null argument");
  mc::ReadOut o;
  o.na = na;
  o.tau0 = tau0;
  o.velp = velp;
  o.taui = taui;
  o.f = f;
  o.c0 = c0;
  o.cm = cm;
  o.ldc = ldc;
  mc::read_restart(path, h, &o);
  MC_API_END
}

int mc_restart_write_dev(mc_ctx* ctx, const char* path, const mc_restart_header* h, const int* na,
                         const double* tau0, const double* velp, const double* taui, const double* f,
                         const void* c0_dev, const void* cm_dev) {
  MC_API_BEGIN
  MC_CHECK(ctx && h, MC_ERR_ARG, "restart: null argument");
  MC_CHECK(h->ngwks == ctx->ngw, MC_ERR_ARG, "restart: ngwks differs from the context's plane-wave count");
  const int n = h->nstate;
  std::vector<double> c0h, cmh;
  if ((h->sections & MC_RST_C0) && n > 0) {
    MC_CHECK(c0_dev, MC_ERR_ARG, "restart: C0 flagged but null");
    c0h.resize(2 * (size_t)ctx->ngw * n);
    mc::wfn_download(ctx, (const double2*)c0_dev, n, c0h.data(), ctx->ngw);
  }
  if ((h->sections & MC_RST_CM) && n > 0) {
    MC_CHECK(cm_dev, MC_ERR_ARG, "restart: CM flagged but null");
    cmh.resize(2 * (size_t)ctx->ngw * n);
    mc::wfn_download(ctx, (const double2*)cm_dev, n, cmh.data(), ctx->ngw);
  }
  MC_HIP(hipStreamSynchronize(ctx->stream));
  mc::write_restart(path, h, na, tau0, velp, taui, f, c0h.empty() ? nullptr : c0h.data(),
                    cmh.empty() ? nullptr : cmh.data(), ctx->ngw);
  MC_API_END
}

int mc_restart_read_dev(mc_ctx* ctx, const char* path, mc_restart_header* h, int* na, double* tau0, double* velp,
                        double* taui, double* f, void* c0_dev, void* cm_dev) {
  MC_API_BEGIN
  MC_CHECK(ctx && path && h, MC_ERR_ARG, "restart: null argument");
  mc_restart_header H{};
  mc::read_restart(path, &H, nullptr);
  MC_CHECK(H.ngwks == ctx->ngw, MC_ERR_ARG, "restart: the file's plane-wave count differs from the context's");
  std::vector<double> c0h, cmh;
  if (c0_dev && (H.sections & MC_RST_C0)) c0h.resize(2 * (size_t)ctx->ngw * std::max(H.nstate, 1));
  if (cm_dev && (H.sections & MC_RST_CM)) cmh.resize(2 * (size_t)ctx->ngw * std::max(H.nstate, 1));
  mc::ReadOut o;
  o.na = na;
  o.tau0 = tau0;
  o.velp = velp;
  o.taui = taui;
  o.f = f;
  o.c0 = c0h.empty() ? nullptr : c0h.data();
  o.cm = cmh.empty() ? nullptr : cmh.data();
  o.ldc = ctx->ngw;
  mc::read_restart(path, h, &o);
  if (!c0h.empty()) mc::wfn_upload(ctx, c0h.data(), ctx->ngw, H.nstate, (double2*)c0_dev);
  if (!cmh.empty()) mc::wfn_upload(ctx, cmh.data(), ctx->ngw, H.nstate, (double2*)cm_dev);
  MC_HIP(hipStreamSynchronize(ctx->stream));
  MC_API_END
}

}  // extern "C"
