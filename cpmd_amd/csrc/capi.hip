// C ABI of the contraction / orthonormalisation / projector entry points.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mc_internal.h"

namespace {

// device copy of host wavefunctions for the host-pointer entry points
struct HostWfn {
  mc::DevBuf buf;
  double2* upload(mc_ctx* ctx, const double* c, int ldc, int n) {
    buf.ensure(sizeof(double2) * (size_t)ctx->ldg * std::max(n, 1));
    mc::wfn_upload(ctx, c, ldc, n, buf.as<double2>());
    return buf.as<double2>();
  }
};

void check_ctx(mc_ctx* ctx) { MC_CHECK(ctx, MC_ERR_ARG, "null context"); }

}  // namespace

extern "C" {

int mc_ovlap_dev(mc_ctx* ctx, int nstate, double* a, const void* c1_dev, const void* c2_dev) {
  MC_API_BEGIN
  check_ctx(ctx);
  MC_CHECK(a && c1_dev && c2_dev && nstate >= 0, MC_ERR_ARG, "ovlap: bad argument");
  if (nstate == 0) return MC_OK;
  ctx->w_small[7].ensure(sizeof(double) * (size_t)nstate * nstate);
  mc::ovlap_dev(ctx, nstate, ctx->w_small[7].as<double>(), nstate, (const double2*)c1_dev, (const double2*)c2_dev);
  MC_HIP(hipMemcpyAsync(a, ctx->w_small[7].p, sizeof(double) * nstate * nstate, hipMemcpyDeviceToHost, ctx->stream));
  MC_HIP(hipStreamSynchronize(ctx->stream));
  MC_API_END
}

int mc_ovlap(mc_ctx* ctx, int nstate, double* a, const double* c1, const double* c2, int ldc) {
  MC_API_BEGIN
  check_ctx(ctx);
  MC_CHECK(a && c1 && c2 && nstate >= 0, MC_ERR_ARG, "ovlap: bad argument");
  if (nstate == 0) return MC_OK;
  HostWfn h1, h2;
  double2* d1 = h1.upload(ctx, c1, ldc, nstate);
  double2* d2 = (c1 == c2) ? d1 : h2.upload(ctx, c2, ldc, nstate);
  ctx->w_small[7].ensure(sizeof(double) * (size_t)nstate * nstate);
  mc::ovlap_dev(ctx, nstate, ctx->w_small[7].as<double>(), nstate, d1, d2);
  MC_HIP(hipMemcpyAsync(a, ctx->w_small[7].p, sizeof(double) * nstate * nstate, hipMemcpyDeviceToHost, ctx->stream));
  MC_HIP(hipStreamSynchronize(ctx->stream));
  MC_API_END
}

int mc_rgs_dev(mc_ctx* ctx, void* c_dev, int nstate, double* smat) {
  MC_API_BEGIN
  check_ctx(ctx);
  MC_CHECK(c_dev && nstate >= 0, MC_ERR_ARG, "rgs: bad argument");
  mc::rgs_dev(ctx, (double2*)c_dev, nstate, smat);
  MC_API_END
}

int mc_rgs(mc_ctx* ctx, double* cp, int ldc, int nstate, double* smat) {
  MC_API_BEGIN
  check_ctx(ctx);
  MC_CHECK(cp && nstate >= 0, MC_ERR_ARG, "rgs: bad argument");
  if (nstate == 0) return MC_OK;
  HostWfn h;
  double2* d = h.upload(ctx, cp, ldc, nstate);
  mc::rgs_dev(ctx, d, nstate, smat);
  mc::wfn_download(ctx, d, nstate, cp, ldc);
  MC_API_END
}

int mc_lowdin_dev(mc_ctx* ctx, void* c_dev, int nstate, int* iters) {
  MC_API_BEGIN
  check_ctx(ctx);
  MC_CHECK(c_dev && nstate >= 0, MC_ERR_ARG, "lowdin: bad argument");
  mc::lowdin_dev(ctx, (double2*)c_dev, nstate, iters);
  MC_API_END
}

int mc_lowdin(mc_ctx* ctx, double* cp, int ldc, int nstate) {
  MC_API_BEGIN
  check_ctx(ctx);
  MC_CHECK(cp && nstate >= 0, MC_ERR_ARG, "lowdin: bad argument");
  if (nstate == 0) return MC_OK;
  HostWfn h;
  double2* d = h.upload(ctx, cp, ldc, nstate);
  mc::lowdin_dev(ctx, d, nstate, nullptr);
  mc::wfn_download(ctx, d, nstate, cp, ldc);
  MC_API_END
}

int mc_rotate_dev(mc_ctx* ctx, double a, const void* c1_dev, double b, void* c2_dev, const double* gam, int nstate) {
  MC_API_BEGIN
  check_ctx(ctx);
  MC_CHECK(c1_dev && c2_dev && gam && nstate >= 0, MC_ERR_ARG, "rotate: bad argument");
  if (nstate == 0) return MC_OK;
  ctx->w_small[7].ensure(sizeof(double) * (size_t)nstate * nstate);
  MC_HIP(hipMemcpyAsync(ctx->w_small[7].p, gam, sizeof(double) * nstate * nstate, hipMemcpyHostToDevice, ctx->stream));
  mc::rotate_dev(ctx, a, (const double2*)c1_dev, b, (double2*)c2_dev, ctx->w_small[7].as<double>(), nstate);
  MC_HIP(hipStreamSynchronize(ctx->stream));
  MC_API_END
}

int mc_rotate(mc_ctx* ctx, double a, const double* c1, double b, double* c2, int ldc, const double* gam, int nstate) {
  MC_API_BEGIN
  check_ctx(ctx);
  MC_CHECK(c1 && c2 && gam && nstate >= 0, MC_ERR_ARG, "rotate: bad argument");
  if (nstate == 0) return MC_OK;
  HostWfn h1, h2;
  double2* d1 = h1.upload(ctx, c1, ldc, nstate);
  double2* d2 = h2.upload(ctx, c2, ldc, nstate);
  ctx->w_small[7].ensure(sizeof(double) * (size_t)nstate * nstate);
  MC_HIP(hipMemcpyAsync(ctx->w_small[7].p, gam, sizeof(double) * nstate * nstate, hipMemcpyHostToDevice, ctx->stream));
  mc::rotate_dev(ctx, a, d1, b, d2, ctx->w_small[7].as<double>(), nstate);
  mc::wfn_download(ctx, d2, nstate, c2, ldc);
  MC_API_END
}

int mc_lagrange_dev(mc_ctx* ctx, const void* c0_dev, void* c2_dev, const double* f, int nstate) {
  MC_API_BEGIN
  check_ctx(ctx);
  MC_CHECK(c0_dev && c2_dev && f && nstate >= 0, MC_ERR_ARG, "lagrange: bad argument");
  mc::lagrange_dev(ctx, (const double2*)c0_dev, (double2*)c2_dev, f, nstate);
  MC_API_END
}

int mc_lagrange(mc_ctx* ctx, const double* c0, double* c2, int ldc, const double* f, int nstate) {
  MC_API_BEGIN
  check_ctx(ctx);
  MC_CHECK(c0 && c2 && f && nstate >= 0, MC_ERR_ARG, "lagrange: bad argument");
  if (nstate == 0) return MC_OK;
  HostWfn h0, h2;
  double2* d0 = h0.upload(ctx, c0, ldc, nstate);
  double2* d2 = h2.upload(ctx, c2, ldc, nstate);
  mc::lagrange_dev(ctx, d0, d2, f, nstate);
  mc::wfn_download(ctx, d2, nstate, c2, ldc);
  MC_API_END
}

int mc_rortog_dev(mc_ctx* ctx, const void* c0_dev, void* cm_dev, double* gam, int tnon, int nstate, double epsog,
                  int maxit, int* iters) {
  MC_API_BEGIN
  check_ctx(ctx);
  MC_CHECK(c0_dev && cm_dev && nstate >= 0 && maxit >= 0, MC_ERR_ARG, "rortog: bad argument");
  mc::rortog_dev(ctx, (const double2*)c0_dev, (double2*)cm_dev, gam, tnon, nstate, epsog, maxit, iters);
  MC_API_END
}

int mc_rortog(mc_ctx* ctx, const double* c0, double* cm, int ldc, double* gam, int tnon, int nstate, double epsog,
              int maxit, int* iters) {
  MC_API_BEGIN
  check_ctx(ctx);
  MC_CHECK(c0 && cm && nstate >= 0 && maxit >= 0, MC_ERR_ARG, "rortog: bad argument");
  if (nstate == 0) return MC_OK;
  HostWfn h0, hm;
  double2* d0 = h0.upload(ctx, c0, ldc, nstate);
  double2* dm = hm.upload(ctx, cm, ldc, nstate);
  mc::rortog_dev(ctx, d0, dm, gam, tnon, nstate, epsog, maxit, iters);
  mc::wfn_download(ctx, dm, nstate, cm, ldc);
  MC_API_END
}

int mc_rortv(mc_ctx* ctx, const double* c0, double* cm, int ldc, double* gamy, int nstate) {
  MC_API_BEGIN
  check_ctx(ctx);
  MC_CHECK(c0 && cm && nstate >= 0, MC_ERR_ARG, "rortv: bad argument");
  if (nstate == 0) return MC_OK;
  HostWfn h0, hm;
  double2* d0 = h0.upload(ctx, c0, ldc, nstate);
  double2* dm = hm.upload(ctx, cm, ldc, nstate);
  mc::rortv_dev(ctx, d0, dm, gamy, nstate);
  mc::dl_finish(ctx);
  mc::wfn_download(ctx, dm, nstate, cm, ldc);
  MC_API_END
}

int mc_set_projectors(mc_ctx* ctx, int nsp, const int* na, int nhxs, const int* ngh, const int* lang,
                      const double* hmat, const double* twnl, int ldtw) {
  MC_API_BEGIN
  check_ctx(ctx);
  MC_CHECK(na && ngh && (nhxs == 0 || (lang && hmat && twnl)), MC_ERR_ARG, "set_projectors: null argument");
  MC_CHECK(ldtw >= ctx->ngw, MC_ERR_ARG, "set_projectors: ldtw < ngw");
  mc::set_projectors(ctx, nsp, na, nhxs, ngh, lang, hmat, twnl, ldtw);
  MC_API_END
}

int mc_set_eigr(mc_ctx* ctx, const double* eigr, int lde, int nat) {
  MC_API_BEGIN
  check_ctx(ctx);
  MC_CHECK(eigr || nat == 0, MC_ERR_ARG, "set_eigr: null argument");
  mc::set_eigr(ctx, eigr, lde, nat);
  MC_API_END
}

int mc_rnlsm1_dev(mc_ctx* ctx, const void* c0_dev, int nstate, double* fnl) {
  MC_API_BEGIN_RO
  check_ctx(ctx);
  MC_CHECK(c0_dev && nstate >= 0, MC_ERR_ARG, "rnlsm1: bad argument");
  mc::rnlsm1_dev(ctx, (const double2*)c0_dev, nstate, fnl);
  MC_API_END
}

int mc_rnlsm1(mc_ctx* ctx, const double* c0, int ldc, int nstate, double* fnl) {
  MC_API_BEGIN
  check_ctx(ctx);
  MC_CHECK(c0 && fnl && nstate >= 0, MC_ERR_ARG, "rnlsm1: bad argument");
  HostWfn h;
  double2* d = h.upload(ctx, c0, ldc, nstate);
  mc::rnlsm1_dev(ctx, d, nstate, fnl);
  MC_API_END
}

int mc_set_gk(mc_ctx* ctx, const double* gk) {
  MC_API_BEGIN
  check_ctx(ctx);
  MC_CHECK(gk || ctx->ngw == 0, MC_ERR_ARG, "set_gk: null argument");
  mc::set_gk(ctx, gk);
  MC_API_END
}

int mc_rnlsm2_dev(mc_ctx* ctx, const void* c0_dev, int nstate, double* dfnl) {
  MC_API_BEGIN_RO
  check_ctx(ctx);
  MC_CHECK(c0_dev && nstate >= 0, MC_ERR_ARG, "rnlsm2: bad argument");
  mc::rnlsm2_dev(ctx, (const double2*)c0_dev, nstate, dfnl);
  MC_API_END
}

int mc_rnlsm2(mc_ctx* ctx, const double* c0, int ldc, int nstate, double* dfnl) {
  MC_API_BEGIN
  check_ctx(ctx);
  MC_CHECK(c0 && dfnl && nstate >= 0, MC_ERR_ARG, "rnlsm2: bad argument");
  HostWfn h;
  double2* d = h.upload(ctx, c0, ldc, nstate);
  mc::rnlsm2_dev(ctx, d, nstate, dfnl);
  MC_API_END
}

int mc_rnlsm_dev(mc_ctx* ctx, const void* c0_dev, int nstate, int tfor, double* fnl, double* dfnl) {
  MC_API_BEGIN_RO
  check_ctx(ctx);
  MC_CHECK(c0_dev && nstate >= 0, MC_ERR_ARG, "rnlsm: bad argument");
  mc::rnlsm_dev(ctx, (const double2*)c0_dev, nstate, tfor != 0, fnl, dfnl);
  MC_API_END
}

int mc_posupa_dev(mc_ctx* ctx, const void* c0_dev, const void* c2_dev, void* cm_dev, void* cp_dev, int nstate,
                  double dtm, double dt, double epsog, int maxit, double* gam, int* iters) {
  MC_API_BEGIN
  check_ctx(ctx);
  MC_CHECK(c0_dev && cm_dev && cp_dev && nstate >= 0 && dt != 0.0 && maxit > 0, MC_ERR_ARG, "posupa: bad argument");
  MC_CHECK(cp_dev != c0_dev && cp_dev != cm_dev && cp_dev != c2_dev, MC_ERR_ARG,
           "posupa: cp must not alias c0, c2 or cm");
  int it = 0;
  mc::posupa_dev(ctx, (const double2*)c0_dev, (const double2*)c2_dev, (double2*)cm_dev, (double2*)cp_dev, nstate, dtm,
                 dt, epsog, maxit, gam, &it);
  if (iters) *iters = it;
  MC_API_END
}

int mc_rnlfor(mc_ctx* ctx, const double* f, int nstate, const double* fnl, const double* dfnl, double* fion) {
  MC_API_BEGIN_RO
  check_ctx(ctx);
  MC_CHECK(f && fion && nstate >= 0, MC_ERR_ARG, "rnlfor: bad argument");
  mc::rnlfor_dev(ctx, f, nstate, fnl, dfnl, fion);
  MC_API_END
}

int mc_rnlrh(mc_ctx* ctx, const double* f, int nstate, double* enl) {
  MC_API_BEGIN_RO
  check_ctx(ctx);
  MC_CHECK(f && enl && nstate >= 0, MC_ERR_ARG, "rnlrh: bad argument");
  mc::rnlrh_dev(ctx, f, nstate, enl);
  MC_API_END
}

int mc_set_density_gset(mc_ctx* ctx, int nhg, const int* inyh, const double* hg, const double* gk) {
  MC_API_BEGIN
  check_ctx(ctx);
  MC_CHECK(inyh && hg && gk, MC_ERR_ARG, "set_density_gset: null argument");
  mc::set_density_gset(ctx, nhg, inyh, hg, gk);
  MC_API_END
}

int mc_set_local_pp(mc_ctx* ctx, int nsp, const int* na, const double* zv, const double* raggio, const double* vps,
                    const double* rhops) {
  MC_API_BEGIN
  check_ctx(ctx);
  MC_CHECK(na && zv && raggio && vps && rhops, MC_ERR_ARG, "set_local_pp: null argument");
  mc::set_local_pp(ctx, nsp, na, zv, raggio, vps, rhops);
  MC_API_END
}

int mc_set_ions(mc_ctx* ctx, const double* tau, const double* ht) {
  MC_API_BEGIN_RO
  check_ctx(ctx);
  MC_CHECK(ht && (tau || ctx->dens.nat == 0), MC_ERR_ARG, "set_ions: null argument");
  mc::set_ions(ctx, tau, ht);
  MC_API_END
}

int mc_vofrho_dev(mc_ctx* ctx, const void* rho_dev, void* v_dev, double* energies, double* fion) {
  MC_API_BEGIN_RO
  check_ctx(ctx);
  MC_CHECK(rho_dev && v_dev && rho_dev != v_dev, MC_ERR_ARG, "vofrho: bad rho / v");
  mc::vofrho_dev(ctx, (const double*)rho_dev, (double*)v_dev, energies, fion);
  MC_API_END
}

int mc_vofrho(mc_ctx* ctx, const double* rho, double* v, double* energies, double* fion) {
  MC_API_BEGIN
  check_ctx(ctx);
  MC_CHECK(rho && v, MC_ERR_ARG, "vofrho: null argument");
  const size_t nnr = (size_t)ctx->kr1 * ctx->kr2 * ctx->kr3;
  mc::DevBuf r, o;
  r.ensure(sizeof(double) * nnr);
  o.ensure(sizeof(double) * nnr);
  MC_HIP(hipMemcpyAsync(r.p, rho, sizeof(double) * nnr, hipMemcpyHostToDevice, ctx->stream));
  mc::vofrho_dev(ctx, r.as<double>(), o.as<double>(), energies, fion);
  MC_HIP(hipMemcpyAsync(v, o.p, sizeof(double) * nnr, hipMemcpyDeviceToHost, ctx->stream));
  MC_HIP(hipStreamSynchronize(ctx->stream));
  MC_API_END
}

int mc_phfac(mc_ctx* ctx) {
  MC_API_BEGIN_RO
  check_ctx(ctx);
  mc::phfac(ctx);
  MC_API_END
}

int mc_ewald(mc_ctx* ctx, int iesr, double* esr, double* eself, double* fion) {
  MC_API_BEGIN_RO
  check_ctx(ctx);
  MC_CHECK(esr && eself, MC_ERR_ARG, "ewald: null argument");
  mc::ewald(ctx, iesr, esr, eself, fion);
  MC_API_END
}

int mc_fnonloc_dev(mc_ctx* ctx, void* c2_dev, const double* f, int nstate, const double* fnl) {
  MC_API_BEGIN
  check_ctx(ctx);
  MC_CHECK(c2_dev && f && nstate >= 0, MC_ERR_ARG, "fnonloc: bad argument");
  mc::fnonloc_dev(ctx, (double2*)c2_dev, f, nstate, fnl);
  MC_API_END
}

int mc_fnonloc(mc_ctx* ctx, double* c2, int ldc, const double* f, int nstate, const double* fnl) {
  MC_API_BEGIN
  check_ctx(ctx);
  MC_CHECK(c2 && f && nstate >= 0, MC_ERR_ARG, "fnonloc: bad argument");
  if (nstate == 0) return MC_OK;
  HostWfn h;
  double2* d = h.upload(ctx, c2, ldc, nstate);
  mc::fnonloc_dev(ctx, d, f, nstate, fnl);
  mc::wfn_download(ctx, d, nstate, c2, ldc);
  MC_API_END
}

}  // extern "C"
