/*
 * micpmd -- MI355X-native plane-wave inner loop for CPMD, C ABI.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes and
 * returns an int status (MC_OK == 0).  On failure mc_last_error() returns a
 * thread-local message; the Fortran shim turns a non-zero status into
 * CPMD's stopgm(procedureN, msg, __LINE__, __FILE__) (the convention of
 * src/cuda_utils.mod.F90:703-710).
 *
 * Conventions (identical to CPMD's serial Gamma-point data):
 *   - complex arrays are interleaved (re,im) float64, i.e. COMPLEX(real_8),
 *     column-major with a leading dimension in complex elements;
 *   - wavefunction rows are CPMD's G order (cppt::inyh / cppt::hg), the first
 *     row being G=0 when geq0;
 *   - real-space arrays have CPMD's layout v(ix + kr1*(iy + kr2*iz)),
 *     nnr1 = kr1*kr2*kr3 (fpar%kr1s.. from leadim, src/loadpa_utils.mod.F90:508-525).
 *
 * Host-pointer entry points ("drop-in") copy their inputs to the device,
 * compute, and copy results back before returning.  *_dev entry points take
 * device pointers to wavefunctions in the library's device layout (rows
 * permuted into stick order, leading dimension mc_wfn_ld()); they run
 * asynchronously on the context's stream except where a host result
 * (energies, matrices) is returned.
 */
#ifndef MICPMD_H
#define MICPMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  MC_OK = 0,
  MC_ERR_ARG = 1,          /* invalid argument / shape                      */
  MC_ERR_HIP = 2,          /* HIP runtime error                             */
  MC_ERR_NOTPOSDEF = 3,    /* Cholesky: leading minor not positive definite */
  MC_ERR_NOCONV = 4,       /* iteration did not converge                    */
  MC_ERR_UNSUPPORTED = 5,  /* transform length or option not supported      */
  MC_ERR_INTERNAL = 6
};

/* kernel ids for the optional per-kernel timer (mc_timer_*) */
enum {
  MC_KID_ZINV = 0,   /* pack + inverse z pass over sticks          */
  MC_KID_YINV = 1,   /* inverse y pass over x-frequency columns    */
  MC_KID_XRHO = 2,   /* inverse x pass + |psi|^2 accumulation      */
  MC_KID_XVPSI = 3,  /* inverse x pass * V(r) * forward x pass     */
  MC_KID_YFWD = 4,   /* forward y pass                             */
  MC_KID_ZFWD = 5,   /* forward z pass + Gamma decode + kinetic    */
  MC_KID_MISC = 6,   /* reductions, permutes, small kernels        */
  MC_KID_GEMM = 7,   /* MFMA contractions (FP64 or split bf16) other
                        than the A^T B ones, bf16 plane builds       */
  MC_KID_VOFRHO = 8, /* dense density transforms, per-G, LDA, Ewald  */
  MC_KID_GEMM_TN = 9,/* FP64 A^T B contractions (overlaps, fnl, dfnl):
                        the split-K MFMA kernel, without reductions  */
  MC_KID_TN_PLANES = 10, /* split-bf16 A^T B from operand planes
                        (k_tn_planes), without reductions            */
  MC_KID_COUNT = 11
};

typedef struct mc_ctx mc_ctx;

/* Layout of the plane-wave set, as CPMD holds it after rggen/loadpa/fftprp
 * (serial run).  Replaces the module state read by vpsi/rhoofr through USE
 * association: spar%nr1s.., fpar%kr1s.., ncpw%ngw, cppt::inyh, cppt::hg,
 * parm%tpiba2, parm%omega, geq0 (src/fftprp_utils.mod.F90:116-285,
 * src/rggen_utils.mod.F90:120-130). */
typedef struct {
  int nr1, nr2, nr3;    /* FFT mesh                                        */
  int kr1, kr2, kr3;    /* leading dimensions of real-space arrays         */
  int ngw;              /* number of wavefunction G vectors (half sphere)  */
  const int* inyh;      /* (3, ngw) 1-based centred indices nr/2+1+m       */
  const double* hg;     /* (ngw) |G|^2 in (2 pi/alat)^2                    */
  double tpiba2;        /* (2 pi/alat)^2                                   */
  double omega;         /* cell volume, bohr^3                             */
  int geq0;             /* 1 if G=0 is row 1                               */
} mc_layout;

const char* mc_last_error(void);
int mc_version(void);
int mc_device_count(int* count);

/* Host-only: geometry of the stick layout the library would build
 * (owner sticks, sticks, x-frequency columns).  Needs no GPU. */
int mc_layout_stats(const mc_layout* lay, int* n_sticks, int* n_columns, int* n_owner);

/* Context: uploads index maps and twiddles (CPMD: fftprp_default_init +
 * setfftn upload, src/fftprp_utils.mod.F90:116-430, src/fftnew_utils.mod.F90:53-175). */
int mc_ctx_create(const mc_layout* lay, mc_ctx** out);
int mc_ctx_destroy(mc_ctx* ctx);
/* Use a caller-owned hipStream_t (e.g. torch's current stream); NULL = own stream. */
int mc_ctx_set_stream(mc_ctx* ctx, void* hip_stream);
int mc_ctx_synchronize(mc_ctx* ctx);
/* Pairs of states transformed per launch in rhoofr / vpsi (workspace size:
 * one pair takes ncols * n2 * n3 complex elements of the transform precision;
 * defaults 128 / 64). */
int mc_ctx_set_batch(mc_ctx* ctx, int pairs_rhoofr, int pairs_vpsi);
/* Column reuse (no counterpart in the reference; for drivers that keep the
 * wavefunctions on the device, such as a CP step calling mc_rhoofr_dev and
 * then mc_forces_c2_dev or mc_vpsi_dev on the same c0_dev).  While on,
 * mc_rhoofr_dev keeps the y-transformed columns of every state pair
 * (npairs * ncols * n2 * n3 complex elements of the transform precision,
 * allocated once and held until off or mc_ctx_destroy) and the next
 * mc_vpsi_dev / mc_forces_c2_dev with the same c0_dev, nstate and precision
 * starts from them, skipping its inverse z and y passes; the results are
 * bit-identical.  Every library call that writes device memory other than
 * rho, V, c2 or the projections, and every host-pointer transform call,
 * invalidates the kept columns.  The caller must not write c0_dev itself
 * (a kernel of its own, a copy) between the two calls while reuse is on.
 * In mixed precision the switch also keeps the bf16 planes of the
 * coefficients across calls: mc_posupa_dev's final rotation writes those of
 * cp_dev (even nstate), mc_rnlsm_dev splits c0_dev when none are held, and
 * the next mc_rnlsm_dev, mc_cp_velupa_rortv_dev (with c2) and mc_posupa_dev
 * on the same pointer use them (no split per CP step instead of three).
 * Not used under a device memory limit (mc_set_mem_limit) or when the
 * allocation fails: the calls then work in batches as before.  Default off. */
int mc_ctx_set_column_reuse(mc_ctx* ctx, int on);
size_t mc_wfn_ld(const mc_ctx* ctx);

/* Arithmetic mode (no counterpart in the reference, which is FP64 only):
 *   MC_PREC_FP64   every transform, contraction and accumulation in FP64 (default);
 *   MC_PREC_FP32   rhoofr/vpsi transforms and stick/column workspaces in FP32;
 *                  wavefunctions, V(r), rho(r), the Gamma decode, the kinetic
 *                  term, c2 and all charge/energy sums stay FP64;
 *   MC_PREC_MIXED  MC_PREC_FP32 plus the tall wavefunction contractions as
 *                  bf16 split MFMA products with FP32 accumulation per K chunk
 *                  and FP64 chunk sums: the A^T B products (ovlap, rnlsm1,
 *                  rnlsm2) split each operand in three bf16 parts (six
 *                  products, FP32-level accuracy), the A X products (rotate,
 *                  fnonloc, the C X of rgs/lowdin) in two (three products,
 *                  hi hi + hi lo + lo hi); the diagonal of square overlaps is
 *                  summed in FP64; n x n algebra (Cholesky, inverse,
 *                  Newton-Schulz) stays FP64; lowdin/rgs apply C X as
 *                  C + C (X - I);
 *   MC_PREC_MIXED_BF16X3  as MC_PREC_MIXED with every split product 3-term. */
enum { MC_PREC_FP64 = 0, MC_PREC_FP32 = 1, MC_PREC_MIXED = 2, MC_PREC_MIXED_BF16X3 = 3 };
int mc_set_precision(mc_ctx* ctx, int fft_precision);

/* Device memory helpers (hipMalloc / hipFree / hipMemcpy).  The copies run on
 * the null stream and do not wait for work queued on a context's stream: call
 * mc_ctx_synchronize first when the buffer was written by a context call. */
int mc_dev_alloc(size_t bytes, void** ptr);
int mc_dev_free(void* ptr);
int mc_memcpy_h2d(void* dst, const void* src, size_t bytes);
int mc_memcpy_d2h(void* dst, const void* src, size_t bytes);
int mc_memset_d(void* dst, int value, size_t bytes);

/* Device memory budget (no counterpart in the reference, whose scratch is
 * sized by maxfft, src/fftprp_utils.mod.F90:308-328).  The library counts
 * every device allocation it makes in this process (context buffers of all
 * contexts and mc_dev_alloc blocks; not pinned host staging, not the fixed
 * array of mc_vpsi_dev_noalloc).  With a limit set (bytes > 0), rhoofr and
 * vpsi transform as many state pairs per batch as fit beside what is already
 * held and a reserve of max(256 MiB, limit / 16) for the buffers allocated
 * later in a step (shrinking their workspaces when needed; the per-pair
 * results do not depend on the batch).  An allocation that would pass the
 * limit first frees the transform workspaces no call is using (the next
 * transform call sizes them again) and, if it still does not fit, fails with
 * MC_ERR_ARG instead of being made.  0 removes the limit. */
int mc_set_mem_limit(size_t bytes);
/* Bytes held now, the most held since the last reset, and the limit. */
int mc_mem_stats(size_t* current, size_t* peak, size_t* limit);
int mc_mem_reset_peak(void);

/* MFMA issue rate (a diagnostic with no counterpart in the reference): the
 * TFLOP/s one matrix instruction reaches on the current GPU with independent
 * accumulator chains, register operands and no memory traffic, best of three
 * timings of about 5 ms.  kind 0: v_mfma_f32_16x16x32_bf16 (the split-bf16
 * contractions), 1: v_mfma_f32_32x32x16_bf16, 2: v_mfma_f64_16x16x4_f64 (the
 * FP64 contractions), 3: v_mfma_f32_16x16x4_f32.  The ceiling the contraction
 * kernels are read against beside the published dense peaks. */
int mc_probe_mfma(int kind, double* tflops);

/* Zero nstate device wavefunctions (stream-ordered; CPMD: zeroing(c2),
 * src/forces_driver.mod.F90:175). */
int mc_wfn_zero(mc_ctx* ctx, void* c_dev, int nstate);

/* y += alpha x over nstate device wavefunctions (the steepest-descent update
 * c0 <- c0 + dt c2 of updwf, src/updwf_utils.mod.F90:116-128). */
int mc_wfn_axpy(mc_ctx* ctx, double alpha, const void* x_dev, void* y_dev, int nstate);
/* y = alpha x + beta y (beta = 0: y is only written) over nstate device
 * wavefunctions (the Verlet position / velocity updates of posupa and velupa,
 * src/posupa_utils.mod.F90:44-187, src/velupa_utils.mod.F90:25-72). */
int mc_wfn_axpby(mc_ctx* ctx, double alpha, const void* x_dev, double beta, void* y_dev, int nstate);

/* One CP step of the electrons before the orthonormality constraint, in one
 * pass: cm += dtm c2 (velupa, src/velupa_utils.mod.F90:25-72, dtm = nnlst
 * dt/(2 emass) in CPMD's notation) and cp = c0 + dt cm (posupa,
 * src/posupa_utils.mod.F90:69-73).  cp must not alias c0, c2 or cm. */
int mc_cp_posupa(mc_ctx* ctx, const void* c0_dev, const void* c2_dev, void* cm_dev, void* cp_dev, int nstate,
                 double dtm, double dt);
/* posupa with velupa's kick folded in, on device wavefunctions: cm += dtm c2
 * (skipped when c2 is NULL), cp = c0 + dt cm, the orthonormality constraint
 * rortog(c0, cp) (cp += c0 gamx, epsog / maxit as in mc_rortog_dev), and the
 * velocities cm = (cp - c0) / dt.  c0 keeps the old coefficients; the caller
 * swaps c0 and cp (CPMD's final dcopy).  gam: the (nstate, nstate) gamx on
 * the host, or NULL; iters: the rortog iterations, or NULL.  In mixed
 * precision the first pass also writes the split-bf16 planes of c0 and cp
 * used by the two overlaps of rortog.  cm is written once, by the final
 * rotation: if rortog fails (MC_ERR_NOCONV) cm may still hold the old
 * velocities in mixed precision and the kicked ones in FP64.
 * Replaces velupa + posupa(c0,cm,c2,gamx,nstate), src/velupa_utils.mod.F90:25-72,
 * src/posupa_utils.mod.F90:41-160 (cntl%tharm, tmass off). */
int mc_posupa_dev(mc_ctx* ctx, const void* c0_dev, const void* c2_dev, void* cm_dev, void* cp_dev, int nstate,
                  double dtm, double dt, double epsog, int maxit, double* gam, int* iters);
/* Velocities after rortog: cm = (cp - c0) / dt, the same as CPMD's
 * cm += c0 gamx / dt (src/posupa_utils.mod.F90:150-158) when rortog returned
 * cp = c0 + dt cm + c0 gamx. */
int mc_cp_velocity(mc_ctx* ctx, const void* c0_dev, const void* cp_dev, void* cm_dev, int nstate, double dt);

/* rortv: the velocity constraint of the velocity Verlet CP step,
 * gamy = ovlap(cm, c0), gamy(i,j) = gamy(j,i) = -(gamy(i,j) + gamy(j,i))/2,
 * cm += c0 gamy, so that <c0|cm> + <cm|c0> = 0 for orthonormal c0; gamy
 * (nstate x nstate, may be NULL) receives the matrix.  Synchronous.
 * Replaces rortv(c0,cm,c2,sc0,gamy,nstate), src/rortv_utils.mod.F90:39-166
 * (overlap constraints; cntl%nonort, tharm, tmass and tdmal off). */
int mc_rortv(mc_ctx* ctx, const double* c0, double* cm, int ldc, double* gamy, int nstate);
int mc_rortv_dev(mc_ctx* ctx, const void* c0_dev, void* cm_dev, double* gamy, int nstate);
/* rekine: ekinc = emass sum_i dotp(cm_i, cm_i), the fictitious kinetic
 * energy of the electrons.  Synchronous.
 * Replaces rekine(cm,nstate,ekinc), src/rekine_utils.mod.F90:34-49 (tmass off). */
int mc_rekine_dev(mc_ctx* ctx, const void* cm_dev, int nstate, double emass, double* ekinc);
/* The end of CPMD's velocity Verlet CP step in one call: velupa (cm += dtm
 * c2, dtm = dt/(2 emass); skipped when c2 is NULL), rortv(c0, cm) (gamy as in
 * mc_rortv_dev, may be NULL) and rekine (ekinc, may be NULL).  With
 * mc_posupa_dev(dtm = dt/(2 emass)) before the force evaluation this is the
 * step of src/mdmain_utils.mod.F90:712-966: velupa + posupa(+rortog), forces,
 * velupa + rortv, rekine.  In mixed precision rekine's norms come from the
 * epilogue of rortv's rotation (summed in another fixed order than
 * mc_rekine_dev's: equal to rounding).  Synchronous. */
int mc_cp_velupa_rortv_dev(mc_ctx* ctx, const void* c0_dev, const void* c2_dev, void* cm_dev, int nstate,
                           double dtm, double emass, double* gamy, double* ekinc);

/* Wavefunction optimisation by preconditioned steepest descent.
 * hesele: the diagonal preconditioner vpp(G), kept in the context.  With
 * prec != 0, vpp = 1 / max(0.5 tpiba2 hg + vploc + sum_is na_is
 * sum_iv,jv hmat(iv,jv,is) twnl(G,iv,is) twnl(G,jv,is), hthrs) over the
 * projectors of mc_set_projectors (none set: no nonlocal term); vploc is
 * the G=0 local potential sum_is na_is vps(is, G=0) (ppener's vploc);
 * prec = 0: vpp = svar2.  Replaces hesele(svar2,vpp),
 * src/hesele_utils.mod.F90:32-142 (Goedecker and wsg branches). */
int mc_hesele(mc_ctx* ctx, int prec, double svar2, double vploc, double hthrs);
/* csize: size of the electronic gradient over the nocc states with
 * f > 1e-5 (taken as the first nocc): gemax = |c2| of the element that
 * izamax picks (largest |Re| + |Im|, first in CPMD order on ties),
 * cnorm = sqrt(sum_i dotp(c2_i, c2_i) / (nocc ngw)).  Synchronous.
 * Replaces csize(c2,nstate,gemax,cnorm), src/csize_utils.mod.F90:100-131
 * (closed shell, Gamma point). */
int mc_csize_dev(mc_ctx* ctx, const void* c2_dev, const double* f, int nstate, double* gemax, double* cnorm);
/* c0(G, i) += vpp(G) pf1 c2(G, i) for all nstate states, vpp from mc_hesele
 * (MC_ERR_ARG before the first mc_hesele).  With prec, pf1 =
 * dt^2 hthrs / emass; without, mc_hesele(ctx, 0, dt^2/emass, ...) and
 * pf1 = 1 give CPMD's daxpy(dt2bye).  Replaces the steepest-descent branch
 * of updwf, src/updwf_utils.mod.F90:117-128. */
int mc_sd_update_dev(mc_ctx* ctx, void* c0_dev, const void* c2_dev, int nstate, double pf1);
/* odiis: one DIIS step on the occupied states (f > 1e-5, the first nocc)
 * with the preconditioner of mc_hesele.  The context keeps the history of
 * mdiis (1..20) coefficient and gradient sets (2 mdiis nstate ldg complex
 * doubles of device memory).  c0 <- sum_i vc_i (c0_i + ff vpp' c2_i) with vc
 * from the DIIS equations over <c2_i| 2 ff^2 vpp'^2 |c2_j> (solved as LAPACK
 * dgelss, rcond 0.2e-15), ff = 1/f with prec (f > 0.1), vpp' = vpp / (1 +
 * vpp gamma), gamma += 1/vpp(G=0) when etot rose since the last call, else
 * gamma *= 0.75; empty states: c0 += svar2 c2.  *reinit is in/out: != 0 on
 * entry (or a new nstate / mdiis) restarts the history; on return it is 1
 * when this call restarted the history for insufficient progress, else 0
 * (CPMD's updwf then counts a DIIS reset, src/updwf_utils.mod.F90:102-109).
 * The progress test (nreset >= 3): more than nreset steps are stored and
 * neither the largest gradient element (grimax) nor the gradient norm fell
 * below (1 - mdiis/100) of the oldest stored one (CPMD sets nreset = mdiis
 * for wavefunction optimisation, src/control_test_utils.mod.F90:452-465).
 * The orthonormalisation (rgs) follows as in updwf.
 * Replaces odiis(c0,c2,vpp,nstate,pme,gde,svar2,reinit),
 * src/odiis_utils.mod.F90:492-740 (closed shell, uncompressed history). */
int mc_odiis_dev(mc_ctx* ctx, void* c0_dev, const void* c2_dev, const double* f, int nstate, double etot,
                 double svar2, int mdiis, int nreset, int* reinit);
/* Frees the ODIIS history (the next mc_odiis_dev starts a new one). */
int mc_odiis_release(mc_ctx* ctx);

/* out[i] = dotp(a_i, b_i), the Gamma-point inner product of state i
 * (src/dotp_utils.mod.F90:58-101); with b = c2 after vpsi + fnonloc,
 * -sum_i out[i] is the band energy sum_i f_i <c_i|H|c_i>. */
int mc_wfn_dotp(mc_ctx* ctx, const void* a_dev, const void* b_dev, int nstate, double* out);

/* Wavefunctions host (CPMD order, leading dim ldc) <-> device layout. */
int mc_wfn_upload(mc_ctx* ctx, const double* c_host, int ldc, int nstate, void* c_dev);
int mc_wfn_download(mc_ctx* ctx, const void* c_dev, int nstate, double* c_host, int ldc);

/* rhoofr: rhoe(nnr1) = sum_i f_i/omega |psi_i(r)|^2 over Gamma state pairs,
 * ekin = tpiba2 sum_i f_i sum_G hg |c_iG|^2, csumg = sum_i f_i <c_i|c_i>,
 * csumr = omega/(nr1 nr2 nr3) sum_r rhoe.
 * Replaces rhoofr(c0,rhoe,psi,nstate), src/rhoofr_utils.mod.F90:122-644
 * (closed-shell, no LSD/LSE/USPP branches), with kin_energy
 * src/kin_energy_utils.mod.F90:28-110 and build_density_sum
 * src/density_utils.mod.F90:61-80.  Pairs with f1 = f2 = 0 are skipped
 * (rhoofr:313-316).  The charge-sum check (rhoofr:625-635) is left to the
 * caller, which receives csumg and csumr.  With FP32 transforms
 * (MC_PREC_FP32 / MC_PREC_MIXED) rhoe is rescaled by csumg / csumr (removing
 * the ~1e-7 systematic charge loss of the FP32 rounding), and csumr is that of
 * the returned rhoe. */
int mc_rhoofr(mc_ctx* ctx, const double* c0, int ldc, const double* f, int nstate,
              double* rhoe, double* ekin, double* csumg, double* csumr);
int mc_rhoofr_dev(mc_ctx* ctx, const void* c0_dev, const double* f, int nstate,
                  void* rhoe_dev, double* ekin, double* csumg, double* csumr);
/* csumg and the real-space charge of the last rhoofr before the FP32 rescale
 * (equal to the returned csumr in FP64).  The FP32 transforms contract
 * norms by a deterministic 1.8e-7 at 192^3 (their rounded DFT constants and
 * twiddles), which this value shows. */
int mc_rhoofr_raw_charge(mc_ctx* ctx, double* csumg, double* csumr_raw);

/* vpsi: c2(:,i) += -(f_i/2)(tpiba2 hg c0(:,i) + 2 P_G[V psi_i]) with f=0 -> 1
 * (Gamma, closed shell, no LSE/CDFT/TDG/MGGA branches).
 * Replaces vpsi(c0,c2,f,vpot,psi,nstate,ikind,ispin,redist_c2),
 * src/vpsi_utils.mod.F90:120-732 (pack :400-435, inverse FFT :443, V :485-493,
 * forward FFT :552, decode :647-672, add_wfn :717). */
int mc_vpsi(mc_ctx* ctx, const double* c0, int ldc, double* c2, const double* f,
            const double* vpot, int nstate);
int mc_vpsi_dev(mc_ctx* ctx, const void* c0_dev, void* c2_dev, const double* f,
                const void* vpot_dev, int nstate);

/* vpsi that allocates nothing: the same computation and kernels as
 * mc_vpsi_dev (c2 += ..., bit-identical results), but the occupations and the
 * transform workspaces live in one fixed array of MC_NOALLOC_SCRATCH_BYTES
 * (128 MiB) of device memory per GPU, which the library holds from its first
 * context on that GPU until the process ends, instead of context buffers
 * that grow with hipMalloc.  The call makes no hipMalloc, hipFree or other
 * allocating call (the per-kernel timer, which creates events, is held off
 * during it).  No counterpart in the reference: it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).
 * Pairs per launch: the batch size of mc_ctx_set_batch, or fewer, as many as
 * fit.  MC_ERR_ARG when the occupations and one state pair's workspaces
 * (16 B in FP64, 8 B in FP32 transforms, times sticks x nr3 + columns x nr2 x
 * nr3) exceed the array (FP64 transforms on a 240^3 mesh with 9,281 sticks
 * in 109 columns need 136 MB).  Every call on one GPU uses the same array:
 * calls that overlap in time on one GPU (other streams or contexts) race. */
#define MC_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int mc_vpsi_dev_noalloc(mc_ctx* ctx, const void* c0_dev, void* c2_dev, const double* f,
                        const void* vpot_dev, int nstate);

/* ------------------------------------------------------------------------
 * Dense contractions (FP64 MFMA).  Wavefunctions are viewed as real
 * (2*ngw x nstate) matrices; matrices are column-major nstate x nstate on the
 * host.
 * ------------------------------------------------------------------------ */

/* ovlap: a = 2 Re(C1^H C2) - Re c1(G=0) Re c2(G=0)^T.
 * Replaces ovlap(nstate,a,c1,c2), src/ovlap_utils.mod.F90:57-190 (dsyrk+dmatc
 * when c1 == c2, dgemm otherwise, dger for G=0). */
int mc_ovlap(mc_ctx* ctx, int nstate, double* a, const double* c1, const double* c2, int ldc);
int mc_ovlap_dev(mc_ctx* ctx, int nstate, double* a, const void* c1_dev, const void* c2_dev);

/* rgs: Gram-Schmidt by Cholesky: S = ovlap(C,C) = U^T U, C <- C U^-1,
 * Im c(G=0) = 0; smat (optional, may be NULL) receives U^-1 (upper).
 * Replaces rgs(cp,nstate,smat), src/rgs_utils.mod.F90:31-93 and uinv :95-162;
 * MC_ERR_NOTPOSDEF mirrors the dpotrf failure stop of uinv :138-158. */
int mc_rgs(mc_ctx* ctx, double* cp, int ldc, int nstate, double* smat);
int mc_rgs_dev(mc_ctx* ctx, void* c_dev, int nstate, double* smat);

/* lowdin: C <- C S^-1/2 with S = ovlap(C,C), Im c(G=0) = 0.
 * Replaces lowdin(cp,cn,nstate), src/lowdin_utils.mod.F90:27-192 (closed shell).
 * iters (may be NULL) returns the Newton-Schulz iteration count. */
int mc_lowdin(mc_ctx* ctx, double* cp, int ldc, int nstate);
int mc_lowdin_dev(mc_ctx* ctx, void* c_dev, int nstate, int* iters);

/* rotate: c2 = a c1 gam + b c2 (gam nstate x nstate, host).
 * Replaces rotate(a,c1,b,c2,gam,nstate,2*ngw,tlsd=.false.,..),
 * src/rotate_utils.mod.F90:63-100.  c1 and c2 must not alias. */
int mc_rotate(mc_ctx* ctx, double a, const double* c1, double b, double* c2, int ldc, const double* gam, int nstate);
int mc_rotate_dev(mc_ctx* ctx, double a, const void* c1_dev, double b, void* c2_dev, const double* gam, int nstate);

/* Lagrange projection of the electronic gradient (iproj = 2):
 * gam = ovlap(c2,c0), gam(i,j) *= fsc(i) fsc(j) (fsc = f != 0), c2 -= c0 gam,
 * Im c2(G=0) = 0.  Replaces the closed-shell branch of addup2 in
 * src/forces_driver.mod.F90:284-370 (gscal :502-518). */
int mc_lagrange(mc_ctx* ctx, const double* c0, double* c2, int ldc, const double* f, int nstate);
int mc_lagrange_dev(mc_ctx* ctx, const void* c0_dev, void* c2_dev, const double* f, int nstate);

/* rortog: Car-Parrinello orthonormality constraint (rortho fixed point,
 * max|dgam| <= epsog/2 within maxit iterations), cm += c0 gam.
 * gam (may be NULL) receives the constraint matrix.  MC_ERR_NOCONV mirrors
 * the stop of src/rortog_utils.mod.F90:464-471.
 * Replaces rortog(c0,cm,gam,tnon,nstate,prteig) -> rortho,
 * src/rortog_utils.mod.F90:44-87, 315-498 (non-distributed, prteig=.false.). */
int mc_rortog(mc_ctx* ctx, const double* c0, double* cm, int ldc, double* gam, int tnon, int nstate,
              double epsog, int maxit, int* iters);
int mc_rortog_dev(mc_ctx* ctx, const void* c0_dev, void* cm_dev, double* gam, int tnon, int nstate,
                  double epsog, int maxit, int* iters);

/* ------------------------------------------------------------------------
 * Nonlocal pseudopotential projectors
 * ------------------------------------------------------------------------ */

/* Projector set of the current cell (CPMD: nlps_com%ngh, nghtol, sgpp2%hlsg /
 * wsg, cppt::twnl; src/nlps.mod.F90, src/sgpp.mod.F90, src/cppt.mod.F90).
 *   na(nsp)            atoms per species (isa ordering as ions0%na)
 *   ngh(nsp)           projectors per species, <= nhxs
 *   lang(nhxs,nsp)     angular momentum l of projector iv (nghtol)
 *   hmat(nhxs,nhxs,nsp) coupling h(iv,jv): Goedecker hlsg(ki,kj,l) for equal
 *                      (l, m), zero otherwise; wsg(iv) on the diagonal for
 *                      BHS/TM potentials (fnonloc.mod.F90:209-357)
 *   twnl(ldtw,nhxs,nsp) form factors in CPMD G order. */
int mc_set_projectors(mc_ctx* ctx, int nsp, const int* na, int nhxs, const int* ngh, const int* lang,
                      const double* hmat, const double* twnl, int ldtw);
/* Structure factors eigr(G, isa) = exp(-i G.R), complex (lde, nat), CPMD G
 * order (phfac, called after every ionic move, src/mdmain_utils.mod.F90:740). */
int mc_set_eigr(mc_ctx* ctx, const double* eigr, int lde, int nat);

/* rnlsm1: fnl(isa, iv, i) = 2 Re sum_G conj((-i)^l twnl eigr) c (G=0 halved);
 * fnl is (nat, nhxs, nstate) float64 (sfac::fnl with imagp = 1).  The
 * library keeps a device copy for mc_fnonloc.
 * Replaces rnlsm1(c0,nstate,ikind), src/rnlsm1_utils.mod.F90:42-193. */
int mc_rnlsm1(mc_ctx* ctx, const double* c0, int ldc, int nstate, double* fnl);
int mc_rnlsm1_dev(mc_ctx* ctx, const void* c0_dev, int nstate, double* fnl);

/* fnonloc: c2(G,i) += -f_i (-i)^l twnl(G) sum_ia eigr(G,ia) dd(ia,i),
 * dd(ia,i) = sum_jv h(iv,jv) fnl(isa,jv,i), f_i < 1e-5 -> 1.  fnl: host
 * array (nat,nhxs,nstate) or NULL for the device copy of the last rnlsm1.
 * Replaces fnonloc(c2,f,nstate,ikind,ispin,redist_c2),
 * src/fnonloc_utils.mod.F90:55-397 (Gamma). */
int mc_fnonloc(mc_ctx* ctx, double* c2, int ldc, const double* f, int nstate, const double* fnl);
int mc_fnonloc_dev(mc_ctx* ctx, void* c2_dev, const double* f, int nstate, const double* fnl);

/* Electronic gradient of forces: c2 <- 0, vpsi, fnonloc (with the device fnl
 * of the last rnlsm1; skipped when no projectors are set) in one call.  The vpsi pass writes c2 instead of adding
 * to it, so c2 is neither cleared nor read first; rows ngw .. ldg-1 are zeroed.
 * Replaces the zeroing(c2) / vpsi / fnonloc sequence of
 * forces(c0,c2,tau0,fion,rhoe,psi,nstate,nkpoint,lproj,tfor),
 * src/forces_driver.mod.F90:175-282 (Gamma, one k point). */
int mc_forces_c2_dev(mc_ctx* ctx, const void* c0_dev, void* c2_dev, const double* f,
                     const void* vpot_dev, int nstate);

/* G vectors gk(3, ngw) in units of 2 pi / alat, CPMD G order (cppt::gk, the
 * first ngw columns; set once per cell).  Needed by mc_rnlsm2. */
int mc_set_gk(mc_ctx* ctx, const double* gk);

/* rnlsm2: dfnl(isa, iv, k, i) = 2 tpiba Re sum_G conj(gk_k twnl (-i)^(l+1) eigr) c
 * (G=0 halved), the derivative of fnl with respect to the position of atom
 * isa; dfnl is (nat, nhxs, 3, nstate) float64 (sfac::dfnl with imagp = 1).
 * The library keeps a device copy for mc_rnlfor.  dfnl may be NULL in
 * mc_rnlsm2_dev.
 * Replaces rnlsm2(c0,nstate,ikpt,ikind), src/rnlsm2_utils.mod.F90:321-482 (Gamma). */
int mc_rnlsm2(mc_ctx* ctx, const double* c0, int ldc, int nstate, double* dfnl);
int mc_rnlsm2_dev(mc_ctx* ctx, const void* c0_dev, int nstate, double* dfnl);

/* rnlsm: rnlsm1 and, if tfor != 0, rnlsm2 of the same device coefficients in
 * one call (in mixed precision the bf16 planes of c0 are built once for all
 * four contractions); fnl / dfnl as in mc_rnlsm1_dev / mc_rnlsm2_dev, each may
 * be NULL.  Replaces rnlsm(c0,nstate,ikpt,ikind,tfor),
 * src/rnlsm_utils.mod.F90:23-49. */
int mc_rnlsm_dev(mc_ctx* ctx, const void* c0_dev, int nstate, int tfor, double* fnl, double* dfnl);

/* rnlfor: fion(k, isa) -= sum_i 2 f_i sum_{iv,jv} h(iv,jv) dfnl(isa,iv,k,i)
 * fnl(isa,jv,i) (states with |f_i| <= 1e-12 skipped), i.e. adds the nonlocal
 * force -dE_NL/dR.  fion is (3, nat) float64 in isa order (CPMD's
 * fion(3, nax, nsp) flattened by species); fnl / dfnl: host arrays or NULL for
 * the device copies of the last mc_rnlsm1 / mc_rnlsm2.
 * Replaces rnlfor(fion,f,wk,nstate,nkpoint), src/rnlfor_utils.mod.F90:41-360
 * (Gamma, wk = 1). */
int mc_rnlfor(mc_ctx* ctx, const double* f, int nstate, const double* fnl, const double* dfnl, double* fion);

/* rnlrh: the nonlocal energy E_NL = sum_i f_i sum_isa sum_{iv,jv}
 * h(iv,jv,is) fnl(isa,iv,i) fnl(isa,jv,i) from the device fnl of the last
 * mc_rnlsm1 / mc_rnlsm (no host copy of fnl needed); summed per atom on the
 * device, over atoms in isa order on the host.
 * Replaces rnlrh(enl,nstate,nkpoint), src/rnlrh_utils.mod.F90:43-200, and
 * rnlrhg, :271-343 (Gamma, imagp = 1, wk = 1). */
int mc_rnlrh(mc_ctx* ctx, const double* f, int nstate, double* enl);

/* ------------------------------------------------------------------------
 * Density -> potential (vofrhoa + vofrhob, closed shell, LDA) and Ewald
 * ------------------------------------------------------------------------ */

/* Density G set: nhg vectors (G = 0 first when geq0), centred 1-based indices
 * inyh(3, nhg), hg(nhg) in (2 pi / alat)^2 and gk(3, nhg) in 2 pi / alat (cppt
 * inyh / hg / gk, all nhg columns).  scg = 4 pi / (tpiba2 hg) is formed here
 * (0 at G = 0). */
int mc_set_density_gset(mc_ctx* ctx, int nhg, const int* inyh, const double* hg, const double* gk);
/* Local pseudopotential per species on the density set: vps(nsp, nhg) and the
 * Gaussian ion charges rhops(nsp, nhg) (cppt::vps / rhops from putps,
 * src/rinforce_utils.mod.F90:414-460), valence charges zv(nsp), Gaussian
 * radii raggio(nsp) and atoms per species na(nsp) (isa order). */
int mc_set_local_pp(mc_ctx* ctx, int nsp, const int* na, const double* zv, const double* raggio, const double* vps,
                    const double* rhops);
/* Ionic positions tau(3, nat) in bohr (isa order; CPMD's tau0(:, ia, is)
 * flattened by species) and the cell ht(3, 3) with ht(i, :) = lattice vector i
 * (metr_com%ht); called after every ionic move. */
int mc_set_ions(mc_ctx* ctx, const double* tau, const double* ht);
/* Projector structure factors eigr(G, isa) = e^{-iG.R} of the ions of
 * mc_set_ions, built on the device (in place of mc_set_eigr after an ionic
 * move).  Requires mc_set_projectors with the same atoms.
 * Replaces the eigr part of phfac(tau0), src/phfac_utils.mod.F90:45-261. */
int mc_phfac(mc_ctx* ctx);
/* vofrhoa + vofrhob for a closed-shell density with the Pade LDA (CPMD's
 * FUNCTIONAL LDA): rho(r) (nnr1) -> v(r) (nnr1) = V_H + V_loc (G != 0) + V_xc;
 * energies[7] = ehep, ehii (ei Omega), ehee, epseu, vploc, exc, vxc
 * (ener_com, src/vofrhoa_utils.mod.F90:105-122, src/vofrhob_utils.mod.F90:459-461);
 * fion (3, nat) += the local / Hartree forces (potfor) when not NULL.
 * Replaces vofrhoa(tau0,fion,rhoe,v,vtemp,tfor,.FALSE.) + vofrhob
 * (src/vofrhoa_utils.mod.F90:41-144, src/vofrhob_utils.mod.F90:80-461). */
int mc_vofrho(mc_ctx* ctx, const double* rho, double* v, double* energies, double* fion);
int mc_vofrho_dev(mc_ctx* ctx, const void* rho_dev, void* v_dev, double* energies, double* fion);
/* Ewald real-space sum over (2 iesr + 1)^3 cells and self energy: esr, eself
 * (ener_com%esr / eself), fion (3, nat) += the esr forces when not NULL.
 * Replaces rpiiint(esr,tau0,fion,iesr,tfor), src/rpiiint_utils.mod.F90:29-200,
 * and the eself sum of src/rinforce_utils.mod.F90:239-249. */
int mc_ewald(mc_ctx* ctx, int iesr, double* esr, double* eself, double* fion);

/* ------------------------------------------------------------------------
 * RESTART files ("STUTTGART VERSION 3.0", Fortran sequential unformatted:
 * 4-byte little-endian record lengths before and after each record).
 * Written: section 1 (header), 2 (cell: ibrav, indpg; celldm), 3 (nsp;
 * na), 4 (tau0), 5 (ionic velocities), 6 (initial coordinates), 7 (ecut,
 * cdual, dual, nel, nr1s..nr3s), 8 (nstate, nkpts = 0, ngwks, ngwls, nhgs,
 * nhgls), 9 (C0 in w_wfnio's layout: iwfntype 1, uncompressed), 10 (CM in
 * wr30wfn's layout, uncompressed), 23 (occupations); sections 11-22 and
 * 24-99 as empty markers.  Read: the same sections (C0/CM in either layout,
 * uncompressed, Gamma only), everything else skipped.  Coefficients are in
 * CPMD's G order (the host layout of every host-pointer entry point); atoms
 * in isa order, positions (3, nat) in bohr.  Host-only: no GPU needed
 * except by the *_dev variants, which download / upload through the
 * context's layout (ngwks must equal the context's ngw).
 * Replaces zhwwf/wv30 (src/wv30_utils.mod.F90:99-757; sections 9 and 10 at
 * :484-535 through w_wfnio src/wfnio_utils.mod.F90:52-233 and wr30wfn
 * src/wr30wfn_utils.mod.F90:569-693) and the matching sections of rv30
 * (src/rv30_utils.mod.F90:286-735).
 * ------------------------------------------------------------------------ */
enum { MC_RST_VELP = 1, MC_RST_TAUI = 2, MC_RST_C0 = 4, MC_RST_CM = 8, MC_RST_OCC = 16 };
typedef struct {
  int ibrav, indpg;         /* section 2: parm%ibrav, symmi%indpg          */
  double celldm[6];         /*            cell_com%celldm                   */
  int nsp, nat;             /* section 3: species; nat = sum(na) (0: any)   */
  double ecut, cdual;       /* section 7: cntr%ecut, dual00%cdual           */
  int dual, nel;            /*            dual00%dual (logical), NINT(nel)  */
  int nr1s, nr2s, nr3s;     /*            mesh                              */
  int nstate, ngwks, ngwls, nhgs, nhgls;   /* section 8 (spar%...)          */
  int sections;             /* MC_RST_* of the optional sections present    */
} mc_restart_header;
/* Writes path: the optional sections flagged in h->sections (their array
 * must be non-NULL); c0 / cm are (ldc, nstate) complex in CPMD G order. */
int mc_restart_write(const char* path, const mc_restart_header* h, const int* na, const double* tau0,
                     const double* velp, const double* taui, const double* f, const double* c0, const double* cm,
                     int ldc);
/* Header and section flags of path (sizes for the buffers of mc_restart_read). */
int mc_restart_read_header(const char* path, mc_restart_header* h);
/* Reads path into the non-NULL buffers: na (nsp), tau0 / velp / taui (3, nat),
 * f (nstate), c0 / cm (ldc, nstate). */
int mc_restart_read(const char* path, mc_restart_header* h, int* na, double* tau0, double* velp, double* taui,
                    double* f, double* c0, double* cm, int ldc);
/* As mc_restart_write / mc_restart_read with device wavefunctions in the
 * library's layout (synchronous). */
int mc_restart_write_dev(mc_ctx* ctx, const char* path, const mc_restart_header* h, const int* na,
                         const double* tau0, const double* velp, const double* taui, const double* f,
                         const void* c0_dev, const void* cm_dev);
int mc_restart_read_dev(mc_ctx* ctx, const char* path, mc_restart_header* h, int* na, double* tau0, double* velp,
                        double* taui, double* f, void* c0_dev, void* cm_dev);

/* Optional per-kernel timing with HIP events on the context stream. */
int mc_timer_enable(mc_ctx* ctx, int on);
/* The kernel ids timed while enabled: bit kid of kid_mask (default all).
 * Each timed launch adds two event records to the stream; timing only the
 * kernels of interest keeps the rest of a measured step unperturbed. */
int mc_timer_select(mc_ctx* ctx, unsigned long long kid_mask);
int mc_timer_reset(mc_ctx* ctx);
/* Synchronizes and returns total milliseconds and launch count of kernel kid. */
int mc_timer_get(mc_ctx* ctx, int kid, double* total_ms, long long* count);

#ifdef __cplusplus
}
#endif

#endif /* MICPMD_H */
