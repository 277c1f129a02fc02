! ==================================================================
! micpmd: CPMD's hot-path routines with their Fortran signatures, bound
! to the micpmd C ABI (include/micpmd.h) through ISO_C_BINDING.
!
! The module keeps one library context (the role of CPMD's fft/cppt
! module state, built by fftprp_default_init / setfftn) and forwards:
!
!   micpmd_vpsi    (c0,c2,f,vpot,psi,nstate,ikind,ispin,redist_c2)
!                  = vpsi,    src/vpsi_utils.mod.F90:120
!   micpmd_rhoofr  (c0,rhoe,psi,nstate,f)  = rhoofr, src/rhoofr_utils.mod.F90:122
!                  (f = crge%f(:,1); ekin/csumg/csumr are left in
!                   micpmd_ekin, micpmd_csumg, micpmd_csumr)
!   micpmd_rgs     (cp,nstate,smat)        = rgs,    src/rgs_utils.mod.F90:31
!   micpmd_lowdin  (cp,cn,nstate)          = lowdin, src/lowdin_utils.mod.F90:27
!   micpmd_ovlap   (nstate,a,c1,c2)        = ovlap,  src/ovlap_utils.mod.F90:57
!   micpmd_rotate  (a,c1,b,c2,gam,nstate)  = rotate, src/rotate_utils.mod.F90:63
!   micpmd_rortog  (c0,cm,gam,tnon,nstate,epsog,maxit)
!                  = rortog -> rortho, src/rortog_utils.mod.F90:44,315
!   micpmd_rnlsm1  (c0,nstate,fnl)         = rnlsm1, src/rnlsm1_utils.mod.F90:42
!   micpmd_fnonloc (c2,f,nstate,fnl)       = fnonloc, src/fnonloc_utils.mod.F90:55
!   micpmd_set_gk  (gk)                    = the cppt::gk G vectors used by rnlsm2
!   micpmd_rnlsm2  (c0,nstate,dfnl)        = rnlsm2, src/rnlsm2_utils.mod.F90:321
!   micpmd_rnlfor  (fion,f,nstate,fnl,dfnl)= rnlfor, src/rnlfor_utils.mod.F90:41
!   micpmd_set_density(nhg,inyh,hg,gk), micpmd_set_local_pp(na,zv,raggio,vps,rhops),
!   micpmd_set_ions(tau,ht)                = the cppt / ions / metr_com state of vofrho
!   micpmd_vofrho  (rhoe,v,ener,fion)      = vofrhoa + vofrhob, src/vofrhoa_utils.mod.F90:41
!   micpmd_ewald   (iesr,esr,eself,fion)   = rpiiint, src/rpiiint_utils.mod.F90:29
!   micpmd_lagrange(c0,c2,f,nstate)        = addup2 projection,
!                                            src/forces_driver.mod.F90:344-370
!   micpmd_rnlsm   (c0,nstate,ikpt,ikind,tfor) = rnlsm, src/rnlsm_utils.mod.F90:23
!                  (fnl / dfnl left in micpmd_fnl / micpmd_dfnl, CPMD's sfac arrays)
!   micpmd_ortho   (nstate,c0,cscr)        = ortho, src/ortho_utils.mod.F90:124
!                  (rgs, or lowdin after micpmd_set_ortho('LOWDIN'))
!   micpmd_rortv   (c0,cm,c2,sc0,gamy,nstate) = rortv, src/rortv_utils.mod.F90:39
!
! Set-up calls carry what CPMD's routines read from module state:
!   micpmd_set_occupations(f)        crge%f(:,1)    (rhoofr)
!   micpmd_set_cp_parameters(epsog, maxit)  cntr%epsog, cnti%maxit (rortog)
!   micpmd_set_precision(mode)       'FP64' (default), 'FP32' or 'MIXED'
!   micpmd_set_column_reuse(on)      keep rhoofr's columns for the next forces_c2
!
! The device-resident path (wavefunctions, rho and V kept in HBM between
! calls) is bound as micpmd_alloc_wfn / micpmd_alloc_real / micpmd_free /
! micpmd_upload / micpmd_download, micpmd_phfac, micpmd_rnlsm_dev,
! micpmd_rhoofr_dev, micpmd_vofrho_dev, micpmd_forces_c2_dev,
! micpmd_rnlfor_dev, micpmd_posupa_dev, micpmd_velupa_rortv_dev,
! micpmd_lagrange_dev, micpmd_rgs_dev, micpmd_hesele, micpmd_csize_dev,
! micpmd_odiis_dev, micpmd_enl and micpmd_rnlrh_dev.
!
! Every non-zero status of the C ABI is turned into a call of the
! procedure registered with micpmd_set_stop (CPMD: stopgm(procedureN,
! msg, __LINE__, __FILE__)); without one the program stops with the
! library's message.
! ==================================================================
MODULE micpmd
  USE, INTRINSIC :: iso_c_binding
  IMPLICIT NONE

  PRIVATE

  INTEGER, PARAMETER :: real_8 = c_double

  TYPE, BIND(C) :: mc_layout_t
     INTEGER(c_int) :: nr1, nr2, nr3, kr1, kr2, kr3, ngw
     TYPE(c_ptr)    :: inyh, hg
     REAL(c_double) :: tpiba2, omega
     INTEGER(c_int) :: geq0
  END TYPE mc_layout_t

  ABSTRACT INTERFACE
     SUBROUTINE stop_proc(procedure_name, msg)
       CHARACTER(*), INTENT(IN) :: procedure_name, msg
     END SUBROUTINE stop_proc
  END INTERFACE

  TYPE(c_ptr), SAVE                       :: ctx = c_null_ptr
  PROCEDURE(stop_proc), POINTER, SAVE     :: stop_hook => NULL()
  INTEGER, SAVE                           :: ngw_s = 0, nat_s = 0, nhxs_s = 0

  REAL(real_8), PUBLIC, SAVE :: micpmd_ekin = 0.0_real_8, micpmd_csumg = 0.0_real_8, &
                                micpmd_csumr = 0.0_real_8
  ! rortog / posupa iterations of the last call
  INTEGER, PUBLIC, SAVE :: micpmd_iters = 0
  ! CPMD's sfac::fnl(nat, nhxs, nstate) and dfnl(nat, nhxs, 3, nstate) (imagp = 1)
  REAL(real_8), ALLOCATABLE, TARGET, PUBLIC, SAVE :: micpmd_fnl(:,:,:), micpmd_dfnl(:,:,:,:)

  REAL(real_8), ALLOCATABLE, SAVE :: f_s(:), hmat_s(:,:,:)
  INTEGER(c_int), ALLOCATABLE, SAVE :: na_s(:), ngh_s(:)
  REAL(real_8), SAVE :: epsog_s = 1.0e-8_real_8
  INTEGER, SAVE :: maxit_s = 30
  LOGICAL, SAVE :: lowdin_s = .FALSE.

  PUBLIC :: micpmd_init, micpmd_finalize, micpmd_set_stop
  PUBLIC :: micpmd_vpsi, micpmd_rhoofr, micpmd_rgs, micpmd_lowdin, micpmd_ovlap
  PUBLIC :: micpmd_rotate, micpmd_rortog, micpmd_lagrange
  PUBLIC :: micpmd_set_projectors, micpmd_set_eigr, micpmd_rnlsm1, micpmd_fnonloc
  PUBLIC :: micpmd_set_gk, micpmd_rnlsm2, micpmd_rnlfor
  PUBLIC :: micpmd_set_density, micpmd_set_local_pp, micpmd_set_ions, micpmd_vofrho, micpmd_ewald
  PUBLIC :: micpmd_set_occupations, micpmd_set_cp_parameters, micpmd_set_precision, micpmd_set_ortho
  PUBLIC :: micpmd_set_column_reuse
  PUBLIC :: micpmd_rnlsm, micpmd_ortho, micpmd_rortv
  PUBLIC :: micpmd_alloc_wfn, micpmd_alloc_real, micpmd_free, micpmd_upload, micpmd_download, micpmd_zero_wfn
  PUBLIC :: micpmd_phfac, micpmd_rnlsm_dev, micpmd_rhoofr_dev, micpmd_vofrho_dev, micpmd_forces_c2_dev
  PUBLIC :: micpmd_rnlfor_dev, micpmd_posupa_dev, micpmd_velupa_rortv_dev, micpmd_lagrange_dev, micpmd_rgs_dev
  PUBLIC :: micpmd_hesele, micpmd_csize_dev, micpmd_odiis_dev, micpmd_enl, micpmd_rnlrh_dev

  INTERFACE
     FUNCTION mc_last_error() BIND(C, name='mc_last_error') RESULT(p)
       IMPORT :: c_ptr
       TYPE(c_ptr) :: p
     END FUNCTION mc_last_error
     FUNCTION mc_ctx_create(lay, out) BIND(C, name='mc_ctx_create') RESULT(ierr)
       IMPORT :: c_int, c_ptr, mc_layout_t
       TYPE(mc_layout_t), INTENT(IN) :: lay
       TYPE(c_ptr), INTENT(OUT) :: out
       INTEGER(c_int) :: ierr
     END FUNCTION mc_ctx_create
     FUNCTION mc_ctx_destroy(c) BIND(C, name='mc_ctx_destroy') RESULT(ierr)
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: c
       INTEGER(c_int) :: ierr
     END FUNCTION mc_ctx_destroy
     FUNCTION mc_vpsi(c, c0, ldc, c2, f, vpot, nstate) BIND(C, name='mc_vpsi') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double_complex, c_double
       TYPE(c_ptr), VALUE :: c
       COMPLEX(c_double_complex), INTENT(IN) :: c0(*)
       INTEGER(c_int), VALUE :: ldc, nstate
       COMPLEX(c_double_complex), INTENT(INOUT) :: c2(*)
       REAL(c_double), INTENT(IN) :: f(*), vpot(*)
       INTEGER(c_int) :: ierr
     END FUNCTION mc_vpsi
     FUNCTION mc_rhoofr(c, c0, ldc, f, nstate, rhoe, ekin, csumg, csumr) BIND(C, name='mc_rhoofr') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double_complex, c_double
       TYPE(c_ptr), VALUE :: c
       COMPLEX(c_double_complex), INTENT(IN) :: c0(*)
       INTEGER(c_int), VALUE :: ldc, nstate
       REAL(c_double), INTENT(IN) :: f(*)
       REAL(c_double), INTENT(OUT) :: rhoe(*), ekin, csumg, csumr
       INTEGER(c_int) :: ierr
     END FUNCTION mc_rhoofr
     FUNCTION mc_rgs(c, cp, ldc, nstate, smat) BIND(C, name='mc_rgs') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double_complex, c_double
       TYPE(c_ptr), VALUE :: c
       COMPLEX(c_double_complex), INTENT(INOUT) :: cp(*)
       INTEGER(c_int), VALUE :: ldc, nstate
       REAL(c_double), INTENT(OUT) :: smat(*)
       INTEGER(c_int) :: ierr
     END FUNCTION mc_rgs
     FUNCTION mc_lowdin(c, cp, ldc, nstate) BIND(C, name='mc_lowdin') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double_complex
       TYPE(c_ptr), VALUE :: c
       COMPLEX(c_double_complex), INTENT(INOUT) :: cp(*)
       INTEGER(c_int), VALUE :: ldc, nstate
       INTEGER(c_int) :: ierr
     END FUNCTION mc_lowdin
     FUNCTION mc_ovlap(c, nstate, a, c1, c2, ldc) BIND(C, name='mc_ovlap') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double_complex, c_double
       TYPE(c_ptr), VALUE :: c
       INTEGER(c_int), VALUE :: nstate, ldc
       REAL(c_double), INTENT(OUT) :: a(*)
       COMPLEX(c_double_complex), INTENT(IN) :: c1(*), c2(*)
       INTEGER(c_int) :: ierr
     END FUNCTION mc_ovlap
     FUNCTION mc_rotate(c, a, c1, b, c2, ldc, gam, nstate) BIND(C, name='mc_rotate') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double_complex, c_double
       TYPE(c_ptr), VALUE :: c
       REAL(c_double), VALUE :: a, b
       COMPLEX(c_double_complex), INTENT(IN) :: c1(*)
       COMPLEX(c_double_complex), INTENT(INOUT) :: c2(*)
       INTEGER(c_int), VALUE :: ldc, nstate
       REAL(c_double), INTENT(IN) :: gam(*)
       INTEGER(c_int) :: ierr
     END FUNCTION mc_rotate
     FUNCTION mc_rortog(c, c0, cm, ldc, gam, tnon, nstate, epsog, maxit, iters) BIND(C, name='mc_rortog') &
          RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double_complex, c_double
       TYPE(c_ptr), VALUE :: c
       COMPLEX(c_double_complex), INTENT(IN) :: c0(*)
       COMPLEX(c_double_complex), INTENT(INOUT) :: cm(*)
       INTEGER(c_int), VALUE :: ldc, tnon, nstate, maxit
       REAL(c_double), INTENT(OUT) :: gam(*)
       REAL(c_double), VALUE :: epsog
       INTEGER(c_int), INTENT(OUT) :: iters
       INTEGER(c_int) :: ierr
     END FUNCTION mc_rortog
     FUNCTION mc_lagrange(c, c0, c2, ldc, f, nstate) BIND(C, name='mc_lagrange') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double_complex, c_double
       TYPE(c_ptr), VALUE :: c
       COMPLEX(c_double_complex), INTENT(IN) :: c0(*)
       COMPLEX(c_double_complex), INTENT(INOUT) :: c2(*)
       INTEGER(c_int), VALUE :: ldc, nstate
       REAL(c_double), INTENT(IN) :: f(*)
       INTEGER(c_int) :: ierr
     END FUNCTION mc_lagrange
     FUNCTION mc_set_projectors(c, nsp, na, nhxs, ngh, lang, hmat, twnl, ldtw) &
          BIND(C, name='mc_set_projectors') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: c
       INTEGER(c_int), VALUE :: nsp, nhxs, ldtw
       INTEGER(c_int), INTENT(IN) :: na(*), ngh(*), lang(*)
       REAL(c_double), INTENT(IN) :: hmat(*), twnl(*)
       INTEGER(c_int) :: ierr
     END FUNCTION mc_set_projectors
     FUNCTION mc_set_eigr(c, eigr, lde, nat) BIND(C, name='mc_set_eigr') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double_complex
       TYPE(c_ptr), VALUE :: c
       COMPLEX(c_double_complex), INTENT(IN) :: eigr(*)
       INTEGER(c_int), VALUE :: lde, nat
       INTEGER(c_int) :: ierr
     END FUNCTION mc_set_eigr
     FUNCTION mc_rnlsm1(c, c0, ldc, nstate, fnl) BIND(C, name='mc_rnlsm1') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double_complex, c_double
       TYPE(c_ptr), VALUE :: c
       COMPLEX(c_double_complex), INTENT(IN) :: c0(*)
       INTEGER(c_int), VALUE :: ldc, nstate
       REAL(c_double), INTENT(OUT) :: fnl(*)
       INTEGER(c_int) :: ierr
     END FUNCTION mc_rnlsm1
     FUNCTION mc_set_gk(c, gk) BIND(C, name='mc_set_gk') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: c
       REAL(c_double), INTENT(IN) :: gk(*)
       INTEGER(c_int) :: ierr
     END FUNCTION mc_set_gk
     FUNCTION mc_rnlsm2(c, c0, ldc, nstate, dfnl) BIND(C, name='mc_rnlsm2') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double_complex, c_double
       TYPE(c_ptr), VALUE :: c
       COMPLEX(c_double_complex), INTENT(IN) :: c0(*)
       INTEGER(c_int), VALUE :: ldc, nstate
       REAL(c_double), INTENT(OUT) :: dfnl(*)
       INTEGER(c_int) :: ierr
     END FUNCTION mc_rnlsm2
     FUNCTION mc_rnlfor(c, f, nstate, fnl, dfnl, fion) BIND(C, name='mc_rnlfor') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: c, fnl, dfnl          ! c_null_ptr: the device copies
       REAL(c_double), INTENT(IN) :: f(*)
       INTEGER(c_int), VALUE :: nstate
       REAL(c_double), INTENT(INOUT) :: fion(*)
       INTEGER(c_int) :: ierr
     END FUNCTION mc_rnlfor
     FUNCTION mc_rnlrh(c, f, nstate, enl) BIND(C, name='mc_rnlrh') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: c
       REAL(c_double), INTENT(IN) :: f(*)
       INTEGER(c_int), VALUE :: nstate
       REAL(c_double), INTENT(OUT) :: enl
       INTEGER(c_int) :: ierr
     END FUNCTION mc_rnlrh
     FUNCTION mc_set_density_gset(c, nhg, inyh, hg, gk) BIND(C, name='mc_set_density_gset') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: c
       INTEGER(c_int), VALUE :: nhg
       INTEGER(c_int), INTENT(IN) :: inyh(*)
       REAL(c_double), INTENT(IN) :: hg(*), gk(*)
       INTEGER(c_int) :: ierr
     END FUNCTION mc_set_density_gset
     FUNCTION mc_set_local_pp(c, nsp, na, zv, raggio, vps, rhops) BIND(C, name='mc_set_local_pp') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: c
       INTEGER(c_int), VALUE :: nsp
       INTEGER(c_int), INTENT(IN) :: na(*)
       REAL(c_double), INTENT(IN) :: zv(*), raggio(*), vps(*), rhops(*)
       INTEGER(c_int) :: ierr
     END FUNCTION mc_set_local_pp
     FUNCTION mc_set_ions(c, tau, ht) BIND(C, name='mc_set_ions') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: c
       REAL(c_double), INTENT(IN) :: tau(*), ht(*)
       INTEGER(c_int) :: ierr
     END FUNCTION mc_set_ions
     FUNCTION mc_vofrho(c, rho, v, ener, fion) BIND(C, name='mc_vofrho') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: c
       REAL(c_double), INTENT(IN) :: rho(*)
       REAL(c_double), INTENT(OUT) :: v(*), ener(*)
       REAL(c_double), INTENT(INOUT) :: fion(*)
       INTEGER(c_int) :: ierr
     END FUNCTION mc_vofrho
     FUNCTION mc_ewald(c, iesr, esr, eself, fion) BIND(C, name='mc_ewald') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: c
       INTEGER(c_int), VALUE :: iesr
       REAL(c_double), INTENT(OUT) :: esr, eself
       REAL(c_double), INTENT(INOUT) :: fion(*)
       INTEGER(c_int) :: ierr
     END FUNCTION mc_ewald
     FUNCTION mc_fnonloc(c, c2, ldc, f, nstate, fnl) BIND(C, name='mc_fnonloc') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double_complex, c_double
       TYPE(c_ptr), VALUE :: c
       COMPLEX(c_double_complex), INTENT(INOUT) :: c2(*)
       INTEGER(c_int), VALUE :: ldc, nstate
       REAL(c_double), INTENT(IN) :: f(*), fnl(*)
       INTEGER(c_int) :: ierr
     END FUNCTION mc_fnonloc
     FUNCTION mc_set_precision(c, mode) BIND(C, name='mc_set_precision') RESULT(ierr)
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: c
       INTEGER(c_int), VALUE :: mode
       INTEGER(c_int) :: ierr
     END FUNCTION mc_set_precision
     FUNCTION mc_ctx_set_column_reuse(c, on) BIND(C, name='mc_ctx_set_column_reuse') RESULT(ierr)
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: c
       INTEGER(c_int), VALUE :: on
       INTEGER(c_int) :: ierr
     END FUNCTION mc_ctx_set_column_reuse
     FUNCTION mc_dev_alloc(bytes, p) BIND(C, name='mc_dev_alloc') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_size_t
       INTEGER(c_size_t), VALUE :: bytes
       TYPE(c_ptr), INTENT(OUT) :: p
       INTEGER(c_int) :: ierr
     END FUNCTION mc_dev_alloc
     FUNCTION mc_dev_free(p) BIND(C, name='mc_dev_free') RESULT(ierr)
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: p
       INTEGER(c_int) :: ierr
     END FUNCTION mc_dev_free
     FUNCTION mc_memset_d(p, v, bytes) BIND(C, name='mc_memset_d') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_size_t
       TYPE(c_ptr), VALUE :: p
       INTEGER(c_int), VALUE :: v
       INTEGER(c_size_t), VALUE :: bytes
       INTEGER(c_int) :: ierr
     END FUNCTION mc_memset_d
     FUNCTION mc_wfn_ld(c) BIND(C, name='mc_wfn_ld') RESULT(ld)
       IMPORT :: c_ptr, c_size_t
       TYPE(c_ptr), VALUE :: c
       INTEGER(c_size_t) :: ld
     END FUNCTION mc_wfn_ld
     FUNCTION mc_wfn_upload(c, host, ldc, nstate, dev) BIND(C, name='mc_wfn_upload') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double_complex
       TYPE(c_ptr), VALUE :: c, dev
       COMPLEX(c_double_complex), INTENT(IN) :: host(*)
       INTEGER(c_int), VALUE :: ldc, nstate
       INTEGER(c_int) :: ierr
     END FUNCTION mc_wfn_upload
     FUNCTION mc_wfn_download(c, dev, nstate, host, ldc) BIND(C, name='mc_wfn_download') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double_complex
       TYPE(c_ptr), VALUE :: c, dev
       COMPLEX(c_double_complex), INTENT(OUT) :: host(*)
       INTEGER(c_int), VALUE :: ldc, nstate
       INTEGER(c_int) :: ierr
     END FUNCTION mc_wfn_download
     FUNCTION mc_wfn_zero(c, dev, nstate) BIND(C, name='mc_wfn_zero') RESULT(ierr)
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: c, dev
       INTEGER(c_int), VALUE :: nstate
       INTEGER(c_int) :: ierr
     END FUNCTION mc_wfn_zero
     FUNCTION mc_phfac(c) BIND(C, name='mc_phfac') RESULT(ierr)
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: c
       INTEGER(c_int) :: ierr
     END FUNCTION mc_phfac
     FUNCTION mc_rnlsm_dev(c, c0, nstate, tfor, fnl, dfnl) BIND(C, name='mc_rnlsm_dev') RESULT(ierr)
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: c, c0, fnl, dfnl
       INTEGER(c_int), VALUE :: nstate, tfor
       INTEGER(c_int) :: ierr
     END FUNCTION mc_rnlsm_dev
     FUNCTION mc_rhoofr_dev(c, c0, f, nstate, rho, ekin, csumg, csumr) BIND(C, name='mc_rhoofr_dev') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: c, c0, rho
       REAL(c_double), INTENT(IN) :: f(*)
       INTEGER(c_int), VALUE :: nstate
       REAL(c_double), INTENT(OUT) :: ekin, csumg, csumr
       INTEGER(c_int) :: ierr
     END FUNCTION mc_rhoofr_dev
     FUNCTION mc_vofrho_dev(c, rho, v, ener, fion) BIND(C, name='mc_vofrho_dev') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: c, rho, v
       REAL(c_double), INTENT(OUT) :: ener(*)
       REAL(c_double), INTENT(INOUT) :: fion(*)
       INTEGER(c_int) :: ierr
     END FUNCTION mc_vofrho_dev
     FUNCTION mc_forces_c2_dev(c, c0, c2, f, vpot, nstate) BIND(C, name='mc_forces_c2_dev') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: c, c0, c2, vpot
       REAL(c_double), INTENT(IN) :: f(*)
       INTEGER(c_int), VALUE :: nstate
       INTEGER(c_int) :: ierr
     END FUNCTION mc_forces_c2_dev
     FUNCTION mc_posupa_dev(c, c0, c2, cm, cp, nstate, dtm, dt, epsog, maxit, gam, iters) &
          BIND(C, name='mc_posupa_dev') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: c, c0, c2, cm, cp, gam
       INTEGER(c_int), VALUE :: nstate, maxit
       REAL(c_double), VALUE :: dtm, dt, epsog
       INTEGER(c_int), INTENT(OUT) :: iters
       INTEGER(c_int) :: ierr
     END FUNCTION mc_posupa_dev
     FUNCTION mc_cp_velupa_rortv_dev(c, c0, c2, cm, nstate, dtm, emass, gamy, ekinc) &
          BIND(C, name='mc_cp_velupa_rortv_dev') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: c, c0, c2, cm, gamy
       INTEGER(c_int), VALUE :: nstate
       REAL(c_double), VALUE :: dtm, emass
       REAL(c_double), INTENT(OUT) :: ekinc
       INTEGER(c_int) :: ierr
     END FUNCTION mc_cp_velupa_rortv_dev
     FUNCTION mc_rortv(c, c0, cm, ldc, gamy, nstate) BIND(C, name='mc_rortv') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double_complex, c_double
       TYPE(c_ptr), VALUE :: c
       COMPLEX(c_double_complex), INTENT(IN) :: c0(*)
       COMPLEX(c_double_complex), INTENT(INOUT) :: cm(*)
       INTEGER(c_int), VALUE :: ldc, nstate
       REAL(c_double), INTENT(OUT) :: gamy(*)
       INTEGER(c_int) :: ierr
     END FUNCTION mc_rortv
     FUNCTION mc_lagrange_dev(c, c0, c2, f, nstate) BIND(C, name='mc_lagrange_dev') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: c, c0, c2
       REAL(c_double), INTENT(IN) :: f(*)
       INTEGER(c_int), VALUE :: nstate
       INTEGER(c_int) :: ierr
     END FUNCTION mc_lagrange_dev
     FUNCTION mc_rgs_dev(c, cdev, nstate, smat) BIND(C, name='mc_rgs_dev') RESULT(ierr)
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: c, cdev, smat
       INTEGER(c_int), VALUE :: nstate
       INTEGER(c_int) :: ierr
     END FUNCTION mc_rgs_dev
     FUNCTION mc_hesele(c, prec, svar2, vploc, hthrs) BIND(C, name='mc_hesele') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: c
       INTEGER(c_int), VALUE :: prec
       REAL(c_double), VALUE :: svar2, vploc, hthrs
       INTEGER(c_int) :: ierr
     END FUNCTION mc_hesele
     FUNCTION mc_csize_dev(c, c2, f, nstate, gemax, cnorm) BIND(C, name='mc_csize_dev') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: c, c2
       REAL(c_double), INTENT(IN) :: f(*)
       INTEGER(c_int), VALUE :: nstate
       REAL(c_double), INTENT(OUT) :: gemax, cnorm
       INTEGER(c_int) :: ierr
     END FUNCTION mc_csize_dev
     FUNCTION mc_odiis_dev(c, c0, c2, f, nstate, etot, svar2, mdiis, nreset, reinit) &
          BIND(C, name='mc_odiis_dev') RESULT(ierr)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: c, c0, c2
       REAL(c_double), INTENT(IN) :: f(*)
       INTEGER(c_int), VALUE :: nstate, mdiis, nreset
       REAL(c_double), VALUE :: etot, svar2
       INTEGER(c_int), INTENT(INOUT) :: reinit
       INTEGER(c_int) :: ierr
     END FUNCTION mc_odiis_dev
     FUNCTION c_strlen(s) BIND(C, name='strlen') RESULT(n)
       IMPORT :: c_ptr, c_size_t
       TYPE(c_ptr), VALUE :: s
       INTEGER(c_size_t) :: n
     END FUNCTION c_strlen
  END INTERFACE

CONTAINS

  ! ------------------------------------------------------------------
  SUBROUTINE micpmd_set_stop(proc)
    PROCEDURE(stop_proc) :: proc
    stop_hook => proc
  END SUBROUTINE micpmd_set_stop

  SUBROUTINE fail(procedure_name, msg)
    CHARACTER(*), INTENT(IN) :: procedure_name, msg
    IF (ASSOCIATED(stop_hook)) THEN
       CALL stop_hook(procedure_name, msg)
    ELSE
       WRITE(6,'(A,A,A,A)') ' micpmd| ', procedure_name, ': ', msg
       ERROR STOP 1
    ENDIF
  END SUBROUTINE fail

  SUBROUTINE check(ierr, procedure_name)
    INTEGER(c_int), INTENT(IN) :: ierr
    CHARACTER(*), INTENT(IN)   :: procedure_name
    CHARACTER(kind=c_char), POINTER :: s(:)
    CHARACTER(len=512) :: msg
    TYPE(c_ptr) :: p
    INTEGER :: i, n
    IF (ierr == 0) RETURN
    p = mc_last_error()
    n = INT(c_strlen(p))
    CALL c_f_pointer(p, s, [n])
    msg = ' '
    DO i = 1, MIN(n, LEN(msg))
       msg(i:i) = s(i)
    ENDDO
    IF (ASSOCIATED(stop_hook)) THEN
       CALL stop_hook(procedure_name, TRIM(msg))
    ELSE
       WRITE(6,'(A,A,A,I0,A,A)') ' micpmd| ', procedure_name, ': status ', ierr, ': ', TRIM(msg)
       ERROR STOP 1
    ENDIF
  END SUBROUTINE check

  ! ------------------------------------------------------------------
  ! context from CPMD's layout: spar%nr1s.., fpar%kr1s.., ncpw%ngw,
  ! cppt::inyh(3,ngw), cppt::hg(ngw), parm%tpiba2, parm%omega, geq0
  SUBROUTINE micpmd_init(nr1, nr2, nr3, kr1, kr2, kr3, ngw, inyh, hg, tpiba2, omega, geq0)
    INTEGER, INTENT(IN)                :: nr1, nr2, nr3, kr1, kr2, kr3, ngw
    INTEGER(c_int), INTENT(IN), TARGET :: inyh(3, ngw)
    REAL(real_8), INTENT(IN), TARGET   :: hg(ngw)
    REAL(real_8), INTENT(IN)           :: tpiba2, omega
    LOGICAL, INTENT(IN)                :: geq0
    TYPE(mc_layout_t) :: lay
    IF (C_ASSOCIATED(ctx)) CALL micpmd_finalize()
    lay%nr1 = nr1; lay%nr2 = nr2; lay%nr3 = nr3
    lay%kr1 = kr1; lay%kr2 = kr2; lay%kr3 = kr3
    lay%ngw = ngw
    lay%inyh = C_LOC(inyh)
    lay%hg = C_LOC(hg)
    lay%tpiba2 = tpiba2
    lay%omega = omega
    lay%geq0 = MERGE(1, 0, geq0)
    CALL check(mc_ctx_create(lay, ctx), 'micpmd_init')
    ngw_s = ngw
  END SUBROUTINE micpmd_init

  SUBROUTINE micpmd_finalize()
    IF (C_ASSOCIATED(ctx)) CALL check(mc_ctx_destroy(ctx), 'micpmd_finalize')
    ctx = c_null_ptr
  END SUBROUTINE micpmd_finalize

  ! ------------------------------------------------------------------
  SUBROUTINE micpmd_vpsi(c0, c2, f, vpot, psi, nstate, ikind, ispin, redist_c2)
    COMPLEX(real_8), INTENT(IN)    :: c0(:,:)
    COMPLEX(real_8), INTENT(INOUT) :: c2(:,:)
    REAL(real_8), INTENT(IN)       :: f(:), vpot(:,:)
    COMPLEX(real_8)                :: psi(:)
    INTEGER, INTENT(IN)            :: nstate, ikind, ispin
    LOGICAL, INTENT(IN)            :: redist_c2
    IF (ikind /= 1 .OR. ispin /= 1) CALL fail('vpsi', 'only the Gamma point with one spin (ikind=1, ispin=1)')
    IF (SIZE(c2, 1) /= SIZE(c0, 1)) CALL fail('vpsi', 'c0 and c2 leading dimensions differ')
    CALL check(mc_vpsi(ctx, c0, INT(SIZE(c0, 1), c_int), c2, f, vpot, INT(nstate, c_int)), 'vpsi')
  END SUBROUTINE micpmd_vpsi

  ! rhoofr(c0,rhoe,psi,nstate) with the occupations of micpmd_set_occupations
  ! (CPMD: crge%f); ekin / csumg / csumr are left in micpmd_ekin / _csumg / _csumr
  SUBROUTINE micpmd_rhoofr(c0, rhoe, psi, nstate)
    COMPLEX(real_8), INTENT(IN) :: c0(:,:)
    REAL(real_8), INTENT(OUT)   :: rhoe(:,:)
    COMPLEX(real_8)             :: psi(:)
    INTEGER, INTENT(IN)         :: nstate
    CALL need_f('rhoofr', nstate)
    CALL check(mc_rhoofr(ctx, c0, INT(SIZE(c0, 1), c_int), f_s, INT(nstate, c_int), rhoe, micpmd_ekin, &
         micpmd_csumg, micpmd_csumr), 'rhoofr')
  END SUBROUTINE micpmd_rhoofr

  SUBROUTINE micpmd_rgs(cp, nstate, smat)
    COMPLEX(real_8), INTENT(INOUT) :: cp(:,:)
    INTEGER, INTENT(IN)            :: nstate
    REAL(real_8), INTENT(OUT)      :: smat(nstate, nstate)
    CALL check(mc_rgs(ctx, cp, INT(SIZE(cp, 1), c_int), INT(nstate, c_int), smat), 'rgs')
  END SUBROUTINE micpmd_rgs

  SUBROUTINE micpmd_lowdin(cp, cn, nstate)
    COMPLEX(real_8), INTENT(INOUT) :: cp(:,:)
    COMPLEX(real_8)                :: cn(:,:)
    INTEGER, INTENT(IN)            :: nstate
    CALL check(mc_lowdin(ctx, cp, INT(SIZE(cp, 1), c_int), INT(nstate, c_int)), 'lowdin')
  END SUBROUTINE micpmd_lowdin

  SUBROUTINE micpmd_ovlap(nstate, a, c1, c2)
    INTEGER, INTENT(IN)         :: nstate
    REAL(real_8), INTENT(OUT)   :: a(nstate, nstate)
    COMPLEX(real_8), INTENT(IN) :: c1(:,:), c2(:,:)
    CALL check(mc_ovlap(ctx, INT(nstate, c_int), a, c1, c2, INT(SIZE(c1, 1), c_int)), 'ovlap')
  END SUBROUTINE micpmd_ovlap

  SUBROUTINE micpmd_rotate(a, c1, b, c2, gam, nstate)
    REAL(real_8), INTENT(IN)       :: a, b
    COMPLEX(real_8), INTENT(IN)    :: c1(:,:)
    COMPLEX(real_8), INTENT(INOUT) :: c2(:,:)
    INTEGER, INTENT(IN)            :: nstate
    REAL(real_8), INTENT(IN)       :: gam(nstate, nstate)
    CALL check(mc_rotate(ctx, a, c1, b, c2, INT(SIZE(c1, 1), c_int), gam, INT(nstate, c_int)), 'rotate')
  END SUBROUTINE micpmd_rotate

  ! rortog(c0,cm,gam,tnon,nstate,prteig) with epsog / maxit of
  ! micpmd_set_cp_parameters (CPMD: cntr%epsog, cnti%maxit)
  SUBROUTINE micpmd_rortog(c0, cm, gam, tnon, nstate, prteig)
    COMPLEX(real_8), INTENT(IN)    :: c0(:,:)
    COMPLEX(real_8), INTENT(INOUT) :: cm(:,:)
    REAL(real_8), INTENT(OUT)      :: gam(nstate, nstate)
    LOGICAL, INTENT(IN)            :: tnon, prteig
    INTEGER, INTENT(IN)            :: nstate
    INTEGER(c_int) :: iters
    CALL check(mc_rortog(ctx, c0, cm, INT(SIZE(c0, 1), c_int), gam, MERGE(1_c_int, 0_c_int, tnon), &
         INT(nstate, c_int), epsog_s, INT(maxit_s, c_int), iters), 'rortog')
    micpmd_iters = iters
  END SUBROUTINE micpmd_rortog

  SUBROUTINE micpmd_lagrange(c0, c2, f, nstate)
    COMPLEX(real_8), INTENT(IN)    :: c0(:,:)
    COMPLEX(real_8), INTENT(INOUT) :: c2(:,:)
    REAL(real_8), INTENT(IN)       :: f(:)
    INTEGER, INTENT(IN)            :: nstate
    CALL check(mc_lagrange(ctx, c0, c2, INT(SIZE(c0, 1), c_int), f, INT(nstate, c_int)), 'lagrange')
  END SUBROUTINE micpmd_lagrange

  ! ------------------------------------------------------------------
  SUBROUTINE micpmd_set_projectors(na, nhxs, ngh, lang, hmat, twnl)
    INTEGER(c_int), INTENT(IN) :: na(:), ngh(:), lang(:,:)
    INTEGER, INTENT(IN)        :: nhxs
    REAL(real_8), INTENT(IN)   :: hmat(:,:,:), twnl(:,:,:)
    CALL check(mc_set_projectors(ctx, INT(SIZE(na), c_int), na, INT(nhxs, c_int), ngh, lang, hmat, twnl, &
         INT(SIZE(twnl, 1), c_int)), 'set_projectors')
    nat_s = SUM(na)
    nhxs_s = nhxs
    na_s = na
    ngh_s = ngh
    hmat_s = hmat
  END SUBROUTINE micpmd_set_projectors

  SUBROUTINE micpmd_set_eigr(eigr)
    COMPLEX(real_8), INTENT(IN) :: eigr(:,:)
    CALL check(mc_set_eigr(ctx, eigr, INT(SIZE(eigr, 1), c_int), INT(SIZE(eigr, 2), c_int)), 'set_eigr')
  END SUBROUTINE micpmd_set_eigr

  SUBROUTINE micpmd_rnlsm1(c0, nstate, fnl)
    COMPLEX(real_8), INTENT(IN) :: c0(:,:)
    INTEGER, INTENT(IN)         :: nstate
    REAL(real_8), INTENT(OUT)   :: fnl(:,:,:)
    CALL check(mc_rnlsm1(ctx, c0, INT(SIZE(c0, 1), c_int), INT(nstate, c_int), fnl), 'rnlsm1')
  END SUBROUTINE micpmd_rnlsm1

  SUBROUTINE micpmd_fnonloc(c2, f, nstate, fnl)
    COMPLEX(real_8), INTENT(INOUT) :: c2(:,:)
    REAL(real_8), INTENT(IN)       :: f(:)
    INTEGER, INTENT(IN)            :: nstate
    REAL(real_8), INTENT(IN)       :: fnl(:,:,:)
    CALL check(mc_fnonloc(ctx, c2, INT(SIZE(c2, 1), c_int), f, INT(nstate, c_int), fnl), 'fnonloc')
  END SUBROUTINE micpmd_fnonloc

  SUBROUTINE micpmd_set_gk(gk)
    REAL(real_8), INTENT(IN) :: gk(:,:)
    CALL check(mc_set_gk(ctx, gk), 'set_gk')
  END SUBROUTINE micpmd_set_gk

  SUBROUTINE micpmd_rnlsm2(c0, nstate, dfnl)
    COMPLEX(real_8), INTENT(IN) :: c0(:,:)
    INTEGER, INTENT(IN)         :: nstate
    REAL(real_8), INTENT(OUT)   :: dfnl(:,:,:,:)
    CALL check(mc_rnlsm2(ctx, c0, INT(SIZE(c0, 1), c_int), INT(nstate, c_int), dfnl), 'rnlsm2')
  END SUBROUTINE micpmd_rnlsm2

  ! fion(3, nat) in isa order (CPMD's fion(:, ia, is) flattened by species)
  SUBROUTINE micpmd_rnlfor(fion, f, nstate, fnl, dfnl)
    REAL(real_8), INTENT(INOUT) :: fion(:,:)
    REAL(real_8), INTENT(IN)    :: f(:)
    INTEGER, INTENT(IN)         :: nstate
    REAL(real_8), INTENT(IN), TARGET, CONTIGUOUS :: fnl(:,:,:), dfnl(:,:,:,:)
    CALL check(mc_rnlfor(ctx, f, INT(nstate, c_int), C_LOC(fnl), C_LOC(dfnl), fion), 'rnlfor')
  END SUBROUTINE micpmd_rnlfor

  SUBROUTINE micpmd_set_density(nhg, inyh, hg, gk)
    INTEGER, INTENT(IN)      :: nhg
    INTEGER, INTENT(IN)      :: inyh(:,:)
    REAL(real_8), INTENT(IN) :: hg(:), gk(:,:)
    CALL check(mc_set_density_gset(ctx, INT(nhg, c_int), inyh, hg, gk), 'set_density_gset')
  END SUBROUTINE micpmd_set_density

  SUBROUTINE micpmd_set_local_pp(na, zv, raggio, vps, rhops)
    INTEGER, INTENT(IN)      :: na(:)
    REAL(real_8), INTENT(IN) :: zv(:), raggio(:), vps(:,:), rhops(:,:)
    CALL check(mc_set_local_pp(ctx, INT(SIZE(na), c_int), na, zv, raggio, vps, rhops), 'set_local_pp')
  END SUBROUTINE micpmd_set_local_pp

  ! tau(3, nat) in isa order; ht(3, 3) with ht(i, :) = lattice vector i
  SUBROUTINE micpmd_set_ions(tau, ht)
    REAL(real_8), INTENT(IN) :: tau(:,:), ht(:,:)
    CALL check(mc_set_ions(ctx, tau, ht), 'set_ions')
  END SUBROUTINE micpmd_set_ions

  ! ener(7) = ehep, ehii, ehee, epseu, vploc, exc, vxc; fion(3, nat) += potfor
  SUBROUTINE micpmd_vofrho(rhoe, v, ener, fion)
    REAL(real_8), INTENT(IN)    :: rhoe(:)
    REAL(real_8), INTENT(OUT)   :: v(:), ener(7)
    REAL(real_8), INTENT(INOUT) :: fion(:,:)
    CALL check(mc_vofrho(ctx, rhoe, v, ener, fion), 'vofrho')
  END SUBROUTINE micpmd_vofrho

  SUBROUTINE micpmd_ewald(iesr, esr, eself, fion)
    INTEGER, INTENT(IN)         :: iesr
    REAL(real_8), INTENT(OUT)   :: esr, eself
    REAL(real_8), INTENT(INOUT) :: fion(:,:)
    CALL check(mc_ewald(ctx, INT(iesr, c_int), esr, eself, fion), 'ewald')
  END SUBROUTINE micpmd_ewald

  ! ------------------------------------------------------------------
  ! set-up of the module state CPMD's routines read by USE association
  SUBROUTINE micpmd_set_occupations(f)
    REAL(real_8), INTENT(IN) :: f(:)
    f_s = f
  END SUBROUTINE micpmd_set_occupations

  SUBROUTINE micpmd_set_cp_parameters(epsog, maxit)
    REAL(real_8), INTENT(IN) :: epsog
    INTEGER, INTENT(IN)      :: maxit
    epsog_s = epsog
    maxit_s = maxit
  END SUBROUTINE micpmd_set_cp_parameters

  ! 'FP64' (CPMD's arithmetic), 'FP32' (FP32 transforms) or 'MIXED' (FP32
  ! transforms and split-bf16 wavefunction contractions)
  SUBROUTINE micpmd_set_precision(mode)
    CHARACTER(*), INTENT(IN) :: mode
    INTEGER(c_int) :: m
    SELECT CASE (TRIM(mode))
    CASE ('FP64'); m = 0
    CASE ('FP32'); m = 1
    CASE ('MIXED'); m = 2
    CASE DEFAULT
       CALL fail('set_precision', 'mode must be FP64, FP32 or MIXED')
    END SELECT
    CALL check(mc_set_precision(ctx, m), 'set_precision')
  END SUBROUTINE micpmd_set_precision

  ! Device-resident drivers: micpmd_rhoofr_dev keeps the y-transformed
  ! columns of every state pair and the next micpmd_forces_c2_dev on the same
  ! wavefunctions starts from them (bit-identical results, one inverse z and
  ! y pass fewer per step).  The driver must not write the device
  ! wavefunctions outside the library between the two calls.
  SUBROUTINE micpmd_set_column_reuse(on)
    LOGICAL, INTENT(IN) :: on
    INTEGER(c_int) :: m
    m = 0
    IF (on) m = 1
    CALL check(mc_ctx_set_column_reuse(ctx, m), 'set_column_reuse')
  END SUBROUTINE micpmd_set_column_reuse

  ! 'LOWDIN' or 'GRAM-SCHMIDT' (CPMD's cntl%tlowd, ORTHOGONALIZATION LOWDIN)
  SUBROUTINE micpmd_set_ortho(method)
    CHARACTER(*), INTENT(IN) :: method
    lowdin_s = TRIM(method) == 'LOWDIN'
  END SUBROUTINE micpmd_set_ortho

  SUBROUTINE need_f(procedure_name, nstate)
    CHARACTER(*), INTENT(IN) :: procedure_name
    INTEGER, INTENT(IN)      :: nstate
    IF (.NOT. ALLOCATED(f_s)) CALL fail(procedure_name, 'occupations not set (micpmd_set_occupations)')
    IF (SIZE(f_s) < nstate) CALL fail(procedure_name, 'fewer occupations than states')
  END SUBROUTINE need_f

  ! ------------------------------------------------------------------
  ! rnlsm(c0,nstate,ikpt,ikind,tfor): fnl (and dfnl with tfor) into
  ! micpmd_fnl / micpmd_dfnl, CPMD's sfac arrays
  SUBROUTINE micpmd_rnlsm(c0, nstate, ikpt, ikind, tfor)
    COMPLEX(real_8), INTENT(IN) :: c0(:,:)
    INTEGER, INTENT(IN)         :: nstate, ikpt, ikind
    LOGICAL, INTENT(IN)         :: tfor
    IF (ikpt /= 1 .OR. ikind /= 1) CALL fail('rnlsm', 'only the Gamma point (ikpt = ikind = 1)')
    CALL size_fnl(nstate, tfor)
    CALL check(mc_rnlsm1(ctx, c0, INT(SIZE(c0, 1), c_int), INT(nstate, c_int), micpmd_fnl), 'rnlsm1')
    IF (tfor) CALL check(mc_rnlsm2(ctx, c0, INT(SIZE(c0, 1), c_int), INT(nstate, c_int), micpmd_dfnl), 'rnlsm2')
  END SUBROUTINE micpmd_rnlsm

  SUBROUTINE size_fnl(nstate, tfor)
    INTEGER, INTENT(IN) :: nstate
    LOGICAL, INTENT(IN) :: tfor
    IF (ALLOCATED(micpmd_fnl)) THEN
       IF (SIZE(micpmd_fnl, 3) /= nstate .OR. SIZE(micpmd_fnl, 1) /= nat_s) DEALLOCATE(micpmd_fnl)
    ENDIF
    IF (.NOT. ALLOCATED(micpmd_fnl)) ALLOCATE(micpmd_fnl(nat_s, nhxs_s, nstate))
    IF (tfor) THEN
       IF (ALLOCATED(micpmd_dfnl)) THEN
          IF (SIZE(micpmd_dfnl, 4) /= nstate .OR. SIZE(micpmd_dfnl, 1) /= nat_s) DEALLOCATE(micpmd_dfnl)
       ENDIF
       IF (.NOT. ALLOCATED(micpmd_dfnl)) ALLOCATE(micpmd_dfnl(nat_s, nhxs_s, 3, nstate))
    ENDIF
  END SUBROUTINE size_fnl

  ! ortho(nstate,c0,cscr): Gram-Schmidt (rgs) or Loewdin per micpmd_set_ortho
  SUBROUTINE micpmd_ortho(nstate, c0, cscr)
    INTEGER, INTENT(IN)            :: nstate
    COMPLEX(real_8), INTENT(INOUT) :: c0(:,:)
    COMPLEX(real_8)                :: cscr(:,:)
    REAL(real_8), ALLOCATABLE :: smat(:,:)
    IF (lowdin_s) THEN
       CALL check(mc_lowdin(ctx, c0, INT(SIZE(c0, 1), c_int), INT(nstate, c_int)), 'ortho')
    ELSE
       ALLOCATE(smat(nstate, nstate))
       CALL check(mc_rgs(ctx, c0, INT(SIZE(c0, 1), c_int), INT(nstate, c_int), smat), 'ortho')
    ENDIF
  END SUBROUTINE micpmd_ortho

  ! rortv(c0,cm,c2,sc0,gamy,nstate): the velocity constraint of the CP step
  SUBROUTINE micpmd_rortv(c0, cm, c2, sc0, gamy, nstate)
    COMPLEX(real_8), INTENT(IN)    :: c0(:,:)
    COMPLEX(real_8), INTENT(INOUT) :: cm(:,:)
    COMPLEX(real_8)                :: c2(:,:), sc0(:,:)
    REAL(real_8), INTENT(OUT)      :: gamy(nstate, nstate)
    INTEGER, INTENT(IN)            :: nstate
    CALL check(mc_rortv(ctx, c0, cm, INT(SIZE(c0, 1), c_int), gamy, INT(nstate, c_int)), 'rortv')
  END SUBROUTINE micpmd_rortv

  ! ------------------------------------------------------------------
  ! device-resident arrays (TYPE(c_ptr) handles in the library's layout)
  FUNCTION micpmd_alloc_wfn(nstate) RESULT(p)
    INTEGER, INTENT(IN) :: nstate
    TYPE(c_ptr) :: p
    INTEGER(c_size_t) :: bytes
    bytes = 16_c_size_t * mc_wfn_ld(ctx) * INT(MAX(nstate, 1), c_size_t)
    CALL check(mc_dev_alloc(bytes, p), 'alloc_wfn')
    CALL check(mc_memset_d(p, 0_c_int, bytes), 'alloc_wfn')
  END FUNCTION micpmd_alloc_wfn

  ! n doubles (rho, V: nnr1 = kr1 kr2 kr3)
  FUNCTION micpmd_alloc_real(n) RESULT(p)
    INTEGER, INTENT(IN) :: n
    TYPE(c_ptr) :: p
    INTEGER(c_size_t) :: bytes
    bytes = 8_c_size_t * INT(MAX(n, 1), c_size_t)
    CALL check(mc_dev_alloc(bytes, p), 'alloc_real')
    CALL check(mc_memset_d(p, 0_c_int, bytes), 'alloc_real')
  END FUNCTION micpmd_alloc_real

  SUBROUTINE micpmd_free(p)
    TYPE(c_ptr), INTENT(INOUT) :: p
    CALL check(mc_dev_free(p), 'free')
    p = c_null_ptr
  END SUBROUTINE micpmd_free

  SUBROUTINE micpmd_upload(c, nstate, p)
    COMPLEX(real_8), INTENT(IN) :: c(:,:)
    INTEGER, INTENT(IN)         :: nstate
    TYPE(c_ptr), INTENT(IN)     :: p
    CALL check(mc_wfn_upload(ctx, c, INT(SIZE(c, 1), c_int), INT(nstate, c_int), p), 'upload')
  END SUBROUTINE micpmd_upload

  SUBROUTINE micpmd_download(p, nstate, c)
    TYPE(c_ptr), INTENT(IN)      :: p
    INTEGER, INTENT(IN)          :: nstate
    COMPLEX(real_8), INTENT(OUT) :: c(:,:)
    CALL check(mc_wfn_download(ctx, p, INT(nstate, c_int), c, INT(SIZE(c, 1), c_int)), 'download')
  END SUBROUTINE micpmd_download

  SUBROUTINE micpmd_zero_wfn(p, nstate)
    TYPE(c_ptr), INTENT(IN) :: p
    INTEGER, INTENT(IN)     :: nstate
    CALL check(mc_wfn_zero(ctx, p, INT(nstate, c_int)), 'zero_wfn')
  END SUBROUTINE micpmd_zero_wfn

  ! phfac: eigr of the ions of micpmd_set_ions, built on the device
  SUBROUTINE micpmd_phfac()
    CALL check(mc_phfac(ctx), 'phfac')
  END SUBROUTINE micpmd_phfac

  ! rnlsm on device coefficients; fnl (and dfnl with tfor) stay on the device
  ! for fnonloc / rnlfor, and a host copy of fnl is left in micpmd_fnl (E_NL)
  SUBROUTINE micpmd_rnlsm_dev(c0, nstate, tfor)
    TYPE(c_ptr), INTENT(IN) :: c0
    INTEGER, INTENT(IN)     :: nstate
    LOGICAL, INTENT(IN)     :: tfor
    CALL size_fnl(nstate, .FALSE.)
    CALL check(mc_rnlsm_dev(ctx, c0, INT(nstate, c_int), MERGE(1_c_int, 0_c_int, tfor), C_LOC(micpmd_fnl), &
         c_null_ptr), 'rnlsm')
  END SUBROUTINE micpmd_rnlsm_dev

  SUBROUTINE micpmd_rhoofr_dev(c0, nstate, rho)
    TYPE(c_ptr), INTENT(IN) :: c0, rho
    INTEGER, INTENT(IN)     :: nstate
    CALL need_f('rhoofr', nstate)
    CALL check(mc_rhoofr_dev(ctx, c0, f_s, INT(nstate, c_int), rho, micpmd_ekin, micpmd_csumg, micpmd_csumr), &
         'rhoofr')
  END SUBROUTINE micpmd_rhoofr_dev

  ! ener(7) = ehep, ehii, ehee, epseu, vploc, exc, vxc; fion(3, nat) += potfor
  SUBROUTINE micpmd_vofrho_dev(rho, v, ener, fion)
    TYPE(c_ptr), INTENT(IN)     :: rho, v
    REAL(real_8), INTENT(OUT)   :: ener(7)
    REAL(real_8), INTENT(INOUT) :: fion(:,:)
    CALL check(mc_vofrho_dev(ctx, rho, v, ener, fion), 'vofrho')
  END SUBROUTINE micpmd_vofrho_dev

  ! c2 <- vpsi + fnonloc gradient of c0 (forces: zeroing(c2), vpsi, fnonloc)
  SUBROUTINE micpmd_forces_c2_dev(c0, c2, v, nstate)
    TYPE(c_ptr), INTENT(IN) :: c0, c2, v
    INTEGER, INTENT(IN)     :: nstate
    CALL need_f('forces', nstate)
    CALL check(mc_forces_c2_dev(ctx, c0, c2, f_s, v, INT(nstate, c_int)), 'forces')
  END SUBROUTINE micpmd_forces_c2_dev

  ! rnlfor with the device fnl / dfnl of the last micpmd_rnlsm_dev
  SUBROUTINE micpmd_rnlfor_dev(fion, nstate)
    REAL(real_8), INTENT(INOUT) :: fion(:,:)
    INTEGER, INTENT(IN)         :: nstate
    CALL need_f('rnlfor', nstate)
    CALL check(mc_rnlfor(ctx, f_s, INT(nstate, c_int), c_null_ptr, c_null_ptr, fion), 'rnlfor')
  END SUBROUTINE micpmd_rnlfor_dev

  ! velupa (cm += dtm c2) + posupa + rortog; cm becomes (cp - c0) / dt
  SUBROUTINE micpmd_posupa_dev(c0, c2, cm, cp, nstate, dtm, dt)
    TYPE(c_ptr), INTENT(IN) :: c0, c2, cm, cp
    INTEGER, INTENT(IN)     :: nstate
    REAL(real_8), INTENT(IN) :: dtm, dt
    INTEGER(c_int) :: iters
    CALL check(mc_posupa_dev(ctx, c0, c2, cm, cp, INT(nstate, c_int), dtm, dt, epsog_s, INT(maxit_s, c_int), &
         c_null_ptr, iters), 'posupa')
    micpmd_iters = iters
  END SUBROUTINE micpmd_posupa_dev

  ! velupa (cm += dtm c2) + rortv + rekine: the end of the velocity Verlet step
  SUBROUTINE micpmd_velupa_rortv_dev(c0, c2, cm, nstate, dtm, emass, ekinc)
    TYPE(c_ptr), INTENT(IN)   :: c0, c2, cm
    INTEGER, INTENT(IN)       :: nstate
    REAL(real_8), INTENT(IN)  :: dtm, emass
    REAL(real_8), INTENT(OUT) :: ekinc
    CALL check(mc_cp_velupa_rortv_dev(ctx, c0, c2, cm, INT(nstate, c_int), dtm, emass, c_null_ptr, ekinc), &
         'velupa_rortv')
  END SUBROUTINE micpmd_velupa_rortv_dev

  SUBROUTINE micpmd_lagrange_dev(c0, c2, nstate)
    TYPE(c_ptr), INTENT(IN) :: c0, c2
    INTEGER, INTENT(IN)     :: nstate
    CALL need_f('lagrange', nstate)
    CALL check(mc_lagrange_dev(ctx, c0, c2, f_s, INT(nstate, c_int)), 'lagrange')
  END SUBROUTINE micpmd_lagrange_dev

  SUBROUTINE micpmd_rgs_dev(c, nstate)
    TYPE(c_ptr), INTENT(IN) :: c
    INTEGER, INTENT(IN)     :: nstate
    CALL check(mc_rgs_dev(ctx, c, INT(nstate, c_int), c_null_ptr), 'rgs')
  END SUBROUTINE micpmd_rgs_dev

  SUBROUTINE micpmd_hesele(prec, svar2, vploc, hthrs)
    LOGICAL, INTENT(IN)      :: prec
    REAL(real_8), INTENT(IN) :: svar2, vploc, hthrs
    CALL check(mc_hesele(ctx, MERGE(1_c_int, 0_c_int, prec), svar2, vploc, hthrs), 'hesele')
  END SUBROUTINE micpmd_hesele

  SUBROUTINE micpmd_csize_dev(c2, nstate, gemax, cnorm)
    TYPE(c_ptr), INTENT(IN)   :: c2
    INTEGER, INTENT(IN)       :: nstate
    REAL(real_8), INTENT(OUT) :: gemax, cnorm
    CALL need_f('csize', nstate)
    CALL check(mc_csize_dev(ctx, c2, f_s, INT(nstate, c_int), gemax, cnorm), 'csize')
  END SUBROUTINE micpmd_csize_dev

  SUBROUTINE micpmd_odiis_dev(c0, c2, nstate, etot, svar2, mdiis, nreset, reinit)
    TYPE(c_ptr), INTENT(IN)  :: c0, c2
    INTEGER, INTENT(IN)      :: nstate, mdiis, nreset
    REAL(real_8), INTENT(IN) :: etot, svar2
    LOGICAL, INTENT(INOUT)   :: reinit
    INTEGER(c_int) :: r
    CALL need_f('odiis', nstate)
    r = MERGE(1_c_int, 0_c_int, reinit)
    CALL check(mc_odiis_dev(ctx, c0, c2, f_s, INT(nstate, c_int), etot, svar2, INT(mdiis, c_int), &
         INT(nreset, c_int), r), 'odiis')
    reinit = r /= 0
  END SUBROUTINE micpmd_odiis_dev

  ! E_NL = sum_i f_i sum_isa sum_iv,jv h(iv,jv) fnl(isa,iv,i) fnl(isa,jv,i)
  ! from micpmd_fnl (CPMD's enl of rnlsm / fnonloc)
  FUNCTION micpmd_enl(nstate) RESULT(e)
    INTEGER, INTENT(IN) :: nstate
    REAL(real_8) :: e, s
    INTEGER :: is, ia, isa0, iv, jv, i
    CALL need_f('enl', nstate)
    e = 0.0_real_8
    isa0 = 0
    DO is = 1, SIZE(na_s)
       DO iv = 1, ngh_s(is)
          DO jv = 1, ngh_s(is)
             IF (hmat_s(iv, jv, is) == 0.0_real_8) CYCLE
             DO i = 1, nstate
                s = 0.0_real_8
                DO ia = isa0 + 1, isa0 + na_s(is)
                   s = s + micpmd_fnl(ia, iv, i) * micpmd_fnl(ia, jv, i)
                ENDDO
                e = e + hmat_s(iv, jv, is) * s * f_s(i)
             ENDDO
          ENDDO
       ENDDO
       isa0 = isa0 + na_s(is)
    ENDDO
  END FUNCTION micpmd_enl

  ! rnlrh(enl,nstate,nkpoint), src/rnlrh_utils.mod.F90:43: E_NL from the device
  ! fnl of the last micpmd_rnlsm_dev, summed on the device
  FUNCTION micpmd_rnlrh_dev(nstate) RESULT(e)
    INTEGER, INTENT(IN) :: nstate
    REAL(real_8) :: e
    CALL need_f('rnlrh', nstate)
    CALL check(mc_rnlrh(ctx, f_s, INT(nstate, c_int), e), 'rnlrh')
  END FUNCTION micpmd_rnlrh_dev

END MODULE micpmd
