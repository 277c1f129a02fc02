! Runs CPMD's velocity Verlet Car-Parrinello step on the device through the
! micpmd Fortran module, in the order of src/mdmain_utils.mod.F90:529-966
! (forcedr at the start, rortv, rekine; then per step velupi + velupa +
! posupi + phfac + posupa, forcedr, velupi + velupa + rortv, rekine), in
! the precision given on the command line, and writes what it computed:
!   cp_driver.x IN_FILE OUT_FILE MODE NSTEPS
! The input (stream file) holds the layout, the density G set, the local and
! nonlocal pseudopotential data, the ions and the coefficients.
PROGRAM cp_driver
  USE, INTRINSIC :: iso_c_binding
  USE micpmd
  IMPLICIT NONE
  INTEGER, PARAMETER :: dp = c_double
  INTEGER(c_int) :: hdr(12)
  INTEGER :: nr1, nr2, nr3, kr1, kr2, kr3, ngw, nhg, nstate, nat, iesr, nsteps, nnr1, u, k, step
  REAL(dp) :: par(8), tpiba2, omega, dt, emass, mass, h11, zv(1), raggio(1)
  INTEGER(c_int), ALLOCATABLE :: inyh(:,:), na(:), ngh(:), lang(:,:)
  REAL(dp), ALLOCATABLE :: hg(:), gk(:,:), vps(:,:), rhops(:,:), twnl(:,:,:), hmat(:,:,:), tau(:,:), f(:)
  REAL(dp), ALLOCATABLE :: fion(:,:), vion(:,:), floc(:,:), fesr(:,:)
  REAL(dp) :: ht(3, 3), ener(7), esr, eself, etot, ekinc
  COMPLEX(dp), ALLOCATABLE :: c0(:,:), cm(:,:)
  TYPE(c_ptr) :: d_c0, d_cm, d_cp, d_c2, d_rho, d_v, t
  CHARACTER(len=512) :: fin, fout, arg

  CALL get_command_argument(1, fin)
  CALL get_command_argument(2, fout)
  CALL get_command_argument(3, arg)
  CALL get_command_argument(4, arg(257:))
  READ(arg(257:), *) nsteps
  OPEN(newunit=u, file=TRIM(fin), access='stream', form='unformatted', status='old')
  READ(u) hdr
  nr1 = hdr(1); nr2 = hdr(2); nr3 = hdr(3); kr1 = hdr(4); kr2 = hdr(5); kr3 = hdr(6)
  ngw = hdr(7); nhg = hdr(8); nstate = hdr(9); nat = hdr(10); iesr = hdr(11)
  READ(u) par
  tpiba2 = par(1); omega = par(2); dt = par(3); emass = par(4); mass = par(5); h11 = par(6)
  zv(1) = par(7); raggio(1) = par(8)
  ALLOCATE(inyh(3, nhg), hg(nhg), gk(3, nhg), vps(1, nhg), rhops(1, nhg), twnl(ngw, 1, 1), tau(3, nat))
  ALLOCATE(c0(ngw, nstate), cm(ngw, nstate), f(nstate))
  READ(u) inyh, hg, gk, vps, rhops, twnl, tau, ht, c0, f
  CLOSE(u)
  nnr1 = kr1 * kr2 * kr3

  ! the set-up CPMD does once per run / cell (fftprp, the pseudopotential
  ! arrays, crge%f, cntr%epsog / cnti%maxit)
  CALL micpmd_init(nr1, nr2, nr3, kr1, kr2, kr3, ngw, inyh(:, 1:ngw), hg(1:ngw), tpiba2, omega, .TRUE.)
  ALLOCATE(na(1), ngh(1), lang(1, 1), hmat(1, 1, 1))
  na = nat; ngh = 1; lang = 0; hmat = h11
  CALL micpmd_set_projectors(na, 1, ngh, lang, hmat, twnl)
  CALL micpmd_set_gk(gk(:, 1:ngw))
  CALL micpmd_set_density(nhg, inyh, hg, gk)
  CALL micpmd_set_local_pp(na, zv, raggio, vps, rhops)
  CALL micpmd_set_occupations(f)
  CALL micpmd_set_cp_parameters(1.0e-8_dp, 60)
  CALL micpmd_set_precision(TRIM(arg(1:256)))
  CALL micpmd_set_column_reuse(.TRUE.)

  d_c0 = micpmd_alloc_wfn(nstate); d_cm = micpmd_alloc_wfn(nstate)
  d_cp = micpmd_alloc_wfn(nstate); d_c2 = micpmd_alloc_wfn(nstate)
  d_rho = micpmd_alloc_real(nnr1); d_v = micpmd_alloc_real(nnr1)
  CALL micpmd_upload(c0, nstate, d_c0)
  ALLOCATE(fion(3, nat), vion(3, nat), floc(3, nat), fesr(3, nat))
  vion = 0.0_dp

  OPEN(newunit=u, file=TRIM(fout), access='stream', form='unformatted', status='replace')
  ! start: forcedr, rortv (cm = 0), rekine
  CALL forces()
  CALL micpmd_velupa_rortv_dev(d_c0, c_null_ptr, d_cm, nstate, 0.0_dp, emass, ekinc)
  WRITE(u) etot, fion, ekinc
  DO step = 1, nsteps
     vion = vion + (0.5_dp * dt / mass) * fion                                ! velupi
     tau = tau + dt * vion                                                     ! posupi
     CALL micpmd_posupa_dev(d_c0, d_c2, d_cm, d_cp, nstate, 0.5_dp * dt / emass, dt)   ! velupa + posupa
     t = d_c0; d_c0 = d_cp; d_cp = t
     CALL forces()                                                             ! forcedr
     vion = vion + (0.5_dp * dt / mass) * fion                                ! velupi
     CALL micpmd_velupa_rortv_dev(d_c0, d_c2, d_cm, nstate, 0.5_dp * dt / emass, emass, ekinc)
     WRITE(u) etot, fion, ekinc, micpmd_iters
  ENDDO
  CALL micpmd_download(d_c0, nstate, c0)
  CALL micpmd_download(d_cm, nstate, cm)
  WRITE(u) tau, vion, c0, cm
  CLOSE(u)
  DO k = 1, 1
     CALL micpmd_free(d_c0); CALL micpmd_free(d_cm); CALL micpmd_free(d_cp); CALL micpmd_free(d_c2)
     CALL micpmd_free(d_rho); CALL micpmd_free(d_v)
  ENDDO
  CALL micpmd_finalize()
  WRITE(6,'(A)') ' cp_driver: done'

CONTAINS

  ! forcedr: phfac, rnlsm (fnl, dfnl), rhoofr, vofrho (+ potfor), Ewald,
  ! vpsi + fnonloc (c2), rnlfor; E_KS = ekin + ehep + epseu + exc + esr - eself + enl
  SUBROUTINE forces()
    REAL(dp) :: enl
    CALL micpmd_set_ions(tau, ht)
    CALL micpmd_phfac()
    CALL micpmd_rnlsm_dev(d_c0, nstate, .TRUE.)
    CALL micpmd_rhoofr_dev(d_c0, nstate, d_rho)
    floc = 0.0_dp
    CALL micpmd_vofrho_dev(d_rho, d_v, ener, floc)
    fesr = 0.0_dp
    CALL micpmd_ewald(iesr, esr, eself, fesr)
    enl = micpmd_rnlrh_dev(nstate)
    IF (ABS(enl - micpmd_enl(nstate)) > 1.0e-10_dp * MAX(1.0_dp, ABS(enl))) ERROR STOP 'rnlrh: device E_NL differs'
    etot = micpmd_ekin + ener(1) + ener(4) + ener(6) + esr - eself + enl
    CALL micpmd_forces_c2_dev(d_c0, d_c2, d_v, nstate)
    fion = floc + fesr
    CALL micpmd_rnlfor_dev(fion, nstate)
  END SUBROUTINE forces

END PROGRAM cp_driver
