! Drives the micpmd Fortran module with CPMD's routine signatures: reads a
! seeded system from a stream file, calls rhoofr / vpsi / rgs / lowdin /
! rnlsm1 / fnonloc through the shim and writes the results for comparison
! with the oracle.
PROGRAM shim_driver
  USE, INTRINSIC :: iso_c_binding
  USE micpmd
  IMPLICIT NONE
  INTEGER, PARAMETER :: dp = c_double
  INTEGER(c_int) :: hdr(9)
  INTEGER :: nr1, nr2, nr3, kr1, kr2, kr3, ngw, nstate, nat, nnr1
  REAL(dp) :: tp(2), h11(1)
  INTEGER(c_int), ALLOCATABLE :: inyh(:,:), na(:), ngh(:), lang(:,:)
  REAL(dp), ALLOCATABLE :: hg(:), f(:), vpot(:,:), rhoe(:,:), smat(:,:), twnl(:,:,:), hmat(:,:,:), fnl(:,:,:)
  COMPLEX(dp), ALLOCATABLE :: c0(:,:), c2(:,:), cp(:,:), cl(:,:), cn(:,:), psi(:), eigr(:,:), c2nl(:,:)
  CHARACTER(len=512) :: fin, fout
  INTEGER :: u

  CALL get_command_argument(1, fin)
  CALL get_command_argument(2, fout)
  OPEN(newunit=u, file=TRIM(fin), access='stream', form='unformatted', status='old')
  READ(u) hdr
  nr1 = hdr(1); nr2 = hdr(2); nr3 = hdr(3); kr1 = hdr(4); kr2 = hdr(5); kr3 = hdr(6)
  ngw = hdr(7); nstate = hdr(8); nat = hdr(9)
  nnr1 = kr1 * kr2 * kr3
  READ(u) tp
  ALLOCATE(inyh(3, ngw), hg(ngw), c0(ngw, nstate), f(nstate), vpot(nnr1, 1), cp(ngw, nstate))
  ALLOCATE(twnl(ngw, 1, 1), eigr(ngw, nat))
  READ(u) inyh, hg, c0, f, vpot, cp, twnl, h11, eigr
  CLOSE(u)

  CALL micpmd_init(nr1, nr2, nr3, kr1, kr2, kr3, ngw, inyh, hg, tp(1), tp(2), .TRUE.)

  ALLOCATE(rhoe(nnr1, 1), psi(1), c2(ngw, nstate), smat(nstate, nstate), cl(ngw, nstate), cn(ngw, nstate))
  CALL micpmd_set_occupations(f)
  CALL micpmd_rhoofr(c0, rhoe, psi, nstate)
  c2 = (0.0_dp, 0.0_dp)
  CALL micpmd_vpsi(c0, c2, f, vpot, psi, nstate, 1, 1, .FALSE.)
  cl = cp
  CALL micpmd_rgs(cp, nstate, smat)
  CALL micpmd_lowdin(cl, cn, nstate)

  ALLOCATE(na(1), ngh(1), lang(1, 1), hmat(1, 1, 1), fnl(nat, 1, nstate), c2nl(ngw, nstate))
  na = nat; ngh = 1; lang = 0; hmat = h11(1)
  CALL micpmd_set_projectors(na, 1, ngh, lang, hmat, twnl)
  CALL micpmd_set_eigr(eigr)
  CALL micpmd_rnlsm1(c0, nstate, fnl)
  c2nl = (0.0_dp, 0.0_dp)
  CALL micpmd_fnonloc(c2nl, f, nstate, fnl)
  CALL micpmd_finalize()

  OPEN(newunit=u, file=TRIM(fout), access='stream', form='unformatted', status='replace')
  WRITE(u) rhoe, micpmd_ekin, micpmd_csumg, micpmd_csumr, c2, cp, smat, cl, fnl, c2nl
  CLOSE(u)
  WRITE(6,'(A)') ' shim_driver: done'
END PROGRAM shim_driver
