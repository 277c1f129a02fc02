! Reads a RESTART file with the Fortran runtime the way CPMD's rv30 /
! r_wfnio / rd30wfn read it (src/rv30_utils.mod.F90:286-735,
! src/wfnio_utils.mod.F90:235-330: sequential unformatted, one READ per
! record, marker records read by their first item only) and copies every
! value it read, in order, to a stream file for the test to compare:
!   restart_read.x RESTART_FILE OUT_FILE
program restart_read
  implicit none
  character(len=512) :: fin, fout
  character(len=80) :: str
  integer :: irecord, ibrav, indpg, nsp, isec, ia, i, k, nat
  integer :: nel, nr1, nr2, nr3, n, nkpts0, ngwks, ngwls, nhgs, nhgls
  integer :: iwfntype, n0, ngw0, icompr, na1, nb1, nx, ios
  logical :: dual, ltlsd
  real(8) :: celldm(6), t(3), ecut, cdual, scale
  integer, allocatable :: na(:), icmp(:)
  real(8), allocatable :: f(:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), form='unformatted', access='sequential', status='old')
  open(11, file=trim(fout), form='unformatted', access='stream', status='replace')
  read(10) str
  if (index(str, 'STUTTGART VERSION 3.0') == 0) stop 2
  ! section 2
  read(10) irecord
  if (irecord /= 2) stop 3
  read(10) ibrav, indpg
  read(10) (celldm(k), k = 1, 6)
  write(11) ibrav, indpg, celldm
  ! section 3
  read(10) irecord
  if (irecord /= 2) stop 4
  read(10) nsp
  allocate(na(nsp))
  read(10) (na(k), k = 1, nsp)
  write(11) nsp, na
  nat = sum(na)
  ! sections 4, 5, 6
  do isec = 4, 6
    read(10) irecord
    write(11) irecord
    do ia = 1, irecord
      read(10) (t(k), k = 1, 3)
      write(11) t
    end do
  end do
  ! section 7
  read(10) irecord
  if (irecord /= 1) stop 5
  read(10) ecut, cdual, dual, nel, nr1, nr2, nr3
  write(11) ecut, cdual, merge(1, 0, dual), nel, nr1, nr2, nr3
  ! section 8
  read(10) irecord
  if (irecord /= 1) stop 6
  read(10) n, nkpts0, ngwks, ngwls, nhgs, nhgls
  write(11) n, nkpts0, ngwks, ngwls, nhgs, nhgls
  ! section 9 (w_wfnio layout) and 10 (wr30wfn layout)
  do isec = 9, 10
    read(10) irecord
    write(11) irecord
    if (irecord == 0) cycle
    if (irecord < 0) then
      read(10) iwfntype
      if (iwfntype /= 1) stop 7
    end if
    read(10) n0, ngw0, icompr, scale
    write(11) n0, ngw0, icompr, scale
    if (irecord < 0) then
      read(10) ltlsd, na1, nb1
      write(11) merge(1, 0, ltlsd), na1, nb1
    end if
    allocate(icmp(4 * ngw0))
    do i = 1, n0
      read(10) nx
      if (nx /= 2 * ngw0) stop 8
      read(10) (icmp(k), k = 1, 2 * nx)
      write(11) icmp
    end do
    deallocate(icmp)
  end do
  ! sections 11..22
  do isec = 11, 22
    read(10) irecord
    do i = 1, irecord
      read(10)
    end do
  end do
  ! section 23
  read(10) irecord
  write(11) irecord
  if (irecord == 1) then
    allocate(f(n))
    read(10) (f(k), k = 1, n)
    write(11) f
  end if
  ! sections 24..99
  do isec = 24, 99
    read(10) irecord
    if (irecord /= 0) stop 9
  end do
  read(10, iostat=ios) irecord
  if (ios >= 0) stop 10
  write(11) -12345
  close(11)
  close(10)
end program restart_read
