! GAMESS-style caller of the drop-in modules (gamess_liberi_wrapper).
!
! usage: test_gamess_wrapper CASE_FILE FOCK_OUT          -> liberi_init, liberi_fock, liberi_finalize
!        test_gamess_wrapper CASE_FILE FOCK_OUT before    -> liberi_fock before liberi_init (must error stop 1)
!        test_gamess_wrapper CASE_FILE FOCK_OUT size      -> liberi_fock with a wrong n (must error stop 1)
! The case file layout is gms_liberi_amd/caseio.py's stream form.
program test_gamess_wrapper
  use liberi_types, only: dp
  use gamess_liberi_wrapper, only: liberi_init, liberi_fock, liberi_finalize
  implicit none
  integer :: u, magic, version, nsh, natoms, num_bas, mxgtot, ntri
  integer, allocatable :: ang(:), con(:), loc(:), anum(:), aloc(:), sb(:), eb(:)
  real(dp), allocatable :: ex(:), cs(:), cp(:), cd(:), cf(:), xyz(:), sw(:), d(:), f(:)
  character(len=1024) :: fin, fout, mode

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  mode = ''
  if (command_argument_count() >= 3) call get_command_argument(3, mode)
  open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
  read (u) magic, version, nsh, natoms, num_bas, mxgtot
  allocate (ang(nsh), con(nsh), loc(nsh), anum(nsh), aloc(nsh), sb(nsh), eb(nsh))
  allocate (ex(mxgtot*nsh), cs(mxgtot*nsh), cp(mxgtot*nsh), cd(mxgtot*nsh), cf(mxgtot*nsh))
  allocate (xyz(3*natoms), sw(nsh*(nsh + 1)/2))
  ntri = num_bas*(num_bas + 1)/2
  allocate (d(ntri), f(ntri))
  read (u) ang, con, loc, anum, aloc, sb, eb
  read (u) ex, cs, cp, cd, cf, xyz, sw, d
  close (u)

  if (trim(mode) == 'before') call liberi_fock(d, f, ntri)
  call liberi_init(natoms, num_bas, nsh, mxgtot, ang, con, loc, anum, aloc, sb, eb, ex, cs, cp, cd, cf, &
                   xyz, sw, 0, 1)
  if (trim(mode) == 'size') call liberi_fock(d, f, ntri - 1)
  f = 0.0_dp
  call liberi_fock(d, f, ntri)
  call liberi_finalize()
  open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
  write (u) f
  close (u)
end program test_gamess_wrapper
