! Driver for the flang build of the reference CPU path (test infrastructure).
!
! Reads one case file (stream-unformatted, little-endian, layout written by
! gms_liberi_amd/caseio.py), calls the reference handle API exactly as
! tests/fortran/test_h2_sonly.F90:160-181 of the reference does
! (liberi_create -> liberi_setup -> liberi_fock_build -> liberi_destroy), and
! writes the packed Fock matrix of the last build as raw float64.
!
! usage: liberi_ref_run CASE_FILE FOCK_OUT [NREP]
!   prints one line per build: "build <k> seconds <t>"
program liberi_ref_run
  use liberi_types, only: dp, int32, int64
  use liberi_interface, only: liberi_handle_t, liberi_create, liberi_setup, &
                              liberi_fock_build, liberi_destroy
  use omp_lib, only: omp_get_wtime, omp_get_max_threads
  implicit none

  type(liberi_handle_t) :: handle
  integer :: u, nsh, natoms, num_bas, mxgtot, ntri, nrep, k, magic, version
  integer, allocatable :: ang_mom(:), contr_num(:), sh_loc(:), atom_num(:), &
                          atom_loc(:), start_bas(:), end_bas(:)
  real(dp), allocatable :: expo(:), cs(:), cp(:), cd(:), cf(:), coords(:), &
                           schwrz(:), dens(:), fock(:)
  character(len=1024) :: fin, fout, arg
  real(dp) :: t0, t1

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  nrep = 1
  if (command_argument_count() >= 3) then
    call get_command_argument(3, arg)
    read (arg, *) nrep
  end if

  open (newunit=u, file=trim(fin), access='stream', form='unformatted', &
        status='old', action='read')
  read (u) magic, version, nsh, natoms, num_bas, mxgtot
  if (magic /= 1230128460) then
    write (*, *) 'bad case file magic'
    error stop 2
  end if
  allocate (ang_mom(nsh), contr_num(nsh), sh_loc(nsh), atom_num(nsh), &
            atom_loc(nsh), start_bas(nsh), end_bas(nsh))
  allocate (expo(mxgtot*nsh), cs(mxgtot*nsh), cp(mxgtot*nsh), &
            cd(mxgtot*nsh), cf(mxgtot*nsh))
  allocate (coords(3*natoms), schwrz(nsh*(nsh + 1)/2))
  ntri = num_bas*(num_bas + 1)/2
  allocate (dens(ntri), fock(ntri))
  read (u) ang_mom, contr_num, sh_loc, atom_num, atom_loc, start_bas, end_bas
  read (u) expo, cs, cp, cd, cf, coords, schwrz, dens
  close (u)

  call liberi_create(handle)
  call liberi_setup(handle, nsh, natoms, num_bas, mxgtot, ang_mom, contr_num, &
                    sh_loc, atom_num, atom_loc, start_bas, end_bas, expo, cs, &
                    cp, cd, cf, coords, schwrz, 0, 1)
  write (*, '(a,i0)') 'threads ', omp_get_max_threads()
  do k = 1, nrep
    t0 = omp_get_wtime()
    call liberi_fock_build(handle, dens, fock, ntri)
    t1 = omp_get_wtime()
    write (*, '(a,i0,a,es16.8)') 'build ', k, ' seconds ', t1 - t0
    flush (6)
  end do
  call liberi_destroy(handle)

  open (newunit=u, file=trim(fout), access='stream', form='unformatted', &
        status='replace', action='write')
  write (u) fock
  close (u)
end program liberi_ref_run
