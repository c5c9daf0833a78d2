! Input-capture stand-in for the reference handle API (fixture tooling only).
!
! The reference's own regression programs (tests/fortran/test_*.F90 of
! gms_libERI) build their basis sets in code.  Compiling one of them against
! these two modules instead of the library records exactly the arrays it
! passes to liberi_setup / liberi_fock_build
! (src/interface/liberi_interface.F90:72-129, :202-217) into a case stream
! (layout: gms_liberi_amd/caseio.py) named by $LIBERI_CAPTURE_FILE.  It
! computes nothing: the program's own comparison then fails, which the
! capture script ignores.
module liberi_types
  use, intrinsic :: iso_fortran_env, only: int32, int64
  implicit none
  integer, parameter :: sp = selected_real_kind(6, 37)
  integer, parameter :: dp = selected_real_kind(15, 307)
end module liberi_types

module liberi_interface
  use liberi_types, only: dp, int32
  implicit none
  private
  public :: liberi_handle_t, liberi_create, liberi_destroy, liberi_setup, &
            liberi_fock_build, liberi_cleanup

  type :: liberi_handle_t
    integer :: nsh = 0, natoms = 0, num_bas = 0, mxgtot = 0
    integer, allocatable :: iv(:, :)
    real(dp), allocatable :: expo(:), cs(:), cp(:), cd(:), cf(:), coords(:), sw(:)
    logical :: captured = .false.
  end type liberi_handle_t

contains

  subroutine liberi_create(handle)
    type(liberi_handle_t), intent(out) :: handle
    handle%captured = .false.
  end subroutine liberi_create

  subroutine liberi_destroy(handle)
    type(liberi_handle_t), intent(inout) :: handle
    handle%nsh = 0
  end subroutine liberi_destroy

  subroutine liberi_cleanup(handle)
    type(liberi_handle_t), intent(inout) :: handle
    handle%nsh = 0
  end subroutine liberi_cleanup

  subroutine liberi_setup(handle, nsh, natoms, num_bas, mxgtot, ang_mom, &
                          contr_num, sh_loc, atom_num, atom_loc, start_bas, &
                          end_bas, exponents, cs, cp, cd, cf, coords, &
                          schwrz_int, my_rank, num_procs)
    type(liberi_handle_t), intent(inout) :: handle
    integer, intent(in) :: nsh, natoms, num_bas, mxgtot, my_rank, num_procs
    integer, intent(in) :: ang_mom(nsh), contr_num(nsh), sh_loc(nsh), &
                           atom_num(nsh), atom_loc(nsh), start_bas(nsh), end_bas(nsh)
    real(dp), intent(in) :: exponents(mxgtot*nsh), cs(mxgtot*nsh), cp(mxgtot*nsh), &
                            cd(mxgtot*nsh), cf(mxgtot*nsh), coords(3*natoms), &
                            schwrz_int(nsh*(nsh + 1)/2)
    handle%nsh = nsh; handle%natoms = natoms
    handle%num_bas = num_bas; handle%mxgtot = mxgtot
    if (allocated(handle%iv)) deallocate (handle%iv)
    allocate (handle%iv(nsh, 7))
    handle%iv(:, 1) = ang_mom; handle%iv(:, 2) = contr_num
    handle%iv(:, 3) = sh_loc; handle%iv(:, 4) = atom_num
    handle%iv(:, 5) = atom_loc; handle%iv(:, 6) = start_bas
    handle%iv(:, 7) = end_bas
    handle%expo = exponents; handle%cs = cs; handle%cp = cp
    handle%cd = cd; handle%cf = cf; handle%coords = coords
    handle%sw = schwrz_int
    if (my_rank < 0 .or. num_procs < 1) stop 3
  end subroutine liberi_setup

  subroutine liberi_fock_build(handle, density, fock, n)
    type(liberi_handle_t), intent(inout) :: handle
    integer, intent(in) :: n
    real(dp), intent(in) :: density(n)
    real(dp), intent(inout) :: fock(n)
    character(len=1024) :: fname
    integer :: u, k
    fock = 0.0_dp
    if (handle%captured) return
    call get_environment_variable('LIBERI_CAPTURE_FILE', fname)
    open (newunit=u, file=trim(fname), access='stream', form='unformatted', &
          status='replace', action='write')
    write (u) int(1230128460, int32), int(1, int32), int(handle%nsh, int32), &
      int(handle%natoms, int32), int(handle%num_bas, int32), int(handle%mxgtot, int32)
    do k = 1, 7
      write (u) int(handle%iv(:, k), int32)
    end do
    write (u) handle%expo, handle%cs, handle%cp, handle%cd, handle%cf, &
      handle%coords, handle%sw, density
    close (u)
    handle%captured = .true.
  end subroutine liberi_fock_build

end module liberi_interface
