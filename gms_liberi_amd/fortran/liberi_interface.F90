! Fortran drop-in for gms_libERI's handle API over the HIP C ABI.
!
! Same module names, procedure names, argument lists and kinds as the
! reference's src/liberi_types.f90 (kinds), src/interface/liberi_interface.F90
! (handle API, :23-262) and src/interface/gms_liberi_wrapper.F90 (GAMESS
! entry points, :29-146), so callers recompile against these modules and link
! libliberi_fortran.so + libliberi_hip.so instead of the reference library.
! Default-kind integers of the caller (e.g. GAMESS built with -i8) are
! converted to the C ABI's int32/int64 here.  As in the reference, a Fock
! build before setup / init and a wrong matrix size stop the program with
! `error stop 1`; any other failure of the GPU library does too, with its
! message.
module liberi_types
  use, intrinsic :: iso_fortran_env, only: int32, int64
  implicit none
  private
  public :: int32, int64
  integer, parameter, public :: sp = selected_real_kind(6, 37)
  integer, parameter, public :: dp = selected_real_kind(15, 307)
  integer, parameter, public :: default_real = dp
end module liberi_types

module liberi_c_abi
  ! bind(C) declarations of include/liberi_hip.h
  use, intrinsic :: iso_c_binding
  implicit none
  public
  interface
    integer(c_int) function liberi_c_create(ctx) bind(C, name='liberi_c_create')
      import :: c_int, c_ptr
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function liberi_c_destroy(ctx) bind(C, name='liberi_c_destroy')
      import :: c_int, c_ptr
      type(c_ptr), intent(inout) :: ctx
    end function
    integer(c_int) function liberi_c_cleanup(ctx) bind(C, name='liberi_c_cleanup')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function liberi_c_setup(ctx, nsh, natoms, num_bas, mxgtot, ang_mom, contr_num, sh_loc, &
                                           atom_num, atom_loc, start_bas, end_bas, exponents, cs, cp, cd, cf, &
                                           coords, schwrz, my_rank, num_procs) bind(C, name='liberi_c_setup')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsh, natoms, num_bas, mxgtot, my_rank, num_procs
      integer(c_int32_t), intent(in) :: ang_mom(*), contr_num(*), sh_loc(*), atom_num(*), atom_loc(*), &
                                        start_bas(*), end_bas(*)
      real(c_double), intent(in) :: exponents(*), cs(*), cp(*), cd(*), cf(*), coords(*), schwrz(*)
    end function
    integer(c_int) function liberi_c_fock_build(ctx, density, fock, n) bind(C, name='liberi_c_fock_build')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: density(*)
      real(c_double), intent(inout) :: fock(*)
      integer(c_int64_t), value :: n
    end function
    type(c_ptr) function liberi_c_last_error() bind(C, name='liberi_c_last_error')
      import :: c_ptr
    end function
  end interface
contains
  function last_error() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: s(:)
    integer :: n
    type(c_ptr) :: p
    p = liberi_c_last_error()
    call c_f_pointer(p, s, [4096])
    n = 0
    do while (n < 4096)
      if (s(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate (character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function last_error

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= 0) then
      write (*, *) 'ERROR: ', what, ': ', last_error()
      error stop 1
    end if
  end subroutine check
end module liberi_c_abi

module liberi_interface
  use, intrinsic :: iso_c_binding
  use liberi_types, only: int32, int64, dp
  use liberi_c_abi
  implicit none
  private

  public :: liberi_handle_t
  public :: liberi_create
  public :: liberi_destroy
  public :: liberi_setup
  public :: liberi_fock_build
  public :: liberi_cleanup

  type :: liberi_handle_t
    private
    type(c_ptr) :: impl = c_null_ptr
    logical :: is_initialized = .false.
  end type liberi_handle_t

contains

  subroutine liberi_create(handle)
    type(liberi_handle_t), intent(out) :: handle
    call check(liberi_c_create(handle%impl), 'liberi_create')
    handle%is_initialized = .false.
  end subroutine liberi_create

  subroutine liberi_destroy(handle)
    type(liberi_handle_t), intent(inout) :: handle
    if (c_associated(handle%impl)) call check(liberi_c_destroy(handle%impl), 'liberi_destroy')
    handle%impl = c_null_ptr
    handle%is_initialized = .false.
  end subroutine liberi_destroy

  subroutine liberi_setup(handle, nsh_in, natoms_in, num_bas_in, mxgtot_in, &
                          ang_mom_in, contr_num_in, sh_loc_in, &
                          atom_num_in, atom_loc_in, &
                          start_bas_in, end_bas_in, &
                          exponents_in, &
                          contr_coef_s_in, contr_coef_p_in, &
                          contr_coef_d_in, contr_coef_f_in, &
                          coords_in, schwrz_int_in, &
                          my_rank_in, num_procs_in)
    type(liberi_handle_t), intent(inout) :: handle
    integer, intent(in) :: nsh_in, natoms_in, num_bas_in, mxgtot_in, my_rank_in, num_procs_in
    integer, intent(in) :: ang_mom_in(nsh_in), contr_num_in(nsh_in), sh_loc_in(nsh_in), &
                           atom_num_in(nsh_in), atom_loc_in(nsh_in), start_bas_in(nsh_in), end_bas_in(nsh_in)
    real(dp), intent(in) :: exponents_in(mxgtot_in*nsh_in)
    real(dp), intent(in) :: contr_coef_s_in(mxgtot_in*nsh_in), contr_coef_p_in(mxgtot_in*nsh_in)
    real(dp), intent(in) :: contr_coef_d_in(mxgtot_in*nsh_in), contr_coef_f_in(mxgtot_in*nsh_in)
    real(dp), intent(in) :: coords_in(3*natoms_in)
    real(dp), intent(in) :: schwrz_int_in((nsh_in*(nsh_in + 1))/2)
    integer(c_int32_t), allocatable :: iv(:, :)

    if (.not. c_associated(handle%impl)) call check(liberi_c_create(handle%impl), 'liberi_setup')
    allocate (iv(nsh_in, 7))
    iv(:, 1) = int(ang_mom_in, c_int32_t)
    iv(:, 2) = int(contr_num_in, c_int32_t)
    iv(:, 3) = int(sh_loc_in, c_int32_t)
    iv(:, 4) = int(atom_num_in, c_int32_t)
    iv(:, 5) = int(atom_loc_in, c_int32_t)
    iv(:, 6) = int(start_bas_in, c_int32_t)
    iv(:, 7) = int(end_bas_in, c_int32_t)
    ! setup on a set-up handle cleans up first (liberi_interface.F90:133-136); the C side does the same
    call check(liberi_c_setup(handle%impl, int(nsh_in, c_int32_t), int(natoms_in, c_int32_t), &
                              int(num_bas_in, c_int32_t), int(mxgtot_in, c_int32_t), iv(:, 1), iv(:, 2), &
                              iv(:, 3), iv(:, 4), iv(:, 5), iv(:, 6), iv(:, 7), exponents_in, &
                              contr_coef_s_in, contr_coef_p_in, contr_coef_d_in, contr_coef_f_in, coords_in, &
                              schwrz_int_in, int(my_rank_in, c_int32_t), int(num_procs_in, c_int32_t)), &
               'liberi_setup')
    handle%is_initialized = .true.
  end subroutine liberi_setup

  subroutine liberi_fock_build(handle, density_in, fock_out, matrix_size)
    type(liberi_handle_t), intent(inout) :: handle
    integer, intent(in) :: matrix_size
    real(dp), intent(in) :: density_in(matrix_size)
    real(dp), intent(inout) :: fock_out(matrix_size)
    if (.not. handle%is_initialized) then
      write (*, *) "ERROR: liberi_fock_build called before liberi_setup"
      error stop 1
    end if
    call check(liberi_c_fock_build(handle%impl, density_in, fock_out, int(matrix_size, c_int64_t)), &
               'liberi_fock_build')
  end subroutine liberi_fock_build

  subroutine liberi_cleanup(handle)
    type(liberi_handle_t), intent(inout) :: handle
    if (.not. handle%is_initialized) return
    call check(liberi_c_cleanup(handle%impl), 'liberi_cleanup')
    handle%is_initialized = .false.
  end subroutine liberi_cleanup

end module liberi_interface

module gamess_liberi_wrapper
  use liberi_interface
  use liberi_types, only: dp
  implicit none
  private

  public :: liberi_init
  public :: liberi_fock
  public :: liberi_finalize

  ! one process-wide handle, as src/interface/gms_liberi_wrapper.F90:23-25
  type(liberi_handle_t), save :: handle
  logical, save :: is_initialized = .false.
  integer, save :: matrix_size = 0

contains

  ! note the argument order (natoms, num_bas, nsh, ...), gms_liberi_wrapper.F90:29
  subroutine liberi_init(natoms, num_bas, nsh, mxgtot, &
                         ang_mom, contr_num, sh_loc, &
                         atom_num, atom_loc, start_bas, end_bas, &
                         exponents, contr_coef_s, contr_coef_p, &
                         contr_coef_d, contr_coef_f, &
                         coords, schwrz_int, my_rank, num_procs)
    integer, intent(in) :: natoms, num_bas, nsh, mxgtot
    integer, intent(in) :: ang_mom(nsh), contr_num(nsh), sh_loc(nsh), atom_num(nsh), atom_loc(nsh), &
                           start_bas(nsh), end_bas(nsh)
    real(dp), intent(in) :: exponents(mxgtot*nsh), contr_coef_s(mxgtot*nsh), contr_coef_p(mxgtot*nsh)
    real(dp), intent(in) :: contr_coef_d(mxgtot*nsh), contr_coef_f(mxgtot*nsh)
    real(dp), intent(in) :: coords(3*natoms), schwrz_int((nsh*(nsh + 1))/2)
    integer, intent(in) :: my_rank, num_procs

    if (is_initialized) then
      call liberi_cleanup(handle)
      call liberi_destroy(handle)
    end if
    call liberi_create(handle)
    call liberi_setup(handle, nsh, natoms, num_bas, mxgtot, ang_mom, contr_num, sh_loc, atom_num, atom_loc, &
                      start_bas, end_bas, exponents, contr_coef_s, contr_coef_p, contr_coef_d, contr_coef_f, &
                      coords, schwrz_int, my_rank, num_procs)
    matrix_size = (num_bas*(num_bas + 1))/2
    is_initialized = .true.
  end subroutine liberi_init

  subroutine liberi_fock(density, fock, n)
    integer, intent(in) :: n
    real(dp), intent(in) :: density(n)
    real(dp), intent(inout) :: fock(n)
    if (.not. is_initialized) then
      write (*, *) "ERROR: liberi_fock called before liberi_init"
      error stop 1
    end if
    if (n /= matrix_size) then
      write (*, *) "ERROR: matrix size mismatch in liberi_fock"
      write (*, *) "  Expected:", matrix_size
      write (*, *) "  Got:", n
      error stop 1
    end if
    call liberi_fock_build(handle, density, fock, n)
  end subroutine liberi_fock

  subroutine liberi_finalize()
    if (is_initialized) then
      call liberi_cleanup(handle)
      call liberi_destroy(handle)
      is_initialized = .false.
    end if
  end subroutine liberi_finalize

end module gamess_liberi_wrapper
