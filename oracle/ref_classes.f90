! Per-class driver for the flang build of the reference (test infrastructure).
!
! Builds the shell pairs with the reference's own shell_pair()
! (src/liberi_shell_pair.F90:39-106) and calls ONE dispatched kernel of
! src/liberi_driver.F90:69-158 on a zeroed Fock matrix, so a class's
! contribution can be compared with the oracle in isolation.
!
! usage: liberi_ref_class CASE_FILE FOCK_OUT KERNEL_NAME
program liberi_ref_class
  use liberi_types, only: dp, basis_t, shell_pair_container_t
  use liberi_shell_pair, only: shell_pair
  use rot_axis_kernels
  use eric_kernels
  use rys_kernels
  implicit none
  type(basis_t) :: b
  type(shell_pair_container_t) :: p
  integer :: u, magic, version, ntri
  real(dp), allocatable :: dens(:), fock(:)
  character(len=1024) :: fin, fout
  character(len=32) :: kname
  integer, allocatable :: iv(:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, kname)
  open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
  read (u) magic, version, b%nsh, b%natoms, b%num_bas, b%mxgtot
  allocate (iv(b%nsh))
  read (u) iv; b%ang_mom = iv
  read (u) iv; b%contr_num = iv
  read (u) iv; b%sh_loc = iv
  read (u) iv; b%atom_num = iv
  read (u) iv; b%atom_loc = iv
  read (u) iv; b%start_bas = iv
  read (u) iv; b%end_bas = iv
  allocate (b%exponents(b%mxgtot*b%nsh), b%contr_coef_s(b%mxgtot*b%nsh), &
            b%contr_coef_p(b%mxgtot*b%nsh), b%contr_coef_d(b%mxgtot*b%nsh), &
            b%contr_coef_f(b%mxgtot*b%nsh), b%coords(3*b%natoms), &
            b%schwrz_int(b%nsh*(b%nsh + 1)/2))
  ntri = b%num_bas*(b%num_bas + 1)/2
  allocate (dens(ntri), fock(ntri))
  read (u) b%exponents, b%contr_coef_s, b%contr_coef_p, b%contr_coef_d, &
    b%contr_coef_f, b%coords, b%schwrz_int, dens
  close (u)
  p%resources%n_rank = 0
  p%resources%n_size = 1
  call shell_pair(b, p)
  fock = 0.0_dp
  associate (r => p%resources)
    select case (trim(kname))
    case ('int0000'); call int0000(p%ss, dens, fock, r)
    case ('int0001'); call int0001(p%ss, p%sp, dens, fock, r)
    case ('int0011'); call int0011(p%ss, p%pp, dens, fock, r)
    case ('int0111'); call int0111(p%sp, p%pp, dens, fock, r)
    case ('int0101'); call int0101(p%sp, dens, fock, r)
    case ('int1111'); call int1111(p%pp, dens, fock, r)
    case ('int0002'); call int0002(p%ss, p%sd, dens, fock, r)
    case ('int0012'); call int0012(p%ss, p%pd, dens, fock, r)
    case ('int0102'); call int0102(p%sp, p%sd, dens, fock, r)
    case ('int0112'); call int0112(p%sp, p%pd, dens, fock, r)
    case ('int0022'); call int0022(p%ss, p%dd, dens, fock, r)
    case ('int0122'); call int0122(p%sp, p%dd, dens, fock, r)
    case ('int1102'); call int1102(p%pp, p%sd, dens, fock, r)
    case ('int0202'); call int0202(p%sd, dens, fock, r)
    case ('int2211'); call int2211(p%dd, p%pp, dens, fock, r)
    case ('int2121'); call int2121(p%pd, dens, fock, r)
    case ('int2120'); call int2120(p%pd, p%sd, dens, fock, r)
    case ('int2111'); call int2111(p%pd, p%pp, dens, fock, r)
    case ('int2220'); call int2220(p%dd, p%sd, dens, fock, r)
    case ('int2221'); call int2221(p%dd, p%pd, dens, fock, r)
    case ('int2222gen'); call int2222gen(p%dd, dens, fock, r)
    case ('int3333'); call int3333(p%ff, dens, fock, r)
    case ('int0030'); call int0030(p%ss, p%sf, dens, fock, r)
    case ('int0031'); call int0031(p%ss, p%pf, dens, fock, r)
    case ('int0032'); call int0032(p%ss, p%df, dens, fock, r)
    case ('int1030'); call int1030(p%sp, p%sf, dens, fock, r)
    case ('int3330'); call int3330(p%ff, p%sf, dens, fock, r)
    case ('int3011'); call int3011(p%sf, p%pp, dens, fock, r)
    case ('int3020'); call int3020(p%sf, p%sd, dens, fock, r)
    case ('int2130'); call int2130(p%pd, p%sf, dens, fock, r)
    case ('int2230'); call int2230(p%dd, p%sf, dens, fock, r)
    case ('int3030'); call int3030(p%sf, dens, fock, r)
    case ('int3110'); call int3110(p%pf, p%sp, dens, fock, r)
    case ('int3111'); call int3111(p%pf, p%pp, dens, fock, r)
    case ('int3120'); call int3120(p%pf, p%sd, dens, fock, r)
    case ('int3121'); call int3121(p%pf, p%pd, dens, fock, r)
    case ('int2231'); call int2231(p%dd, p%pf, dens, fock, r)
    case ('int3130'); call int3130(p%pf, p%sf, dens, fock, r)
    case ('int3131'); call int3131(p%pf, dens, fock, r)
    case ('int3210'); call int3210(p%df, p%sp, dens, fock, r)
    case ('int3331'); call int3331(p%ff, p%pf, dens, fock, r)
    case ('int3211'); call int3211(p%df, p%pp, dens, fock, r)
    case ('int3220'); call int3220(p%df, p%sd, dens, fock, r)
    case ('int3221'); call int3221(p%df, p%pd, dens, fock, r)
    case ('int3222'); call int3222(p%df, p%dd, dens, fock, r)
    case ('int3230'); call int3230(p%df, p%sf, dens, fock, r)
    case ('int3231'); call int3231(p%df, p%pf, dens, fock, r)
    case ('int3232'); call int3232(p%df, dens, fock, r)
    case ('int3300'); call int3300(p%ff, p%ss, dens, fock, r)
    case ('int3310'); call int3310(p%ff, p%sp, dens, fock, r)
    case ('int3311'); call int3311(p%ff, p%pp, dens, fock, r)
    case ('int3320'); call int3320(p%ff, p%sd, dens, fock, r)
    case ('int3321'); call int3321(p%ff, p%pd, dens, fock, r)
    case ('int3322'); call int3322(p%ff, p%dd, dens, fock, r)
    case ('int3332'); call int3332(p%ff, p%df, dens, fock, r)
    case default
      write (*, *) 'unknown kernel ', trim(kname)
      error stop 2
    end select
  end associate
  open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
  write (u) fock
  close (u)
end program liberi_ref_class
