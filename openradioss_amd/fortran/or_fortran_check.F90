! or_fortran_check.F90 -- a Fortran caller of the C ABI through or_gpu_mod.
!
! Without arguments: prints the library version and the number of HIP devices.
! With <in> <out>: reads one SFORC3 group in the Engine's layouts from the
! unformatted stream file <in>, calls OR_SFORC3_GROUP exactly as FORINT would
! call SFORC3 for the group, and writes the status, MSTOP, DT2T, FSKY and the
! updated GBUF arrays to <out> (unformatted stream).
!
! <in>:  int32 NUMNOD, NEL, LSKY; or_hex_params_t; X(3,NUMNOD), V(3,NUMNOD);
!        IXS(11,NEL), IADS(8,NEL) int32; SIG(NEL*6), EINT, RHO, VOL, PLA, EPSD, OFF,
!        QVIS (NEL each); DT1.
! <out>: int32 status, MSTOP; DT2T; FSKY(8,LSKY); SIG(NEL*6), EINT, RHO, PLA, EPSD,
!        OFF, QVIS.
program or_fortran_check
  use, intrinsic :: iso_c_binding
  use or_gpu_mod
  implicit none
  integer(c_int) :: ndev, ierr
  integer(c_int32_t) :: numnod, nel, lsky, mstop
  type(or_hex_params_t) :: hp
  real(c_double), allocatable :: x(:, :), v(:, :), fsky(:, :)
  integer(c_int32_t), allocatable :: ixs(:, :), iads(:, :)
  real(c_double), allocatable :: sig(:), eint(:), rho(:), vol(:), pla(:), epsd(:), off(:), qvis(:)
  real(c_double) :: dt1, dt2t
  character(len=4096) :: fin, fout
  integer :: u

  if (command_argument_count() < 2) then
    ierr = or_device_count(ndev)
    write (*, '(a,i0)') 'or_version ', or_version()
    write (*, '(a,i0)') 'devices ', ndev
  else
    call run_group()
  end if

contains

  subroutine run_group()
    call get_command_argument(1, fin)
    call get_command_argument(2, fout)

    open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
    read (u) numnod, nel, lsky
    read (u) hp
    allocate (x(3, numnod), v(3, numnod), ixs(11, nel), iads(8, nel), fsky(8, lsky))
    allocate (sig(6*nel), eint(nel), rho(nel), vol(nel), pla(nel), epsd(nel), off(nel), qvis(nel))
    read (u) x, v, ixs, iads, sig, eint, rho, vol, pla, epsd, off, qvis, dt1
    close (u)

    fsky = 0.0d0
    dt2t = 1.0d6
    mstop = 0
    ierr = or_sforc3_group(0_c_int, hp, int(numnod, c_int64_t), nel, 0_c_int64_t, x, v, ixs, sig, eint, rho, &
                           vol, pla, epsd, off, qvis, int(lsky, c_int64_t), fsky, iads, dt1, dt2t, mstop)
    if (ierr /= OR_OK) write (*, '(a,i0,2a)') 'or_sforc3_group status ', ierr, ': ', or_error_message()

    open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
    write (u) int(ierr, c_int32_t), mstop, dt2t, fsky, sig, eint, rho, pla, epsd, off, qvis
    close (u)
  end subroutine run_group
end program or_fortran_check
