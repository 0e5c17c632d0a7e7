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
!
! With <mode> <in> <out>, the other calls an Engine makes (layouts in the subroutines below):
!   q4group  FORINTC's call of one CFORC3 group: the group's IPARG column, PM and GEO words and
!            the law's parameters go through OR_GPU_ELIGIBLE_SHELL and OR_PACK_Q4_PARAMS, and
!            the group through OR_CFORC3_GROUP;
!   hexgroup the same for one SFORC3 group (OR_GPU_ELIGIBLE_SOLID, OR_PACK_HEX_PARAMS,
!            OR_SFORC3_GROUP);
!   nodal    RESOL's nodal half of a cycle through OR_NODAL_UPDATE;
!   model    a resident model: create, upload, OR_MODEL_TH_SETUP, OR_MODEL_RUN, OR_MODEL_RUN_TH, download.
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
  character(len=4096) :: fin, fout, mode
  integer :: u

  if (command_argument_count() < 2) then
    ierr = or_device_count(ndev)
    write (*, '(a,i0)') 'or_version ', or_version()
    write (*, '(a,i0)') 'devices ', ndev
  else if (command_argument_count() == 2) then
    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    call run_group()
  else
    call get_command_argument(1, mode)
    call get_command_argument(2, fin)
    call get_command_argument(3, fout)
    select case (trim(mode))
    case ('q4group')
      call run_q4group()
    case ('hexgroup')
      call run_hexgroup()
    case ('nodal')
      call run_nodal()
    case ('model')
      call run_model()
    case default
      write (*, '(2a)') 'unknown mode ', trim(mode)
    end select
  end if

contains

  subroutine run_group()
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

  ! <in>:  int32 NUMNOD, NEL, LSKY; IPARG(100) int32, IPARAM(4) int32; mat RHO, BULK, SHEAR,
  !        UPARAM(10); GEO(40); DTFAC1(1), DTMX; X, V (3,NUMNOD); IXS(11,NEL), IADS(8,NEL) int32;
  !        SIG(NEL*6), EINT, RHO, VOL, PLA, EPSD, OFF, QVIS; DT1.
  ! <out>: int32 status, eligible, MSTOP; DT2T; FSKY(8,LSKY); SIG, EINT, RHO, PLA, EPSD, OFF, QVIS.
  subroutine run_hexgroup()
    integer(c_int32_t) :: iparg(100), iparam(4), ok
    real(c_double) :: mrho, mbulk, mshear, uparam(10), geo(40), dtfac, dtmx
    open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
    read (u) numnod, nel, lsky, iparg, iparam, mrho, mbulk, mshear, uparam, geo, dtfac, dtmx
    allocate (x(3, numnod), v(3, numnod), ixs(11, nel), iads(8, nel), fsky(8, lsky))
    allocate (sig(6*nel), eint(nel), rho(nel), vol(nel), pla(nel), epsd(nel), off(nel), qvis(nel))
    read (u) x, v, ixs, iads, sig, eint, rho, vol, pla, epsd, off, qvis, dt1
    close (u)
    fsky = 0.0d0
    dt2t = 1.0d6
    mstop = 0
    ierr = OR_ERR_ARG
    ok = 0
    if (or_gpu_eligible_solid(iparg, iparam, uparam(8))) then
      ok = 1
      call or_pack_hex_params(mrho, mbulk, mshear, uparam, iparam, geo, dtfac, dtmx, hp)
      ierr = or_sforc3_group(0_c_int, hp, int(numnod, c_int64_t), nel, int(iparg(3), c_int64_t), x, v, ixs, &
                             sig, eint, rho, vol, pla, epsd, off, qvis, int(lsky, c_int64_t), fsky, iads, dt1, &
                             dt2t, mstop)
      if (ierr /= OR_OK) write (*, '(a,i0,2a)') 'or_sforc3_group status ', ierr, ': ', or_error_message()
    end if
    open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
    write (u) int(ierr, c_int32_t), ok, mstop, dt2t, fsky, sig, eint, rho, pla, epsd, off, qvis
    close (u)
  end subroutine run_hexgroup

  ! <in>:  int32 NUMNOD, NEL, LSKY; IPARG(100), IPARAM(4) int32; PM(30), UPARAM(10), GEO(40),
  !        DTFAC1(3), DTMIN1(3); X, V, VR (3,NUMNOD); IXC(7,NEL), IADC(4,NEL) int32; GBUF OFF, THK,
  !        THKE (NEL), FOR(NEL*5), MOM(NEL*3), STRA(NEL*8), HOURG(NEL*5), EINT(NEL*2), EPSD(NEL);
  !        LBUF SIG(NPT*5*NEL), PLA(NPT*NEL), EPSD(NPT*NEL) with NPT = IPARG(6); DT1.
  ! <out>: int32 status, eligible; DT2T; FSKY(8,LSKY); OFF, THK, FOR, MOM, STRA, HOURG, EINT,
  !        EPSD, SIG, PLA, EPSD per point.
  subroutine run_q4group()
    integer(c_int32_t) :: iparg(100), iparam(4), ok, npt
    real(c_double) :: pm(30), uparam(10), geo(40), dtfac, dtmin
    real(c_double), allocatable :: vr(:, :), thk(:), thke(:), forc(:), mom(:), gstr(:), hourg(:), qeint(:)
    real(c_double), allocatable :: qepsd(:), lsig(:), lpla(:), lepsd(:)
    integer(c_int32_t), allocatable :: ixc(:, :), iadc(:, :)
    type(or_q4_params_t) :: qp
    open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
    read (u) numnod, nel, lsky, iparg, iparam, pm, uparam, geo, dtfac, dtmin
    npt = iparg(6)
    allocate (x(3, numnod), v(3, numnod), vr(3, numnod), ixc(7, nel), iadc(4, nel), fsky(8, lsky))
    allocate (off(nel), thk(nel), thke(nel), forc(5*nel), mom(3*nel), gstr(8*nel), hourg(5*nel), qeint(2*nel))
    allocate (qepsd(nel), lsig(npt*5*nel), lpla(npt*nel), lepsd(npt*nel))
    read (u) x, v, vr, ixc, iadc, off, thk, thke, forc, mom, gstr, hourg, qeint, qepsd, lsig, lpla, lepsd, dt1
    close (u)
    fsky = 0.0d0
    dt2t = 1.0d6
    ierr = OR_ERR_ARG
    ok = 0
    if (or_gpu_eligible_shell(iparg, iparam)) then
      ok = 1
      call or_pack_q4_params(pm, uparam, iparam, geo, iparg, dtfac, dtmin, qp)
      ierr = or_cforc3_group(0_c_int, qp, int(numnod, c_int64_t), nel, int(iparg(3), c_int64_t), x, v, vr, ixc, &
                             off, thk, thke, forc, mom, gstr, hourg, qeint, qepsd, lsig, lpla, lepsd, &
                             int(lsky, c_int64_t), fsky, iadc, dt1, dt2t)
      if (ierr /= OR_OK) write (*, '(a,i0,2a)') 'or_cforc3_group status ', ierr, ': ', or_error_message()
    end if
    open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
    write (u) int(ierr, c_int32_t), ok, dt2t, fsky, off, thk, forc, mom, gstr, hourg, qeint, qepsd, lsig, lpla, lepsd
    close (u)
  end subroutine run_q4group

  ! <in>:  int64 NUMNOD, LSKY; ADSKY(NUMNOD+1) int64 (1-based); FSKY(8,LSKY); MS, IN (NUMNOD);
  !        node codes (NUMNOD) int32; VIMP, X, V, VR, D (3,NUMNOD); DT2T, DTMX; or_time_state_t.
  ! <out>: int32 status; X, V, VR, D (3,NUMNOD); STIFN, STIFR (NUMNOD); or_time_state_t.
  subroutine run_nodal()
    integer(c_int64_t) :: nn, ls
    integer(c_int64_t), allocatable, target :: adsky(:)
    real(c_double), allocatable, target :: fs(:, :), ms(:), in(:), vimp(:, :), xx(:, :), vv(:, :), vr(:, :), d(:, :)
    real(c_double), allocatable, target :: stifn(:), stifr(:)
    integer(c_int32_t), allocatable, target :: code(:)
    real(c_double) :: dtmx
    type(or_time_state_t) :: ts
    open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
    read (u) nn, ls
    allocate (adsky(nn + 1), fs(8, ls), ms(nn), in(nn), code(nn), vimp(3, nn), xx(3, nn), vv(3, nn), vr(3, nn))
    allocate (d(3, nn), stifn(nn), stifr(nn))
    read (u) adsky, fs, ms, in, code, vimp, xx, vv, vr, d, dt2t, dtmx, ts
    close (u)
    stifn = 0.0d0
    stifr = 0.0d0
    ierr = or_nodal_update(0_c_int, nn, ls, c_loc(adsky), c_loc(fs), c_loc(ms), c_loc(in), c_loc(code), &
                           c_loc(vimp), c_loc(xx), c_loc(vv), c_loc(vr), c_loc(d), c_loc(stifn), c_loc(stifr), &
                           dt2t, dtmx, ts)
    if (ierr /= OR_OK) write (*, '(a,i0,2a)') 'or_nodal_update status ', ierr, ': ', or_error_message()
    open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
    write (u) int(ierr, c_int32_t), xx, vv, vr, d, stifn, stifr, ts
    close (u)
  end subroutine run_nodal

  ! <in>:  int64 NN, NEH, NEQ; int32 NPT, PRECISION, NCYCLES; or_hex_params_t; or_q4_params_t;
  !        X, V, VR (3,NN); MS, IN (NN); node codes (NN) int32; VIMP (3,NN); hex connectivity
  !        (8,NEH) int32 0-based, VOL0 and RHO (NEH); shell connectivity (4,NEQ) int32 0-based,
  !        THKE (NEQ); ADSKY (NN+1), IADS (8,NEH), IADC (4,NEQ) int64 0-based.
  ! <out>: int32 status; X, V, VR, D (3,NN); or_time_state_t; or_ecrit_t; the time-history record of
  !        node 1, hex 1 and shell 1: XMOMT, YMOMT, ZMOMT, EHOUR, 18 node, 13 hex and 12+NPT shell values.
  subroutine run_model()
    integer(c_int64_t) :: nn, neh, neq
    integer(c_int32_t) :: npt, prec, ncyc
    type(or_q4_params_t) :: qp
    type(c_ptr) :: h
    real(c_double), allocatable, target :: xx(:, :), vv(:, :), vr(:, :), d(:, :), ms(:), in(:), vimp(:, :)
    real(c_double), allocatable, target :: vol0(:), hrho(:), thke(:)
    integer(c_int32_t), allocatable, target :: code(:), connh(:, :), connq(:, :)
    integer(c_int64_t), allocatable, target :: adsky(:), iadsh(:, :), iadcq(:, :)
    type(or_time_state_t) :: ts
    type(or_ecrit_t) :: ec
    ! time history of the first node, hex and shell (HIST2's values of the output cycle)
    integer(c_int32_t), target :: th_first(1) = [0_c_int32_t]
    real(c_double) :: glob(4), nodeo(18), hexo(13)
    real(c_double), allocatable :: q4o(:)
    open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
    read (u) nn, neh, neq, npt, prec, ncyc, hp, qp
    allocate (q4o(12 + max(npt, 0_c_int32_t)))
    hexo = 0.0_c_double
    q4o = 0.0_c_double
    allocate (xx(3, nn), vv(3, nn), vr(3, nn), d(3, nn), ms(nn), in(nn), code(nn), vimp(3, nn))
    allocate (connh(8, max(neh, 1_c_int64_t)), vol0(max(neh, 1_c_int64_t)), hrho(max(neh, 1_c_int64_t)))
    allocate (connq(4, max(neq, 1_c_int64_t)), thke(max(neq, 1_c_int64_t)))
    allocate (adsky(nn + 1), iadsh(8, max(neh, 1_c_int64_t)), iadcq(4, max(neq, 1_c_int64_t)))
    read (u) xx, vv, vr, ms, in, code, vimp
    if (neh > 0) read (u) connh(:, 1:neh), vol0(1:neh), hrho(1:neh)
    if (neq > 0) read (u) connq(:, 1:neq), thke(1:neq)
    read (u) adsky
    if (neh > 0) read (u) iadsh(:, 1:neh)
    if (neq > 0) read (u) iadcq(:, 1:neq)
    close (u)
    ierr = or_model_create(h, 0_c_int, nn, neh, neq, npt, int(prec, c_int))
    if (ierr == OR_OK .and. neh > 0) ierr = or_model_set_hex_params(h, hp)
    if (ierr == OR_OK .and. neq > 0) ierr = or_model_set_q4_params(h, qp)
    if (ierr == OR_OK) ierr = or_model_upload_nodes(h, c_loc(xx), c_loc(vv), c_loc(vr), c_loc(ms), c_loc(in), &
                                                    c_loc(code), c_loc(vimp))
    if (ierr == OR_OK .and. neh > 0) ierr = or_model_upload_hex(h, c_loc(connh), c_null_ptr, c_null_ptr, &
                                                                c_loc(hrho), c_loc(vol0), c_null_ptr, &
                                                                c_null_ptr, c_null_ptr, c_null_ptr)
    if (ierr == OR_OK .and. neq > 0) ierr = or_model_upload_q4(h, c_loc(connq), c_loc(thke), c_null_ptr, &
                                                               c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                                                               c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                                                               c_null_ptr, c_null_ptr)
    if (ierr == OR_OK) ierr = or_model_upload_skyline(h, c_loc(adsky), c_loc(iadsh), c_loc(iadcq))
    if (ierr == OR_OK) ierr = or_model_th_setup(h, 1_c_int64_t, th_first, min(neh, 1_c_int64_t), th_first, &
                                                min(neq, 1_c_int64_t), th_first)
    ! ncyc cycles, the last one an output cycle (ECRIT's line and the time-history record)
    if (ierr == OR_OK) ierr = or_model_run(h, int(ncyc - 1, c_int64_t))
    if (ierr == OR_OK) ierr = or_model_run_th(h, ec, glob, nodeo, hexo, q4o)
    if (ierr == OR_OK) ierr = or_model_sync(h)
    if (ierr == OR_OK) ierr = or_model_download_nodes(h, c_loc(xx), c_loc(vv), c_loc(vr), c_loc(d))
    if (ierr == OR_OK) ierr = or_model_get_time(h, ts)
    if (ierr /= OR_OK) write (*, '(a,i0,2a)') 'resident model status ', ierr, ': ', or_error_message()
    if (c_associated(h)) then
      if (or_model_destroy(h) /= OR_OK) write (*, '(a)') 'or_model_destroy failed'
    end if
    open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
    write (u) int(ierr, c_int32_t), xx, vv, vr, d, ts, ec, glob, nodeo, hexo, q4o
    close (u)
  end subroutine run_model
end program or_fortran_check
