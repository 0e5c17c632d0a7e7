! or_gpu_mod.F90 -- Fortran ISO_C_BINDING interface of the MI355X path (include/or_gpu.h).
!
! The Engine side of the boundary: RESOL / FORINT / FORINTC call these in place
! of SFORC3 (forint.F:1159), CFORC3 (forintc.F:452), or, for a whole resident
! model, in place of the element and nodal sequence of one cycle.  Every
! function returns an integer status (OR_OK = 0); OR_ERROR_MESSAGE() returns the
! text of the last failure of the calling thread.  Arrays are the Engine's own:
! X(3,NUMNOD), V(3,NUMNOD), VR(3,NUMNOD), IXS(NIXS,*), IXC(NIXC,*), GBUF / LBUF
! component-major arrays of one group, FSKY(8,LSKY), IADS(8,*), IADC(4,*).
module or_gpu_mod
  use, intrinsic :: iso_c_binding
  implicit none
  private
  public :: or_hex_params_t, or_q4_params_t, or_time_state_t
  public :: or_version, or_device_count, or_device_memory, or_error_message
  public :: or_model_create, or_model_destroy, or_model_set_hex_params, or_model_set_q4_params
  public :: or_model_upload_nodes, or_model_upload_hex, or_model_upload_q4, or_model_upload_skyline
  public :: or_model_upload_displacement
  public :: or_model_set_time, or_model_get_time, or_model_run, or_model_run_output, or_model_sync
  public :: or_ecrit_t, or_model_run_ecrit, or_model_wfext
  public :: or_model_th_setup, or_model_run_th, or_model_ehour, or_model_set_ehour
  public :: or_model_download_nodes, or_model_download_hex, or_model_download_q4
  public :: or_model_energy, or_model_kernel_times, or_model_set_timing, or_model_footprint
  public :: or_model_footprint_estimate
  public :: or_batch_create, or_batch_destroy, or_batch_add, or_batch_remove, or_batch_run, or_batch_sync
  public :: or_batch_active, or_model_set_variant, or_model_autotune, or_batch_autotune
  public :: or_sforc3_group, or_sforc3_group_static, or_cforc3_group, or_nodal_update, or_math_eval
  public :: or_gpu_eligible_solid, or_gpu_eligible_shell, or_pack_hex_params, or_pack_q4_params

  integer(c_int), parameter, public :: OR_OK = 0, OR_ERR_ARG = 1, OR_ERR_HIP = 2, OR_ERR_NODEV = 3, &
                                       OR_ERR_STATE = 4, OR_ERR_MSTOP = 5
  integer(c_int), parameter, public :: OR_PREC_FP64 = 0, OR_PREC_SPLIT = 1
  ! MSTOP bits of or_time_state_t: non-positive Jacobian, time step <= 0
  integer(c_int32_t), parameter, public :: OR_MSTOP_JACOBIAN = 1, OR_MSTOP_TIME_STEP = 2
  integer(c_int32_t), parameter, public :: OR_BC_TX = 4, OR_BC_TY = 2, OR_BC_TZ = 1, OR_BC_RX = 32, &
                                           OR_BC_RY = 16, OR_BC_RZ = 8, OR_IMP_X = 256, OR_IMP_Y = 128, &
                                           OR_IMP_Z = 64

  ! law 2 + /PROP/SOLID Isolid=1 + run control (or_hex_params)
  type, bind(c) :: or_hex_params_t
    real(c_double) :: rho0, bulk, shear
    real(c_double) :: ca, cb, cn, epmx, sigmx, cc, epdr, fisokin, asrate, m_exp
    integer(c_int32_t) :: icc, vp, israte, iform
    real(c_double) :: qa, qb, qh, vns1, vns2
    real(c_double) :: dtfac, dtmx
  end type

  ! law 2 for shells + /PROP/SHELL Ishell=1 (or_q4_params)
  type, bind(c) :: or_q4_params_t
    real(c_double) :: rho, young, nu, shear, ssp
    real(c_double) :: ca, cb, cn, epmx, ymax, cc, epdr, m_exp
    real(c_double) :: asrate
    integer(c_int32_t) :: icc, vp, israte, iform
    integer(c_int32_t) :: npt, ipla, ithk, ishear
    real(c_double) :: hm, hf, hr, fsh, dtfac, dtmin
  end type

  ! TT, DT1, DT2, DT12, DT2OLD, NCYCLE, MSTOP (or_time_state)
  type, bind(c) :: or_time_state_t
    real(c_double) :: tt, dt1, dt2, dt12, dt2old
    integer(c_int64_t) :: ncycle
    integer(c_int32_t) :: mstop
  end type

  ! what ECRIT prints for one output cycle (include/or_gpu.h or_ecrit): ENINT, ENCIN, ENROT,
  ! WFEXT, ITYPTS and NELTS (0-based index of the controlling element in its family)
  type, bind(c) :: or_ecrit_t
    real(c_double) :: enint, encin, enrot, wfext
    integer(c_int32_t) :: itypts, pad
    integer(c_int64_t) :: nelts
  end type

  interface
    integer(c_int) function or_version() bind(c, name="or_version")
      import :: c_int
    end function

    type(c_ptr) function or_last_error_c() bind(c, name="or_last_error")
      import :: c_ptr
    end function

    integer(c_int) function or_device_count(count) bind(c, name="or_device_count")
      import :: c_int
      integer(c_int), intent(out) :: count
    end function

    integer(c_int) function or_device_memory(device, free_bytes, total_bytes) bind(c, name="or_device_memory")
      import :: c_int, c_int64_t
      integer(c_int), value :: device
      integer(c_int64_t), intent(out) :: free_bytes, total_bytes
    end function

    integer(c_int) function or_model_create(handle, device, nn, ne_hex, ne_q4, npt, precision) &
        bind(c, name="or_model_create")
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      type(c_ptr), intent(out) :: handle
      integer(c_int), value :: device, precision
      integer(c_int64_t), value :: nn, ne_hex, ne_q4
      integer(c_int32_t), value :: npt
    end function

    integer(c_int) function or_model_destroy(handle) bind(c, name="or_model_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
    end function

    integer(c_int) function or_model_set_hex_params(handle, p) bind(c, name="or_model_set_hex_params")
      import :: c_int, c_ptr, or_hex_params_t
      type(c_ptr), value :: handle
      type(or_hex_params_t), intent(in) :: p
    end function

    integer(c_int) function or_model_set_q4_params(handle, p) bind(c, name="or_model_set_q4_params")
      import :: c_int, c_ptr, or_q4_params_t
      type(c_ptr), value :: handle
      type(or_q4_params_t), intent(in) :: p
    end function

    ! optional arrays (vr, in, code, vimp) may be passed as c_null_ptr through the c_ptr forms
    integer(c_int) function or_model_upload_nodes(handle, x, v, vr, ms, in, code, vimp) &
        bind(c, name="or_model_upload_nodes")
      import :: c_int, c_ptr
      type(c_ptr), value :: handle, x, v, vr, ms, in, code, vimp
    end function
    ! D(3,NUMNOD) of a restart (wrrestp.F:1188-1194 writes X, D, V, VR)
    integer(c_int) function or_model_upload_displacement(handle, d) &
        bind(c, name="or_model_upload_displacement")
      import :: c_int, c_ptr
      type(c_ptr), value :: handle, d
    end function

    integer(c_int) function or_model_upload_hex(handle, conn, sig, eint, rho, vol0, pla, epsd, offg, qvis) &
        bind(c, name="or_model_upload_hex")
      import :: c_int, c_ptr
      type(c_ptr), value :: handle, conn, sig, eint, rho, vol0, pla, epsd, offg, qvis
    end function

    integer(c_int) function or_model_upload_q4(handle, conn, thke, offg, thk, forc, mom, gstr, hourg, eint, &
        epsd, sig, pla, epsd_ip) bind(c, name="or_model_upload_q4")
      import :: c_int, c_ptr
      type(c_ptr), value :: handle, conn, thke, offg, thk, forc, mom, gstr, hourg, eint, epsd, sig, pla, epsd_ip
    end function

    integer(c_int) function or_model_upload_skyline(handle, adsky, iads, iadc) &
        bind(c, name="or_model_upload_skyline")
      import :: c_int, c_ptr
      type(c_ptr), value :: handle, adsky, iads, iadc
    end function

    integer(c_int) function or_model_set_time(handle, ts) bind(c, name="or_model_set_time")
      import :: c_int, c_ptr, or_time_state_t
      type(c_ptr), value :: handle
      type(or_time_state_t), intent(in) :: ts
    end function

    integer(c_int) function or_model_get_time(handle, ts) bind(c, name="or_model_get_time")
      import :: c_int, c_ptr, or_time_state_t
      type(c_ptr), value :: handle
      type(or_time_state_t), intent(out) :: ts
    end function

    integer(c_int) function or_model_run(handle, ncycles) bind(c, name="or_model_run")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: handle
      integer(c_int64_t), value :: ncycles
    end function

    ! one cycle with ECRIT's energies: e(1) internal, e(2) translational, e(3) rotational kinetic
    integer(c_int) function or_model_run_output(handle, e) bind(c, name="or_model_run_output")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      real(c_double), intent(out) :: e(3)
    end function

    integer(c_int) function or_model_run_ecrit(handle, e) bind(c, name="or_model_run_ecrit")
      import :: c_int, c_ptr, or_ecrit_t
      type(c_ptr), value :: handle
      type(or_ecrit_t), intent(out) :: e
    end function

    integer(c_int) function or_model_wfext(handle, w) bind(c, name="or_model_wfext")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      real(c_double), intent(out) :: w
    end function

    ! HIST2's values (hist2.F:299-330, 959-999): the listed nodes, hexes and shells (0-based)
    integer(c_int) function or_model_th_setup(handle, n_node, node, n_hex, hex, n_q4, q4) &
        bind(c, name="or_model_th_setup")
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      type(c_ptr), value :: handle
      integer(c_int64_t), value :: n_node, n_hex, n_q4
      integer(c_int32_t), intent(in) :: node(*), hex(*), q4(*)
    end function

    ! one output cycle: ECRIT's record, XMOMT/YMOMT/ZMOMT/EHOUR, node (18, n), hex (13, n) and
    ! shell (12 + NPT, n) values
    integer(c_int) function or_model_run_th(handle, e, glob, node_out, hex_out, q4_out) &
        bind(c, name="or_model_run_th")
      import :: c_int, c_ptr, c_double, or_ecrit_t
      type(c_ptr), value :: handle
      type(or_ecrit_t), intent(out) :: e
      real(c_double), intent(out) :: glob(4), node_out(*), hex_out(*), q4_out(*)
    end function

    integer(c_int) function or_model_ehour(handle, ehour) bind(c, name="or_model_ehour")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      real(c_double), intent(out) :: ehour
    end function

    integer(c_int) function or_model_set_ehour(handle, ehour) bind(c, name="or_model_set_ehour")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      real(c_double), value :: ehour
    end function

    integer(c_int) function or_model_sync(handle) bind(c, name="or_model_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
    end function

    integer(c_int) function or_model_download_nodes(handle, x, v, vr, d) bind(c, name="or_model_download_nodes")
      import :: c_int, c_ptr
      type(c_ptr), value :: handle, x, v, vr, d
    end function

    integer(c_int) function or_model_download_hex(handle, sig, eint, rho, pla, epsd, offg, qvis) &
        bind(c, name="or_model_download_hex")
      import :: c_int, c_ptr
      type(c_ptr), value :: handle, sig, eint, rho, pla, epsd, offg, qvis
    end function

    integer(c_int) function or_model_download_q4(handle, offg, thk, forc, mom, gstr, hourg, eint, epsd, sig, &
        pla, epsd_ip) bind(c, name="or_model_download_q4")
      import :: c_int, c_ptr
      type(c_ptr), value :: handle, offg, thk, forc, mom, gstr, hourg, eint, epsd, sig, pla, epsd_ip
    end function

    integer(c_int) function or_model_energy(handle, e) bind(c, name="or_model_energy")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      real(c_double), intent(out) :: e(3)
    end function

    integer(c_int) function or_model_kernel_times(handle, reset, hex_ms, q4_ms, node_ms, cycles) &
        bind(c, name="or_model_kernel_times")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: handle
      integer(c_int), value :: reset
      real(c_double), intent(out) :: hex_ms, q4_ms, node_ms
      integer(c_int64_t), intent(out) :: cycles
    end function

    integer(c_int) function or_model_set_timing(handle, enable) bind(c, name="or_model_set_timing")
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
      integer(c_int), value :: enable
    end function

    integer(c_int) function or_model_footprint(handle, nbytes) bind(c, name="or_model_footprint")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: handle
      integer(c_int64_t), intent(out) :: nbytes
    end function

    integer(c_int) function or_model_footprint_estimate(nn, ne_hex, ne_q4, npt, precision, strain_rate_filter, &
        rows_hex, rows_q4, nbytes) bind(c, name="or_model_footprint_estimate")
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      integer(c_int64_t), value :: nn, ne_hex, ne_q4
      integer(c_int32_t), value :: npt
      integer(c_int), value :: precision, strain_rate_filter
      type(c_ptr), value :: rows_hex, rows_q4
      integer(c_int64_t), intent(out) :: nbytes
    end function

    ! batches: many resident models advanced together, each to its own end time
    integer(c_int) function or_batch_create(batch, device) bind(c, name="or_batch_create")
      import :: c_int, c_ptr
      type(c_ptr), intent(out) :: batch
      integer(c_int), value :: device
    end function

    integer(c_int) function or_batch_destroy(batch) bind(c, name="or_batch_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
    end function

    integer(c_int) function or_batch_add(batch, handle, t_end) bind(c, name="or_batch_add")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: batch, handle
      real(c_double), value :: t_end
    end function

    integer(c_int) function or_batch_remove(batch, handle) bind(c, name="or_batch_remove")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch, handle
    end function

    integer(c_int) function or_batch_run(batch, ncycles) bind(c, name="or_batch_run")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: batch
      integer(c_int64_t), value :: ncycles
    end function

    integer(c_int) function or_batch_sync(batch) bind(c, name="or_batch_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
    end function

    integer(c_int) function or_batch_active(batch, nactive) bind(c, name="or_batch_active")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: batch
      integer(c_int32_t), intent(out) :: nactive
    end function

    ! launch variants of the element kernels (speed only, results bit-identical)
    integer(c_int) function or_model_set_variant(handle, family, variant) bind(c, name="or_model_set_variant")
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
      integer(c_int), value :: family, variant
    end function

    integer(c_int) function or_model_autotune(handle, cycles_per_variant, chosen) bind(c, name="or_model_autotune")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: handle
      integer(c_int32_t), value :: cycles_per_variant
      integer(c_int32_t), intent(out) :: chosen(2)
    end function

    integer(c_int) function or_batch_autotune(batch, cycles_per_variant, chosen) bind(c, name="or_batch_autotune")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: batch
      integer(c_int32_t), value :: cycles_per_variant
      integer(c_int32_t), intent(out) :: chosen(4)
    end function

    ! SFORC3 on one group (sforc3.F:131-166 argument meaning): GBUF%SIG, EINT, RHO, VOL, PLA,
    ! EPSD, OFF, QVIS of the group; FSKY rows written like SCUMU3P; DT2T min-folded.
    integer(c_int) function or_sforc3_group(device, p, numnod, nel, nft, x, v, ixs, sig, eint, rho, vol, pla, &
        epsd, off, qvis, lsky, fsky, iads, dt1, dt2t, mstop) bind(c, name="or_sforc3_group")
      import :: c_int, c_int32_t, c_int64_t, c_double, or_hex_params_t
      integer(c_int), value :: device
      type(or_hex_params_t), intent(in) :: p
      integer(c_int64_t), value :: numnod, nft, lsky
      integer(c_int32_t), value :: nel
      real(c_double), intent(in) :: x(3, *), v(3, *), vol(*)
      integer(c_int32_t), intent(in) :: ixs(11, *), iads(8, *)
      real(c_double), intent(inout) :: sig(*), eint(*), rho(*), pla(*), epsd(*), off(*), qvis(*)
      real(c_double), intent(inout) :: fsky(8, *)
      real(c_double), value :: dt1
      real(c_double), intent(inout) :: dt2t
      integer(c_int32_t), intent(inout) :: mstop
    end function

    ! the same without any device allocation (one file-scope device array per device; calls
    ! that overlap on one device race), for callers that cannot allocate
    integer(c_int) function or_sforc3_group_static(device, p, numnod, nel, nft, x, v, ixs, sig, eint, rho, vol, &
        pla, epsd, off, qvis, lsky, fsky, iads, dt1, dt2t, mstop) bind(c, name="or_sforc3_group_static")
      import :: c_int, c_int32_t, c_int64_t, c_double, or_hex_params_t
      integer(c_int), value :: device
      type(or_hex_params_t), intent(in) :: p
      integer(c_int64_t), value :: numnod, nft, lsky
      integer(c_int32_t), value :: nel
      real(c_double), intent(in) :: x(3, *), v(3, *), vol(*)
      integer(c_int32_t), intent(in) :: ixs(11, *), iads(8, *)
      real(c_double), intent(inout) :: sig(*), eint(*), rho(*), pla(*), epsd(*), off(*), qvis(*)
      real(c_double), intent(inout) :: fsky(8, *)
      real(c_double), value :: dt1
      real(c_double), intent(inout) :: dt2t
      integer(c_int32_t), intent(inout) :: mstop
    end function

    ! CFORC3 on one group (cforc3.F:85-113 argument meaning): GBUF%OFF, THK, FOR, MOM, STRA,
    ! HOURG, EINT, EPSD; LBUF SIG / PLA / EPSD per point; FSKY rows written like CUPDT3P.
    integer(c_int) function or_cforc3_group(device, p, numnod, nel, nft, x, v, vr, ixc, offg, thk, thke, forc, &
        mom, gstr, hourg, eint, epsd, sig, pla, epsd_ip, lsky, fsky, iadc, dt1, dt2t) &
        bind(c, name="or_cforc3_group")
      import :: c_int, c_int32_t, c_int64_t, c_double, or_q4_params_t
      integer(c_int), value :: device
      type(or_q4_params_t), intent(in) :: p
      integer(c_int64_t), value :: numnod, nft, lsky
      integer(c_int32_t), value :: nel
      real(c_double), intent(in) :: x(3, *), v(3, *), vr(3, *), thke(*)
      integer(c_int32_t), intent(in) :: ixc(7, *), iadc(4, *)
      real(c_double), intent(inout) :: offg(*), thk(*), forc(*), mom(*), gstr(*), hourg(*), eint(*), epsd(*)
      real(c_double), intent(inout) :: sig(*), pla(*), epsd_ip(*)
      real(c_double), intent(inout) :: fsky(8, *)
      real(c_double), value :: dt1
      real(c_double), intent(inout) :: dt2t
    end function

    ! ASSPAR4 -> DEPLA on the Engine's arrays: ADSKY(NUMNOD+1), FSKY(8,LSKY), MS, IN, node codes,
    ! imposed velocities, X/V/VR/D(3,NUMNOD) updated, STIFN/STIFR out (vr, in, stifn, stifr
    ! may be c_null_ptr through the c_ptr forms), time state advanced
    integer(c_int) function or_nodal_update(device, numnod, lsky, adsky, fsky, ms, in, code, vimp, x, v, vr, d, &
        stifn, stifr, dt2t, dtmx, ts) bind(c, name="or_nodal_update")
      import :: c_int, c_int64_t, c_double, c_ptr, or_time_state_t
      integer(c_int), value :: device
      integer(c_int64_t), value :: numnod, lsky
      type(c_ptr), value :: adsky, fsky, ms, in, code, vimp, x, v, vr, d, stifn, stifr
      real(c_double), value :: dt2t, dtmx
      type(or_time_state_t), intent(inout) :: ts
    end function

    integer(c_int) function or_math_eval(device, fn, n, x, y, out) bind(c, name="or_math_eval")
      import :: c_int, c_int64_t, c_double
      integer(c_int), value :: device, fn
      integer(c_int64_t), value :: n
      real(c_double), intent(in) :: x(*), y(*)
      real(c_double), intent(out) :: out(*)
    end function

    integer(c_size_t) function c_strlen(s) bind(c, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! Engine glue.  The predicates decide, per element group, whether the group runs on this
  ! path; they are written from the path's scope (the SFORC3 / CFORC3 branches this library
  ! implements), after the pattern of the reference's own GPU predicate (forintc.F:301-318).
  ! Their arguments are the group's IPARG column (IPARG(NPARG,NG)), the material law's
  ! integer parameters (mat_param%iparam(1:4): IFORM, ICC, VP, ISRATE) and the group's
  ! property words the branches depend on.

  ! An 8-node solid group of law 2 on the SFORC3 branch of this path: ITY = 1 (IPARG(5)),
  ! MLW = 2 (IPARG(1)), group active (IPARG(8) /= 1), Isolid = 1 (JHBE = IPARG(23) = 1),
  ! ISMSTR = 4 (IPARG(9)), Lagrangian (JLAG = IPARG(14) = 1, JALE = IPARG(7) = 0, JEUL =
  ! IPARG(11) = 0), no failure
  ! model (IPARG(43) = 0), no thermal coupling (IPARG(13) = 0), global frame (JCVT = IPARG(37)
  ! = 0), and a Johnson-Cook form with isotropic hardening (IFORM = 0, FISOKIN = 0) whose
  ! strain rate is the total or deviatoric one (VP = 2 or 3).
  logical function or_gpu_eligible_solid(iparg, iparam, fisokin) result(ok)
    integer, intent(in) :: iparg(*), iparam(4)
    real(c_double), intent(in) :: fisokin
    ok = iparg(5) == 1 .and. iparg(1) == 2 .and. iparg(8) /= 1 .and. iparg(23) == 1 .and. iparg(9) == 4 &
         .and. iparg(14) == 1 .and. iparg(7) == 0 .and. iparg(11) == 0 .and. iparg(43) == 0 &
         .and. iparg(13) == 0 &
         .and. iparg(37) == 0 .and. iparam(1) == 0 .and. fisokin == 0.0d0 &
         .and. (iparam(3) == 2 .or. iparam(3) == 3)
  end function

  ! A 4-node shell group of law 2 on the CFORC3 branch of this path: ITY = 3, MLW = 2, group
  ! active, Ishell = 1 (IHBE = IPARG(23) = 1), 2..10 points through the thickness (NPT =
  ! IPARG(6)), Iplas 0..2 (IPARG(29)), ISMSTR = 2 (IPARG(9)), property type 1 (IGTYP =
  ! IPARG(38)), no failure model (IPARG(43) = 0), no sub-stack (IPARG(71) = 0), no XFEM
  ! (IPARG(54) = 0), no seat belt (IPARG(91) = 0), and a Johnson-Cook form with the total
  ! strain rate (IFORM = 0, VP = 2).
  logical function or_gpu_eligible_shell(iparg, iparam) result(ok)
    integer, intent(in) :: iparg(*), iparam(4)
    ok = iparg(5) == 3 .and. iparg(1) == 2 .and. iparg(8) /= 1 .and. iparg(23) == 1 &
         .and. iparg(6) >= 2 .and. iparg(6) <= 10 .and. iparg(29) >= 0 .and. iparg(29) <= 2 &
         .and. iparg(9) == 2 .and. iparg(38) == 1 .and. iparg(43) == 0 .and. iparg(71) == 0 &
         .and. iparg(54) == 0 .and. iparg(91) == 0 .and. iparam(1) == 0 .and. iparam(3) == 2
  end function

  ! or_hex_params_t from the values SFORC3's callees read: the law-2 material (mat_param%rho,
  ! %bulk, %shear; %uparam(1:10) = CA, CB, CN, EPMX, SIGMX, CC, EPDR, FISOKIN, ASRATE, M_EXP
  ! and %iparam(1:4) = IFORM, ICC, VP, ISRATE, m2law.F:135-155), the /PROP/SOLID words
  ! GEO(13:17) = QH, QA, QB, VNS1, VNS2 (shvis3.F:167, mqviscb.F:201-204), DTFAC1(1) and DTMX.
  subroutine or_pack_hex_params(rho, bulk, shear, uparam, iparam, geo, dtfac, dtmx, p)
    real(c_double), intent(in) :: rho, bulk, shear, uparam(10), geo(*), dtfac, dtmx
    integer, intent(in) :: iparam(4)
    type(or_hex_params_t), intent(out) :: p
    p%rho0 = rho
    p%bulk = bulk
    p%shear = shear
    p%ca = uparam(1)
    p%cb = uparam(2)
    p%cn = uparam(3)
    p%epmx = uparam(4)
    p%sigmx = uparam(5)
    p%cc = uparam(6)
    p%epdr = uparam(7)
    p%fisokin = uparam(8)
    p%asrate = uparam(9)
    p%m_exp = uparam(10)
    p%iform = iparam(1)
    p%icc = iparam(2)
    p%vp = iparam(3)
    p%israte = iparam(4)
    p%qh = geo(13)
    p%qa = geo(14)
    p%qb = geo(15)
    p%vns1 = geo(16)
    p%vns2 = geo(17)
    p%dtfac = dtfac
    p%dtmx = dtmx
  end subroutine

  ! or_q4_params_t from the values CFORC3's callees read: PM(1) RHO, PM(20) YOUNG, PM(21) NU,
  ! PM(22) G, PM(27) SSP (ccoef3.F:125-131); the law's %uparam(1:10) and %iparam(1:4) as for
  ! solids (sigeps02c.F:91-123); the /PROP/SHELL words GEO(13:15) = Hm, Hf, Hr (chvis3),
  ! GEO(37) = ISHEAR and GEO(38) = shear factor (ccoef3.F:161-163); the group's NPT, ITHK,
  ! IPLA (IPARG(6), IPARG(28), IPARG(29)); DTFAC1(3) and DTMIN1(3).
  subroutine or_pack_q4_params(pm, uparam, iparam, geo, iparg, dtfac, dtmin, p)
    real(c_double), intent(in) :: pm(*), uparam(10), geo(*), dtfac, dtmin
    integer, intent(in) :: iparam(4), iparg(*)
    type(or_q4_params_t), intent(out) :: p
    p%rho = pm(1)
    p%young = pm(20)
    p%nu = pm(21)
    p%shear = pm(22)
    p%ssp = pm(27)
    p%ca = uparam(1)
    p%cb = uparam(2)
    p%cn = uparam(3)
    p%epmx = uparam(4)
    p%ymax = uparam(5)
    p%cc = uparam(6)
    p%epdr = uparam(7)
    p%asrate = uparam(9)
    p%m_exp = uparam(10)
    p%iform = iparam(1)
    p%icc = iparam(2)
    p%vp = iparam(3)
    p%israte = iparam(4)
    p%npt = iparg(6)
    p%ithk = iparg(28)
    p%ipla = iparg(29)
    p%ishear = nint(geo(37))
    p%hm = geo(13)
    p%hf = geo(14)
    p%hr = geo(15)
    p%fsh = geo(38)
    p%dtfac = dtfac
    p%dtmin = dtmin
  end subroutine

  ! Text of the last failure of the calling thread.
  function or_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: i, n
    p = or_last_error_c()
    if (.not. c_associated(p)) then
      msg = ""
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate (character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function

end module or_gpu_mod
