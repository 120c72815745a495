! ISO_C_BINDING interface of libtiger_chem (include/tiger_chem.h) for a Fortran
! host, plus the drop-in body of finitechem_source_chemistry.
!
! The reference's chemistry step is the external subroutine
! finitechem_source_chemistry(RHO_, SC_) (combustion/finitechem.f90:962-1071),
! called from combustion_source_scalar_full (combustion/combustion.f90:1048).
! It reads its other inputs from modules (dt, P_thermo, P_thermo_old, mask,
! isc_sc, isc_T, rtol, atol) and writes HR_gp and chemsrc_gp.  The procedures
! here take those inputs as arguments, so the replacement body is one call
! (see tiger_chem_source_chemistry below); the fields stay in the caller's
! layout and are advanced in place.
module tiger_chem
  use iso_c_binding
  implicit none
  private

  integer, parameter, public :: TGR_E_ARG = -1, TGR_E_HIP = -2, TGR_E_MECH = -3, TGR_E_NOMEM = -4

  ! block extents and scalar slots (1-based), tgr_grid in the C header
  type, bind(C), public :: tgr_grid
     integer(c_int) :: imin_, imax_, jmin_, jmax_, kmin_, kmax_
     integer(c_int) :: imino_, imaxo_, jmino_, jmaxo_, kmino_, kmaxo_
     integer(c_int) :: imino, imaxo, jmino, jmaxo
     integer(c_int) :: nscalar, isc_sc, isc_T
  end type tgr_grid

  ! metric operators of the flow solver (device arrays), tgr_metric in the C
  ! header; TGR_DIFF_* are the flags of tgr_chem_diffusion_device
  type, bind(C), public :: tgr_metric
     integer(c_int) :: st1, st2
     type(c_ptr) :: interp_sc_x, interp_sc_y, interp_sc_z
     type(c_ptr) :: grad_x, grad_y, grad_z
     type(c_ptr) :: div_u, div_v, div_w
     type(c_ptr) :: interp_u_xm, interp_v_ym, interp_w_zm
     type(c_ptr) :: grad_xm, grad_ym, grad_zm
  end type tgr_metric
  integer, parameter, public :: TGR_DIFF_MOLAR = 1, TGR_DIFF_GASRAD = 2, TGR_DIFF_CV = 4

  ! per-call integrator statistics, tgr_stats in the C header
  type, bind(C), public :: tgr_stats
     integer(c_int64_t) :: cells, nst, nfe, nfe_dq, nje, nsetups, nni, netf, ncfn, nlu, failed
     real(c_double) :: kernel_ms
  end type tgr_stats

  public :: tgr_chem_create, tgr_chem_destroy, tgr_chem_source_chemistry, tgr_chem_source_chemistry_device
  public :: tgr_chem_molar_mass, tgr_chem_species_name, tgr_chem_set_option, tgr_chem_set_option_real
  public :: tgr_chem_last_error
  public :: tgr_chem_properties_device, tgr_chem_temperature_device, tgr_chem_diffusion_device
  public :: tgr_chem_load_surrogate, tgr_chem_source_chemistry_surrogate
  public :: tiger_chem_load_surrogate, tiger_chem_source_chemistry_surrogate
  public :: tiger_chem_init, tiger_chem_final, tiger_chem_source_chemistry, tiger_chem_message
  public :: tiger_chem_ctx

  interface
     integer(c_int) function tgr_chem_create(mech, rtol, atol, max_steps, ctx) bind(C, name="tgr_chem_create")
       import :: c_int, c_char, c_double, c_ptr
       character(kind=c_char), dimension(*), intent(in) :: mech
       real(c_double), value :: rtol, atol
       integer(c_int), value :: max_steps
       type(c_ptr), intent(out) :: ctx
     end function tgr_chem_create

     integer(c_int) function tgr_chem_destroy(ctx) bind(C, name="tgr_chem_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function tgr_chem_destroy

     integer(c_int) function tgr_chem_molar_mass(ctx, W) bind(C, name="tgr_chem_molar_mass")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), dimension(*) :: W
     end function tgr_chem_molar_mass

     integer(c_int) function tgr_chem_species_name(ctx, k, buf, len) bind(C, name="tgr_chem_species_name")
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: ctx
       integer(c_int), value :: k, len
       character(kind=c_char), dimension(*) :: buf
     end function tgr_chem_species_name

     integer(c_int) function tgr_chem_set_option(ctx, key, val) bind(C, name="tgr_chem_set_option")
       import :: c_int, c_ptr, c_char, c_int64_t
       type(c_ptr), value :: ctx
       character(kind=c_char), dimension(*), intent(in) :: key
       integer(c_int64_t), value :: val
     end function tgr_chem_set_option

     integer(c_int) function tgr_chem_set_option_real(ctx, key, val) bind(C, name="tgr_chem_set_option_real")
       import :: c_int, c_ptr, c_char, c_double
       type(c_ptr), value :: ctx
       character(kind=c_char), dimension(*), intent(in) :: key
       real(c_double), value :: val
     end function tgr_chem_set_option_real

     ! host arrays; mask, HR_gp, chemsrc_gp, stats may be c_null_ptr
     integer(c_int) function tgr_chem_source_chemistry(ctx, dt, p_thermo, p_thermo_old, g, rho, sc, mask, hr, &
          chemsrc, stats) bind(C, name="tgr_chem_source_chemistry")
       import :: c_int, c_ptr, c_double, tgr_grid
       type(c_ptr), value :: ctx
       real(c_double), value :: dt, p_thermo, p_thermo_old
       type(tgr_grid), intent(in) :: g
       type(c_ptr), value :: rho, sc, mask, hr, chemsrc, stats
     end function tgr_chem_source_chemistry

     ! device (HBM-resident) arrays, enqueued on stream
     integer(c_int) function tgr_chem_source_chemistry_device(ctx, dt, p_thermo, p_thermo_old, g, rho, sc, mask, &
          hr, chemsrc, stream, stats) bind(C, name="tgr_chem_source_chemistry_device")
       import :: c_int, c_ptr, c_double, tgr_grid
       type(c_ptr), value :: ctx
       real(c_double), value :: dt, p_thermo, p_thermo_old
       type(tgr_grid), intent(in) :: g
       type(c_ptr), value :: rho, sc, mask, hr, chemsrc, stream, stats
     end function tgr_chem_source_chemistry_device

     ! finitechem_density / finitechem_viscosity / finitechem_diffusivity on
     ! device arrays (rho, visc, diff, lewis, mask may be c_null_ptr)
     integer(c_int) function tgr_chem_properties_device(ctx, p_thermo, g, sc, mask, lewis, isc_zmix, rho, visc, &
          diff, stream) bind(C, name="tgr_chem_properties_device")
       import :: c_int, c_ptr, c_double, tgr_grid
       type(c_ptr), value :: ctx
       real(c_double), value :: p_thermo
       type(tgr_grid), intent(in) :: g
       type(c_ptr), value :: sc, mask, lewis
       integer(c_int), value :: isc_zmix
       type(c_ptr), value :: rho, visc, diff, stream
     end function tgr_chem_properties_device

     ! finitechem_diffusion (and with src the update of
     ! finitechem_source_transport) on device arrays; diff_corr, diff_gp,
     ! src and mask may be c_null_ptr
     integer(c_int) function tgr_chem_diffusion_device(ctx, dt, g, m, flags, rho, sc, diff, mask, diff_corr, &
          diff_gp, src, stream) bind(C, name="tgr_chem_diffusion_device")
       import :: c_int, c_ptr, c_double, tgr_grid, tgr_metric
       type(c_ptr), value :: ctx
       real(c_double), value :: dt
       type(tgr_grid), intent(in) :: g
       type(tgr_metric), intent(in) :: m
       integer(c_int), value :: flags
       type(c_ptr), value :: rho, sc, diff, mask, diff_corr, diff_gp, src, stream
     end function tgr_chem_diffusion_device

     ! temperature from mixture enthalpy, device arrays
     integer(c_int) function tgr_chem_temperature_device(ctx, n, Y, hmix, T0, T, iters, stream) &
          bind(C, name="tgr_chem_temperature_device")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: n
       type(c_ptr), value :: Y, hmix, T0, T, iters, stream
     end function tgr_chem_temperature_device

     ! surrogate-guided step: load a trained network file, then the step on
     ! host arrays (sstats: tgr_surrogate_stats or c_null_ptr)
     integer(c_int) function tgr_chem_load_surrogate(ctx, path) bind(C, name="tgr_chem_load_surrogate")
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char), dimension(*), intent(in) :: path
     end function tgr_chem_load_surrogate

     integer(c_int) function tgr_chem_source_chemistry_surrogate(ctx, dt, p_thermo, p_thermo_old, g, rho, sc, &
          mask, hr, chemsrc, stats, sstats) bind(C, name="tgr_chem_source_chemistry_surrogate")
       import :: c_int, c_ptr, c_double, tgr_grid
       type(c_ptr), value :: ctx
       real(c_double), value :: dt, p_thermo, p_thermo_old
       type(tgr_grid), intent(in) :: g
       type(c_ptr), value :: rho, sc, mask, hr, chemsrc, stats, sstats
     end function tgr_chem_source_chemistry_surrogate

     type(c_ptr) function tgr_chem_last_error() bind(C, name="tgr_chem_last_error")
       import :: c_ptr
     end function tgr_chem_last_error
  end interface

  ! the context of this rank (finitechem_init creates it once)
  type(c_ptr), save :: tiger_chem_ctx = c_null_ptr

contains

  ! message of the last failing call
  function tiger_chem_message() result(msg)
    character(len=256) :: msg
    character(kind=c_char), dimension(:), pointer :: p
    type(c_ptr) :: cp
    integer :: i
    msg = ''
    cp = tgr_chem_last_error()
    if (.not. c_associated(cp)) return
    call c_f_pointer(cp, p, [256])
    do i = 1, 256
       if (p(i) == c_null_char) exit
       msg(i:i) = p(i)
    end do
  end function tiger_chem_message

  ! replaces the set-up part of finitechem_init (finitechem.f90:184-319) and
  ! the CVODE set-up of finitechem.f90:1014-1028; ierr = 0 on success
  subroutine tiger_chem_init(mech, rtol, atol, ierr)
    character(len=*), intent(in) :: mech
    real(c_double), intent(in) :: rtol, atol
    integer, intent(out) :: ierr
    ierr = tgr_chem_create(trim(mech) // c_null_char, rtol, atol, 10000_c_int, tiger_chem_ctx)
  end subroutine tiger_chem_init

  subroutine tiger_chem_final()
    integer(c_int) :: ierr
    if (c_associated(tiger_chem_ctx)) ierr = tgr_chem_destroy(tiger_chem_ctx)
    tiger_chem_ctx = c_null_ptr
  end subroutine tiger_chem_final

  ! finitechem_source_chemistry(RHO_, SC_) with its module inputs as
  ! arguments.  RHO_(imino_:imaxo_, jmino_:jmaxo_, kmino_:kmaxo_) and
  ! SC_(..., nscalar) are advanced in place over dt at P = (P_thermo +
  ! P_thermo_old)/2 on every interior cell with mask /= 1; HR_gp and
  ! chemsrc_gp(..., 1:3) (H2O, CO, CO2) are written where present.
  subroutine tiger_chem_source_chemistry(RHO_, SC_, g, dt, P_thermo, P_thermo_old, ierr, mask, HR_gp, &
       chemsrc_gp, stats)
    type(tgr_grid), intent(in) :: g
    real(c_double), intent(inout), target, contiguous :: RHO_(:,:,:)
    real(c_double), intent(inout), target, contiguous :: SC_(:,:,:,:)
    real(c_double), intent(in) :: dt, P_thermo, P_thermo_old
    integer, intent(out) :: ierr
    integer(c_int), intent(in), target, contiguous, optional :: mask(:,:)
    real(c_double), intent(inout), target, contiguous, optional :: HR_gp(:,:,:)
    real(c_double), intent(inout), target, contiguous, optional :: chemsrc_gp(:,:,:,:)
    type(tgr_stats), intent(out), target, optional :: stats
    type(c_ptr) :: pm, ph, pc, ps
    pm = c_null_ptr
    ph = c_null_ptr
    pc = c_null_ptr
    ps = c_null_ptr
    if (present(mask)) pm = c_loc(mask)
    if (present(HR_gp)) ph = c_loc(HR_gp)
    if (present(chemsrc_gp)) pc = c_loc(chemsrc_gp)
    if (present(stats)) ps = c_loc(stats)
    if (.not. extents_ok(g, RHO_, SC_, mask, HR_gp, chemsrc_gp)) then
       ierr = TGR_E_ARG
       return
    end if
    ierr = tgr_chem_source_chemistry(tiger_chem_ctx, dt, P_thermo, P_thermo_old, g, c_loc(RHO_), c_loc(SC_), &
         pm, ph, pc, ps)
  end subroutine tiger_chem_source_chemistry

  ! attach a trained network file (e.g. tiger_nga_amd/models/gri30.bin)
  subroutine tiger_chem_load_surrogate(path, ierr)
    character(len=*), intent(in) :: path
    integer, intent(out) :: ierr
    ierr = tgr_chem_load_surrogate(tiger_chem_ctx, trim(path) // c_null_char)
  end subroutine tiger_chem_load_surrogate

  ! tiger_chem_source_chemistry by the surrogate-guided path: cells the
  ! guided solve does not accept are integrated by the plain step in the same
  ! call
  subroutine tiger_chem_source_chemistry_surrogate(RHO_, SC_, g, dt, P_thermo, P_thermo_old, ierr, mask, HR_gp, &
       chemsrc_gp, stats)
    type(tgr_grid), intent(in) :: g
    real(c_double), intent(inout), target, contiguous :: RHO_(:,:,:)
    real(c_double), intent(inout), target, contiguous :: SC_(:,:,:,:)
    real(c_double), intent(in) :: dt, P_thermo, P_thermo_old
    integer, intent(out) :: ierr
    integer(c_int), intent(in), target, contiguous, optional :: mask(:,:)
    real(c_double), intent(inout), target, contiguous, optional :: HR_gp(:,:,:)
    real(c_double), intent(inout), target, contiguous, optional :: chemsrc_gp(:,:,:,:)
    type(tgr_stats), intent(out), target, optional :: stats
    type(c_ptr) :: pm, ph, pc, ps
    pm = c_null_ptr
    ph = c_null_ptr
    pc = c_null_ptr
    ps = c_null_ptr
    if (present(mask)) pm = c_loc(mask)
    if (present(HR_gp)) ph = c_loc(HR_gp)
    if (present(chemsrc_gp)) pc = c_loc(chemsrc_gp)
    if (present(stats)) ps = c_loc(stats)
    if (.not. extents_ok(g, RHO_, SC_, mask, HR_gp, chemsrc_gp)) then
       ierr = TGR_E_ARG
       return
    end if
    ierr = tgr_chem_source_chemistry_surrogate(tiger_chem_ctx, dt, P_thermo, P_thermo_old, g, c_loc(RHO_), &
         c_loc(SC_), pm, ph, pc, ps, c_null_ptr)
  end subroutine tiger_chem_source_chemistry_surrogate

  ! the extents of every array the C side reads or writes: RHO_ and the
  ! leading extents of SC_, HR_gp and chemsrc_gp are the ghosted block,
  ! chemsrc_gp has 3 planes, mask is (imino:imaxo, jmino:jmaxo); a wrong
  ! shape would otherwise be read or written past its end
  logical function extents_ok(g, RHO_, SC_, mask, HR_gp, chemsrc_gp)
    type(tgr_grid), intent(in) :: g
    real(c_double), intent(in) :: RHO_(:,:,:)
    real(c_double), intent(in) :: SC_(:,:,:,:)
    integer(c_int), intent(in), optional :: mask(:,:)
    real(c_double), intent(in), optional :: HR_gp(:,:,:)
    real(c_double), intent(in), optional :: chemsrc_gp(:,:,:,:)
    integer :: n1, n2, n3
    n1 = g%imaxo_ - g%imino_ + 1
    n2 = g%jmaxo_ - g%jmino_ + 1
    n3 = g%kmaxo_ - g%kmino_ + 1
    extents_ok = size(RHO_, 1) == n1 .and. size(RHO_, 2) == n2 .and. size(RHO_, 3) == n3 .and. &
         size(SC_, 1) == n1 .and. size(SC_, 2) == n2 .and. size(SC_, 3) == n3 .and. size(SC_, 4) == g%nscalar
    if (present(mask)) extents_ok = extents_ok .and. size(mask, 1) == g%imaxo - g%imino + 1 .and. &
         size(mask, 2) == g%jmaxo - g%jmino + 1
    if (present(HR_gp)) extents_ok = extents_ok .and. size(HR_gp, 1) == n1 .and. size(HR_gp, 2) == n2 .and. &
         size(HR_gp, 3) == n3
    if (present(chemsrc_gp)) extents_ok = extents_ok .and. size(chemsrc_gp, 1) == n1 .and. &
         size(chemsrc_gp, 2) == n2 .and. size(chemsrc_gp, 3) == n3 .and. size(chemsrc_gp, 4) == 3
  end function extents_ok

end module tiger_chem
