! Fortran caller of the chemistry step through the ISO_C_BINDING module, as a
! Tiger_NGA rank would call it in place of finitechem_source_chemistry.
! Usage: drive_step <mech> <in.bin> <out.bin> [network.bin | --check-shapes]
!   with a network file the surrogate-guided step is called instead;
!   --check-shapes calls both steps with wrongly shaped optional arrays and
!   prints the error codes (needs no device: the module checks the extents
!   before it calls the library)
!   in.bin  (stream): int32 nx ny nz nover nscalar isc_sc isc_T; float64 dt P rtol atol;
!           RHO_(nxo,nyo,nzo); SC_(nxo,nyo,nzo,nscalar); int32 mask(nxo,nyo)
!   out.bin (stream): int32 ierr; SC_; HR_gp; chemsrc_gp(...,3); int64 stats(11)
program drive_step
  use iso_c_binding
  use tiger_chem
  implicit none
  character(len=64) :: mech
  character(len=512) :: fin, fout, fnet
  integer(c_int) :: hdr(7)
  integer :: nx, ny, nz, nover, nscalar, ierr, u
  real(c_double) :: dt, P, rtol, atol
  real(c_double), allocatable, target :: RHO_(:,:,:), SC_(:,:,:,:), HR_gp(:,:,:), chemsrc_gp(:,:,:,:)
  integer(c_int), allocatable, target :: mask(:,:)
  type(tgr_grid) :: g
  type(tgr_stats) :: st

  call get_command_argument(1, mech)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)
  fnet = ''
  if (command_argument_count() >= 4) call get_command_argument(4, fnet)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) hdr
  nx = hdr(1); ny = hdr(2); nz = hdr(3); nover = hdr(4); nscalar = hdr(5)
  read(u) dt, P, rtol, atol
  allocate(RHO_(1-nover:nx+nover, 1-nover:ny+nover, 1-nover:nz+nover))
  allocate(SC_(1-nover:nx+nover, 1-nover:ny+nover, 1-nover:nz+nover, nscalar))
  allocate(HR_gp(1-nover:nx+nover, 1-nover:ny+nover, 1-nover:nz+nover))
  allocate(chemsrc_gp(1-nover:nx+nover, 1-nover:ny+nover, 1-nover:nz+nover, 3))
  allocate(mask(1-nover:nx+nover, 1-nover:ny+nover))
  read(u) RHO_, SC_, mask
  close(u)
  HR_gp = 0.0_c_double
  chemsrc_gp = 0.0_c_double

  g = tgr_grid(1, nx, 1, ny, 1, nz, 1-nover, nx+nover, 1-nover, ny+nover, 1-nover, nz+nover, &
       1-nover, nx+nover, 1-nover, ny+nover, nscalar, hdr(6), hdr(7))
  if (trim(fnet) == '--check-shapes') then
     call check_shapes()
     stop
  end if
  call tiger_chem_init(mech, rtol, atol, ierr)
  if (len_trim(fnet) > 0) then
     if (ierr == 0) call tiger_chem_load_surrogate(fnet, ierr)
     if (ierr == 0) call tiger_chem_source_chemistry_surrogate(RHO_, SC_, g, dt, P, P, ierr, mask, HR_gp, &
          chemsrc_gp, st)
  else
     if (ierr == 0) call tiger_chem_source_chemistry(RHO_, SC_, g, dt, P, P, ierr, mask, HR_gp, chemsrc_gp, st)
  end if
  if (ierr /= 0) then
     write(*, '(a, i0, 2a)') 'tiger_chem error ', ierr, ': ', trim(tiger_chem_message())
     st = tgr_stats(0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0.0_c_double)
  end if
  call tiger_chem_final()

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) int(ierr, c_int), SC_, HR_gp, chemsrc_gp
  write(u) st%cells, st%nst, st%nfe, st%nfe_dq, st%nje, st%nsetups, st%nni, st%netf, st%ncfn, st%nlu, st%failed
  close(u)
  if (ierr /= 0) stop 1

contains

  ! every optional array with one extent off by one: each call must return
  ! TGR_E_ARG without touching the fields
  subroutine check_shapes()
    real(c_double), allocatable, target :: hr_bad(:,:,:), cs_bad(:,:,:,:)
    integer(c_int), allocatable, target :: mask_bad(:,:)
    integer :: e(6)
    allocate(hr_bad(nx + 2 * nover, ny + 2 * nover, nz + 2 * nover - 1))
    allocate(cs_bad(nx + 2 * nover, ny + 2 * nover, nz + 2 * nover, 2))
    allocate(mask_bad(nx + 2 * nover - 1, ny + 2 * nover))
    call tiger_chem_source_chemistry(RHO_, SC_, g, dt, P, P, e(1), mask_bad)
    call tiger_chem_source_chemistry(RHO_, SC_, g, dt, P, P, e(2), mask, hr_bad)
    call tiger_chem_source_chemistry(RHO_, SC_, g, dt, P, P, e(3), mask, HR_gp, cs_bad)
    call tiger_chem_source_chemistry_surrogate(RHO_, SC_, g, dt, P, P, e(4), mask_bad)
    call tiger_chem_source_chemistry_surrogate(RHO_, SC_, g, dt, P, P, e(5), mask, hr_bad)
    call tiger_chem_source_chemistry_surrogate(RHO_, SC_, g, dt, P, P, e(6), mask, HR_gp, cs_bad)
    write(*, '(a, 6(1x, i0))') 'shape errors:', e
  end subroutine check_shapes

end program drive_step
