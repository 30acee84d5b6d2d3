! config_case.f90 -- a config plug-in (the four hooks of a reference config,
! e.g. for_config/gm1.f90:7-121) that takes its integration from a case file
! written by raytracer_3d_amd/cases.py: int32[16] header, float64[8]
! parameters, six interior T-spacing vectors, Nsqr/U/V halo'd, interior E.
! The file name is the driver's first command-line argument; an argument
! snap=N asks for diagnose every N steps.  The forcing of a case, if any,
! is set on the device (rt6d_set_forcing), so E stays resident there.
module config_case_state
  implicit none
  integer :: case_unit = -1
  integer :: case_steps = 0
end module config_case_state

subroutine set_parameter
  use main_module
  use config_case_state
  implicit none
  integer :: hdr(16), i, snapsteps
  real*8 :: rh(8)
  character(len=1024) :: fname, arg
  call get_command_argument(1, fname)
  if (len_trim(fname) == 0) call halt_stop(' usage: rt6d_driver CASE [snap=N] [out=FILE] [fp32] [restart]')
  snapsteps = 0
  do i = 2, command_argument_count()
    call get_command_argument(i, arg)
    if (arg(1:5) == 'snap=') read(arg(6:), *) snapsteps
  enddo
  open(newunit=case_unit, file=trim(fname), access='stream', form='unformatted', status='old', action='read')
  read(case_unit) hdr
  if (hdr(1) /= 1144149074) call halt_stop(' not a case file')
  read(case_unit) rh
  nx = hdr(3); ny = hdr(4); nz = hdr(5); nk = hdr(6); nl = hdr(7); nm = hdr(8)
  enable_upper_reflection = hdr(9) /= 0
  enable_lower_reflection = hdr(10) /= 0
  case_steps = hdr(11)
  dt = rh(1); f0 = rh(2); sign_of_omega_i = rh(3)
  k_fixed = rh(4); l_fixed = rh(5); m_fixed = rh(6)
  ! main.f90:51 takes int(runlen/dt) steps and main.f90:74 diagnoses every
  ! int(snapint/dt) steps: half a step more keeps both whole step counts
  runlen = (case_steps + 0.5d0)*dt
  snapint = 0
  if (snapsteps > 0) snapint = (snapsteps + 0.5d0)*dt
  host_forcing = .false.
end subroutine set_parameter

subroutine set_grid
  use main_module
  use config_case_state
  implicit none
  read(case_unit) dxt(1:nx)
  read(case_unit) dyt(1:ny)
  read(case_unit) dzt(1:nz)
  read(case_unit) dkt(1:nk)
  read(case_unit) dlt(1:nl)
  read(case_unit) dmt(1:nm)
end subroutine set_grid

subroutine set_initial_conditions
  use main_module
  use config_case_state
  implicit none
  real*8, allocatable :: buf(:, :, :, :, :, :)
  read(case_unit) Nsqr
  read(case_unit) U
  read(case_unit) V
  ! a contiguous buffer: amdflang moves a strided section element by element
  allocate(buf(nx, ny, nz, nk, nl, nm))
  read(case_unit) buf
  E(1:nx, 1:ny, 1:nz, 1:nk, 1:nl, 1:nm) = buf
  deallocate(buf)
  close(case_unit)
end subroutine set_initial_conditions

subroutine set_forcing
end subroutine set_forcing
