! rt6d_driver.f90 -- Fortran driver with the reference's set-up order and
! time loop (for_src/main.f90:36-79), running the hot path on the GPU through
! the drop-in routines of rt6d_hot.f90.  Linked with a config plug-in: this
! package's config_case.f90 (a case file from raytracer_3d_amd/cases.py), or
! a reference plug-in (for_config/*.f90) through the Makefile's PLUGIN build.
!
! usage: rt6d_driver [plug-in arguments] [out=FILE] [fp32] [restart]
!   out=FILE   the final interior E is written to FILE as a raw float64 stream
!   fp32       run the FP32 form
!   restart    read_restart after calc_parameter (main.f90:49) and
!              write_restart after the loop (main.f90:81), in the working
!              directory
! The run takes int(runlen/dt) steps from itt (main.f90:51); with snapint > 0
! it calls diagnose every int(snapint/dt) steps and at step 0 (main.f90:74),
! writing snap3D.cdf and snap6D.cdf in the working directory.  With
! host_forcing on (the default), E is brought to the host before every
! set_forcing and back to the device after it (main.f90:69).
! The environment variable RT6D_RESTART_MAX_SUBRECORD, when set, replaces
! gfortran's subrecord limit of the restart records (bytes).
program main
  use main_module
  use rt6d_abi
  implicit none
  integer :: enditt, snapsteps, iu, itt0, i, ios
  logical :: use_restart
  integer(8) :: c0, c1, crate, lim
  character(len=1024) :: arg, outfile
  real*8 :: secs
  real*8, allocatable :: buf(:, :, :, :, :, :)

  outfile = '-'
  use_restart = .false.
  do i = 1, command_argument_count()
    call get_command_argument(i, arg)
    if (arg(1:4) == 'out=') outfile = arg(5:)
    if (trim(arg) == 'fp32') step_precision = 1
    if (trim(arg) == 'restart') use_restart = .true.
  enddo
  call get_environment_variable('RT6D_RESTART_MAX_SUBRECORD', arg, status=ios)
  if (ios == 0) then
    read(arg, *) lim
    if (lim < 1 .or. lim > 2147483639_8) call halt_stop(' RT6D_RESTART_MAX_SUBRECORD out of range')
    restart_max_subrecord = lim
  endif
  if (rt6d_device_count() < 1) call halt_stop(' no HIP device')

  call set_parameter
  call pe_decomposition
  call allocate_main_module
  call set_grid
  call calc_grid
  call set_initial_conditions
  call calc_parameter
  if (use_restart) call read_restart

  itt0 = itt
  enditt = itt + int(runlen/dt)
  snapsteps = 0
  if (snapint > 0) snapsteps = int(snapint/dt)
  if (snapsteps > 0) call init_diag
  print '(a,i8,a,i8,a,es12.5,a)', ' integrating from step ', itt, ' to ', enditt, ' with dt = ', dt, ' s'
  call system_clock(c0, crate)
  do while (itt < enditt)
    if (host_forcing) call e_to_host
    call set_forcing
    if (host_forcing) call e_to_device
    call integrate
    if (snapsteps > 0) then
      if (mod(itt, snapsteps) == 0 .or. itt == 0) call diagnose
    endif
    itt = itt + 1
  enddo
  call rt6d_check(rt6d_synchronize(rt6d_handle), 'main loop')
  call system_clock(c1)
  secs = dble(c1 - c0)/dble(crate)
  print '(a,f12.4,a,es12.4,a)', ' main loop ', secs, ' s, ', &
        dble(nx)*ny*nz*nk*nl*nm*(enditt - itt0)/max(secs, 1d-12), ' cell-updates/s'

  if (use_restart) call write_restart
  call e_to_host
  if (trim(outfile) /= '-') then
    ! through a contiguous buffer: amdflang writes a strided section element by element
    allocate(buf(nx, ny, nz, nk, nl, nm))
    buf = E(1:nx, 1:ny, 1:nz, 1:nk, 1:nl, 1:nm)
    open(newunit=iu, file=trim(outfile), access='stream', form='unformatted', status='replace')
    write(iu) buf
    close(iu)
    deallocate(buf)
  endif
  call rt6d_check(rt6d_destroy(rt6d_handle), 'shutdown')
end program main
