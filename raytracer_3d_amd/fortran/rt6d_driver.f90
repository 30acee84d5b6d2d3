! rt6d_driver.f90 -- Fortran driver with the reference's set-up order and
! time loop (for_src/main.f90:36-79), running the hot path on the GPU through
! the drop-in routines of rt6d_hot.f90.
!
! usage: rt6d_driver CASE [OUT] [SNAPSTEPS] [fp32|fp64] [restart]
!   CASE       case file (config_case.f90)
!   OUT        if given and not '-', the final interior E is written there
!              as a raw float64 stream
!   SNAPSTEPS  diagnose every SNAPSTEPS steps and at step 0 (0: never),
!              writing snap3D.cdf and snap6D.cdf in the working directory
!   fp32       run the FP32 form
!   restart    read_restart after calc_parameter (main.f90:49) and
!              write_restart after the loop (main.f90:81), in the working
!              directory
program main
  use main_module
  use rt6d_abi
  implicit none
  integer :: enditt, snapsteps, iu, itt0
  logical :: use_restart
  integer(8) :: c0, c1, crate
  character(len=1024) :: arg, outfile
  real*8 :: secs

  outfile = '-'
  snapsteps = 0
  if (command_argument_count() >= 2) call get_command_argument(2, outfile)
  if (command_argument_count() >= 3) then
    call get_command_argument(3, arg)
    read(arg, *) snapsteps
  endif
  if (command_argument_count() >= 4) then
    call get_command_argument(4, arg)
    if (trim(arg) == 'fp32') step_precision = 1
  endif
  use_restart = .false.
  if (command_argument_count() >= 5) then
    call get_command_argument(5, arg)
    use_restart = trim(arg) == 'restart'
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
  enditt = itt + int(runlen/dt + 0.5d0)
  if (snapsteps > 0) call init_diag
  print '(a,i8,a,i8,a,es12.5,a)', ' integrating from step ', itt, ' to ', enditt, ' with dt = ', dt, ' s'
  call system_clock(c0, crate)
  do while (itt < enditt)
    call set_forcing
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
    open(newunit=iu, file=trim(outfile), access='stream', form='unformatted', status='replace')
    write(iu) E(1:nx, 1:ny, 1:nz, 1:nk, 1:nl, 1:nm)
    close(iu)
  endif
  call rt6d_check(rt6d_destroy(rt6d_handle), 'shutdown')
end program main
