! A config plug-in in the form of for_config/gm1.f90 and lee1.f90 -- the
! four hooks over main_module, nothing else -- for the drop-in tests.  The
! integration comes from files the test writes in the working directory:
!   plugin.txt      nx ny nz nk nl nm / nsteps snapsteps forcing /
!                   dt f0 sign_of_omega_i k_fixed l_fixed m_fixed /
!                   upper and lower reflection (0/1)
!   spacing.bin     dxt(1:nx) .. dmt(1:nm), float64
!   background.bin  Nsqr, U, V over their halo'd extents, float64
!   E0.bin          interior E, float64
!   forcing.bin     (forcing 1 or 2) a field over (x,y,k,l,m), float64
! set_forcing changes E on the host -- forcing 1 injects a bottom flux as
! for_config/lee1.f90:193-212 does, forcing 2 relaxes the bottom cells
! towards the field with the host zdot as critical_gm1.f90:119-139 does --
! so it runs with the driver's default host_forcing.
module config_module
  implicit none
  integer :: nsteps_ = 0, snapsteps_ = 0, forcing_ = 0
  real*8, allocatable :: flux_bot(:,:,:,:,:)
end module config_module

subroutine set_parameter
  use main_module
  use config_module
  implicit none
  integer :: iu, up, down
  open(newunit=iu, file='plugin.txt', status='old', action='read')
  read(iu, *) nx, ny, nz, nk, nl, nm
  read(iu, *) nsteps_, snapsteps_, forcing_
  read(iu, *) dt, f0, sign_of_omega_i, k_fixed, l_fixed, m_fixed
  read(iu, *) up, down
  close(iu)
  enable_upper_reflection = up /= 0
  enable_lower_reflection = down /= 0
  runlen = (nsteps_ + 0.5d0)*dt
  snapint = 0
  if (snapsteps_ > 0) snapint = (snapsteps_ + 0.5d0)*dt
  enable_write_6D_single = .true.
  enable_write_6D_single_himem = .true.
end subroutine set_parameter

subroutine set_grid
  use main_module
  implicit none
  integer :: iu
  open(newunit=iu, file='spacing.bin', access='stream', form='unformatted', status='old', action='read')
  read(iu) dxt(1:nx)
  read(iu) dyt(1:ny)
  read(iu) dzt(1:nz)
  read(iu) dkt(1:nk)
  read(iu) dlt(1:nl)
  read(iu) dmt(1:nm)
  close(iu)
end subroutine set_grid

subroutine set_initial_conditions
  use main_module
  use config_module
  implicit none
  integer :: iu
  real*8, allocatable :: buf(:, :, :, :, :, :)
  open(newunit=iu, file='background.bin', access='stream', form='unformatted', status='old', action='read')
  read(iu) Nsqr
  read(iu) U
  read(iu) V
  close(iu)
  allocate(buf(nx, ny, nz, nk, nl, nm))
  open(newunit=iu, file='E0.bin', access='stream', form='unformatted', status='old', action='read')
  read(iu) buf
  close(iu)
  E(1:nx, 1:ny, 1:nz, 1:nk, 1:nl, 1:nm) = buf
  deallocate(buf)
  if (forcing_ /= 0) then
    allocate(flux_bot(nx, ny, nk, nl, nm))
    open(newunit=iu, file='forcing.bin', access='stream', form='unformatted', status='old', action='read')
    read(iu) flux_bot
    close(iu)
  endif
end subroutine set_initial_conditions

subroutine set_forcing
  use main_module
  use config_module
  implicit none
  integer :: i, j, k, l, n, m
  if (forcing_ == 0) return
  if (my_blk_z == 1) then
    k = 1
    do m = ms_pe, me_pe
      do n = ls_pe, le_pe
        do l = ks_pe, ke_pe
          do j = ys_pe, ye_pe
            do i = xs_pe, xe_pe
              if (forcing_ == 1) then
                E(i,j,k,l,n,m) = E(i,j,k,l,n,m) + dt*flux_bot(i,j,l,n,m)/dzt(k)
              else
                E(i,j,k,l,n,m) = E(i,j,k,l,n,m) - dt*max(0d0,zdot(i,j,k,l,n,m))* &
                                 (E(i,j,k,l,n,m)-flux_bot(i,j,l,n,m))/dzt(k)
              endif
            enddo
          enddo
        enddo
      enddo
    enddo
  endif
end subroutine set_forcing
