! rt6d_pe.f90 -- the processor-level routines of for_src/parallel_mpi.f90
! that config plug-ins call (for_config/gm1.f90:87,93 global_sum,
! brava.f90:91-92 pe0_bcast, lee1.f90:178 pe0_recv_arr_mdim, ...), for the
! single processor that owns the whole grid here: every reduction over PEs
! is the value itself, every broadcast from or gather to PE 0 leaves the
! data where it is, and a barrier has nothing to wait for.

subroutine fortran_barrier          ! parallel_mpi.f90:240
end subroutine fortran_barrier

subroutine pe0_bcast_int(a, len)    ! :303
  implicit none
  integer, intent(in) :: len
  integer, intent(inout) :: a(len)
end subroutine pe0_bcast_int

subroutine pe0_bcast(a, len)        ! :317
  implicit none
  integer, intent(in) :: len
  real*8, intent(inout) :: a(len)
end subroutine pe0_bcast

subroutine pe0_bcast_complex(a, len)   ! :332
  implicit none
  integer, intent(in) :: len
  complex*16, intent(inout) :: a(len)
end subroutine pe0_bcast_complex

subroutine bcast_real(x, len, pe)   ! :346
  implicit none
  integer :: len, pe
  real*8 :: x(len)
end subroutine bcast_real

subroutine bcast_integer(x, len, pe)   ! :360
  implicit none
  integer :: len, pe
  integer :: x(len)
end subroutine bcast_integer

subroutine global_max(x)            ! :375
  implicit none
  real*8, intent(inout) :: x
end subroutine global_max

subroutine global_min(x)            ! :391
  implicit none
  real*8, intent(inout) :: x
end subroutine global_min

subroutine global_sum(x)            ! :407
  implicit none
  real*8, intent(inout) :: x
end subroutine global_sum

subroutine global_max_int(x)        ! :427
  implicit none
  integer, intent(inout) :: x
end subroutine global_max_int

subroutine global_min_int(x)        ! :442
  implicit none
  integer, intent(inout) :: x
end subroutine global_min_int

subroutine global_sum_int(x)        ! :457
  implicit none
  integer, intent(inout) :: x
end subroutine global_sum_int

subroutine wavenumber_sum(x)        ! :472 (sum over the k,l,m blocks of one x,y,z block)
  use main_module
  implicit none
  real*8, intent(inout) :: x(1-onx:nx+onx, 1-onx:ny+onx, 1-onx:nz+onx)
end subroutine wavenumber_sum

subroutine pe0_recv_arr_phys(a)     ! :624 (gather of a 3D field on PE 0)
  use main_module
  implicit none
  real*8, intent(inout) :: a(nx, ny, nz)
end subroutine pe0_recv_arr_phys

subroutine pe0_recv_arr_mdim(a, i)  ! :682 (gather of one m slice of a 5D field on PE 0)
  use main_module
  implicit none
  integer :: i
  real*8, intent(inout) :: a(nx, ny, nz, nk, nl)
end subroutine pe0_recv_arr_mdim

subroutine pe0_recv_arr(a)          ! :742 (gather of a 6D field on PE 0)
  use main_module
  implicit none
  real*8, intent(inout) :: a(nx, ny, nz, nk, nl, nm)
end subroutine pe0_recv_arr
