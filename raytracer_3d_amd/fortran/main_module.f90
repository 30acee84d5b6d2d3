! main_module.f90 -- model state of the Fortran host side, under the
! reference's module and variable names (for_src/main_module.f90:3-127), so
! that config plug-ins written against the reference's four hooks
! (set_parameter, set_grid, set_initial_conditions, set_forcing) compile
! unchanged.  The 6D coefficient fields, fluxes and omegas live on the GPU;
! the host keeps E, dE and meanfl (halo'd, as the reference allocates them),
! the 1D grids and the 3D background.
module main_module
  use iso_c_binding, only: c_ptr, c_null_ptr
  implicit none

  ! The reference writes pi as a default-real literal (main_module.f90:10),
  ! so its value is the binary32 rounding of pi; config plug-ins that build
  ! their grids from pi see the same value here.
  real*8, parameter :: pi = real(3.14159265358979323846_4, 8)

  integer :: nx = 1, ny = 1, nz = 1, nk = 1, nl = 1, nm = 1
  real*8 :: dt = 0d0, Lx = 0d0, Ly = 0d0, Lz = 0d0
  real*8 :: f0 = 1d-4
  real*8 :: k_fixed = 0d0, l_fixed = 0d0, m_fixed = 0d0
  real*8 :: sign_of_omega_i = 1d0
  logical :: enable_upper_reflection = .true.
  logical :: enable_lower_reflection = .true.
  integer :: itt = 0
  real*8 :: runlen = 0d0, snapint = 0d0
  logical :: enable_write_6D_single = .true.   ! snap6D.cdf next to snap3D.cdf (main_module.f90:122)

  ! host mirrors of the state (halo'd: (1-onx : n+onx) per axis)
  real*8, allocatable :: E(:,:,:,:,:,:), dE(:,:,:,:,:,:), meanfl(:,:,:,:,:,:)
  real*8, allocatable :: Nsqr(:,:,:), U(:,:,:), V(:,:,:)
  real*8, allocatable :: xt(:), xu(:), dxt(:), dxu(:)
  real*8, allocatable :: yt(:), yu(:), dyt(:), dyu(:)
  real*8, allocatable :: zt(:), zu(:), dzt(:), dzu(:)
  real*8, allocatable :: kt(:), ku(:), dkt(:), dku(:)
  real*8, allocatable :: lt(:), lu(:), dlt(:), dlu(:)
  real*8, allocatable :: mt(:), mu(:), dmt(:), dmu(:)

  ! one GPU holds the whole grid: the decomposition is a single block
  integer :: n_pes = 1, my_pe = 0
  integer :: n_pes_x = 1, n_pes_y = 1, n_pes_z = 1, n_pes_k = 1, n_pes_l = 1, n_pes_m = 1
  integer :: my_blk_x = 1, my_blk_y = 1, my_blk_z = 1, my_blk_k = 1, my_blk_l = 1, my_blk_m = 1
  integer :: x_blk, y_blk, z_blk, k_blk, l_blk, m_blk
  integer :: xs_pe, xe_pe, ys_pe, ye_pe, zs_pe, ze_pe, ks_pe, ke_pe, ls_pe, le_pe, ms_pe, me_pe
  integer :: onx = 2

  ! device side (no reference counterpart)
  type(c_ptr) :: rt6d_handle = c_null_ptr
  integer :: step_precision = 0       ! 0 binary64, 1 the FP32 form
  logical :: host_E_current = .true.  ! host E equals the device E
end module main_module
