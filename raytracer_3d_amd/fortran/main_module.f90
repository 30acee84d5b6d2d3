! main_module.f90 -- model state of the Fortran host side, under the
! reference's module and variable names (for_src/main_module.f90:3-127), so
! that config plug-ins written against the reference's four hooks
! (set_parameter, set_grid, set_initial_conditions, set_forcing) compile
! unchanged.  The 6D coefficient fields, fluxes and omegas live on the GPU;
! the host keeps E, dE and meanfl (halo'd, as the reference allocates them),
! the 1D grids and the 3D background.  The host copies of the face
! velocities (xdot .. mdot) are allocated and filled by calc_parameter only
! when host_forcing is on, for set_forcing routines that read them
! (for_config/critical_gm1.f90:132 reads zdot); the other 6D fields of the
! reference are declared for plug-ins that name them and stay unallocated.
module main_module
  use iso_c_binding, only: c_ptr, c_null_ptr
  implicit none

  real*8, parameter :: version = 0.1
  ! The reference writes pi as a default-real literal (main_module.f90:8),
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

  ! host mirrors of the state (halo'd: (1-onx : n+onx) per axis); target for
  ! the restart record I/O, which addresses them as bytes
  real*8, allocatable, target :: E(:,:,:,:,:,:), dE(:,:,:,:,:,:), meanfl(:,:,:,:,:,:)
  ! face velocities on the host (faces 0..n of their axis; host_forcing only)
  real*8, allocatable :: xdot(:,:,:,:,:,:), ydot(:,:,:,:,:,:), zdot(:,:,:,:,:,:)
  real*8, allocatable :: kdot(:,:,:,:,:,:), ldot(:,:,:,:,:,:), mdot(:,:,:,:,:,:)
  ! declared for plug-ins that name them; the device holds these fields
  real*8, allocatable :: omega_i(:,:,:,:,:,:), omega_e(:,:,:,:,:,:)
  real*8, allocatable :: flux_x(:,:,:,:,:,:), flux_y(:,:,:,:,:,:), flux_z(:,:,:,:,:,:)
  real*8, allocatable :: flux_k(:,:,:,:,:,:), flux_l(:,:,:,:,:,:), flux_m(:,:,:,:,:,:)
  real*8, allocatable :: lambda_x(:,:,:,:,:,:), lambda_y(:,:,:,:,:,:), lambda_z(:,:,:,:,:,:)
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
  integer :: mpi_comm_phys = 0

  ! diagnostic output switches (main_module.f90:118-125); snap6D.cdf is one
  ! file written next to snap3D.cdf, so the single/himem choices have no
  ! further effect here
  logical :: enable_show_grid_details = .false.
  logical :: enable_write_6D_single = .true.
  logical :: enable_write_6D_single_himem = .true.
  character*160 :: netcdf_6D_file = 'snap6D.cdf'

  ! device side (no reference counterpart)
  type(c_ptr) :: rt6d_handle = c_null_ptr
  integer :: step_precision = 0       ! 0 binary64, 1 the FP32 form
  logical :: host_E_current = .true.  ! host E equals the device E
  ! set_forcing acts on the host E: the driver brings E to the host before
  ! it and back to the device after it, and calc_parameter fills the host
  ! face velocities.  On by default, since a reference set_forcing changes
  ! E on the host; a plug-in whose set_forcing leaves E alone (or forces on
  ! the device through rt6d_set_forcing) turns it off in set_parameter and
  ! keeps E resident on the device between steps.
  logical :: host_forcing = .true.
  ! restart records: gfortran's subrecord length limit (bytes)
  integer*8 :: restart_max_subrecord = 2147483639_8
end module main_module
