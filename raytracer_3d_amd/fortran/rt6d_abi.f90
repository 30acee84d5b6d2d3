! rt6d_abi.f90 -- Fortran (iso_c_binding) interfaces of the C ABI in
! include/rt6d.h.  One interface per entry point; constants mirror the
! header's #defines.
module rt6d_abi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: RT6D_FP64 = 0, RT6D_FP32 = 1
  integer(c_int), parameter :: RT6D_LAYOUT_INTERIOR = 0, RT6D_LAYOUT_FACES = 1, RT6D_LAYOUT_HALO = 2
  integer(c_int), parameter :: RT6D_FIELD_E = 0, RT6D_FIELD_DE = 1, RT6D_FIELD_MEANFL = 2
  integer(c_int), parameter :: RT6D_FIELD_OMEGA_I = 3, RT6D_FIELD_OMEGA_E = 4
  integer(c_int), parameter :: RT6D_FIELD_XDOT = 5, RT6D_FIELD_FLUX_X = 11, RT6D_FIELD_LAMBDA_X = 17
  integer(c_int), parameter :: RT6D_OPT_EXACT = 1, RT6D_OPT_PATH = 2, RT6D_OPT_BLOCK = 3, RT6D_OPT_KERNEL = 4, &
                                RT6D_OPT_SYNC = 5
  integer(c_int), parameter :: RT6D_Q_HOMOGENEOUS = 1, RT6D_Q_DEVICE_BYTES = 2

  ! scalars of the diagnose reductions (rt6d_diag)
  type, bind(C) :: rt6d_diag_t
    real(c_double) :: eint, mflint, mfl_max
    real(c_double) :: vmax(6)
    real(c_double) :: bfl(12)
    real(c_double) :: cfl(6)
    real(c_double) :: residual
  end type rt6d_diag_t

  interface
    function rt6d_last_error() bind(C, name='rt6d_last_error')
      import :: c_ptr
      type(c_ptr) :: rt6d_last_error
    end function
    function rt6d_device_count() bind(C, name='rt6d_device_count')
      import :: c_int
      integer(c_int) :: rt6d_device_count
    end function
    function rt6d_create(n, up, down, prec, ctx) bind(C, name='rt6d_create')
      import :: c_int, c_ptr
      integer(c_int), intent(in) :: n(6)
      integer(c_int), value :: up, down, prec
      type(c_ptr), intent(out) :: ctx
      integer(c_int) :: rt6d_create
    end function
    function rt6d_destroy(ctx) bind(C, name='rt6d_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rt6d_destroy
    end function
    function rt6d_set_grid(ctx, axis, t, dt) bind(C, name='rt6d_set_grid')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: axis
      real(c_double), intent(in) :: t(*), dt(*)
      integer(c_int) :: rt6d_set_grid
    end function
    function rt6d_set_background(ctx, Nsqr, U, V, f0, sgn) bind(C, name='rt6d_set_background')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: Nsqr(*), U(*), V(*)
      real(c_double), value :: f0, sgn
      integer(c_int) :: rt6d_set_background
    end function
    function rt6d_calc_parameter(ctx) bind(C, name='rt6d_calc_parameter')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rt6d_calc_parameter
    end function
    function rt6d_cfl_dt(ctx, cfl, dt) bind(C, name='rt6d_cfl_dt')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: cfl
      real(c_double), intent(out) :: dt
      integer(c_int) :: rt6d_cfl_dt
    end function
    function rt6d_upload_E(ctx, E, layout) bind(C, name='rt6d_upload_E')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: E(*)
      integer(c_int), value :: layout
      integer(c_int) :: rt6d_upload_E
    end function
    function rt6d_step(ctx, dt, nsteps) bind(C, name='rt6d_step')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
      integer(c_int), value :: nsteps
      integer(c_int) :: rt6d_step
    end function
    function rt6d_fluxes(ctx, dt) bind(C, name='rt6d_fluxes')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
      integer(c_int) :: rt6d_fluxes
    end function
    function rt6d_download(ctx, field, layout, out) bind(C, name='rt6d_download')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: field, layout
      real(c_double), intent(inout) :: out(*)
      integer(c_int) :: rt6d_download
    end function
    function rt6d_diagnose(ctx, dt, d, E_int, meanfl_int, diss6) bind(C, name='rt6d_diagnose')
      import :: c_int, c_ptr, c_double, rt6d_diag_t
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
      type(rt6d_diag_t), intent(out) :: d
      real(c_double), intent(out) :: E_int(*), meanfl_int(*), diss6(*)
      integer(c_int) :: rt6d_diagnose
    end function
    function rt6d_set_option(ctx, option, value) bind(C, name='rt6d_set_option')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: option, value
      integer(c_int) :: rt6d_set_option
    end function
    function rt6d_query(ctx, what, value) bind(C, name='rt6d_query')
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: ctx
      integer(c_int), value :: what
      integer(c_long_long), intent(out) :: value
      integer(c_int) :: rt6d_query
    end function
    function rt6d_synchronize(ctx) bind(C, name='rt6d_synchronize')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rt6d_synchronize
    end function
    ! set_forcing on the device (kind 0 none, 1 bottom flux, 2 bottom relaxation)
    function rt6d_set_forcing(ctx, kind, field) bind(C, name='rt6d_set_forcing')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: kind
      real(c_double), intent(in) :: field(*)
      integer(c_int) :: rt6d_set_forcing
    end function
    function rt6d_init_separable(ctx, factors, peak, noise, seed, member) bind(C, name='rt6d_init_separable')
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: factors(*)
      real(c_double), value :: peak, noise
      integer(c_long_long), value :: seed, member
      integer(c_int) :: rt6d_init_separable
    end function
    ! ensembles of independent integrations
    function rt6d_ensemble_create(prec, ens) bind(C, name='rt6d_ensemble_create')
      import :: c_int, c_ptr
      integer(c_int), value :: prec
      type(c_ptr), intent(out) :: ens
      integer(c_int) :: rt6d_ensemble_create
    end function
    function rt6d_ensemble_destroy(ens) bind(C, name='rt6d_ensemble_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ens
      integer(c_int) :: rt6d_ensemble_destroy
    end function
    function rt6d_ensemble_add(ens, member, dt, slot) bind(C, name='rt6d_ensemble_add')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ens, member
      real(c_double), value :: dt
      integer(c_int), intent(out) :: slot
      integer(c_int) :: rt6d_ensemble_add
    end function
    function rt6d_ensemble_remove(ens, slot) bind(C, name='rt6d_ensemble_remove')
      import :: c_int, c_ptr
      type(c_ptr), value :: ens
      integer(c_int), value :: slot
      integer(c_int) :: rt6d_ensemble_remove
    end function
    function rt6d_ensemble_step(ens, nsteps, ms) bind(C, name='rt6d_ensemble_step')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ens
      integer(c_int), value :: nsteps
      real(c_double), intent(out) :: ms
      integer(c_int) :: rt6d_ensemble_step
    end function
    function rt6d_ensemble_tune(ens, trial_steps, blocks_per_cu, order) bind(C, name='rt6d_ensemble_tune')
      import :: c_int, c_ptr
      type(c_ptr), value :: ens
      integer(c_int), value :: trial_steps
      integer(c_int), intent(out) :: blocks_per_cu, order
      integer(c_int) :: rt6d_ensemble_tune
    end function
  end interface

contains

  ! Message of the last failing call as a Fortran string.
  function rt6d_error_message() result(msg)
    character(len=512) :: msg
    character(kind=c_char), pointer :: cs(:)
    type(c_ptr) :: p
    integer :: i
    msg = ' '
    p = rt6d_last_error()
    if (.not. c_associated(p)) return
    call c_f_pointer(p, cs, [512])
    do i = 1, 512
      if (cs(i) == c_null_char) exit
      msg(i:i) = cs(i)
    enddo
  end function rt6d_error_message

end module rt6d_abi
