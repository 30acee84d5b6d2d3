! rt6d_hot.f90 -- the reference's hot-path routines under their own names
! and argument-less signatures, forwarding to the GPU through rt6d_abi:
!   calc_parameter   for_src/calc_parameter.f90:2
!   integrate        for_src/integrate.f90:4
!   superbee_fluxes  for_src/superbee_fluxes.f90:4
!   diagnose         for_src/diagnose.f90:20 (reductions and printout; no netCDF)
! plus the host set-up the driver keeps (pe_decomposition for one device,
! allocate_main_module, calc_grid) and halt_stop.

subroutine halt_stop(string)
  implicit none
  character*(*) :: string
  print *, ' halting: ', trim(string)
  stop 99
end subroutine halt_stop

subroutine rt6d_check(rc, where)
  use iso_c_binding, only: c_int
  use rt6d_abi, only: rt6d_error_message
  implicit none
  integer(c_int), intent(in) :: rc
  character*(*), intent(in) :: where
  if (rc /= 0) then
    print *, ' rt6d error in ', trim(where), ': ', trim(rt6d_error_message())
    call halt_stop(where)
  endif
end subroutine rt6d_check

! One GPU owns the whole grid (parallel_mpi.f90:153-162, the n_pes = 1 branch).
subroutine pe_decomposition
  use main_module
  implicit none
  if (nk > 1 .and. mod(nk, 2) /= 0) call halt_stop(' nk must be even or one ')
  if (nl > 1 .and. mod(nl, 2) /= 0) call halt_stop(' nl must be even or one ')
  if (nm > 1 .and. mod(nm, 2) /= 0) call halt_stop(' nm must be even or one')
  x_blk = nx; y_blk = ny; z_blk = nz; k_blk = nk; l_blk = nl; m_blk = nm
  xs_pe = 1; xe_pe = nx; ys_pe = 1; ye_pe = ny; zs_pe = 1; ze_pe = nz
  ks_pe = 1; ke_pe = nk; ls_pe = 1; le_pe = nl; ms_pe = 1; me_pe = nm
end subroutine pe_decomposition

! Host arrays (halo'd, zero-initialised) and the device context.
subroutine allocate_main_module
  use main_module
  use rt6d_abi
  implicit none
  integer(c_int) :: n(6), up, down
  allocate(E(1-onx:nx+onx, 1-onx:ny+onx, 1-onx:nz+onx, 1-onx:nk+onx, 1-onx:nl+onx, 1-onx:nm+onx))
  allocate(dE, meanfl, mold=E)
  allocate(Nsqr(1-onx:nx+onx, 1-onx:ny+onx, 1-onx:nz+onx))
  allocate(U, V, mold=Nsqr)
  allocate(xt(1-onx:nx+onx), xu(1-onx:nx+onx), dxt(1-onx:nx+onx), dxu(1-onx:nx+onx))
  allocate(yt(1-onx:ny+onx), yu(1-onx:ny+onx), dyt(1-onx:ny+onx), dyu(1-onx:ny+onx))
  allocate(zt(1-onx:nz+onx), zu(1-onx:nz+onx), dzt(1-onx:nz+onx), dzu(1-onx:nz+onx))
  allocate(kt(1-onx:nk+onx), ku(1-onx:nk+onx), dkt(1-onx:nk+onx), dku(1-onx:nk+onx))
  allocate(lt(1-onx:nl+onx), lu(1-onx:nl+onx), dlt(1-onx:nl+onx), dlu(1-onx:nl+onx))
  allocate(mt(1-onx:nm+onx), mu(1-onx:nm+onx), dmt(1-onx:nm+onx), dmu(1-onx:nm+onx))
  E = 0; dE = 0; meanfl = 0; Nsqr = 0; U = 0; V = 0
  xt = 0; xu = 0; dxt = 0; dxu = 0; yt = 0; yu = 0; dyt = 0; dyu = 0
  zt = 0; zu = 0; dzt = 0; dzu = 0; kt = 0; ku = 0; dkt = 0; dku = 0
  lt = 0; lu = 0; dlt = 0; dlu = 0; mt = 0; mu = 0; dmt = 0; dmu = 0
  n = [nx, ny, nz, nk, nl, nm]
  up = merge(1, 0, enable_upper_reflection)
  down = merge(1, 0, enable_lower_reflection)
  call rt6d_check(rt6d_create(n, up, down, int(step_precision, c_int), rt6d_handle), 'allocate_main_module')
  ! integrate enqueues its step and returns; the next call that reads the
  ! device state (a download, diagnose, rt6d_synchronize) waits for it
  call rt6d_check(rt6d_set_option(rt6d_handle, RT6D_OPT_SYNC, 0_c_int), 'allocate_main_module')
end subroutine allocate_main_module

! u-centred grid from T spacings (calc_grid.f90:300-322 semantics): the U
! points sit half way between T points; all arrays have n entries.
subroutine u_centered_grid(dyt_, dyu_, yt_, yu_, n)
  implicit none
  integer, intent(in) :: n
  real*8, intent(in) :: dyt_(n)
  real*8, intent(out) :: dyu_(n), yt_(n), yu_(n)
  integer :: i
  yu_(1) = 0
  do i = 2, n
    yu_(i) = yu_(i-1) + dyt_(i)
  enddo
  yt_(1) = yu_(1) - dyt_(1)*0.5d0
  do i = 2, n
    yt_(i) = 2*yu_(i-1) - yt_(i-1)
  enddo
  do i = 1, n-1
    dyu_(i) = yt_(i+1) - yt_(i)
  enddo
  dyu_(n) = 2*dyt_(n) - dyu_(max(1, n-1))
end subroutine u_centered_grid

! T/U positions and spacings of all six axes from the T spacings set by
! set_grid (calc_grid.f90:104-177) and the monotonicity checks of :260-293.
subroutine calc_grid
  use main_module
  implicit none
  integer :: i
  call axis_grid(dxt, dxu, xt, xu, nx, 1, 0d0)
  call axis_grid(dyt, dyu, yt, yu, ny, 1, 0d0)
  call axis_grid(dzt, dzu, zt, zu, nz, 2, 0d0)
  call axis_grid(dkt, dku, kt, ku, nk, 3, k_fixed)
  call axis_grid(dlt, dlu, lt, lu, nl, 3, l_fixed)
  call axis_grid(dmt, dmu, mt, mu, nm, 3, m_fixed)
  do i = 2, nz
    if (dzu(i) - dzu(i-1) > 1d-12 .or. dzt(i) - dzt(i-1) > 1d-12) &
      call halt_stop(' in calc_grid: dzu/dzt do not decrease monotonically')
  enddo
  do i = 2, nm/2
    if (dmu(i) - dmu(i-1) > 1d-12 .or. dmt(i) - dmt(i-1) > 1d-12) &
      call halt_stop(' in calc_grid: dmu/dmt do not decrease monotonically')
  enddo
  do i = 2, nl/2
    if (dlu(i) - dlu(i-1) > 1d-12 .or. dlt(i) - dlt(i-1) > 1d-12) &
      call halt_stop(' in calc_grid: dlu/dlt do not decrease monotonically')
  enddo
contains
  ! kind 1: shift by du(1) (x, y); 2: u(n) = 0 (z); 3: u(n/2) = 0 or fixed (wavenumbers)
  subroutine axis_grid(d, du, t, u, n, kind, fixed)
    integer, intent(in) :: n, kind
    real*8, intent(inout) :: d(1-onx:n+onx)
    real*8, intent(out) :: du(1-onx:n+onx), t(1-onx:n+onx), u(1-onx:n+onx)
    real*8, intent(in) :: fixed
    real*8 :: sh
    integer :: ii
    do ii = 1, onx
      d(n+ii) = d(n); d(1-ii) = d(1)
    enddo
    call u_centered_grid(d, du, t, u, n + 2*onx)
    select case (kind)
    case (1)
      sh = du(1)
    case (2)
      sh = u(n)
    case default
      if (n == 1) then
        t = fixed; u = fixed
        return
      endif
      sh = u(n/2)
    end select
    t = t - sh
    u = u - sh
  end subroutine axis_grid
end subroutine calc_grid

! Host E -> device (after set_initial_conditions, a host set_forcing or a
! restart).  Refused while the host E is older than the device E (steps
! taken since the last e_to_host): uploading it would throw those steps away.
subroutine e_to_device
  use main_module
  use rt6d_abi
  implicit none
  if (.not. host_E_current) call halt_stop(' e_to_device: the host E is stale (call e_to_host first)')
  call rt6d_check(rt6d_upload_E(rt6d_handle, E, RT6D_LAYOUT_HALO), 'e_to_device')
  host_E_current = .true.
end subroutine e_to_device

! Device E -> host interior (ghost entries are left as they were).
subroutine e_to_host
  use main_module
  use rt6d_abi
  implicit none
  if (host_E_current) return
  call rt6d_check(rt6d_download(rt6d_handle, RT6D_FIELD_E, RT6D_LAYOUT_HALO, E), 'e_to_host')
  host_E_current = .true.
end subroutine e_to_host

! calc_parameter.f90:2 on the device: grids, background and E go over, the
! coefficient fields are computed there and stay there.
subroutine calc_parameter
  use main_module
  use rt6d_abi
  implicit none
  type(c_ptr) :: h
  h = rt6d_handle
  call rt6d_check(rt6d_set_grid(h, 0_c_int, xt, dxt), 'calc_parameter/x grid')
  call rt6d_check(rt6d_set_grid(h, 1_c_int, yt, dyt), 'calc_parameter/y grid')
  call rt6d_check(rt6d_set_grid(h, 2_c_int, zt, dzt), 'calc_parameter/z grid')
  call rt6d_check(rt6d_set_grid(h, 3_c_int, kt, dkt), 'calc_parameter/k grid')
  call rt6d_check(rt6d_set_grid(h, 4_c_int, lt, dlt), 'calc_parameter/l grid')
  call rt6d_check(rt6d_set_grid(h, 5_c_int, mt, dmt), 'calc_parameter/m grid')
  call rt6d_check(rt6d_set_background(h, Nsqr, U, V, f0, sign_of_omega_i), 'calc_parameter/background')
  call rt6d_check(rt6d_calc_parameter(h), 'calc_parameter')
  if (host_forcing) then
    call velocity_to_host(0, xdot)
    call velocity_to_host(1, ydot)
    call velocity_to_host(2, zdot)
    call velocity_to_host(3, kdot)
    call velocity_to_host(4, ldot)
    call velocity_to_host(5, mdot)
  endif
  call e_to_device
contains
  ! face velocities of axis d on the host, halo'd as the reference allocates
  ! them; faces 0..n of the axis are filled (face f lies between cells f and
  ! f+1), the ghost faces stay zero
  subroutine velocity_to_host(d, a)
    integer, intent(in) :: d
    real*8, allocatable, intent(inout) :: a(:,:,:,:,:,:)
    real*8, allocatable :: tmp(:,:,:,:,:,:)
    integer :: lo(6)
    lo = 1
    lo(d + 1) = 0
    if (.not. allocated(a)) allocate(a, mold=E)
    a = 0
    allocate(tmp(lo(1):nx, lo(2):ny, lo(3):nz, lo(4):nk, lo(5):nl, lo(6):nm))
    call rt6d_check(rt6d_download(h, RT6D_FIELD_XDOT + d, RT6D_LAYOUT_FACES, tmp), 'calc_parameter/velocities')
    a(lo(1):nx, lo(2):ny, lo(3):nz, lo(4):nk, lo(5):nl, lo(6):nm) = tmp
    deallocate(tmp)
  end subroutine velocity_to_host
end subroutine calc_parameter

! integrate.f90:4 -- one forward-Euler step of E on the device, enqueued on
! the context's stream (allocate_main_module turns the wait off).
subroutine integrate
  use main_module
  use rt6d_abi
  implicit none
  call rt6d_check(rt6d_step(rt6d_handle, dt, 1_c_int), 'integrate')
  host_E_current = .false.
end subroutine integrate

! superbee_fluxes.f90:4 -- face fluxes of the current E, kept on the device
! (read them with rt6d_download).
subroutine superbee_fluxes
  use main_module
  use rt6d_abi
  implicit none
  call rt6d_check(rt6d_fluxes(rt6d_handle, dt), 'superbee_fluxes')
end subroutine superbee_fluxes

! diagnose.f90:45-329 -- energy budget of the last step, printed as the
! reference prints it (:316-329).
subroutine diagnose
  use main_module
  use rt6d_abi
  implicit none
  type(rt6d_diag_t) :: d
  real*8, allocatable :: Ei(:), Mi(:), Di(:)
  real*8 :: b(12)
  allocate(Ei(nx*ny*nz), Mi(nx*ny*nz), Di(6*nx*ny*nz))
  call rt6d_check(rt6d_diagnose(rt6d_handle, dt, d, Ei, Mi, Di), 'diagnose')
  b = d%bfl
  print '(a,f15.2,a)', ' snapshot at ', itt*dt, ' s'
  print '(a,6f8.2)', ' max CFL in x,y,z,k,l,m-dir = ', d%cfl
  if (d%mfl_max > 0) print '(a,es12.4)', ' mean flow interaction scale ', dt/d%mfl_max
  print '(a,es14.6)', ' total dE          : ', d%eint
  print '(a,es14.6)', ' dE by mean-flow   : ', d%mflint
  print '(a,2es14.6)', ' by flux at min/max x  : ', b(1), b(2)
  print '(a,2es14.6)', ' by flux at min/max y  : ', b(3), b(4)
  print '(a,2es14.6)', ' by flux at min/max k  : ', b(7), b(8)
  print '(a,2es14.6)', ' by flux at min/max l  : ', b(9), b(10)
  print '(a,2es14.6)', ' by flux at min/max m  : ', b(11), b(12)
  print '(a,es14.6)', ' residual : ', d%residual
  call write_snapshots(d, Ei, Mi, Di)   ! after init_diag; a no-op without it
  deallocate(Ei, Mi, Di)
end subroutine diagnose

! Restart file name of this PE, 'restart_pe=' with the PE number in i9 and
! blanks written as zeros, as for_src/restart.f90:7-8 names it.
subroutine restart_file_name(fname)
  use main_module
  implicit none
  character(len=80), intent(out) :: fname
  integer :: i
  write(fname, '(a,i9,a)') 'restart_pe=', my_pe, '.dta'
  do i = 1, len_trim(fname)
    if (fname(i:i) == ' ') fname(i:i) = '0'
  enddo
end subroutine restart_file_name

! Records of gfortran's unformatted sequential layout (the reference's
! compiler), written and read through stream access so that records past
! 2 GiB keep that layout: subrecords of at most restart_max_subrecord bytes
! between 4-byte markers; the leading marker is negative when the record
! continues in the next subrecord, the trailing marker when the subrecord
! continues a previous one.  A record's payload is the bytes of np pieces.
subroutine put_record(io, np, p, nb)
  use iso_c_binding
  use main_module, only: restart_max_subrecord
  implicit none
  integer, intent(in) :: io, np
  type(c_ptr), intent(in) :: p(np)
  integer(c_int64_t), intent(in) :: nb(np)
  integer(c_int8_t), pointer :: b(:)
  integer(c_int64_t) :: left, sub, need, take, off
  integer(c_int32_t) :: mk
  integer :: ip
  logical :: first, more
  left = sum(nb)
  first = .true.
  ip = 1
  off = 0
  do
    sub = min(left, restart_max_subrecord)
    more = left > sub
    mk = int(sub, c_int32_t)
    if (more) mk = -mk
    write(io) mk
    need = sub
    do while (need > 0)
      do while (nb(ip) - off == 0)
        ip = ip + 1
        off = 0
      enddo
      take = min(need, nb(ip) - off)
      call c_f_pointer(p(ip), b, [nb(ip)])
      write(io) b(off + 1:off + take)
      off = off + take
      need = need - take
    enddo
    mk = int(sub, c_int32_t)
    if (.not. first) mk = -mk
    write(io) mk
    left = left - sub
    first = .false.
    if (.not. more) exit
  enddo
end subroutine put_record

! One record of that layout into np pieces; ok is false when the framing is
! broken or the record's length is not the pieces' total.
subroutine get_record(io, np, p, nb, ok)
  use iso_c_binding
  implicit none
  integer, intent(in) :: io, np
  type(c_ptr), intent(in) :: p(np)
  integer(c_int64_t), intent(in) :: nb(np)
  logical, intent(out) :: ok
  integer(c_int8_t), pointer :: b(:)
  integer(c_int64_t) :: total, got, sub, need, take, off
  integer(c_int32_t) :: lead, trail
  integer :: ip, ios
  logical :: first
  ok = .false.
  total = sum(nb)
  got = 0
  first = .true.
  ip = 1
  off = 0
  do
    read(io, iostat=ios) lead
    if (ios /= 0) return
    sub = abs(int(lead, c_int64_t))
    if (got + sub > total) return
    need = sub
    do while (need > 0)
      do while (nb(ip) - off == 0)
        ip = ip + 1
        off = 0
      enddo
      take = min(need, nb(ip) - off)
      call c_f_pointer(p(ip), b, [nb(ip)])
      read(io, iostat=ios) b(off + 1:off + take)
      if (ios /= 0) return
      off = off + take
      need = need - take
    enddo
    read(io, iostat=ios) trail
    if (ios /= 0) return
    if (abs(int(trail, c_int64_t)) /= sub .or. ((trail < 0) .eqv. first)) return
    got = got + sub
    first = .false.
    if (lead >= 0) exit
  enddo
  ok = got == total
end subroutine get_record

! write_restart (for_src/restart.f90:1-30): the same four records --
! dimensions and itt, the PE bounds (two records), E, dE, meanfl over their
! halo'd extents -- and the 'ritt' file.  E comes from the device; dE and
! meanfl are the last step's by-products (zero before the first step).
subroutine write_restart
  use iso_c_binding
  use main_module
  use rt6d_abi
  implicit none
  character(len=80) :: fname
  integer :: io, ios
  integer(c_int32_t), target :: h1(7), h2(6), h3(6)
  integer(c_int64_t) :: nbf
  interface
    subroutine put_record(io, np, p, nb)
      import :: c_ptr, c_int64_t
      integer, intent(in) :: io, np
      type(c_ptr), intent(in) :: p(np)
      integer(c_int64_t), intent(in) :: nb(np)
    end subroutine
  end interface
  call e_to_host
  call rt6d_check(rt6d_download(rt6d_handle, RT6D_FIELD_DE, RT6D_LAYOUT_HALO, dE), 'write_restart/dE')
  call rt6d_check(rt6d_download(rt6d_handle, RT6D_FIELD_MEANFL, RT6D_LAYOUT_HALO, meanfl), 'write_restart/meanfl')
  call restart_file_name(fname)
  print '(a,a)', ' writing to restart file ', trim(fname)
  print '(a,i12)', ' at itt = ', itt
  open(newunit=io, file=trim(fname), access='stream', form='unformatted', status='replace', iostat=ios)
  if (ios /= 0) goto 10
  h1 = [nx, ny, nz, nk, nl, nm, itt]
  h2 = [xs_pe, xe_pe, ys_pe, ye_pe, zs_pe, ze_pe]
  h3 = [ks_pe, ke_pe, ls_pe, le_pe, ms_pe, me_pe]
  nbf = 8_c_int64_t*size(E, kind=c_int64_t)
  call put_record(io, 1, [c_loc(h1)], [28_c_int64_t])
  call put_record(io, 1, [c_loc(h2)], [24_c_int64_t])
  call put_record(io, 1, [c_loc(h3)], [24_c_int64_t])
  call put_record(io, 3, [c_loc(E), c_loc(dE), c_loc(meanfl)], [nbf, nbf, nbf])
  close(io)
  open(newunit=io, file='ritt', form='formatted', status='unknown')
  write(io, *) itt
  close(io)
  return
10 continue
  print '(a)', ' Warning: error writing restart file'
end subroutine write_restart

! read_restart (for_src/restart.f90:35-93): nothing happens without a
! restart file; dimensions and PE bounds must match; E goes to the device,
! dE and meanfl stay on the host (the device recomputes its by-products at
! the next step).
subroutine read_restart
  use iso_c_binding
  use main_module
  implicit none
  character(len=80) :: fname
  logical :: file_exists, ok
  integer :: io, ios
  integer(c_int32_t), target :: h1(7), b(12)
  integer :: bb(12)
  integer(c_int64_t) :: nbf
  interface
    subroutine get_record(io, np, p, nb, ok)
      import :: c_ptr, c_int64_t
      integer, intent(in) :: io, np
      type(c_ptr), intent(in) :: p(np)
      integer(c_int64_t), intent(in) :: nb(np)
      logical, intent(out) :: ok
    end subroutine
  end interface
  call restart_file_name(fname)
  inquire(file=trim(fname), exist=file_exists)
  if (.not. file_exists) then
    print '(a,a)', ' found no restart file ', trim(fname)
    return
  endif
  print '(a,a)', ' reading restart file ', trim(fname)
  open(newunit=io, file=trim(fname), access='stream', form='unformatted', status='old', action='read', iostat=ios)
  if (ios /= 0) goto 10
  call get_record(io, 1, [c_loc(h1)], [28_c_int64_t], ok)
  if (.not. ok) goto 10
  if (nx /= h1(1) .or. ny /= h1(2) .or. nz /= h1(3) .or. nk /= h1(4) .or. nl /= h1(5) .or. nm /= h1(6)) then
    print *, ' read from restart dimensions: ', h1(1:6)
    print *, ' does not match dimensions   : ', nx, ny, nz, nk, nl, nm
    call halt_stop(' read_restart')
  endif
  itt = h1(7)
  call get_record(io, 1, [c_loc(b(1))], [24_c_int64_t], ok)
  if (.not. ok) goto 10
  call get_record(io, 1, [c_loc(b(7))], [24_c_int64_t], ok)
  if (.not. ok) goto 10
  bb = [xs_pe, xe_pe, ys_pe, ye_pe, zs_pe, ze_pe, ks_pe, ke_pe, ls_pe, le_pe, ms_pe, me_pe]
  if (any(b /= bb)) then
    print *, ' read from restart PE boundaries: ', b
    print *, ' which does not match           : ', bb
    call halt_stop(' read_restart')
  endif
  nbf = 8_c_int64_t*size(E, kind=c_int64_t)
  call get_record(io, 3, [c_loc(E), c_loc(dE), c_loc(meanfl)], [nbf, nbf, nbf], ok)
  if (.not. ok) goto 10
  close(io)
  print *, ' itt = ', itt
  host_E_current = .true.   ! the file's E is the state from here on
  call e_to_device
  return
10 continue
  print '(a)', ' Warning: error reading restart file'
  call halt_stop(' read_restart')
end subroutine read_restart
