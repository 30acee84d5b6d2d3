! rt6d_netcdf.f90 -- snapshot files snap3D.cdf and snap6D.cdf for the Fortran
! host, in the schema of for_src/netcdf_io_3D.f90:1-260 and
! for_src/netcdf_io_6D.f90:240-656, with the records diagnose appends
! (for_src/diagnose.f90:335-530).
!
! The netCDF library is not part of this build, so the module writes the
! published netCDF "64-bit offset" layout (CDF-2) itself: a header, the
! fixed-size variables, then the records, all big-endian (stream access with
! convert='big_endian'). Only what the two schemas use is supported: doubles,
! one record dimension, text attributes on variables and double global
! attributes. A record is appended by writing its slabs first and the record
! count last, so a reader never sees half a record.
!
! The layout (definition order, offsets) is the one raytracer_3d_amd/
! snapshot.py writes, so the Fortran and Python hosts write the same bytes
! for the same run.
module rt6d_netcdf
  implicit none
  private
  public :: cdf_file, cdf_reset, cdf_dim, cdf_var, cdf_gatt, cdf_create, cdf_put, cdf_put_rec, cdf_end_rec

  integer, parameter :: MAXDIM = 16, MAXVAR = 40, MAXATT = 4, MAXG = 8
  integer, parameter :: NC_DIMENSION = 10, NC_VARIABLE = 11, NC_ATTRIBUTE = 12, NC_CHAR = 2, NC_DOUBLE = 6

  type :: cdf_variable
    character(len=32) :: name = ''
    integer :: ndims = 0
    integer :: dimid(8) = 0                 ! 0-based dimension ids, file order (record dimension first)
    integer :: natt = 0
    character(len=16) :: attname(MAXATT) = ''
    character(len=32) :: atttext(MAXATT) = ''
    integer :: attlen(MAXATT) = 0
    logical :: rec = .false.
    integer(8) :: count = 0                 ! values per variable (per record for record variables)
    integer(8) :: begin = 0
  end type cdf_variable

  type :: cdf_file
    character(len=512) :: path = ''
    integer :: ndims = 0, nvars = 0, ngatt = 0
    character(len=32) :: dimname(MAXDIM) = ''
    integer :: dimlen(MAXDIM) = 0           ! 0: the record dimension
    type(cdf_variable) :: var(MAXVAR)
    character(len=16) :: gattname(MAXG) = ''
    real(8) :: gattval(MAXG) = 0d0
    integer :: numrecs = 0
    integer(8) :: recsize = 0
    logical :: open = .false.
  end type cdf_file

contains

  integer function pad4(n)
    integer, intent(in) :: n
    pad4 = (n + 3) / 4 * 4
  end function pad4

  subroutine cdf_reset(f)
    type(cdf_file), intent(inout) :: f
    integer :: i
    f%path = ''
    f%ndims = 0; f%nvars = 0; f%ngatt = 0; f%numrecs = 0; f%recsize = 0
    f%open = .false.
    do i = 1, MAXVAR
      f%var(i) = cdf_variable()
    enddo
  end subroutine cdf_reset

  subroutine cdf_dim(f, name, length)
    type(cdf_file), intent(inout) :: f
    character(len=*), intent(in) :: name
    integer, intent(in) :: length           ! 0: the record dimension
    f%ndims = f%ndims + 1
    f%dimname(f%ndims) = name
    f%dimlen(f%ndims) = length
  end subroutine cdf_dim

  integer function dim_index(f, name)
    type(cdf_file), intent(in) :: f
    character(len=*), intent(in) :: name
    integer :: i
    do i = 1, f%ndims
      if (trim(f%dimname(i)) == trim(name)) then
        dim_index = i
        return
      endif
    enddo
    print '(a,a)', ' rt6d_netcdf: unknown dimension ', trim(name)
    stop 99
  end function dim_index

  ! A double variable over dims (file order: the record dimension first,
  ! then the Fortran order reversed), with blank-padded long_name (24) and
  ! units (16) as ncaptc stores them, and optionally one more text attribute.
  subroutine cdf_var(f, name, dims, long_name, units, extra_name, extra_text)
    type(cdf_file), intent(inout) :: f
    character(len=*), intent(in) :: name, dims(:), long_name, units
    character(len=*), intent(in), optional :: extra_name, extra_text
    type(cdf_variable) :: v
    integer :: i, d
    v%name = name
    v%ndims = size(dims)
    v%count = 1
    do i = 1, size(dims)
      d = dim_index(f, dims(i))
      v%dimid(i) = d - 1
      if (f%dimlen(d) == 0) then
        v%rec = .true.
      else
        v%count = v%count * f%dimlen(d)
      endif
    enddo
    v%natt = 2
    v%attname(1) = 'long_name'; v%atttext(1) = long_name; v%attlen(1) = 24
    v%attname(2) = 'units';     v%atttext(2) = units;     v%attlen(2) = 16
    if (present(extra_name)) then
      v%natt = 3
      v%attname(3) = extra_name; v%atttext(3) = extra_text; v%attlen(3) = len(extra_text)
    endif
    f%nvars = f%nvars + 1
    f%var(f%nvars) = v
  end subroutine cdf_var

  subroutine cdf_gatt(f, name, value)
    type(cdf_file), intent(inout) :: f
    character(len=*), intent(in) :: name
    real(8), intent(in) :: value
    f%ngatt = f%ngatt + 1
    f%gattname(f%ngatt) = name
    f%gattval(f%ngatt) = value
  end subroutine cdf_gatt

  integer function name_bytes(s)
    character(len=*), intent(in) :: s
    name_bytes = 4 + pad4(len_trim(s))
  end function name_bytes

  subroutine put_name(u, s)
    integer, intent(in) :: u
    character(len=*), intent(in) :: s
    integer :: n
    n = len_trim(s)
    write(u) int(n, 4), s(1:n), repeat(char(0), pad4(n) - n)
  end subroutine put_name

  ! the header's size: every field has a fixed width, begin offsets included
  integer(8) function header_bytes(f)
    type(cdf_file), intent(in) :: f
    integer :: i, a
    integer(8) :: h
    h = 4 + 4 + 8
    do i = 1, f%ndims
      h = h + name_bytes(f%dimname(i)) + 4
    enddo
    h = h + 8
    do i = 1, f%ngatt
      h = h + name_bytes(f%gattname(i)) + 8 + 8
    enddo
    h = h + 8
    do i = 1, f%nvars
      h = h + name_bytes(f%var(i)%name) + 4 + 4 * f%var(i)%ndims + 8
      do a = 1, f%var(i)%natt
        h = h + name_bytes(f%var(i)%attname(a)) + 8 + pad4(f%var(i)%attlen(a))
      enddo
      h = h + 4 + 4 + 8
    enddo
    header_bytes = h
  end function header_bytes

  subroutine cdf_create(f, path)
    type(cdf_file), intent(inout) :: f
    character(len=*), intent(in) :: path
    integer :: u, i, a, d
    integer(8) :: off
    f%path = path
    off = header_bytes(f)
    do i = 1, f%nvars
      if (.not. f%var(i)%rec) then
        f%var(i)%begin = off
        off = off + 8 * f%var(i)%count
      endif
    enddo
    f%recsize = 0
    do i = 1, f%nvars
      if (f%var(i)%rec) then
        f%var(i)%begin = off
        off = off + 8 * f%var(i)%count
        f%recsize = f%recsize + 8 * f%var(i)%count
      endif
    enddo
    f%numrecs = 0
    open(newunit=u, file=trim(path), access='stream', form='unformatted', status='replace', &
         convert='big_endian')
    write(u) 'CDF', char(2), int(0, 4)
    write(u) int(NC_DIMENSION, 4), int(f%ndims, 4)
    do d = 1, f%ndims
      call put_name(u, f%dimname(d))
      write(u) int(f%dimlen(d), 4)
    enddo
    if (f%ngatt > 0) then
      write(u) int(NC_ATTRIBUTE, 4), int(f%ngatt, 4)
    else
      write(u) int(0, 4), int(0, 4)          ! ABSENT
    endif
    do i = 1, f%ngatt
      call put_name(u, f%gattname(i))
      write(u) int(NC_DOUBLE, 4), int(1, 4), f%gattval(i)
    enddo
    write(u) int(NC_VARIABLE, 4), int(f%nvars, 4)
    do i = 1, f%nvars
      call put_name(u, f%var(i)%name)
      write(u) int(f%var(i)%ndims, 4)
      do d = 1, f%var(i)%ndims
        write(u) int(f%var(i)%dimid(d), 4)
      enddo
      write(u) int(NC_ATTRIBUTE, 4), int(f%var(i)%natt, 4)
      do a = 1, f%var(i)%natt
        call put_name(u, f%var(i)%attname(a))
        ! ncaptc stores the CHARACTER*len value: blank-padded to its length
        write(u) int(NC_CHAR, 4), int(f%var(i)%attlen(a), 4), &
                 f%var(i)%atttext(a)(1:f%var(i)%attlen(a)), &
                 repeat(char(0), pad4(f%var(i)%attlen(a)) - f%var(i)%attlen(a))
      enddo
      write(u) int(NC_DOUBLE, 4), int(min(8 * f%var(i)%count, 4294967295_8), 4), f%var(i)%begin
    enddo
    close(u)
    f%open = .true.
  end subroutine cdf_create

  integer function var_index(f, name)
    type(cdf_file), intent(in) :: f
    character(len=*), intent(in) :: name
    integer :: i
    do i = 1, f%nvars
      if (trim(f%var(i)%name) == trim(name)) then
        var_index = i
        return
      endif
    enddo
    print '(a,a)', ' rt6d_netcdf: unknown variable ', trim(name)
    stop 99
  end function var_index

  ! a fixed-size variable; data in file order (a Fortran array's storage)
  subroutine cdf_put(f, name, data)
    type(cdf_file), intent(in) :: f
    character(len=*), intent(in) :: name
    real(8), intent(in) :: data(:)
    integer :: i, u
    i = var_index(f, name)
    if (size(data, kind=8) /= f%var(i)%count) then
      print '(a,a)', ' rt6d_netcdf: wrong size for ', trim(name)
      stop 99
    endif
    open(newunit=u, file=trim(f%path), access='stream', form='unformatted', status='old', &
         action='readwrite', convert='big_endian')
    write(u, pos=f%var(i)%begin + 1) data
    close(u)
  end subroutine cdf_put

  ! one record variable's slab of the record being appended (index numrecs)
  subroutine cdf_put_rec(f, name, data)
    type(cdf_file), intent(in) :: f
    character(len=*), intent(in) :: name
    real(8), intent(in) :: data(:)
    integer :: i, u
    i = var_index(f, name)
    if (.not. f%var(i)%rec .or. size(data, kind=8) /= f%var(i)%count) then
      print '(a,a)', ' rt6d_netcdf: wrong record slab for ', trim(name)
      stop 99
    endif
    open(newunit=u, file=trim(f%path), access='stream', form='unformatted', status='old', &
         action='readwrite', convert='big_endian')
    write(u, pos=f%var(i)%begin + int(f%numrecs, 8) * f%recsize + 1) data
    close(u)
  end subroutine cdf_put_rec

  ! the record is complete: write the new record count
  subroutine cdf_end_rec(f)
    type(cdf_file), intent(inout) :: f
    integer :: u
    f%numrecs = f%numrecs + 1
    open(newunit=u, file=trim(f%path), access='stream', form='unformatted', status='old', &
         action='readwrite', convert='big_endian')
    write(u, pos=5) int(f%numrecs, 4)
    close(u)
  end subroutine cdf_end_rec

end module rt6d_netcdf


! The schema and the snapshot records, with the reference's routine names.
module rt6d_snapshots
  use rt6d_netcdf
  implicit none
  type(cdf_file), save :: snap3d, snap6d
  logical, save :: snaps_on = .false.
  character(len=8), parameter :: BFLUX(10) = [character(len=8) :: 'xs', 'xe', 'ys', 'ye', 'ks', 'ke', &
                                                'ls', 'le', 'ms', 'me']
  character(len=8), parameter :: DISS(6) = [character(len=8) :: 'diss_ks', 'diss_ke', 'diss_ls', &
                                               'diss_le', 'diss_ms', 'diss_me']
end module rt6d_snapshots

! init_diag (for_src/diagnose.f90:2-13): both files, the 6D one when
! enable_write_6D_single is set (the reference's default).
subroutine init_diag
  use main_module
  use rt6d_snapshots
  implicit none
  print '(a)', ' initialize snapshot file'
  call init_snap3D_cdf
  if (enable_write_6D_single) call init_snap6D_cdf_single
  snaps_on = .true.
end subroutine init_diag

! init_snap3D_cdf (for_src/netcdf_io_3D.f90:1-260)
subroutine init_snap3D_cdf
  use main_module
  use rt6d_snapshots
  implicit none
  character(len=4) :: t3(3), t4(4)
  character(len=32) :: side
  integer :: i
  call cdf_reset(snap3d)
  call cdf_dim(snap3d, 'xt', nx); call cdf_dim(snap3d, 'xu', nx)
  call cdf_dim(snap3d, 'yt', ny); call cdf_dim(snap3d, 'yu', ny)
  call cdf_dim(snap3d, 'zt', nz); call cdf_dim(snap3d, 'zu', nz)
  call cdf_dim(snap3d, 'Time', 0)
  call cdf_var(snap3d, 'Time', ['Time'], 'Time ', 'seconds', 'time_origin', '01-JAN-1900 00:00:00')
  call cdf_var(snap3d, 'xt', ['xt'], 'distance', 'm'); call cdf_var(snap3d, 'xu', ['xu'], 'distance', 'm')
  call cdf_var(snap3d, 'yt', ['yt'], 'distance', 'm'); call cdf_var(snap3d, 'yu', ['yu'], 'distance', 'm')
  call cdf_var(snap3d, 'zt', ['zt'], 'distance', 'm'); call cdf_var(snap3d, 'zu', ['zu'], 'distance', 'm')
  t3 = [character(len=4) :: 'zt', 'yt', 'xt']
  t4 = [character(len=4) :: 'Time', 'zt', 'yt', 'xt']
  call cdf_var(snap3d, 'Nsqr', t3, 'buoyancy frequency', '1/s')
  call cdf_var(snap3d, 'U', t3, 'mean flow', 'm/s')
  call cdf_var(snap3d, 'V', t3, 'mean flow', 'm/s')
  call cdf_var(snap3d, 'E', t4, 'energy density', 'm^2/s^2')
  call cdf_var(snap3d, 'meanfl', t4, 'energy change', 'm^2/s^3')
  do i = 1, 6
    call cdf_var(snap3d, trim(DISS(i)), t4, 'energy change', 'm^2/s^3')
  enddo
  do i = 1, 10
    side = 'Energy flux across min ' // BFLUX(i)(1:1)
    if (BFLUX(i)(2:2) == 'e') side = 'Energy flux across max ' // BFLUX(i)(1:1)
    call cdf_var(snap3d, 'bflux_' // trim(BFLUX(i)), ['Time'], trim(side), ' ')
  enddo
  call cdf_var(snap3d, 'mflint', ['Time'], 'mean flow energy flux', ' ')
  call cdf_gatt(snap3d, 'f0', f0)
  call cdf_gatt(snap3d, 'k_fixed', k_fixed)
  call cdf_gatt(snap3d, 'l_fixed', l_fixed)
  call cdf_gatt(snap3d, 'm_fixed', m_fixed)
  call cdf_create(snap3d, 'snap3D.cdf')
  call cdf_put(snap3d, 'Nsqr', reshape(Nsqr(1:nx, 1:ny, 1:nz), [nx * ny * nz]))
  call cdf_put(snap3d, 'U', reshape(U(1:nx, 1:ny, 1:nz), [nx * ny * nz]))
  call cdf_put(snap3d, 'V', reshape(V(1:nx, 1:ny, 1:nz), [nx * ny * nz]))
  call cdf_put(snap3d, 'xt', xt(1:nx)); call cdf_put(snap3d, 'xu', xu(1:nx))
  call cdf_put(snap3d, 'yt', yt(1:ny)); call cdf_put(snap3d, 'yu', yu(1:ny))
  call cdf_put(snap3d, 'zt', zt(1:nz)); call cdf_put(snap3d, 'zu', zu(1:nz))
end subroutine init_snap3D_cdf

! The face array of axis d (n_d + 1 faces, 0..n_d) at the reference's
! indices 1..n_d (the right face of each cell), flat in Fortran order.
subroutine faces_at_cells(n, d, faces, out)
  implicit none
  integer, intent(in) :: n(6), d
  real(8), intent(in) :: faces(*)
  real(8), intent(out) :: out(*)
  integer(8) :: s(6), sf(6), i, o
  integer :: a, c(6)
  s(1) = 1; sf(1) = 1
  do a = 2, 6
    s(a) = s(a - 1) * n(a - 1)
    sf(a) = sf(a - 1) * merge(n(a - 1) + 1, n(a - 1), a - 1 == d)
  enddo
  do i = 0, product(int(n, 8)) - 1
    o = i
    do a = 6, 1, -1
      c(a) = int(o / s(a))
      o = o - c(a) * s(a)
    enddo
    c(d) = c(d) + 1               ! face c_d + 1 (0-based) is the right face of cell c_d
    o = 0
    do a = 1, 6
      o = o + c(a) * sf(a)
    enddo
    out(i + 1) = faces(o + 1)
  enddo
end subroutine faces_at_cells

! init_snap6D_cdf_single (for_src/netcdf_io_6D.f90:240-656)
subroutine init_snap6D_cdf_single
  use main_module
  use rt6d_abi
  use rt6d_snapshots
  implicit none
  character(len=2), parameter :: ax(6) = ['x', 'y', 'z', 'k', 'l', 'm']
  character(len=24), parameter :: vln(6) = [character(len=24) :: 'hor. group velocity', 'hor. group velocity', &
      'vertical group velocity', 'hor. wavenr. change', 'hor. wavenr. change', 'vert. wavenr. change']
  character(len=4), parameter :: vun(6) = [character(len=4) :: 'm/s', 'm/s', 'm/s', '1/ms', '1/ms', '1/ms']
  character(len=4) :: td(6), dd(6)
  character(len=16) :: ln, un
  integer :: a, n(6)
  integer(8) :: cells
  real(8), allocatable :: buf(:), fb(:)
  n = [nx, ny, nz, nk, nl, nm]
  cells = product(int(n, 8))
  call cdf_reset(snap6d)
  do a = 1, 6
    call cdf_dim(snap6d, trim(ax(a)) // 't', n(a))
    call cdf_dim(snap6d, trim(ax(a)) // 'u', n(a))
  enddo
  call cdf_dim(snap6d, 'Time', 0)
  call cdf_var(snap6d, 'Time', ['Time'], 'Time ', 'seconds', 'time_origin', '01-JAN-1900 00:00:00')
  do a = 1, 6
    ln = 'distance'; un = 'm'
    if (a > 3) then
      ln = 'wavenumber'; un = '1/m'
    endif
    call cdf_var(snap6d, trim(ax(a)) // 't', [trim(ax(a)) // 't'], trim(ln), trim(un))
    call cdf_var(snap6d, trim(ax(a)) // 'u', [trim(ax(a)) // 'u'], trim(ln), trim(un))
    call cdf_var(snap6d, 'd' // trim(ax(a)) // 't', [trim(ax(a)) // 'u'], trim(ln), trim(un))
  enddo
  do a = 1, 6
    td(a) = trim(ax(7 - a)) // 't'
  enddo
  do a = 1, 6
    dd = td
    dd(7 - a) = trim(ax(a)) // 'u'
    call cdf_var(snap6d, trim(ax(a)) // 'dot', dd, trim(vln(a)), trim(vun(a)))
  enddo
  call cdf_var(snap6d, 'omega_i', td, 'intrinsic frequency', '1/s')
  call cdf_var(snap6d, 'omega_e', td, 'extrinsic frequency', '1/s')
  call cdf_var(snap6d, 'E', ['Time', td], 'energy density', 'm^5/s^2')
  call cdf_var(snap6d, 'meanfl', ['Time', td], 'energy change', 'm^5/s^3')
  call cdf_create(snap6d, 'snap6D.cdf')
  call cdf_put(snap6d, 'xt', xt(1:nx)); call cdf_put(snap6d, 'xu', xu(1:nx)); call cdf_put(snap6d, 'dxt', dxt(1:nx))
  call cdf_put(snap6d, 'yt', yt(1:ny)); call cdf_put(snap6d, 'yu', yu(1:ny)); call cdf_put(snap6d, 'dyt', dyt(1:ny))
  call cdf_put(snap6d, 'zt', zt(1:nz)); call cdf_put(snap6d, 'zu', zu(1:nz)); call cdf_put(snap6d, 'dzt', dzt(1:nz))
  call cdf_put(snap6d, 'kt', kt(1:nk)); call cdf_put(snap6d, 'ku', ku(1:nk)); call cdf_put(snap6d, 'dkt', dkt(1:nk))
  call cdf_put(snap6d, 'lt', lt(1:nl)); call cdf_put(snap6d, 'lu', lu(1:nl)); call cdf_put(snap6d, 'dlt', dlt(1:nl))
  call cdf_put(snap6d, 'mt', mt(1:nm)); call cdf_put(snap6d, 'mu', mu(1:nm)); call cdf_put(snap6d, 'dmt', dmt(1:nm))
  allocate(buf(cells))
  do a = 1, 6
    allocate(fb(cells / n(a) * (n(a) + 1)))
    call rt6d_check(rt6d_download(rt6d_handle, int(RT6D_FIELD_XDOT + a - 1, c_int), RT6D_LAYOUT_FACES, fb), &
                    'init_snap6D_cdf_single')
    call faces_at_cells(n, a, fb, buf)
    call cdf_put(snap6d, trim(ax(a)) // 'dot', buf)
    deallocate(fb)
  enddo
  call rt6d_check(rt6d_download(rt6d_handle, RT6D_FIELD_OMEGA_I, RT6D_LAYOUT_INTERIOR, buf), 'init_snap6D_cdf_single')
  call cdf_put(snap6d, 'omega_i', buf)
  call rt6d_check(rt6d_download(rt6d_handle, RT6D_FIELD_OMEGA_E, RT6D_LAYOUT_INTERIOR, buf), 'init_snap6D_cdf_single')
  call cdf_put(snap6d, 'omega_e', buf)
  deallocate(buf)
end subroutine init_snap6D_cdf_single

! The file part of diagnose (for_src/diagnose.f90:335-530): one record in
! each file at Time = itt*dt. Ei, Mi, Di: the 3D fields of rt6d_diagnose.
subroutine write_snapshots(d, Ei, Mi, Di)
  use main_module
  use rt6d_abi
  use rt6d_snapshots
  implicit none
  type(rt6d_diag_t), intent(in) :: d
  real(8), intent(in) :: Ei(nx * ny * nz), Mi(nx * ny * nz), Di(6 * nx * ny * nz)   ! explicit shape: called without an interface
  integer :: i, n3, bi(10)
  integer(8) :: cells
  real(8), allocatable :: buf(:)
  if (.not. snaps_on) return
  n3 = nx * ny * nz
  if (enable_write_6D_single) then
    cells = int(nx, 8) * ny * nz * nk * nl * nm
    allocate(buf(cells))
    print '(a,i6)', ' time steps in file snap6D.cdf : ', snap6d%numrecs + 1
    call cdf_put_rec(snap6d, 'Time', [itt * dt])
    call rt6d_check(rt6d_download(rt6d_handle, RT6D_FIELD_E, RT6D_LAYOUT_INTERIOR, buf), 'diagnose')
    call cdf_put_rec(snap6d, 'E', buf)
    call rt6d_check(rt6d_download(rt6d_handle, RT6D_FIELD_MEANFL, RT6D_LAYOUT_INTERIOR, buf), 'diagnose')
    call cdf_put_rec(snap6d, 'meanfl', buf)
    call cdf_end_rec(snap6d)
    deallocate(buf)
  endif
  print '(a,i6)', ' time steps in file snap3D.cdf : ', snap3d%numrecs + 1
  call cdf_put_rec(snap3d, 'Time', [itt * dt])
  call cdf_put_rec(snap3d, 'E', Ei)
  call cdf_put_rec(snap3d, 'meanfl', Mi)
  do i = 1, 6
    call cdf_put_rec(snap3d, trim(DISS(i)), Di((i - 1) * n3 + 1:i * n3))
  enddo
  call cdf_put_rec(snap3d, 'mflint', [d%mflint])
  ! rt6d_diag_t%bfl order: xs xe ys ye zs ze ks ke ls le ms me (z has no flux variables)
  bi = [1, 2, 3, 4, 7, 8, 9, 10, 11, 12]
  do i = 1, 10
    call cdf_put_rec(snap3d, 'bflux_' // trim(BFLUX(i)), [d%bfl(bi(i))])
  enddo
  call cdf_end_rec(snap3d)
end subroutine write_snapshots
