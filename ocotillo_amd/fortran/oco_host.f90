! Drop-in host for the GPU path: keeps ocotillo's user surface -- input.in
! namelists, resolution.in, argv snapshot tag, the analytic disk or the
! Athena snapshot reader, diagnostics.txt and the per-block binary outputs --
! and replaces the column loop and the block's binary outputs
! (source/flux_feautrier.f90:160-228, 243) by one call per processor block
! into the HIP library (oco_solve_block_files), which streams U, V and
! absorp from the device into the three files; the Athena field arithmetic
! runs on the device too (oco_athena_slab).
!
! Differences from the reference, all additive:
!   * resolution.in is read at run time (the reference compiles it in,
!     source/common.f90:9), so one executable serves every shape;
!   * optional namelist &oco_input: lwave_log (log-spaced wavelengths,
!     default .false. = the reference's linear grid), oco_mode (0 FP64,
!     1 mixed, 2 accurate), oco_device, hbm_budget_gb, check_finite (stop at
!     a NaN or infinity in the outputs, default .false.);
!   * several snapshot tags or ranges a:b on the command line run each
!     snapshot as its own run would, into output_snapshot<tag>/ (what
!     samples/multiInTest/snapshot_looper.sh meant to do), in one process.

!------------------------------------------------------------------------------
module oco_shapes
  implicit none
  integer :: nw = 0, nx = 0, ny = 0, nz = 0, nprocx = 1, nprocy = 1, nprocz = 1, nxloc = 0, nyloc = 0
  integer, parameter :: ng = 3
  integer :: mz, n1, n2
  real(8), parameter :: epsi = epsilon(1d0)
  real(8), parameter :: amu = 1.66053906660d-24, k1_cgs = 1d0 / 1.380649d-16
  real(8), parameter :: G_Newton_cgs = 6.6742d-8, c_light_cgs = 2.99792458d10
  real(8), parameter :: SolarMass = 1.988409870698051d+33
  logical :: lroot = .true.

  integer, parameter, private :: maxvars = 64
  character(len=32), private :: vname(maxvars)
  integer, private :: vval(maxvars), nvars = 0
  character(len=512), private :: expr_s
  integer, private :: expr_p

contains

  ! resolution.in: lines "integer, parameter :: a=expr, b=expr ! comment".
  ! Expressions use integers, names defined earlier, + - * / and ().
  subroutine read_resolution(path)
    character(len=*), intent(in) :: path
    character(len=512) :: line, item
    integer :: u, ios, k, c0, depth, eq
    open(newunit=u, file=trim(path), status='old', action='read', iostat=ios)
    if (ios /= 0) then
      print*, "Cannot open ", trim(path)
      stop 1
    endif
    do
      read(u, '(A)', iostat=ios) line
      if (ios /= 0) exit
      k = index(line, '!')
      if (k > 0) line(k:) = ' '
      k = index(line, '::')
      if (k == 0) cycle
      line = line(k+2:)
      c0 = 1
      depth = 0
      do k = 1, len_trim(line) + 1
        if (k <= len_trim(line)) then
          if (line(k:k) == '(') depth = depth + 1
          if (line(k:k) == ')') depth = depth - 1
          if (.not. (line(k:k) == ',' .and. depth == 0)) cycle
        endif
        item = adjustl(line(c0:k-1))
        c0 = k + 1
        eq = index(item, '=')
        if (eq == 0) cycle
        call define(lower(trim(adjustl(item(1:eq-1)))), eval_expr(item(eq+1:)))
      enddo
    enddo
    close(u)
    nw = lookup('nw'); nx = lookup('nx'); ny = lookup('ny'); nz = lookup('nz')
    nprocx = lookup('nprocx'); nprocy = lookup('nprocy'); nprocz = lookup('nprocz')
    nxloc = lookup('nxloc'); nyloc = lookup('nyloc')
    mz = nz + 2*ng
    n1 = ng + 1
    n2 = mz - ng
  end subroutine read_resolution

  function lower(s) result(r)
    character(len=*), intent(in) :: s
    character(len=len(s)) :: r
    integer :: i
    r = s
    do i = 1, len(s)
      if (r(i:i) >= 'A' .and. r(i:i) <= 'Z') r(i:i) = achar(iachar(r(i:i)) + 32)
    enddo
  end function lower

  subroutine define(name, val)
    character(len=*), intent(in) :: name
    integer, intent(in) :: val
    integer :: i
    do i = 1, nvars
      if (vname(i) == name) then
        vval(i) = val
        return
      endif
    enddo
    nvars = nvars + 1
    vname(nvars) = name
    vval(nvars) = val
  end subroutine define

  integer function lookup(name)
    character(len=*), intent(in) :: name
    integer :: i
    do i = 1, nvars
      if (vname(i) == name) then
        lookup = vval(i)
        return
      endif
    enddo
    print*, "resolution.in does not define ", name
    stop 1
  end function lookup

  integer function eval_expr(s)
    character(len=*), intent(in) :: s
    expr_s = lower(s)
    expr_p = 1
    eval_expr = parse_sum()
  end function eval_expr

  subroutine skip_blank()
    do while (expr_p <= len_trim(expr_s))
      if (expr_s(expr_p:expr_p) /= ' ') exit
      expr_p = expr_p + 1
    enddo
  end subroutine skip_blank

  recursive integer function parse_sum() result(v)
    character :: op
    v = parse_term()
    do
      call skip_blank()
      if (expr_p > len_trim(expr_s)) exit
      op = expr_s(expr_p:expr_p)
      if (op /= '+' .and. op /= '-') exit
      expr_p = expr_p + 1
      if (op == '+') then
        v = v + parse_term()
      else
        v = v - parse_term()
      endif
    enddo
  end function parse_sum

  recursive integer function parse_term() result(v)
    character :: op
    v = parse_factor()
    do
      call skip_blank()
      if (expr_p > len_trim(expr_s)) exit
      op = expr_s(expr_p:expr_p)
      if (op /= '*' .and. op /= '/') exit
      expr_p = expr_p + 1
      if (op == '*') then
        v = v * parse_factor()
      else
        v = v / parse_factor()
      endif
    enddo
  end function parse_term

  recursive integer function parse_factor() result(v)
    integer :: p0
    call skip_blank()
    if (expr_s(expr_p:expr_p) == '(') then
      expr_p = expr_p + 1
      v = parse_sum()
      call skip_blank()
      expr_p = expr_p + 1
    else if (expr_s(expr_p:expr_p) == '-') then
      expr_p = expr_p + 1
      v = -parse_factor()
    else if (index('0123456789', expr_s(expr_p:expr_p)) > 0) then
      p0 = expr_p
      do while (expr_p <= len_trim(expr_s))
        if (index('0123456789', expr_s(expr_p:expr_p)) == 0) exit
        expr_p = expr_p + 1
      enddo
      read(expr_s(p0:expr_p-1), *) v
    else
      p0 = expr_p
      do while (expr_p <= len_trim(expr_s))
        if (index('abcdefghijklmnopqrstuvwxyz0123456789_', expr_s(expr_p:expr_p)) == 0) exit
        expr_p = expr_p + 1
      enddo
      v = lookup(expr_s(p0:expr_p-1))
    endif
  end function parse_factor

  logical function file_exists(filename)
    character(len=*), intent(in) :: filename
    inquire(file=trim(filename), exist=file_exists)
  end function file_exists

  character(len=21) function itoa(n)
    integer, intent(in) :: n
    write(itoa, '(I21)') n
    itoa = adjustl(itoa)
  end function itoa

end module oco_shapes

!------------------------------------------------------------------------------
! The analytic disk (source/disk.f90:9-129): &temperature_input, &density_input.
module oco_disk
  use oco_shapes
  implicit none
  real(8) :: isoTemp, sigma, midTemp, floorTemp, switchTemp
  real(8) :: rho0, rho_floor, H
  real(8), allocatable, private :: Tgauss(:)
  namelist /temperature_input/ isoTemp, sigma, midTemp, floorTemp
  namelist /density_input/ rho0, rho_floor, H
contains
  subroutine read_disk_input(inputfile)
    character(len=*), intent(in) :: inputfile
    integer :: u
    open(newunit=u, file=trim(inputfile))
    read(u, nml=temperature_input)
    close(u)
    switchTemp = (1 + epsi) * isoTemp
    open(newunit=u, file=trim(inputfile))
    read(u, nml=density_input)
    close(u)
  end subroutine read_disk_input

  subroutine calc_grid(z1, z0, z, dz)
    real(8), intent(in) :: z1, z0
    real(8), intent(out) :: z(mz), dz
    integer :: i
    dz = (z1 - z0) / (nz - 1)
    do i = 1, nz
      z(n1 + i - 1) = z0 + (i - 1) * dz
    enddo
    print*, z(n1), z(n2)
    do i = 1, ng
      z(n1 - i) = z(n1) - i * dz
      z(n2 + i) = z(n2) + i * dz
    enddo
  end subroutine calc_grid

  ! Forget the kept profile: each snapshot of a multi-snapshot run starts
  ! as the reference's separate process would.
  subroutine reset_profile()
    if (allocated(Tgauss)) deallocate(Tgauss)
  end subroutine reset_profile

  ! The Gaussian temperature profile, built on the first call and kept
  ! (source/disk.f90:87-90).
  subroutine gaussian_profile(z, tg)
    real(8), intent(in) :: z(mz)
    real(8), intent(out) :: tg(nz)
    if (.not. allocated(Tgauss)) then
      allocate(Tgauss(nz))
      Tgauss = midTemp * exp(-0.5d0 * z(n1:n2)**2 / sigma**2)
    endif
    tg = Tgauss
  end subroutine gaussian_profile

  ! T rule of the analytic disk (source/disk.f90:64-109).
  subroutine calc_temperature(T, z, lathena)
    real(8), intent(inout) :: T(nz)
    real(8), intent(in) :: z(mz)
    logical, intent(in) :: lathena
    real(8) :: tg(nz)
    integer :: i
    if (.not. lathena) T = midTemp
    call gaussian_profile(z, tg)
    do i = 1, nz
      if (T(i) < switchTemp) then
        if (tg(i) > floorTemp) then
          T(i) = tg(i)
        else
          T(i) = floorTemp
        endif
      endif
    enddo
    if (.not. lathena) print*, "maxval(T), minval(T)", maxval(T), minval(T)
  end subroutine calc_temperature

  subroutine calc_density(rho, z)
    real(8), intent(out) :: rho(nz)
    real(8), intent(in) :: z(mz)
    integer :: i
    rho = rho0 * exp(-0.5d0 * z(n1:n2)**2 / H**2)
    do i = 1, nz
      if (rho(i) < rho_floor) rho(i) = rho_floor
    enddo
    print*, "maxval(rho), minval(rho)", maxval(rho), minval(rho)
  end subroutine calc_density
end module oco_disk

!------------------------------------------------------------------------------
! Athena-4.2 .bin reader (file format and units of source/read_athena.f90:33-291).
! The host reads the files and sets up the z grid, the units and the Gaussian
! profile; the field arithmetic (kinetic and internal energy, unit
! conversion, density floor, temperature rule) runs on the device, one
! oco_athena_slab call per z slab.
module oco_athena
  use oco_shapes
  use oco_disk, only: gaussian_profile, switchTemp, floorTemp
  use oco_capi, only: oco_athena_units, oco_athena_slab, oco_message, OCO_OK
  use, intrinsic :: iso_c_binding, only: c_int, c_double, c_ptr
  use, intrinsic :: ieee_arithmetic
  implicit none
  character(len=90) :: RunName = ' ', datadir = "./input_athena/distributed"
  real(8) :: Mbh_SolarMasses, r0ref_rg, aspect_ratio, mean_molecular_weight, rho0, rho_floor
  logical :: lavoid_momentum_underflow = .true.
  real(8), private :: unit_length, unit_density, unit_temperature
  namelist /athena_input/ RunName, Mbh_SolarMasses, r0ref_rg, aspect_ratio, mean_molecular_weight, rho0, &
       datadir, rho_floor, lavoid_momentum_underflow
contains
  subroutine read_athena_input(inputfile)
    character(len=*), intent(in) :: inputfile
    integer :: u
    real(8) :: Mbh, rg, rr, g0, Omega, Hs, unit_time, unit_velocity
    rho_floor = ieee_value(rho_floor, ieee_quiet_nan)
    open(newunit=u, file=trim(inputfile))
    read(u, nml=athena_input)
    close(u)
    Mbh = Mbh_SolarMasses * SolarMass
    rg = 2 * (G_Newton_cgs / c_light_cgs**2) * Mbh
    rr = r0ref_rg * rg
    g0 = G_Newton_cgs * Mbh
    Omega = sqrt(g0) * rr**(-1.5d0)
    Hs = aspect_ratio * rr
    unit_density = rho0
    unit_time = 1d0 / Omega
    unit_length = Hs
    unit_velocity = unit_length / unit_time
    unit_temperature = mean_molecular_weight * amu * k1_cgs * unit_velocity**2
    print*, 'unit_time        = ', unit_time, ' s'
    print*, 'unit_length      = ', unit_length, ' cm'
    print*, 'unit_velocity    = ', unit_velocity, ' cm/s'
    print*, 'unit_density     = ', unit_density, ' g/cm3'
    print*, 'unit_temperature = ', unit_temperature, ' K'
  end subroutine read_athena_input

  function slab_file(iprocx, iprocy, iprocz, snapshot) result(filename)
    integer, intent(in) :: iprocx, iprocy, iprocz
    character(len=*), intent(in) :: snapshot
    character(len=256) :: filename, head, sproc
    integer :: iproc
    iproc = modulo(iprocz, nprocz) * nprocx * nprocy + modulo(iprocy, nprocy) * nprocx + modulo(iprocx, nprocx)
    sproc = itoa(iproc)
    head = trim(datadir) // "/id" // trim(sproc) // "/" // trim(RunName)
    if (iproc == 0) then
      filename = trim(head) // "." // trim(snapshot) // ".bin"
    else
      filename = trim(head) // "-id" // trim(sproc) // "." // trim(snapshot) // ".bin"
    endif
    if (.not. file_exists(filename)) then
      print*, "iproc=", iproc
      print*, "File ", trim(filename), " does not exist"
      stop 1
    endif
  end function slab_file

  ! Header of one slab file up to and including the z coordinates
  ! (source/read_athena.f90:123-150); leaves the unit positioned at the fields.
  subroutine read_header(u, nzl, gamma1, zloc)
    integer, intent(in) :: u
    integer(c_int), intent(out) :: nzl
    real(c_double), intent(out) :: gamma1
    real(c_double), allocatable, intent(out) :: zloc(:)
    integer(c_int) :: hdr(4), nxl, nyl, coordsys
    real(c_double) :: misc(3)
    real(c_double), allocatable :: xloc(:), yloc(:)
    read(u) coordsys
    read(u) nxl
    if (nxl /= nxloc) then
      print*, "Resolution in x from Athena = ", nxl
      print*, "RT x-dimensionality from resolution.in = ", nxloc
      print*, "The Athena x dimensionality does not match the chosen for the RT post-processing. Fix."
      stop 1
    endif
    read(u) nyl
    if (nyl /= nyloc) then
      print*, "Resolution in y from Athena = ", nyl
      print*, "RT y-dimensionality from resolution.in = ", nyloc
      print*, "The Athena y dimensionality does not match the chosen for the RT post-processing. Fix."
      stop 1
    endif
    read(u) nzl
    read(u) hdr
    read(u) gamma1
    read(u) misc
    allocate(xloc(nxloc), yloc(nyloc), zloc(nzl))
    read(u) xloc
    read(u) yloc
    read(u) zloc
  end subroutine read_header

  subroutine read_from_athena(ctx, z, dz, rho, temp, iprocx, iprocy, snapshot)
    type(c_ptr), intent(in) :: ctx
    real(8), intent(out) :: z(mz), dz, rho(nz, nyloc, nxloc), temp(nz, nyloc, nxloc)
    integer, intent(in) :: iprocx, iprocy
    character(len=*), intent(in) :: snapshot
    integer(c_int) :: nzl, st
    real(c_double) :: gamma1
    real(c_double), allocatable :: zloc(:), f(:,:,:,:)
    real(8) :: zn(nz), z0, tg(nz)
    integer :: iprocz, u, iz0(0:nprocz - 1), nzls(0:nprocz - 1), i
    type(oco_athena_units) :: units

    ! pass 1: headers -> the block's z grid and each slab's place in it
    z0 = 0; dz = 0
    do iprocz = 0, nprocz - 1
      open(newunit=u, file=trim(slab_file(iprocx, iprocy, iprocz, snapshot)), form='unformatted', &
           access='stream', status='old', action='read')
      call read_header(u, nzl, gamma1, zloc)
      close(u)
      if (iprocz == 0) then
        z0 = zloc(1)
        dz = zloc(2) - zloc(1)
      endif
      iz0(iprocz) = nint((zloc(1) - z0) / dz) + 1
      nzls(iprocz) = nzl
      if (iz0(iprocz) < 1 .or. iz0(iprocz) + nzl - 1 > nz) then
        print*, "Athena slab", iprocz, " does not fit nz =", nz
        stop 1
      endif
      zn(iz0(iprocz):iz0(iprocz) + nzl - 1) = zloc
    enddo
    ! postprocess_athena_values (source/read_athena.f90:232-268): the grid
    ! and the scalars here, the fields on the device
    if (ieee_is_nan(rho_floor)) then
      print*, ""
      print*, "Must set `rho_floor` (cgs) value in namelist: `input.in/athena_input`."
      print*, ""
      stop 1
    endif
    zn = zn * unit_length
    dz = (zn(nz) - zn(1)) / (nz - 1)
    z(n1:n2) = zn
    do i = 1, ng
      z(n1 - i) = z(n1) - i * dz
      z(n2 + i) = z(n2) + i * dz
    enddo
    call gaussian_profile(z, tg)
    ! gamma1 is the last slab's, as in the reference
    units = oco_athena_units(unit_density, unit_temperature, gamma1, rho_floor, switchTemp, floorTemp, &
         merge(1_c_int, 0_c_int, lavoid_momentum_underflow))
    ! pass 2: the fields of each slab, converted on the device
    do iprocz = 0, nprocz - 1
      open(newunit=u, file=trim(slab_file(iprocx, iprocy, iprocz, snapshot)), form='unformatted', &
           access='stream', status='old', action='read')
      call read_header(u, nzl, gamma1, zloc)
      allocate(f(nxloc, nyloc, nzl, 5))
      do i = 1, 5
        read(u) f(:, :, :, i)
      enddo
      close(u)
      st = oco_athena_slab(ctx, int(nz, c_int), int(nyloc, c_int), int(nxloc, c_int), nzl, &
           int(iz0(iprocz) - 1, c_int), f(:, :, :, 1), f(:, :, :, 2), f(:, :, :, 3), f(:, :, :, 4), &
           f(:, :, :, 5), tg, units, rho, temp)
      deallocate(f)
      if (st /= OCO_OK) then
        print*, "ocotillo GPU library: ", oco_message(st)
        stop 2
      endif
    enddo
    if (lroot) then
      print*, 'minval(rho),maxval(rho),mean(rho)', minval(rho), maxval(rho), sum(rho) / (nxloc * nyloc * nz)
      print*, 'minval(T),maxval(T),mean(T)', minval(temp), maxval(temp), sum(temp) / (nxloc * nyloc * nz)
    endif
  end subroutine read_from_athena
end module oco_athena

!------------------------------------------------------------------------------
! Writers (source/io.f90:31-73).  The per-block binary outputs
! (output_binary, :76-101) are written by the library as the block is solved
! (oco_solve_block_files).
module oco_output
  use oco_shapes
  use oco_capi, only: oco_write_record, oco_message, OCO_OK
  use, intrinsic :: iso_c_binding, only: c_int, c_size_t, c_null_char
  implicit none
  ! directory of the run's outputs: output/ (the reference's) for one
  ! snapshot, output_snapshot<tag>/ per snapshot when several are given
  character(len=256) :: outroot = 'output'
contains
  subroutine output_ascii(U, absorp)
    real(8), intent(in) :: U(mz, nw), absorp(nz, nw)
    character(len=32) :: fmt
    integer :: i, iw, u10
    write(fmt, '("(I6,",I0,"(2e15.5))")') nw
    open(newunit=u10, file=trim(outroot) // "/diagnostics.txt", status="replace", action='write')
    do i = 1, nz
      write(u10, fmt) i, (U(n1 + i - 1, iw), iw = 1, nw), (absorp(i, iw), iw = 1, nw)
    enddo
    close(u10)
  end subroutine output_ascii

  subroutine output_grid(z, waves)
    real(8), intent(in) :: z(mz), waves(nw)
    integer :: i, u
    open(newunit=u, file=trim(outroot) // "/zgrid.txt", status="replace", action='write')
    do i = 1, nz
      write(u, *) i, z(n1 + i - 1)
    enddo
    close(u)
    call write_record(trim(outroot) // "/zgrid.bin", z, size(z))
    open(newunit=u, file=trim(outroot) // "/wavegrid.txt", status="replace", action="write")
    do i = 1, nw
      write(u, *) i, waves(i)
    enddo
    close(u)
    call write_record(trim(outroot) // "/wavegrid.bin", waves, size(waves))
  end subroutine output_grid

  ! One array as a one-record file in the reference's (gfortran's) layout.
  subroutine write_record(path, a, n)
    character(len=*), intent(in) :: path
    integer, intent(in) :: n
    real(8), intent(in) :: a(n)
    integer(c_int) :: st
    st = oco_write_record(trim(path) // c_null_char, a, int(8 * n, c_size_t))
    if (st /= OCO_OK) then
      print*, "cannot write ", trim(path), ": ", oco_message(st)
      stop 2
    endif
  end subroutine write_record

end module oco_output

!------------------------------------------------------------------------------
program flux_feautrier_gpu
  use, intrinsic :: iso_c_binding
  use oco_capi
  use oco_shapes
  use oco_disk
  use oco_athena, only: read_athena_input, read_from_athena
  use oco_output
  implicit none
  real(8), allocatable :: rho3d(:,:,:), temp3d(:,:,:), Ucol(:,:), Acol(:,:)
  real(8), allocatable :: z(:), waves(:), rho(:), T(:)
  real(8) :: z0, z1, w0 = 3000, w1 = 5000, sigma_grey, dz, dz1 = 0, dz2 = 0, dw
  real(8) :: fully_ionized_T, hbm_budget_gb = 0
  real :: start, finish, start_loop = 0, finish_loop = 0
  integer(8) :: c0, c1, crate
  logical :: lgrey = .false., lread_athena = .true., lwave_log = .false., check_finite = .false.
  integer :: oco_mode = 0, oco_device = 0, iprocx, iprocy, ix, iy, i, err, iu
  integer(c_int) :: st
  character(len=90) :: snapshot, inputfile = './input.in', arg
  character(len=90), allocatable :: snaps(:)
  character(len=256) :: outdir
  integer :: isnap, nsnap
  type(c_ptr) :: ctx
  type(oco_diag) :: diag
  real(8) :: loop_wall = 0, dev_ms = 0
  namelist /input/ z0, z1, w0, w1, sigma_grey, lgrey, lread_athena
  namelist /gas_state_input/ fully_ionized_T
  namelist /oco_input/ lwave_log, oco_mode, oco_device, hbm_budget_gb, check_finite

  call cpu_time(start)
  call get_command_argument(1, snapshot)
  if (snapshot == '') snapshot = "0000"
  if (snapshot /= '--print-config') call snapshot_list(snaps)
  nsnap = 1
  if (allocated(snaps)) nsnap = size(snaps)

  call read_resolution('./resolution.in')
  open(newunit=iu, file=trim(inputfile))
  read(iu, nml=input)
  close(iu)
  open(newunit=iu, file=trim(inputfile))
  read(iu, nml=oco_input, iostat=err)
  close(iu)
  if (snapshot == '--print-config') then
    ! inputs as parsed, without touching the GPU
    print '(A,9I8)', 'shapes nw nx ny nz nprocx nprocy nprocz nxloc nyloc:', nw, nx, ny, nz, nprocx, nprocy, &
         nprocz, nxloc, nyloc
    print '(A,L2,L2,L2,I3)', 'lgrey lread_athena lwave_log oco_mode:', lgrey, lread_athena, lwave_log, oco_mode
    print '(A,4ES24.16)', 'z0 z1 w0 w1:', z0, z1, w0, w1
    stop
  endif
  if (lgrey .and. (nw /= 1)) then
    print*, "For Grey RT use only one wavelength. Switch nw=1 in resolution.in"
    stop
  endif

  ! U, V and absorp stay on the device and go to the files in chunks; the
  ! host holds the inputs and the column printed to diagnostics.txt
  allocate(rho3d(nz, nyloc, nxloc), temp3d(nz, nyloc, nxloc), z(mz), waves(nw), rho(nz), T(nz), &
       Ucol(mz, nw), Acol(nz, nw), stat=err)
  if (err /= 0) then
    print*, "Could not allocate the input arrays."
    stop
  endif

  call read_disk_input(inputfile)
  open(newunit=iu, file=trim(inputfile))
  read(iu, nml=gas_state_input)
  close(iu)

  ! calc_wavelength (source/continuous_opacity.f90:37-62), plus a log grid
  if (nw > 1) then
    if (lwave_log) then
      do i = 1, nw
        waves(i) = exp(log(w0) + (i - 1) * (log(w1) - log(w0)) / (nw - 1))
      enddo
    else
      dw = (w1 - w0) / (nw - 1)
      do i = 1, nw
        waves(i) = w0 + (i - 1) * dw
      enddo
    endif
  else
    waves = w0
  endif
  print*, 'waves_angstrom min/max', minval(waves), maxval(waves)
  print*, 'number of waves', nw

  st = oco_create(int(oco_device, c_int), int(hbm_budget_gb * 1073741824d0, c_size_t), ctx)
  if (st /= OCO_OK) then
    print*, "ocotillo GPU library: ", oco_message(st)
    stop 2
  endif
  st = oco_set_wavelengths(ctx, int(nw, c_int), waves)
  if (st == OCO_OK) st = oco_set_physics(ctx, fully_ionized_T, merge(1_c_int, 0_c_int, lgrey), sigma_grey)
  if (st == OCO_OK .and. check_finite) st = oco_set_checks(ctx, OCO_CHECK_FINITE)
  if (st /= OCO_OK) then
    print*, "ocotillo GPU library: ", oco_message(st)
    stop 2
  endif

  if (lread_athena) call read_athena_input(inputfile)

  ! one snapshot: the reference's run; several (the replacement of
  ! samples/multiInTest/snapshot_looper.sh): each as its own run would be,
  ! into output_snapshot<tag>/, with one context and resident tables
  do isnap = 1, nsnap
  snapshot = snaps(isnap)
  if (nsnap > 1) then
    outroot = 'output_snapshot' // trim(snapshot)
    call execute_command_line('mkdir -p ' // trim(outroot))
    call reset_profile()
  endif
  if (lread_athena) print*, "snapshot: ", snapshot
  do iprocx = 0, nprocx - 1
    do iprocy = 0, nprocy - 1
      lroot = (iprocx == 0) .and. (iprocy == 0)
      if (lread_athena) then
        call read_from_athena(ctx, z, dz, rho3d, temp3d, iprocx, iprocy, snapshot)
      else
        if (lroot) call calc_grid(z1, z0, z, dz)
        call calc_density(rho, z)
        call calc_temperature(T, z, .false.)
        do ix = 1, nxloc
          do iy = 1, nyloc
            rho3d(:, iy, ix) = rho
            temp3d(:, iy, ix) = T
          enddo
        enddo
      endif
      if (lroot) then
        dz1 = 1d0 / dz
        dz2 = dz**2
      endif
      if (lroot) call cpu_time(start_loop)
      outdir = trim(outroot) // '/procx' // trim(itoa(iprocx)) // '_procy' // trim(itoa(iprocy))
      call execute_command_line('mkdir -p ' // trim(outdir))
      call system_clock(c0, crate)
      st = oco_solve_block_files(ctx, int(nz, c_int), int(nyloc, c_int), int(nxloc, c_int), dz, dz1, dz2, &
           rho3d, temp3d, int(oco_mode, c_int), trim(outdir) // c_null_char, trim(snapshot) // c_null_char, &
           0_c_size_t, Ucol, Acol, diag)
      call system_clock(c1)
      if (st == OCO_E_PIVOT0_STAGE1) then
        print*, "ERROR tridiag stage 1: bet=b(1)) = ", 0d0
        print*, "(column ix,iy =", diag%fail_ix, diag%fail_iy, " wavelength", diag%fail_iw, ")"
        stop
      else if (st == OCO_E_PIVOT0_STAGE2) then
        print*, "ERROR tridiag stage 2: bet = b(j)-a(j)*gam(j) =", 0d0
        print*, "(column ix,iy =", diag%fail_ix, diag%fail_iy, " wavelength", diag%fail_iw, " row", diag%fail_j, ")"
        stop
      else if (st == OCO_E_NONFINITE) then
        ! the device counterpart of the reference's -ffpe-trap debug build
        print*, "ERROR: NaN or infinity in the outputs (check_finite)"
        print*, "(column ix,iy =", diag%fail_ix, diag%fail_iy, " wavelength", diag%fail_iw, " cell", diag%fail_j, ")"
        stop 2
      else if (st /= OCO_OK) then
        print*, "ocotillo GPU library: ", oco_message(st)
        stop 2
      endif
      if (lroot) then
        call cpu_time(finish_loop)
        loop_wall = real(c1 - c0, 8) / real(crate, 8)
        dev_ms = diag%cells_ms + diag%sweep_ms
        print*, 'sum(a), sum(b), sum(c), sum(d)=', diag%sum_a, diag%sum_b, diag%sum_c, diag%sum_d
        print*, 'min/max(U)', diag%minU, diag%maxU
        print*, 'min/max(V)', diag%minV, diag%maxV
        print*, 'min/max(absorp_coeff)', diag%minA, diag%maxA
        call output_grid(z, waves)
        call output_ascii(Ucol, Acol)
      endif
    enddo
  enddo
  if (nsnap > 1) print*, "snapshot ", trim(snapshot), " done: ", trim(outroot)
  enddo
  st = oco_destroy(ctx)

  call cpu_time(finish)
  print*, ''
  print*, "processor column time = ", finish_loop - start_loop, " seconds."
  print*, "Execution time =", (finish_loop - start_loop) / (1d0 * nz * nyloc * nxloc * nw) * 1e6, &
       " micro-seconds per wavelength point per mesh point."
  print*, "GPU block and its output files (wall) =", loop_wall, " s; device kernels =", dev_ms * 1d-3, " s"
  print*, ''
  print*, "Wall time = ", finish - start, " seconds."
  print*, "Execution time =", (finish - start) / (1d0 * nz * ny * nx * nw) * 1e6, &
       " micro-seconds per wavelength point per mesh point."
contains

  ! Snapshot tags from the command line: none -> 0000; each argument a tag
  ! or a range a:b of snapshot numbers (tags zero-padded to four digits, as
  ! Athena names its dumps).
  subroutine snapshot_list(tags)
    character(len=90), allocatable, intent(out) :: tags(:)
    character(len=90), allocatable :: tmp(:)
    integer :: iarg, narg, n, a0, a1, k, colon, ios
    narg = command_argument_count()
    allocate(tags(0))
    if (narg == 0) then
      tags = [character(len=90) :: "0000"]
      return
    endif
    do iarg = 1, narg
      call get_command_argument(iarg, arg)
      colon = index(arg, ':')
      if (colon > 0) then
        read(arg(1:colon - 1), *, iostat=ios) a0
        if (ios == 0) read(arg(colon + 1:), *, iostat=ios) a1
        if (ios /= 0 .or. a1 < a0 .or. a0 < 0) then
          print*, "bad snapshot range: ", trim(arg)
          stop 1
        endif
        n = size(tags)
        allocate(tmp(n + a1 - a0 + 1))
        tmp(1:n) = tags
        do k = a0, a1
          if (k < 10000) then
            write(tmp(n + k - a0 + 1), '(I4.4)') k
          else
            write(tmp(n + k - a0 + 1), '(I0)') k
          endif
        enddo
        call move_alloc(tmp, tags)
      else
        tags = [tags, arg]
      endif
    enddo
  end subroutine snapshot_list

end program flux_feautrier_gpu
