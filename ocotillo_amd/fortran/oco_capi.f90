! iso_c_binding interface to the C ABI of include/ocotillo.h.
! Arrays are assumed-size real(c_double) in the reference's Fortran layouts
! (source/flux_feautrier.f90:36-42); scalars pass by value.
module oco_capi
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: OCO_OK = 0, OCO_E_PIVOT0_STAGE1 = 1, OCO_E_PIVOT0_STAGE2 = 2, &
       OCO_E_ARG = 3, OCO_E_OOM = 4, OCO_E_HIP = 5, OCO_E_STATE = 6, OCO_E_NODEVICE = 7, OCO_E_BUSY = 8, &
       OCO_E_IO = 9, OCO_E_NONFINITE = 10
  integer(c_int), parameter, public :: OCO_MODE_FP64 = 0, OCO_MODE_MIXED = 1
  integer(c_int), parameter, public :: OCO_CHECK_FINITE = 1

  type, bind(C), public :: oco_diag
    integer(c_int) :: status, fail_ix, fail_iy, fail_iw, fail_j
    real(c_double) :: sum_a, sum_b, sum_c, sum_d
    real(c_double) :: minU, maxU, minV, maxV, minA, maxA
    real(c_double) :: cells_ms, sweep_ms, total_ms
  end type oco_diag

  ! oco_athena_units of include/ocotillo.h
  type, bind(C), public :: oco_athena_units
    real(c_double) :: unit_density, unit_temperature, gamma1, rho_floor, switch_temp, floor_temp
    integer(c_int) :: avoid_momentum_underflow
  end type oco_athena_units

  public :: oco_create, oco_destroy, oco_set_wavelengths, oco_set_physics, oco_set_checks, oco_solve_block, &
            oco_message
  public :: oco_athena_slab, oco_solve_block_files, oco_write_record

  interface
    integer(c_int) function oco_create(device, hbm_budget, ctx) bind(C, name='oco_create')
      import :: c_int, c_size_t, c_ptr
      integer(c_int), value :: device
      integer(c_size_t), value :: hbm_budget
      type(c_ptr), intent(out) :: ctx
    end function oco_create

    integer(c_int) function oco_destroy(ctx) bind(C, name='oco_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function oco_destroy

    integer(c_int) function oco_set_wavelengths(ctx, nw, waves) bind(C, name='oco_set_wavelengths')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nw
      real(c_double), intent(in) :: waves(*)
    end function oco_set_wavelengths

    integer(c_int) function oco_set_physics(ctx, fully_ionized_T, lgrey, sigma_grey) &
         bind(C, name='oco_set_physics')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: fully_ionized_T, sigma_grey
      integer(c_int), value :: lgrey
    end function oco_set_physics

    integer(c_int) function oco_set_checks(ctx, flags) bind(C, name='oco_set_checks')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: flags
    end function oco_set_checks

    integer(c_int) function oco_solve_block(ctx, nz, nyloc, nxloc, dz, dz1, dz2, rho3d, t3d, u, v, absorp, &
         mode, diag) bind(C, name='oco_solve_block')
      import :: c_int, c_ptr, c_double, oco_diag
      type(c_ptr), value :: ctx
      integer(c_int), value :: nz, nyloc, nxloc, mode
      real(c_double), value :: dz, dz1, dz2
      real(c_double), intent(in) :: rho3d(*), t3d(*)
      real(c_double), intent(inout) :: u(*), v(*), absorp(*)
      type(oco_diag), intent(out) :: diag
    end function oco_solve_block

    integer(c_int) function oco_solve_block_files(ctx, nz, nyloc, nxloc, dz, dz1, dz2, rho3d, t3d, mode, dir, &
         snapshot, stage_bytes, u_col, a_col, diag) bind(C, name='oco_solve_block_files')
      import :: c_int, c_ptr, c_double, c_char, c_size_t, oco_diag
      type(c_ptr), value :: ctx
      integer(c_int), value :: nz, nyloc, nxloc, mode
      real(c_double), value :: dz, dz1, dz2
      real(c_double), intent(in) :: rho3d(*), t3d(*)
      character(kind=c_char), intent(in) :: dir(*), snapshot(*)
      integer(c_size_t), value :: stage_bytes
      real(c_double), intent(out) :: u_col(*), a_col(*)
      type(oco_diag), intent(out) :: diag
    end function oco_solve_block_files

    integer(c_int) function oco_write_record(path, data, nbytes) bind(C, name='oco_write_record')
      import :: c_int, c_double, c_char, c_size_t
      character(kind=c_char), intent(in) :: path(*)
      real(c_double), intent(in) :: data(*)
      integer(c_size_t), value :: nbytes
    end function oco_write_record

    integer(c_int) function oco_athena_slab(ctx, nz, nyloc, nxloc, nzl, iz0, rho_raw, mx, my, mz, e, tgauss, &
         units, rho3d, t3d) bind(C, name='oco_athena_slab')
      import :: c_int, c_ptr, c_double, oco_athena_units
      type(c_ptr), value :: ctx
      integer(c_int), value :: nz, nyloc, nxloc, nzl, iz0
      real(c_double), intent(in) :: rho_raw(*), mx(*), my(*), mz(*), e(*), tgauss(*)
      type(oco_athena_units), intent(in) :: units
      real(c_double), intent(inout) :: rho3d(*), t3d(*)
    end function oco_athena_slab

    type(c_ptr) function oco_strerror_c(status) bind(C, name='oco_strerror')
      import :: c_int, c_ptr
      integer(c_int), value :: status
    end function oco_strerror_c

    integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function c_strlen
  end interface

contains

  ! Message text of a status code, as a Fortran string.
  function oco_message(status) result(msg)
    integer(c_int), intent(in) :: status
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = oco_strerror_c(status)
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    enddo
  end function oco_message

end module oco_capi
