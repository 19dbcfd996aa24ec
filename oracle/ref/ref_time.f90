! Timing driver for the ocotillo reference (test infrastructure only: the CPU
! baseline beside the GPU).
!
! Links against the reference's own modules, compiled unchanged from
! source/{common,gas_state,continuous_opacity,auxiliary}.f90, and runs the
! column loop of source/flux_feautrier.f90:160-228 -- for each column
! wavelength_independent_pillars, then per wavelength calc_opacity_and_albedo
! (or grey_parameters), get_tridag_coefficients, tridag, update_ghosts,
! calc_flux and the stores into U, absorp_coeff and V -- over ncol columns
! read at run time, into arrays laid out as the reference's
! (source/flux_feautrier.f90:36-42: U(mz, column, nw), V and absorp(nz, column,
! nw)).  nz and nw come from the resolution.in the recipe writes next to the
! objects (compile-time shapes, source/common.f90:9).
!
! Usage: ref_time.x <input.bin> <gas_state namelist file> [<dump.bin> <ndump>]
!   input.bin (stream, little endian):
!     int32 ncol, int32 lgrey, float64 dz, float64 sigma_grey,
!     float64 waves(nw), float64 rho(nz,ncol), float64 T(nz,ncol)
! Prints one line: ncol nz nw wall_s cpu_s sum(U) sum(V) sum(absorp).
! With a dump file, after the timed loop, writes U(:,1:ndump,:),
! V(:,1:ndump,:) and absorp(:,1:ndump,:) to it (stream, in that order) for
! a point-by-point comparison.
program ref_time
  use Common
  use GasState
  use ContinuousOpacity
  use Auxiliary
  implicit none
  type (pillar_case) :: p
  real, dimension(nz) :: a, b, c, d
  real, dimension(nw) :: waves
  real, allocatable :: rho(:,:), T(:,:), U(:,:,:), V(:,:,:), absorp(:,:,:)
  real :: dz, dz1, dz2, sigma_grey, c0, c1
  integer(4) :: ncol, lgrey_i
  integer :: icol, iw, iu
  integer(8) :: t0, t1, rate
  character(len=256) :: fin, fnml, fdump, sdump
  character(len=90) :: nml90
  integer :: ndump

  call get_command_argument(1, fin)
  call get_command_argument(2, fnml)
  nml90 = trim(fnml)
  call read_gas_state_input(nml90)

  open(newunit=iu, file=trim(fin), form='unformatted', access='stream', status='old')
  read(iu) ncol, lgrey_i, dz, sigma_grey
  read(iu) waves
  allocate(rho(nz,ncol), T(nz,ncol))
  read(iu) rho
  read(iu) T
  close(iu)
  allocate(U(mz,ncol,nw), V(nz,ncol,nw), absorp(nz,ncol,nw))
  U = 0.
  V = 0.
  absorp = 0.

  lfirst = .false.
  lroot = .true.
  call pre_calc_opacity_quantities(waves)
  dz1 = 1./dz
  dz2 = dz**2

  call system_clock(t0, rate)
  call cpu_time(c0)
  do icol = 1, ncol
    p%rho = rho(:,icol)
    p%T = T(:,icol)
    call wavelength_independent_pillars(p)
    do iw = 1, nw
      if (lgrey_i /= 0) then
        call grey_parameters(p, sigma_grey)
      else
        call calc_opacity_and_albedo(p, iw)
      endif
      call get_tridag_coefficients(a, b, c, d, p, dz, dz2)
      call tridag(a, b, c, d, U(n1:n2,icol,iw))
      call update_ghosts(U(:,icol,iw))
      call calc_flux(U(:,icol,iw), p, dz1)
      absorp(:,icol,iw) = p%opacity
      V(:,icol,iw) = p%flux
    enddo
  enddo
  call cpu_time(c1)
  call system_clock(t1)
  print '(3I8,2ES16.8,3ES24.16)', ncol, nz, nw, real(t1 - t0) / real(rate), c1 - c0, sum(U), sum(V), sum(absorp)
  if (command_argument_count() >= 4) then
    call get_command_argument(3, fdump)
    call get_command_argument(4, sdump)
    read(sdump, *) ndump
    ndump = min(ndump, int(ncol))
    open(newunit=iu, file=trim(fdump), form='unformatted', access='stream', status='replace')
    write(iu) U(:,1:ndump,:)
    write(iu) V(:,1:ndump,:)
    write(iu) absorp(:,1:ndump,:)
    close(iu)
  endif
endprogram ref_time
