! Golden-vector driver for the ocotillo reference (test infrastructure only).
!
! Links against the reference's own modules, compiled unchanged from
! source/{common,gas_state,continuous_opacity,auxiliary}.f90, and runs the
! per-column body of source/flux_feautrier.f90:174-216 on arbitrary
! (rho, T) columns and an arbitrary wavelength array, dumping every
! intermediate the GPU path must reproduce.  nz and nw come from the
! resolution.in the recipe writes next to the objects (compile-time shapes,
! source/common.f90:9).
!
! Usage: ref_dump.x <input.bin> <gas_state namelist file> <output.bin>
!   input.bin (stream, little endian):
!     int32 ncol, int32 lgrey, float64 dz, float64 sigma_grey,
!     float64 waves(nw), float64 rho(nz,ncol), float64 T(nz,ncol)
!   output.bin (stream): 32 float64 constants, then per column
!     x, fn, e_scatter, hm_bf_factor, theta, lgtheta   (nz each)
!     and per wavelength: alpha, albedo, S, a, b, c, d (nz each), U (mz), V (nz)
program ref_dump
  use Common
  use GasState
  use ContinuousOpacity
  use Auxiliary
  implicit none
  type (pillar_case) :: p
  real, dimension(nz) :: a, b, c, d
  real, dimension(mz) :: U
  real, dimension(nw) :: waves
  real, allocatable :: rho(:,:), T(:,:)
  real :: dz, dz1, dz2, sigma_grey
  integer(4) :: ncol, lgrey_i
  integer :: icol, iw, iu, ou
  character(len=256) :: fin, fnml, fout
  character(len=90) :: nml90
  real, dimension(32) :: consts

  call get_command_argument(1, fin)
  call get_command_argument(2, fnml)
  call get_command_argument(3, fout)
  nml90 = trim(fnml)
  call read_gas_state_input(nml90)

  open(newunit=iu, file=trim(fin), form='unformatted', access='stream', status='old')
  read(iu) ncol, lgrey_i, dz, sigma_grey
  read(iu) waves
  allocate(rho(nz,ncol), T(nz,ncol))
  read(iu) rho
  read(iu) T
  close(iu)

  lfirst = .false.
  lroot = .true.
  call pre_calc_opacity_quantities(waves)
  dz1 = 1./dz
  dz2 = dz**2

  consts = 0.
  consts(1) = sigma_sb;  consts(2) = epsi;   consts(3) = log_overflow_limit
  consts(4) = mp;        consts(5) = mp1;    consts(6) = amu
  consts(7) = h_planck;  consts(8) = pi;     consts(9) = pi1
  consts(10) = sqrt3;    consts(11) = exp1;  consts(12) = k_cgs
  consts(13) = k_eV;     consts(14) = k1_cgs; consts(15) = k1_eV
  consts(16) = hydrogen_ionization_eV; consts(17) = c_light_cgs
  consts(18) = RydbergEnergy; consts(19) = RydbergEnergy1; consts(20) = log10e
  consts(21) = Rcm;      consts(22) = Iev;   consts(23) = Iev1
  consts(24) = bremsstrahlung_constant; consts(25) = hcR1
  consts(26) = one_third; consts(27) = two_thirds; consts(28) = dz1; consts(29) = dz2

  open(newunit=ou, file=trim(fout), form='unformatted', access='stream', status='replace')
  write(ou) consts
  do icol = 1, ncol
    p%rho = rho(:,icol)
    p%T = T(:,icol)
    call wavelength_independent_pillars(p)
    write(ou) p%NHII_NHINHII, p%ionization_factor, p%e_scatter, p%hm_bf_factor, p%theta, p%lgtheta
    do iw = 1, nw
      if (lgrey_i /= 0) then
        call grey_parameters(p, sigma_grey)
      else
        call calc_opacity_and_albedo(p, iw)
      endif
      call get_tridag_coefficients(a, b, c, d, p, dz, dz2)
      call tridag(a, b, c, d, U(n1:n2))
      call update_ghosts(U)
      call calc_flux(U, p, dz1)
      write(ou) p%opacity, p%albedo, p%source_function, a, b, c, d, U, p%flux
    enddo
  enddo
  close(ou)
endprogram ref_dump
