"""Single-point opacity query on the GPU: the command line of
tools/ocotillo-py/solve_opacity.py (same options, same table), answered by
the library's opacity code (oco_opacity_points) with the Fortran program's
semantics.

    python -m ocotillo_amd.opacity -d 1e-9 -t 5000 -n 5 --lambdaMin 3000 --lambdaMax 10000

The Python tool diverges from the Fortran it accompanies; this query follows
the Fortran (source/gas_state.f90, source/continuous_opacity.f90):
  * Saha: the proton mass m_p (the tool uses the atomic mass unit) and the
    fully_ionized_T threshold of &gas_state_input, 2e4 K by default (the
    tool hard-codes 20000 K); --fully-ionized-T sets it;
  * k = 1.380649e-16 erg/K (the tool: 1.380658e-16 in places), and the
    Fortran's constants where the tool takes astropy's;
  * H bound-free summed over n = 1..6 (the tool: 1..5);
  * negative H- bound-free values are kept (the tool clamps them to 0);
  * fully ionised gas (no neutral hydrogen) gives f_n = 0 (the tool divides
    by zero);
  * the wavelength grid is calc_wavelength's w0 + (i-1) dw
    (source/continuous_opacity.f90:37-62), not numpy.linspace;
  * the grey line is grey_parameters' (alpha_grey + sigma_grey) rho
    (:377-390) with sigma_grey = 0.4, as the tool prints it.
Both share the reference's free-free quirk (the neutral branch adds the
updated kappa rho to alpha).  --mixed answers in the mixed-precision mode.
There is no CPU path: without a GPU the query fails.
"""
import argparse
import sys

import numpy as np


def wavelengths(w0, w1, nw):
    """calc_wavelength (source/continuous_opacity.f90:37-62)."""
    if nw == 1:
        return np.array([float(w0)])
    dw = (w1 - w0) / (nw - 1)
    return np.array([w0 + i * dw for i in range(nw)])


def query(rho, T, waves, mode=None, fully_ionized_T=2e4, device=0):
    """alpha (1/cm), albedo, S of (rho, T) points at `waves` (Angstrom):
    arrays [npts, nw]."""
    import ocotillo_amd as oco
    mode = oco.MODE_FP64 if mode is None else mode
    with oco.Context(device) as ctx:
        ctx.set_wavelengths(waves)
        ctx.set_physics(fully_ionized_T, False, 0.4)
        return ctx.opacity(rho, T, mode)


def grey(rho, T, sigma_grey=0.4, device=0):
    """grey_parameters' absorption coefficient of (rho, T) points."""
    import ocotillo_amd as oco
    with oco.Context(device) as ctx:
        ctx.set_wavelengths(np.array([1.0]))
        ctx.set_physics(2e4, True, sigma_grey)
        return ctx.opacity(rho, T)[0][:, 0]


def parse(argv=None):
    p = argparse.ArgumentParser(prog="python -m ocotillo_amd.opacity",
                                description="Continuous absorption coefficient of one (density, temperature) "
                                            "point (GPU, Fortran semantics)")
    p.add_argument("-d", "--density", type=np.float64, default=1e-9,
                   help="Gas density in cgs units (g cm^-3). Default: 1e-9 g cm^-3")
    p.add_argument("-t", "--temperature", type=np.float64, default=1000, help="Gas temperature in Kelvin")
    p.add_argument("-n", "--nlambda", type=int, default=2,
                   help="Integer number of wavelengths to be used. Default: nlambda = 2")
    p.add_argument("--lambdaMin", type=np.float64, default=3e3,
                   help="Minimum wavelength in Angstroms. Default: lambdaMin = 3,000 A")
    p.add_argument("--lambdaMax", type=np.float64, default=1e4,
                   help="Maximum wavelength in Angstroms. Default: lambdaMax = 10,000 A")
    p.add_argument("-v", "--verbose", action="store_true",
                   help="If set, also print the albedo and the source function")
    p.add_argument("--fully-ionized-T", type=np.float64, default=2e4,
                   help="fully_ionized_T of &gas_state_input (K). Default: 2e4")
    p.add_argument("--mixed", action="store_true", help="mixed-precision mode instead of FP64")
    p.add_argument("--device", type=int, default=0)
    args = p.parse_args(argv)
    if args.nlambda < 1:
        p.error("nlambda must be at least 1")
    return args


def main(argv=None):
    import ocotillo_amd as oco
    args = parse(argv)
    waves = wavelengths(float(args.lambdaMin), float(args.lambdaMax), args.nlambda)
    print(f'Calculating for {args.nlambda} wavelength bins over the range '
          f'[{args.lambdaMin:.0f}, {args.lambdaMax:.0f}] Angstroms.')
    rho, T = float(args.density), float(args.temperature)
    try:
        alpha, albedo, S = query(np.array([rho]), np.array([T]), waves,
                                 oco.MODE_MIXED if args.mixed else oco.MODE_FP64, float(args.fully_ionized_T),
                                 args.device)
        ag = grey(np.array([rho]), np.array([T]), device=args.device)[0]
    except oco.OcoError as e:
        print("ocotillo GPU library: %s" % e, file=sys.stderr)
        return 2
    print("")
    print("Gas thermal properties:")
    print(f"Temperature: {T:.2f} K")
    print(f"Density: {rho:.2e} g cm^-3")
    print("")
    print("Continuous absorption coefficients by wavelength:")
    print(f"{'Wavelength (A)':>14s} {'Absorp. Coeff. (1/cm)':>25s}")
    for iw in range(waves.size):
        print(f"{waves[iw]:14.2f} {alpha[0, iw]:25.15e}")
    print(f"{'Grey':>14s} {ag:25.15e}")
    if args.verbose:
        print("")
        print(f"{'Wavelength (A)':>14s} {'Albedo':>25s} {'Source function':>25s}")
        for iw in range(waves.size):
            print(f"{waves[iw]:14.2f} {albedo[0, iw]:25.15e} {S[0, iw]:25.15e}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
