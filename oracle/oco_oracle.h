/*
 * oco_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement (plain C, FP64) of ocotillo's per-column, per-wavelength
 * hot path, used as the checker for the HIP path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product never links or calls it.
 *
 * Every routine follows the reference's operation order (Fortran evaluates
 * left to right; no contraction, no reassociation).  Build with
 * -ffp-contract=off.  Two builds exist:
 *   liboco_oracle.so     elementary functions from glibc libm, as the
 *                        reference (gfortran/flang -> libm) uses them;
 *   liboco_oracle_pm.so  elementary functions from the product's portable
 *                        header ocotillo_amd/csrc/oco_math.h, which the GPU
 *                        kernels also use, so the GPU must match this build
 *                        bit for bit.
 */
#ifndef OCO_ORACLE_H
#define OCO_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OO_MDIM 6

/* Wavelength tables: source/continuous_opacity.f90:20-33 (module globals),
 * filled by pre_calc_opacity_quantities (:64-141). */
typedef struct {
    int nw;
    double *wa, *wcm1, *nu, *chi, *chi1, *wa3, *w1cm5, *gff, *s_coeff;
    double *f1, *f2, *f3;           /* f_coeff(nw,1..3) */
    double *g;                      /* g(nw,6) stored [iw*6 + n-1] */
    double n1[OO_MDIM], chi_n[OO_MDIM], chi_m;
    double damping_factor_const, source_function_const;
} oo_tables;

/* Per-cell wavelength-independent state: source/gas_state.f90:30-46. */
typedef struct {
    double rho, rho1, T, T1, x, n, ninv, nHI, nHII, ne, fn, Pe, es;
    double theta, theta1, lgtheta, lgtheta2, hmbf;
} oo_cell;

/* Per-point output of calc_opacity_and_albedo / grey_parameters. */
typedef struct {
    double alpha, albedo, alpha1, alpha2, S, stim;
} oo_point;

int  oo_tables_init(oo_tables* t, int nw, const double* waves_angstrom);
void oo_tables_free(oo_tables* t);
void oo_calc_wavelength(double w1, double w0, int nw, double* wa);

void oo_cell_state(double rho, double T, double fully_ionized_T, oo_cell* c);
void oo_opacity_point(const oo_tables* t, const oo_cell* c, int iw, oo_point* p);
void oo_grey_point(const oo_cell* c, double sigma_grey, oo_point* p);

/* Coefficients (source/auxiliary.f90:111-201) for one system. */
void oo_tridag_coefficients(int nz, const double* alpha, const double* alpha1,
                            const double* alpha2, const double* albedo, const double* S,
                            double dz, double dz2, double* a, double* b, double* c, double* d);
/* Thomas solve (source/auxiliary.f90:14-47).  Returns 0, or 1 / 2 for a zero
 * pivot in stage 1 / 2, with *jfail the 1-based row. */
int  oo_tridag(int n, const double* a, const double* b, const double* c, const double* r,
               double* u, double* gam, int* jfail);
/* Ghosts + 6th-order derivative + flux (source/auxiliary.f90:49-78,96-109). */
void oo_ghosts_flux(int nz, double* U /* mz */, const double* alpha1, double dz1, double* V);

/* Diagnostics of one block (first failure in reference loop order, and the
 * first-pillar canaries of source/flux_feautrier.f90:198-224). */
typedef struct {
    int status;          /* 0 ok, 1/2 zero pivot stage 1/2 */
    int fail_ix, fail_iy, fail_iw, fail_j;   /* 1-based, reference loop order */
    double sum_a, sum_b, sum_c, sum_d;
    double minU, maxU, minV, maxV, minA, maxA;
} oo_diag;

/* One block: rho3d,T3d (nz,nyloc,nxloc) Fortran order; outputs
 * U(mz,nyloc,nxloc,nw), V and alpha (nz,nyloc,nxloc,nw).  dz1, dz2 as the
 * reference sets them on the root block (source/flux_feautrier.f90:151-154).
 * nthreads <= 0 means all OpenMP threads. */
int oo_solve_block(int nz, int nyloc, int nxloc, const oo_tables* t,
                   double dz, double dz1, double dz2,
                   const double* rho3d, const double* T3d,
                   double fully_ionized_T, int lgrey, double sigma_grey,
                   double* U, double* V, double* alpha, oo_diag* diag, int nthreads);
/* As oo_solve_block with rh != 0: rows solved in difference form
 * (oo_rh_coefficients, oo_tridag_rh) -- the accurate solution of the same
 * discrete problem, the reference for the mixed mode's elimination. */
int oo_solve_block_ex(int nz, int nyloc, int nxloc, const oo_tables* t, double dz, double dz1, double dz2,
                      const double* rho3d, const double* T3d, double fully_ionized_T, int lgrey,
                      double sigma_grey, double* U, double* V, double* alpha, oo_diag* diag, int nthreads,
                      int rh);
/* rh == 2: every system solved exactly (oo_exact_system). */
/* The reference's rows, Thomas, ghosts, der and flux of one system evaluated
 * in binary128 from FP64 alpha, albedo, S: the exact solution of the
 * reference's discrete problem, rounded once to FP64.  U has nz + 6 values. */
void oo_exact_system(int nz, const double* al, const double* om, const double* S, double dz, double dz1,
                     double dz2, double* U, double* V);
void oo_rh_coefficients(int nz, const double* al, const double* om, const double* S, double dz, double dz2,
                        double* A, double* C, double* H, double* R);
void oo_tridag_rh(int n, const double* A, const double* C, const double* H, const double* R, double* u,
                  double* Dw);

/* Analytic disk input: source/disk.f90:42-129. */
void oo_calc_grid(int nz, double z1, double z0, double* z /* mz */, double* dz);
void oo_calc_density(int nz, const double* z, double rho0, double H, double rho_floor, double* rho);
void oo_calc_temperature(int nz, const double* z, double midTemp, double sigma, double floorTemp,
                         double switchTemp, const double* Tgauss_or_null, int from_athena, double* T);

/* Athena snapshot input: source/read_athena.f90:33-268.
 * oo_athena_units: read_athena_input's code units (:42-56) -> unit_length,
 * unit_density, unit_temperature.
 * oo_athena_slab_raw: one slab file's fields (file order, (nxloc, nyloc,
 * nzl)) into the block's code-unit rho, temp (nz, nyloc, nxloc) at z rows
 * iz0..iz0+nzl-1 (0-based), :163-205 and :213-228.
 * oo_athena_postprocess: the whole block's unit conversion, density floor and
 * temperature rule, :232-268 with source/disk.f90:96-104 (tgauss = the
 * Gaussian profile of the first calc_temperature call). */
void oo_athena_units(double Mbh_SolarMasses, double r0ref_rg, double aspect_ratio, double mean_molecular_weight,
                     double rho0, double* unit_length, double* unit_density, double* unit_temperature);
void oo_athena_slab_raw(int nz, int nyloc, int nxloc, int nzl, int iz0, const double* f_rho, const double* f_mx,
                        const double* f_my, const double* f_mz, const double* f_e, int avoid_momentum_underflow,
                        double* rho, double* temp);
void oo_athena_postprocess(int nz, int nyloc, int nxloc, double unit_density, double rho_floor, double gamma1,
                           double unit_temperature, const double* tgauss, double switchTemp, double floorTemp,
                           double* rho, double* temp);

#ifdef __cplusplus
}
#endif
#endif
