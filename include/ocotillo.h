/*
 * ocotillo.h -- C ABI of the MI355X implementation of ocotillo's
 * per-column, per-wavelength hot path (Saha state, hydrogen continuous
 * opacity, Feautrier tridiagonal solve, 6th-order flux).
 *
 * The reference has no FFI: its hot path is the Fortran call sequence inside
 * the column loop of source/flux_feautrier.f90:160-228, with state in module
 * globals.  Each entry point below names the reference code it replaces.
 * Arrays are plain pointers in the reference's Fortran (column-major)
 * layouts; scalars are passed by value, so the Fortran host binds them with
 * bind(C) / value and the Python tools with ctypes.
 *
 * Threading: a context is used by one host thread at a time.  Every call is
 * synchronous unless its name ends in _async.  No global state but the
 * per-GPU static scratch of oco_solve_block_device_noalloc.
 */
#ifndef OCOTILLO_H
#define OCOTILLO_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  OCO_E_PIVOT0_STAGE1/2 reproduce the reference's two
 * "ERROR tridiag stage" stops (source/auxiliary.f90:27-30, 36-39). */
enum {
    OCO_OK = 0,
    OCO_E_PIVOT0_STAGE1 = 1,
    OCO_E_PIVOT0_STAGE2 = 2,
    OCO_E_ARG = 3,
    OCO_E_OOM = 4,
    OCO_E_HIP = 5,
    OCO_E_STATE = 6,
    OCO_E_NODEVICE = 7,
    OCO_E_BUSY = 8,
    OCO_E_IO = 9,
    OCO_E_NONFINITE = 10   /* an output is NaN or infinite (OCO_CHECK_FINITE) */
};

/* Precision modes of oco_solve_block*.  MIXED and ACCURATE solve the same
 * rows (source/auxiliary.f90:147-197) in difference form, which keeps the
 * row excess zeta that the reference's b = -((a + c) + zeta) rounds away in
 * optically thin cells, and form the emergent flux V(1), V(nz) from the
 * rows' interface differences instead of differences of U (:64-78, 96-109);
 * both are held to the exact solution of the reference's discrete problem.
 * Their columns take 4 to 1088 cells (batches and blocks; the server takes
 * any nz from 4 up to its max_nz): their emergent flux uses the 4 cells
 * next to each boundary.  The FP64 drop-in takes any nz >= 2 in a block
 * solve, as the reference does (nz = 1 leaves the reference's dz and its
 * ghost update undefined); OCO_E_ARG otherwise. */
enum {
    OCO_MODE_FP64 = 0,     /* drop-in: the reference's FP64 arithmetic, operation by operation */
    OCO_MODE_MIXED = 1,    /* FP32-core opacity with exponent scaling, FP64 Saha and elimination */
    OCO_MODE_ACCURATE = 2  /* the drop-in's FP64 opacity (alpha bit for bit), difference-form solve */
};

typedef struct oco_ctx oco_ctx;

/* Diagnostics of one block solve.  fail_* are the 1-based (ix, iy, iw, row)
 * of the first zero pivot in the reference's loop order; sum_* and min/max
 * are the first-pillar canaries the reference prints for (ix,iy,iw)=(1,1,1)
 * (source/flux_feautrier.f90:198-224); *_ms are device times of the kernels
 * (HIP events on the context's stream). */
typedef struct {
    int status;
    int fail_ix, fail_iy, fail_iw, fail_j;
    double sum_a, sum_b, sum_c, sum_d;
    double minU, maxU, minV, maxV, minA, maxA;
    double cells_ms, sweep_ms, total_ms;
} oco_diag;

/* Create a context on HIP device `device`, allowed to hold at most
 * `hbm_budget` bytes of device memory (0 = no limit).  Replaces nothing in
 * the reference (its state is module globals, source/common.f90:1-83). */
int oco_create(int device, size_t hbm_budget, oco_ctx** ctx);
int oco_destroy(oco_ctx* ctx);

/* Wavelength grid in Angstrom.  Replaces pre_calc_opacity_quantities
 * (source/continuous_opacity.f90:64-141): builds the per-wavelength tables
 * and keeps them resident on the device. */
int oco_set_wavelengths(oco_ctx* ctx, int nw, const double* waves_angstrom);

/* Numeric checks of the block solves (oco_solve_block, _device, _noalloc,
 * _files), off by default; also set by the environment variable
 * OCO_CHECK_FINITE=1 at oco_create.  OCO_CHECK_FINITE: after the solve, every
 * U, V and absorp value is checked, and a NaN or infinity anywhere turns the
 * result into OCO_E_NONFINITE with fail_ix, fail_iy, fail_iw of the first
 * such (column, wavelength) system in the reference's loop order and fail_j
 * its first such cell (1-based; its absorp first, then its U, whose ghosts
 * count as -2..0 and nz+1..nz+3, then its V); oco_solve_block_files then
 * writes no file.  The device counterpart of the reference's debug build
 * flags -ffpe-trap=invalid,zero,overflow (source/Makefile.source:8), which
 * stop the run at the first invalid, divide-by-zero or overflow operation;
 * it costs one read of the outputs (about 20 ms on a 256 x 256 x 512 block
 * of 128 wavelengths). */
#define OCO_CHECK_FINITE 1
int oco_set_checks(oco_ctx* ctx, int flags);

/* Replaces the &gas_state_input namelist value fully_ionized_T
 * (source/gas_state.f90:14-28) and the &input values lgrey, sigma_grey
 * (source/flux_feautrier.f90:51-58). */
int oco_set_physics(oco_ctx* ctx, double fully_ionized_T, int lgrey, double sigma_grey);

/* One processor block: replaces the column loop body
 * source/flux_feautrier.f90:160-228 (wavelength_independent_pillars,
 * calc_opacity_and_albedo / grey_parameters, get_tridag_coefficients,
 * tridag, update_ghosts, calc_flux and the stores at :215-216).
 * Host buffers, Fortran layouts as in source/flux_feautrier.f90:36-42:
 *   rho3d, T3d   (nz, nyloc, nxloc)        read
 *   U            (nz+6, nyloc, nxloc, nw)  written, 3 ghost cells each side
 *   V, absorp    (nz, nyloc, nxloc, nw)    written
 * dz, dz1, dz2 are the grid step and the 1/dz, dz**2 the reference sets on
 * the root block (:151-154).  Returns OCO_OK or a status code; *diag (may
 * be NULL) receives the details. */
int oco_solve_block(oco_ctx* ctx, int nz, int nyloc, int nxloc, double dz, double dz1, double dz2,
                    const double* rho3d, const double* T3d, double* U, double* V, double* absorp,
                    int mode, oco_diag* diag);

/* Same as oco_solve_block with every array already in device memory
 * (same layouts).  No host<->device copies except the diagnostics. */
int oco_solve_block_device(oco_ctx* ctx, int nz, int nyloc, int nxloc, double dz, double dz1,
                           double dz2, const double* d_rho3d, const double* d_T3d, double* d_U,
                           double* d_V, double* d_absorp, int mode, oco_diag* diag);

/* oco_solve_block_device with no device allocation at all, for callers
 * that cannot allocate: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  No counterpart in the reference.  Its
 * scratch (cell records, queue tickets, failure key, canaries) is one
 * static device array per GPU of OCO_STATIC_SCRATCH_MIB MiB, held by each
 * GPU the library runs on until the process ends: blocks of at most about
 * OCO_STATIC_SCRATCH_MIB * 2^20 / 184 cells (1.45M: 64 x 64 columns of 356
 * cells; 184 bytes of cell record each), else OCO_E_ARG.  Modes OCO_MODE_MIXED and OCO_MODE_ACCURATE only
 * (the drop-in's elimination scratch grows with the grid): OCO_E_ARG for
 * OCO_MODE_FP64.  The wavelength table must already be set
 * (oco_set_wavelengths, which allocates at set-up).  Every call on one GPU
 * uses that GPU's one array, whatever the context: calls that overlap on a
 * GPU (two host threads, two contexts' streams) race and must be
 * serialised by the caller.  Results equal oco_solve_block_device's bit for
 * bit. */
#define OCO_STATIC_SCRATCH_MIB 256
int oco_solve_block_device_noalloc(oco_ctx* ctx, int nz, int nyloc, int nxloc, double dz, double dz1,
                                   double dz2, const double* d_rho3d, const double* d_T3d, double* d_U,
                                   double* d_V, double* d_absorp, int mode, oco_diag* diag);

/* ---- Opacity query ----
 * Absorption coefficient alpha (1/cm), single-scattering albedo and source
 * function S of npts (rho, T) points at the context's wavelengths, in `mode`
 * precision: wavelength_independent_pillars then calc_opacity_and_albedo per
 * point (source/gas_state.f90:30-110, source/continuous_opacity.f90:143-341,
 * 392-409), or grey_parameters (:377-390) on a grey context.  This is the
 * query tools/ocotillo-py/solve_opacity.py:63-226 answers in Python, with the
 * Fortran code's semantics instead of the Python's (its divergences: the
 * Saha mass and threshold, k, the H bound-free sum to n = 5, the clamp of
 * negative H- bound-free, astropy constants, a division by zero when fully
 * ionised).  Host arrays; outputs (nw, npts) in Fortran order (wavelength
 * fastest); albedo and S may be NULL. */
int oco_opacity_points(oco_ctx* ctx, int npts, const double* rho, const double* T, int mode, double* alpha,
                       double* albedo, double* S);

/* ---- Block outputs ----
 * One block solved and written as the reference's three output files
 * (source/io.f90:76-101, output_binary):
 *   <dir>/mean_intensity_<snapshot>.bin           U       (nz+6, nyloc, nxloc, nw)
 *   <dir>/absorption_coefficients_<snapshot>.bin  absorp  (nz, nyloc, nxloc, nw)
 *   <dir>/flux_<snapshot>.bin                     V       (nz, nyloc, nxloc, nw)
 * each one Fortran unformatted sequential record in the layout of the
 * reference's compiler, gfortran (records over 2147483639 bytes as
 * subrecords with negated markers).  Replaces the column loop plus
 * output_binary (source/flux_feautrier.f90:160-228, 243) without holding
 * the outputs in host memory: the block is solved in wavelength chunks of
 * about `stage_bytes` (0: 512 MiB), and the solve of chunk k+1, the
 * device-to-host copy of chunk k and the file writes of chunk k-1 overlap.
 * The files appear under their names only when complete; on a zero pivot
 * (the reference stops before output_binary) none is written.  `dir` must
 * exist.  U_col (nz+6, nw) and A_col (nz, nw), may be NULL, receive the
 * column (ix, iy) = (nxloc, nyloc) the reference prints to diagnostics.txt
 * (source/flux_feautrier.f90:241).  diag as for oco_solve_block, except
 * sweep_ms (device time from the first chunk's solve to the last's) and
 * total_ms (host wall time of the whole call, files closed). */
int oco_solve_block_files(oco_ctx* ctx, int nz, int nyloc, int nxloc, double dz, double dz1, double dz2,
                          const double* rho3d, const double* T3d, int mode, const char* dir, const char* snapshot,
                          size_t stage_bytes, double* U_col, double* A_col, oco_diag* diag);

/* One array as a one-record Fortran unformatted sequential file, the
 * layout above (what `write(u) array` gives under gfortran,
 * source/io.f90:57-61, 71-73, 88-99).  No GPU involved. */
int oco_write_record(const char* path, const void* data, size_t nbytes);

/* ---- Athena input construction ----
 * The arithmetic that turns one Athena z slab (the file id<iproc>.<snap>.bin
 * of one iprocz: raw density, three momenta and total energy in code units,
 * each (nxloc, nyloc, nzl) in the file's order) into the block's rho3d, T3d
 * (nz, nyloc, nxloc) at z offset iz0 (0-based): replaces
 * source/read_athena.f90:163-205 (kinetic and internal energy, with the
 * momentum-underflow guard of :213-228), :232-258 (density unit and floor,
 * temperature unit) and the temperature rule source/disk.f90:96-104.  The
 * host reads the file and provides the units (read_athena_input,
 * source/read_athena.f90:33-62), switchTemp and the Gaussian profile
 * tgauss(z) = midTemp exp(-z^2/(2 sigma^2)) (source/disk.f90:28, 89). */
typedef struct {
    double unit_density, unit_temperature, gamma1, rho_floor, switch_temp, floor_temp;
    int avoid_momentum_underflow;   /* lavoid_momentum_underflow */
} oco_athena_units;

/* Host arrays: the slab's z range of rho3d, T3d is written, the rest kept. */
int oco_athena_slab(oco_ctx* ctx, int nz, int nyloc, int nxloc, int nzl, int iz0, const double* rho_raw,
                    const double* mx, const double* my, const double* mz, const double* e, const double* tgauss,
                    const oco_athena_units* units, double* rho3d, double* T3d);
/* Device arrays (tgauss too), enqueued and completed before return. */
int oco_athena_slab_device(oco_ctx* ctx, int nz, int nyloc, int nxloc, int nzl, int iz0, const double* d_rho_raw,
                           const double* d_mx, const double* d_my, const double* d_mz, const double* d_e,
                           const double* d_tgauss, const oco_athena_units* units, double* d_rho3d,
                           double* d_T3d);

/* ---- Batched uneven columns ----
 * Independent columns, each with its own nz, grid step dz and wavelength
 * grid, solved in one persistent launch.  Replaces running the reference
 * once per column shape (its shapes are compile-time, source/common.f90:9)
 * over the column loop body source/flux_feautrier.f90:160-228.
 * Layout: columns concatenated; rho, T hold sum(nz) values; column c's U is
 * (nz_c+6, nw_c) and its V, absorp (nz_c, nw_c), each Fortran order,
 * concatenated in column order (oco_batch_layout gives the totals): the
 * host arrays of oco_batch_run_host.
 * The plan keeps device-resident inputs and outputs (oco_batch_buffers);
 * on the device each system's row is padded to a multiple of 16 doubles
 * (one 128-byte line): column c's U is (ldu_c, nw_c) at u_off_c and its V,
 * absorp (ldva_c, nw_c) at va_off_c, ldu_c = 16 ceil((nz_c + 6) / 16),
 * ldva_c = 16 ceil(nz_c / 16) (oco_batch_column); rho and T are packed.
 * The padding lets the lane sweep write whole lines (4096 uneven columns,
 * FP64: 8.47 -> 7.26 ms per run on MI355X);
 * its first run autotunes the kernel and grid and captures the solve as a
 * hipGraph that later runs replay.  fail_ix in the diagnostics is the
 * 1-based column, fail_iw its wavelength. */
typedef struct oco_batch oco_batch;
int oco_batch_create(oco_ctx* ctx, int ncol, const int* nz, const int* nw, const double* waves_angstrom,
                     const double* dz, oco_batch** plan);
int oco_batch_destroy(oco_batch* plan);
int oco_batch_layout(oco_batch* plan, size_t* ncells, size_t* nsys, size_t* nU, size_t* nVA);
int oco_batch_buffers(oco_batch* plan, double** d_rho, double** d_T, double** d_U, double** d_V,
                      double** d_absorp);
/* Where column `col` (0-based) sits in the device outputs: offsets in
 * doubles and leading dimensions of its U and of its V and absorp. */
int oco_batch_column(oco_batch* plan, int col, size_t* u_off, int* ldu, size_t* va_off, int* ldva);
int oco_batch_run(oco_batch* plan, int mode, oco_diag* diag);
int oco_batch_run_host(oco_batch* plan, const double* rho, const double* T, double* U, double* V,
                       double* absorp, int mode, oco_diag* diag);
/* Autotuner outcome for a precision mode: kernel variant, workgroups, and
 * whether tuning ran (each mode is tuned and captured on its first run). */
int oco_batch_tuning(oco_batch* plan, int mode, int* variant, int* nblocks, int* tuned);
/* The autotuner's timings for a mode, in the order it tried the candidates
 * (median of three launches each, ms): FP64, the lane kernel's 3 variants x
 * 3 grids (variant-major; 2, 3, 4 wavefronts per SIMD); mixed, per column
 * class that holds columns, 1, 2, 4 wavefronts per workgroup (fewer where a
 * size does not fit).  Writes at most cap values and returns how many there
 * are (0 before the first run in the mode), or a negative OCO_E_ARG. */
int oco_batch_tune_times(oco_batch* plan, int mode, float* ms, int cap);

/* ---- Serving ----
 * A request loop over one device: jobs (a snapshot block or column set,
 * each with its own shape and wavelength grid) are submitted as they arrive
 * and return the emergent flux V(1), V(nz) of every (column, wavelength)
 * system.  The server admits queued jobs in arrival order while its HBM
 * budget and job slots allow, keeps wavelength tables resident, and feeds
 * one persistent launch per iteration with tiles of every admitted job
 * (continuous batching: a job that arrives while others run joins at the
 * next iteration, and its systems share launches with theirs).  Each system
 * is solved exactly as it would be alone.  No reference counterpart (the
 * reference runs one block per process, source/flux_feautrier.f90:160-228);
 * the physics per system is the column loop body, in `mode` precision. */
typedef struct oco_server oco_server;

typedef struct {
    int nx, ny, nz, nw;     /* nxloc, nyloc, cells per column, wavelengths */
    double dz;              /* grid step (1/dz and dz**2 derived as the reference does) */
    const double* rho3d;    /* (nz, ny, nx) Fortran order; read until the job completes */
    const double* T3d;      /* (nz, ny, nx) */
    const double* waves;    /* nw wavelengths in Angstrom */
    double* emergent;       /* out: (2, ny, nx, nw): V(1), V(nz) of every system */
} oco_job;

typedef struct {
    int status;             /* OCO_OK or a pivot status (FP64 mode) */
    int fail_ix, fail_iy, fail_iw, fail_j;
    double queue_ms;        /* submit -> admitted */
    double latency_ms;      /* submit -> outputs in host memory */
    double points;          /* nz*nx*ny*nw */
} oco_job_result;

typedef struct {
    long long submitted, completed, queued, active;
    double p50_ms, p99_ms, mean_ms, max_ms;   /* latency of completed jobs */
    double points;                            /* points of completed jobs */
    double wall_s;                            /* first submit -> last completion */
    double points_per_s;
    long long iterations;                     /* persistent launches */
    size_t hbm_budget, hbm_peak;              /* job arena: budget and high-water mark */
} oco_serve_stats;

/* hbm_budget: bytes of device memory for admitted jobs and resident tables;
 * max_nz: longest column accepted; mode: OCO_MODE_FP64 or OCO_MODE_MIXED. */
int oco_server_create(oco_ctx* ctx, size_t hbm_budget, int max_nz, int mode, oco_server** srv);
/* Queue a job; returns at once with a ticket.  OCO_E_OOM if the job alone
 * exceeds the budget. */
int oco_server_submit(oco_server* srv, const oco_job* job, long long* ticket);
/* Block until the job has completed (outputs written); then forget it. */
int oco_server_wait(oco_server* srv, long long ticket, oco_job_result* res);
/* OCO_OK and the result if the job has completed (then forgotten),
 * OCO_E_BUSY if not yet. */
int oco_server_poll(oco_server* srv, long long ticket, oco_job_result* res);
int oco_server_stats(oco_server* srv, oco_serve_stats* st);
/* Finish every submitted job, then release the server. */
int oco_server_destroy(oco_server* srv);

/* Message text for a status code. */
const char* oco_strerror(int status);

/* Library version string. */
const char* oco_version(void);

#ifdef __cplusplus
}
#endif
#endif
