/* Internal interface between the C-ABI layer (oco_api.cpp) and the kernels
 * (oco_kernels.hip). */
#ifndef OCO_INTERNAL_H
#define OCO_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ocotillo.h"

struct oco_sweep_args {
    int nz;              /* cells per column */
    int nw;              /* wavelengths */
    long ncol;           /* columns in the block (nxloc*nyloc) */
    double dz, dz1, dz2;
    double sigma_grey;
    double fully_ionized_T;
    const double* rho;   /* (nz, ncol) */
    const double* T;     /* (nz, ncol) */
    const double* cells; /* cell records: (ncol*nz, OCO_CELL_STRIDE), or field-major for the wave sweep */
    const double* wtab;  /* (nw, OCO_WT_STRIDE) */
    double* U;           /* (nz+6, ncol, nw) */
    double* V;           /* (nz, ncol, nw) */
    double* A;           /* (nz, ncol, nw) */
    double* scratch;     /* per-block slots of nz*3*64 doubles */
    unsigned long long* fail;  /* first zero pivot key (atomicMin) */
    double* canary;      /* 4 + 5*nz doubles: sum(a..d) of system 0, then its point values */
    unsigned long long* tickets;  /* wavefront sweep: OCO_TICKET_QUEUES counters, OCO_TICKET_STRIDE apart */
};

/* Work queues of the wavefront sweep: one ticket counter per XCD, each in a
 * 256-byte line of its own. */
#define OCO_TICKET_QUEUES 8
#define OCO_TICKET_STRIDE 32
#define OCO_TICKET_BYTES (OCO_TICKET_QUEUES * OCO_TICKET_STRIDE * 8)

/* Leading dimensions of a batched column's device outputs: U rows of nz + 6
 * doubles and V, absorp rows of nz, each rounded up to 16 doubles, so every
 * system's row starts on a 128-byte line and the lane sweep's output runs
 * are whole lines */
__host__ __device__ static inline int oco_batch_ldu(int nz) { return (nz + 6 + 15) & ~15; }
__host__ __device__ static inline int oco_batch_ldv(int nz) { return (nz + 15) & ~15; }
/* wavefronts per workgroup of the FP64 batch kernel, at most (its grid is
 * the tuned wavefront count rounded up to whole workgroups) */
#define OCO_BATCH_WG_MAX 4

/* Batched uneven columns (per-column shapes and wavelength grids). */
struct oco_batch_args {
    int ntiles;
    const int* tiles;      /* ntiles x (column, first wavelength) */
    const int* nz;         /* per column */
    const int* nw;
    const double* dz;      /* per column: dz, 1/dz, dz**2 */
    const long* cell_off;  /* per column: first cell (inputs and cell records) */
    const long* wt_off;    /* per column: first wavelength-table row (= first system) */
    const long* u_off;     /* per column: offset of U (nz+6, nw) */
    const long* v_off;     /* per column: offset of V and alpha (nz, nw) */
    const double* cells;
    const double* rho;
    const double* T;
    const double* wtab;
    double* U;
    double* V;
    double* A;
    double* scratch;
    long scr_stride;       /* doubles per wavefront slot: max_nz * 3 * 64 */
    unsigned int* counter;
    unsigned long long* fail;
    /* mixed mode, wavefront-per-system kernel: items (column, first of the
     * group's wavelengths) in OCO_TICKET_QUEUES queues, queue q holding
     * items [wq_off[q], wq_off[q + 1]); records field-major per column
     * (field k of cell z of column c at cell_off[c]*CR_NFIELD + k*nz[c] + z) */
    const int* witems;
    const int* wq_off;
    unsigned long long* wtickets;   /* OCO_TICKET_QUEUES counters, OCO_TICKET_STRIDE apart */
    const double* cells_soa;
    int max_nz;
};

/* Serving: one admitted job's device descriptor. */
struct oco_serve_job {
    const double* cells;   /* cell records: FP64 mode cell-major (ncol*nz, OCO_CELL_STRIDE); mixed mode
                            * chunk-major per column (oco_launch_cells_chunked) */
    const double* wtab;    /* (nw, OCO_WT_STRIDE) */
    double* em;            /* emergent V: (2, ncol, nw) -- V(1), V(nz) per system */
    long ncol;
    int nz, nw;
    double dz, dz1, dz2;
};

struct oco_serve_args {
    int ntiles;
    const int* tiles;             /* ntiles x (job slot, column, first wavelength) */
    const oco_serve_job* jobs;    /* per slot */
    double* scratch;
    long scr_stride;              /* doubles per wavefront slot: max_nz * 3 * 64 */
    unsigned int* counter;
    unsigned long long* fail;     /* per slot: first zero pivot key (FP64 mode) */
    int win_rows;                 /* mixed mode: boundary-window rows per system (oco_serve_win_rows(max_nz)) */
};

/* One-record Fortran unformatted sequential file in gfortran's layout
 * (oco_output.cpp): created as <path>.partial, data written at any record
 * offset, renamed to <path> by close(keep=1), removed by close(keep=0). */
struct oco_record_file;
int oco_record_open(const char* path, unsigned long long nbytes, oco_record_file** f);
int oco_record_write(oco_record_file* f, unsigned long long off, const void* data, size_t n);
int oco_record_close(oco_record_file* f, int keep);

/* Opacity query of (point, wavelength) pairs (see opacity_points). */
struct oco_points_args {
    int npts, nw;
    const double* rho;    /* (npts) */
    const double* T;
    const double* cells;  /* cell records of the points (non-grey) */
    const double* wtab;   /* (nw, OCO_WT_STRIDE) */
    double sigma_grey;
    double* alpha;        /* (nw, npts) Fortran order: wavelength fastest */
    double* albedo;       /* may be null */
    double* S;            /* may be null */
};
hipError_t oco_launch_points(const oco_points_args& a, int lgrey, int mixed, hipStream_t stream);

/* Athena slab to the block's rho, T (see athena_slab in oco_kernels.hip). */
struct oco_athena_args {
    int nz, nyloc, nxloc, nzl, iz0;
    const double *rho, *mx, *my, *mz, *e;   /* slab fields, file order (nxloc, nyloc, nzl) */
    const double* tgauss;                     /* (nz) */
    double unit_density, unit_temperature, gamma1, rho_floor, switch_temp, floor_temp;
    int avoid_momentum_underflow;
    double* rho3d;                            /* (nz, nyloc, nxloc) */
    double* T3d;
};
hipError_t oco_launch_athena(const oco_athena_args& a, hipStream_t stream);

/* Launchers (enqueue on `stream`, no synchronisation). */
hipError_t oco_launch_serve(const oco_serve_args& a, int mixed, int nblocks, hipStream_t stream);
/* mixed-mode serving records: column c's cell z = l*m + j (m = rows per lane
 * of the wavefront solve, l the lane, j the row in its chunk) has field k at
 * c*(m*CR_NFIELD*64) + (j*CR_NFIELD + k)*64 + l, so the 64 lanes of a
 * wavefront load one 512-byte line per field, each lane from its own chunk */
hipError_t oco_launch_cells_chunked(const double* rho, const double* T, long ncells, int nz, double fully_ionized_T,
                                    double* cells, hipStream_t stream);
size_t oco_cells_chunked_doubles(long ncol, int nz);
int oco_serve_win_rows(int max_nz);
int oco_serve_grid(int mixed, int max_nz, int device);
hipError_t oco_launch_fail_reset(unsigned long long* fail, int n, hipStream_t stream);
/* Non-finite scan of one output array of a block solve, laid out (inner,
 * nyloc, nxloc, nw_run) with its first wavelength iw0 of nw: the first NaN
 * or infinity in the reference's loop order as the key
 * ((system * 3 + which) * (nz + 6) + e), system = (ix * nyloc + iy) * nw + iw,
 * atomicMin into *key (OCO_FAIL_NONE when none); which: 0 absorp, 1 U, 2 V. */
hipError_t oco_launch_finite_scan(const double* p, long n, int inner, int which, int nyloc, int nxloc, int iw0, int nw,
                                  int nz, unsigned long long* key, hipStream_t stream);
/* mixed != 0 selects the mixed-precision kernels (non-grey only). */
/* the FP64 lane-per-system batch kernel, variant 0..2, on nblocks
 * wavefronts (rounded up to whole workgroups of OCO_BATCH_WG_MAX at most) */
hipError_t oco_launch_batch(const oco_batch_args& a, int variant, int nblocks, hipStream_t stream);
/* batched mixed mode, difference form: the per-column field-major records
 * (and the failure key), then the sweep of one class of columns (its items,
 * tickets and longest column in a) with k (1, 2, 4) wavefronts per workgroup
 * and `lanes` (8, 16, 32, 64) lanes per system */
hipError_t oco_launch_batch_wave_records(const oco_batch_args& a, int ncol, double fully_ionized_T,
                                         hipStream_t stream);
hipError_t oco_launch_batch_wave(const oco_batch_args& a, int k, int lanes, int nblocks, hipStream_t stream);
/* resident workgroups of that sweep for columns of at most max_nz cells; 0 if
 * such columns do not fit it */
int oco_batch_wave_grid(int max_nz, int k, int lanes, int device);
int oco_batch_variants(void);
/* soa_nz > 0: field-major records per column of soa_nz cells (field k of
 * cell z of column c at (c*CR_NFIELD + k)*soa_nz + z); 0: cell-major */
hipError_t oco_launch_cells(const double* rho, const double* T, long ncells, double fully_ionized_T,
                            double* cells, int soa_nz, hipStream_t stream);
/* wave == 0: the FP64 drop-in's lane-per-system sweep; wave = k (1, 2 or 4):
 * the difference-form sweep with k wavefronts per workgroup and `lanes` (8,
 * 16, 32, 64) lanes per system on points of kind pk (OCO_PK_*), which reads
 * field-major cell records (grey: rho and T) and needs no scratch */
hipError_t oco_launch_sweep(const oco_sweep_args& a, int lgrey, int wave, int lanes, int pk, int nblocks,
                            hipStream_t stream);
/* point kinds of the difference-form sweep */
enum { OCO_PK_MIXED = 0, OCO_PK_FP64 = 1, OCO_PK_GREY = 2 };
/* lanes per system for nz cells (want: 8, 16, 32, 64, or 0 for the default) */
int oco_wave_lanes(int nz, int want);
int oco_wave_fits(int nz, int lanes, int pk, int device);
/* wavefronts per workgroup of the difference-form sweep (want: 1, 2, 4, or 0 for the default) */
int oco_wave_group(int nz, int lanes, int pk, int want, int device);
int oco_wave_grid(int nz, int k, int lanes, int pk, int device);
/* workgroups resident on the device, and wavefronts per workgroup, of the
 * FP64 drop-in's lane-per-system sweep */
int oco_sweep_grid(int lgrey, int uniform, int device);
int oco_sweep_wg(void);

/* host helpers shared by the C-ABI files */
#ifdef __cplusplus
#include <vector>
struct oco_ctx;
void oco_build_wtab(const std::vector<double>& waves, std::vector<double>& wt);
int oco_ctx_device(const oco_ctx* c);
double oco_ctx_fully_ionized_T(const oco_ctx* c);
#endif

/* The allocation-free block solve's scratch: this many bytes per GPU in a
 * file-scope device array; oco_static_scratch gives the calling device's
 * address (looked up once per device, with that device current). */
#define OCO_STATIC_SCRATCH_BYTES ((size_t)OCO_STATIC_SCRATCH_MIB << 20)
int oco_static_scratch(int device, void** ptr, size_t* bytes);
hipError_t oco_launch_pillar_minmax(const double* U, const double* V, const double* A, int nz, double* out,
                                    hipStream_t stream);

/* failure key layout: system index << 32 | stage << 28 | row */
#define OCO_FAIL_NONE 0xffffffffffffffffULL

#endif
