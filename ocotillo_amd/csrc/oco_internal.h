/* Internal interface between the C-ABI layer (oco_api.cpp) and the kernels
 * (oco_kernels.hip). */
#ifndef OCO_INTERNAL_H
#define OCO_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

struct oco_sweep_args {
    int nz;              /* cells per column */
    int nw;              /* wavelengths */
    long ncol;           /* columns in the block (nxloc*nyloc) */
    double dz, dz1, dz2;
    double sigma_grey;
    const double* rho;   /* (nz, ncol) */
    const double* T;     /* (nz, ncol) */
    const double* cells; /* cell records, (ncol*nz, OCO_CELL_STRIDE) */
    const double* wtab;  /* (nw, OCO_WT_STRIDE) */
    double* U;           /* (nz+6, ncol, nw) */
    double* V;           /* (nz, ncol, nw) */
    double* A;           /* (nz, ncol, nw) */
    double* scratch;     /* per-block slots of nz*3*64 doubles */
    unsigned long long* fail;  /* first zero pivot key (atomicMin) */
    double* canary;      /* 4 + 5*nz doubles: sum(a..d) of system 0, then its point values */
};

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
};

/* Launchers (enqueue on `stream`, no synchronisation). */
/* mixed != 0 selects the mixed-precision kernels (non-grey only). */
hipError_t oco_launch_batch(const oco_batch_args& a, int mixed, int variant, int nblocks, hipStream_t stream);
int oco_batch_variants(void);
hipError_t oco_launch_cells(const double* rho, const double* T, long ncells, double fully_ionized_T,
                            double* cells, hipStream_t stream);
hipError_t oco_launch_sweep(const oco_sweep_args& a, int lgrey, int mixed, int nblocks, hipStream_t stream);
int oco_sweep_grid(int lgrey, int mixed, int uniform, int device);

/* failure key layout: system index << 32 | stage << 28 | row */
#define OCO_FAIL_NONE 0xffffffffffffffffULL

#endif
