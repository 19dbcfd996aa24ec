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

/* Launchers (enqueue on `stream`, no synchronisation). */
hipError_t oco_launch_cells(const double* rho, const double* T, long ncells, double fully_ionized_T,
                            double* cells, hipStream_t stream);
hipError_t oco_launch_sweep(const oco_sweep_args& a, int lgrey, int nblocks, hipStream_t stream);
int oco_sweep_max_blocks(void);

/* failure key layout: system index << 32 | stage << 28 | row */
#define OCO_FAIL_NONE 0xffffffffffffffffULL

#endif
