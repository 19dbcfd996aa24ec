/*
 * oco_kernels.hip -- gfx950 kernels of the FP64 drop-in path.
 *
 *   cells_fp64  one lane per cell: Saha state, number densities and every
 *               temperature-only factor of the opacity (the reference's
 *               wavelength_independent_pillars plus the T-only parts of
 *               calc_opacity_and_albedo), stored as a 192-byte record.
 *   sweep_fp64  one lane per (column, wavelength) system; a wavefront takes
 *               64 consecutive systems (wavelength fastest, so for nw >= 64
 *               all lanes share one column and its records are wave-uniform
 *               scalar loads).  Forward pass: opacity, albedo and source
 *               function recomputed in registers for cell i+1, rows of the
 *               tridiagonal system, Thomas elimination; gam, u and 1/alpha
 *               go to a per-wavefront scratch slot laid out [z][3][lane] so
 *               every access is one coalesced 512-byte line.  Backward pass:
 *               back-substitution fused with the ghost extrapolation, the
 *               6th-order derivative (7-point register window) and V.
 *
 * The FP64 path is compiled with -ffp-contract=off and its elementary
 * functions are oco_math.h's, read from an LDS copy of their tables.
 * sweep_mixed / batch_mixed are the same kernels with the mixed-precision
 * opacity of oco_mixed.h (OCO_MODE_MIXED).
 */
#include "oco_internal.h"
#include "oco_physics.h"
#include "oco_mixed.h"

__constant__ oco_mtab oco_mtab_dev = OCO_MTAB_INIT;

#define OCO_WAVE 64
#ifndef OCO_BACK_CHUNK
#define OCO_BACK_CHUNK 2   /* back-substitution rows whose scratch loads are batched */
#endif

__device__ static inline void load_mtab(oco_mtab* smt)
{
    const double* src = (const double*)&oco_mtab_dev;
    double* dst = (double*)smt;
    for (int i = threadIdx.x; i < (int)(sizeof(oco_mtab) / sizeof(double)); i += blockDim.x) dst[i] = src[i];
    __syncthreads();
}

__global__ __launch_bounds__(256) void cells_fp64(const double* __restrict__ rho, const double* __restrict__ T,
                                                  long ncells, double fully_ionized_T, double* __restrict__ cells)
{
    __shared__ oco_mtab smt;
    load_mtab(&smt);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < ncells; i += (long)gridDim.x * blockDim.x) {
        double rec[CR_NFIELD];
        oco_cell_record(rho[i], T[i], fully_ionized_T, &smt, rec);
        double* out = cells + i * OCO_CELL_STRIDE;
#pragma unroll
        for (int k = 0; k < CR_NFIELD; k++) out[k] = rec[k];
    }
}

/* Cell records are read-only during the sweep.  When a wavefront's lanes
 * share one column the record address is wave-uniform; reading it through
 * the constant address space turns the loads into scalar loads (one fetch
 * per wavefront instead of 64 lanes of vector loads). */
typedef const __attribute__((address_space(4))) double* oco_cptr;

template <bool UNIFORM>
__device__ static inline void load_rec(const double* cells, int z, double* r)
{
    if (UNIFORM) {
        oco_cptr p = (oco_cptr)(cells + (long)z * OCO_CELL_STRIDE);
#pragma unroll
        for (int k = 0; k < CR_NFIELD; k++) r[k] = p[k];
    } else {
        const double* p = cells + (long)z * OCO_CELL_STRIDE;
#pragma unroll
        for (int k = 0; k < CR_NFIELD; k++) r[k] = p[k];
    }
}

template <bool UNIFORM>
__device__ static inline void load_record(const oco_sweep_args& a, long col, int z, double* r)
{
    load_rec<UNIFORM>(a.cells + col * a.nz * OCO_CELL_STRIDE, z, r);
}

/* One (column, wavelength) system as the solver sees it. */
struct oco_sys {
    int nz;
    double dz, dz1, dz2, sigma_grey;
    const double* cells;   /* the column's cell records */
    const double* rho;     /* the column's rho, T (grey case) */
    const double* T;
    double* U;             /* this system's U (nz+6), V (nz), alpha (nz) */
    double* V;
    double* A;
    long key;              /* position in the reference's loop order */
};

__device__ static inline void record_fail(unsigned long long* fail, long sys, int stage, int row)
{
    unsigned long long key = ((unsigned long long)sys << 32) | ((unsigned long long)stage << 28) |
                             (unsigned long long)row;
    atomicMin(fail, key);
}

/* x / y: the IEEE divide of the FP64 path, or the reciprocal-based divide of
 * the mixed path (MIX is a compile-time switch; the FP64 path compiles to
 * exactly the reference's operations). */
template <bool MIX>
__device__ static inline double qdiv(double x, double y)
{
    if constexpr (MIX) return mx_div(x, y);
    else return x / y;
}

/* ---- coalesced output stores ----
 * A lane owns one system, and its outputs U(:), V(:), alpha(:) are contiguous
 * runs, but consecutive lanes' runs lie (nz+6)*ncol doubles apart in the
 * reference's (z, column, wavelength) layout, so a plain per-row store sends
 * 64 lines per instruction.  Rows are staged in LDS instead, R per lane, and
 * every R rows the wavefront writes them out cooperatively: each store
 * instruction then covers 64/R whole runs of R consecutive doubles.
 * Every lane of the wavefront takes part in a flush, so lanes without a
 * system of their own run a duplicate one with their stores switched off. */
#ifndef OCO_STAGE_ROWS
#define OCO_STAGE_ROWS 8
#endif
struct oco_stage {
    double ub[OCO_WAVE * (OCO_STAGE_ROWS + 1)];   /* U rows; alpha rows during the forward pass */
    double vb[OCO_WAVE * (OCO_STAGE_ROWS + 1)];   /* V rows */
    double* base[3][OCO_WAVE];                    /* per lane: U, V, alpha; null = lane stores nothing */
};

/* LDS traffic of one wavefront is performed in issue order, so ordering the
 * wavefront's own accesses needs only a compiler barrier at wavefront scope */
__device__ static inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ static inline void stage_put(double* buf, int idx, double v)
{
    buf[threadIdx.x * (OCO_STAGE_ROWS + 1) + (idx & (OCO_STAGE_ROWS - 1))] = v;
}

/* write block b (elements b*R .. b*R+R-1, clipped to [0, len)) of every lane:
 * thread t handles element r = t mod R of lanes t/R + k*(64/R), k = 0..R-1,
 * so R consecutive threads store one lane's run */
__device__ static inline void stage_flush(const double* buf, double* const* base, int b, int len)
{
    constexpr int R = OCO_STAGE_ROWS, L = OCO_WAVE / OCO_STAGE_ROWS;
    const int r = (int)threadIdx.x & (R - 1);
    const int s0 = (int)threadIdx.x / R;
    const double* bp = buf + s0 * (R + 1) + r;
    double* const* pp = base + s0;
    const int idx = b * R + r;
    const bool inb = idx < len;
#pragma unroll
    for (int k = 0; k < R; k++) {
        double* p = pp[k * L];
        double v = bp[k * L * (R + 1)];
        if (p != nullptr && inb) p[idx] = v;
    }
}

/* Row coefficients (source/auxiliary.f90:147-197) -- interior row from the
 * cell and its neighbours' alpha, boundary rows from the cell alone. */
template <bool MIX = false>
__device__ static inline void interior_row(double am, const oco_pt& pc, double ap, double dz2, double& av,
                                           double& bv, double& cv, double& dv)
{
    double kp = 0.5 * (ap + pc.alpha);
    double km = 0.5 * (pc.alpha + am);
    av = qdiv<MIX>(pc.alpha2, km);
    cv = qdiv<MIX>(pc.alpha2, kp);
    double ze = (dz2 * ((pc.alpha * pc.alpha) * pc.alpha)) * (1.0 - pc.albedo);
    bv = -((av + cv) + ze);
    dv = -(pc.S * ze);
}

__device__ static inline void boundary_row(const oco_pt& pc, double dz, double dz2, double& bv, double& dv)
{
    double ze = ((pc.alpha * dz2) * (1.0 - pc.albedo)) * 0.25;
    bv = -((pc.alpha1 + dz) + ze) * pc.alpha2;
    dv = -((pc.S * ze) * pc.alpha2);
}

template <bool GREY, bool UNIFORM, int WS = 1>
__device__ static inline oco_pt point_rec(const oco_sweep_args& a, long col, int z, const double* rec,
                                          const double* w, const oco_mtab* M)
{
    if (GREY) {
        long c = col * a.nz + z;
        return oco_grey_point(a.rho[c], a.T[c], a.sigma_grey, M);
    }
    return oco_point_s<WS>(rec, w, M);
}

template <bool GREY, int WS = 1, bool MIX = false>
__device__ static inline oco_pt point_sys(const oco_sys& d, int z, const double* rec, const double* w,
                                          const oco_mtab* M)
{
    if (GREY) return oco_grey_point(d.rho[z], d.T[z], d.sigma_grey, M);
    if constexpr (MIX) return oco_point_mixed<WS>(rec, w);
    else return oco_point_s<WS>(rec, w, M);
}

/* Tuning knob: OCO_SWEEP_WAVES_PER_EU=n asks the compiler to fit n waves per
 * SIMD (register budget 512/n VGPRs). */
#ifdef OCO_SWEEP_WAVES_PER_EU
#define OCO_SWEEP_ATTR __attribute__((amdgpu_waves_per_eu(OCO_SWEEP_WAVES_PER_EU)))
#else
#define OCO_SWEEP_ATTR
#endif

/* One (column, wavelength) system, one lane: forward elimination with the
 * opacity of the cells recomputed in registers, then back-substitution fused
 * with ghosts, der and V.  `w` is the lane's wavelength-table row with
 * stride WS; `scr` its scratch column (stride 64 doubles per entry).
 * MIX selects the mixed-precision opacity and reciprocal-based divides; the
 * elimination itself stays FP64 either way. */
template <bool GREY, bool UNIFORM, int WS, int BC = OCO_BACK_CHUNK, bool MIX = false>
__device__ static inline void solve_system(const oco_sys& d, bool valid, const double* w, double* scr,
                                           unsigned long long* fail, const oco_mtab* M, oco_stage* S)
{
    const int nz = d.nz;
    const double dz = d.dz, dz1 = d.dz1, dz2 = d.dz2;
    constexpr int R = OCO_STAGE_ROWS;
    const int lane = threadIdx.x;
    wave_sync();   /* the previous system's flushes have read the buffers */
    S->base[0][lane] = valid ? d.U : nullptr;
    S->base[1][lane] = valid ? d.V : nullptr;
    S->base[2][lane] = valid ? d.A : nullptr;

    /* ---- forward elimination (source/auxiliary.f90:25-41) ----
     * Two-deep pipeline: iteration j eliminates row j (which needs the
     * points j-1, j, j+1, all computed earlier) and evaluates point j+2,
     * whose work is independent of the elimination chain. */
    double rec[CR_NFIELD];
    if (!GREY) load_rec<UNIFORM>(d.cells, 0, rec);
    oco_pt pB = point_sys<GREY, WS, MIX>(d, 0, rec, w, M);
    if (!GREY) load_rec<UNIFORM>(d.cells, 1, rec);
    oco_pt pC = point_sys<GREY, WS, MIX>(d, 1, rec, w, M);
    if (!GREY) load_rec<UNIFORM>(d.cells, 2, rec);

    double bv, dv, av, cv;
    boundary_row(pB, dz, dz2, bv, dv);      /* row 0: a = 0, c = alpha */
    cv = pB.alpha;
    int fail_stage = (bv == 0.0) ? 1 : 0;
    int fail_row = 1;
    double bet = bv;
    double u = qdiv<MIX>(dv, bet);
    scr[0 * OCO_WAVE] = 0.0;
    scr[1 * OCO_WAVE] = u;
    scr[2 * OCO_WAVE] = pB.alpha1;
    stage_put(S->ub, 0, pB.alpha);
    double aA = pB.alpha;   /* alpha of row j-1 */
    pB = pC;                /* row j = 1 */
    pC = point_sys<GREY, WS, MIX>(d, 2, rec, w, M);

    for (int j = 1; j <= nz - 3; j++) {
        if (!GREY) load_rec<UNIFORM>(d.cells, j + 2, rec);
        /* row j and its elimination step */
        double cprev = cv;
        interior_row<MIX>(aA, pB, pC.alpha, dz2, av, bv, cv, dv);
        double gam = qdiv<MIX>(cprev, bet);
        bet = bv - av * gam;
        int first = (bet == 0.0) && (fail_stage == 0);
        fail_row = first ? j + 1 : fail_row;
        fail_stage = first ? 2 : fail_stage;
        u = qdiv<MIX>(dv - av * u, bet);
        double* sp = scr + (long)j * 3 * OCO_WAVE;
        sp[0] = gam;
        sp[OCO_WAVE] = u;
        sp[2 * OCO_WAVE] = pB.alpha1;
        stage_put(S->ub, j, pB.alpha);
        if ((j & (R - 1)) == R - 1) {
            wave_sync();
            stage_flush(S->ub, S->base[2], j / R, nz);
            wave_sync();
        }
        /* point j+2 (independent of the chain above) */
        oco_pt pD = point_sys<GREY, WS, MIX>(d, j + 2, rec, w, M);
        aA = pB.alpha;
        pB = pC;
        pC = pD;
    }
    {   /* row nz-2 (interior, no new point) and row nz-1 (boundary) */
        int j = nz - 2;
        double cprev = cv;
        interior_row<MIX>(aA, pB, pC.alpha, dz2, av, bv, cv, dv);
        double gam = qdiv<MIX>(cprev, bet);
        bet = bv - av * gam;
        int first = (bet == 0.0) && (fail_stage == 0);
        fail_row = first ? j + 1 : fail_row;
        fail_stage = first ? 2 : fail_stage;
        u = qdiv<MIX>(dv - av * u, bet);
        double* sp = scr + (long)j * 3 * OCO_WAVE;
        sp[0] = gam;
        sp[OCO_WAVE] = u;
        sp[2 * OCO_WAVE] = pB.alpha1;
        stage_put(S->ub, j, pB.alpha);
        if ((j & (R - 1)) == R - 1) {
            wave_sync();
            stage_flush(S->ub, S->base[2], j / R, nz);
            wave_sync();
        }
        j = nz - 1;
        cprev = cv;
        boundary_row(pC, dz, dz2, bv, dv);
        av = pC.alpha;                       /* a(nz) = alpha, c(nz) = 0 */
        gam = qdiv<MIX>(cprev, bet);
        bet = bv - av * gam;
        first = (bet == 0.0) && (fail_stage == 0);
        fail_row = first ? j + 1 : fail_row;
        fail_stage = first ? 2 : fail_stage;
        u = qdiv<MIX>(dv - av * u, bet);
        sp = scr + (long)j * 3 * OCO_WAVE;
        sp[0] = gam;
        sp[OCO_WAVE] = u;
        sp[2 * OCO_WAVE] = pC.alpha1;
        stage_put(S->ub, j, pC.alpha);
        wave_sync();
        stage_flush(S->ub, S->base[2], j / R, nz);   /* last (possibly partial) block */
    }
    if (fail_stage != 0) {
        /* the reference stops at this system (source/auxiliary.f90:27-39):
         * its U and V are not written */
        if (valid) record_fail(fail, d.key, fail_stage, fail_row);
        S->base[0][lane] = nullptr;
        S->base[1][lane] = nullptr;
    }
    wave_sync();

    /* ---- back-substitution (:43-45) fused with update_ghosts (:49-62),
     *      der (:64-78) and calc_flux (:96-109) ---- */
    const double c60 = 1.0 / 60.0;
    /* window w0..w6 = extended U at e = j .. j+6 (e = z + 3) */
    double Un = u;                       /* U(nz-1) */
    double gnext = scr[(long)(nz - 1) * 3 * OCO_WAVE];
    double a1q0 = scr[(long)(nz - 1) * 3 * OCO_WAVE + 2 * OCO_WAVE];   /* 1/alpha of the V row pending */
    double a1q1 = 0.0, a1q2 = 0.0, a1q3 = 0.0;
    double t[4];
    t[0] = Un;
    for (int k = 1; k < 4; k++) {
        int j = nz - 1 - k;
        const double* sp = scr + (long)j * 3 * OCO_WAVE;
        double Uj = sp[OCO_WAVE] - gnext * t[k - 1];
        gnext = sp[0];
        double a1 = sp[2 * OCO_WAVE];
        a1q3 = a1q2; a1q2 = a1q1; a1q1 = a1q0; a1q0 = a1;
        t[k] = Uj;
    }
    /* top ghosts (f(n2+i) = 2 f(n2) - f(n2-i)) */
    double g1 = 2.0 * t[0] - t[1], g2 = 2.0 * t[0] - t[2], g3 = 2.0 * t[0] - t[3];
    /* U(e) for e = nz+5 down to nz-1, in descending order: a block is
     * flushed when its lowest element is staged */
    {
        const double top[7] = {g3, g2, g1, t[0], t[1], t[2], t[3]};
#pragma unroll
        for (int k = 0; k < 7; k++) {
            const int e = nz + 5 - k;
            stage_put(S->ub, e, top[k]);
            if ((e & (R - 1)) == 0) {
                wave_sync();
                stage_flush(S->ub, S->base[0], e / R, nz + 6);
                wave_sync();
            }
        }
    }
    double w0 = t[3], w1 = t[2], w2 = t[1], w3 = t[0], w4 = g1, w5 = g2, w6 = g3;
    {
        double df = c60 * (((45.0 * (w4 - w2)) - (9.0 * (w5 - w1))) + (w6 - w0));
        stage_put(S->vb, nz - 1, ((-1.0 * a1q3) * df) * dz1);
        if (((nz - 1) & (R - 1)) == 0) {
            wave_sync();
            stage_flush(S->vb, S->base[1], (nz - 1) / R, nz);
            wave_sync();
        }
    }
    /* rows nz-5 .. 0 in chunks of BC: each chunk's scratch
     * entries are loaded together, one chunk ahead of their use, so the loop
     * pays one memory latency per chunk instead of one per row */
    {
        double cg[BC], cu[BC], ca[BC];
        double ng[BC], nu[BC], na[BC];
        int jc = nz - 5;
#pragma unroll
        for (int k = 0; k < BC; k++) {
            int j = jc - k;
            const double* sp = scr + (long)(j < 0 ? 0 : j) * 3 * OCO_WAVE;
            cg[k] = sp[0]; cu[k] = sp[OCO_WAVE]; ca[k] = sp[2 * OCO_WAVE];
        }
        for (; jc >= 0; jc -= BC) {
#pragma unroll
            for (int k = 0; k < BC; k++) {      /* next chunk, in flight */
                int j = jc - BC - k;
                const double* sp = scr + (long)(j < 0 ? 0 : j) * 3 * OCO_WAVE;
                ng[k] = sp[0]; nu[k] = sp[OCO_WAVE]; na[k] = sp[2 * OCO_WAVE];
            }
#pragma unroll
            for (int k = 0; k < BC; k++) {
                int j = jc - k;
                if (j >= 0) {
                    double Uj = cu[k] - gnext * w0;
                    gnext = cg[k];
                    a1q3 = a1q2; a1q2 = a1q1; a1q1 = a1q0; a1q0 = ca[k];
                    const int e = j + 3;   /* U(e) and V(z = e) */
                    stage_put(S->ub, e, Uj);
                    w6 = w5; w5 = w4; w4 = w3; w3 = w2; w2 = w1; w1 = w0; w0 = Uj;
                    double df = c60 * (((45.0 * (w4 - w2)) - (9.0 * (w5 - w1))) + (w6 - w0));
                    stage_put(S->vb, e, ((-1.0 * a1q3) * df) * dz1);
                    if ((e & (R - 1)) == 0) {
                        wave_sync();
                        stage_flush(S->ub, S->base[0], e / R, nz + 6);
                        stage_flush(S->vb, S->base[1], e / R, nz);
                        wave_sync();
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < BC; k++) { cg[k] = ng[k]; cu[k] = nu[k]; ca[k] = na[k]; }
        }
    }
    /* bottom ghosts (f(n1-i) = 2 f(n1) - f(n1+i)); window now e = 3..9 */
    double u0 = w0, G1 = 2.0 * u0 - w1, G2 = 2.0 * u0 - w2, G3 = 2.0 * u0 - w3;
    stage_put(S->ub, 2, G1);
    stage_put(S->ub, 1, G2);
    stage_put(S->ub, 0, G3);
    const double gh[3] = {G1, G2, G3};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        w6 = w5; w5 = w4; w4 = w3; w3 = w2; w2 = w1; w1 = w0; w0 = gh[k];
        a1q3 = a1q2; a1q2 = a1q1; a1q1 = a1q0;
        double df = c60 * (((45.0 * (w4 - w2)) - (9.0 * (w5 - w1))) + (w6 - w0));
        stage_put(S->vb, 2 - k, ((-1.0 * a1q3) * df) * dz1);
    }
    wave_sync();
    stage_flush(S->ub, S->base[0], 0, nz + 6);
    stage_flush(S->vb, S->base[1], 0, nz);
}

/* One wavefront per workgroup, grid-stride over tiles of 64 systems; the
 * lane's wavelength-table row lives in registers. */
template <bool GREY, bool UNIFORM, bool MIX>
__device__ static inline void sweep_tiles(const oco_sweep_args& a, const oco_mtab* M, oco_stage* S)
{
    const int lane = threadIdx.x;
    const int nw = a.nw;
    const long nsys = a.ncol * nw;
    const long ntiles = (nsys + OCO_WAVE - 1) / OCO_WAVE;
    double* scr = a.scratch + (long)blockIdx.x * a.nz * 3 * OCO_WAVE + lane;
    for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        long sys0 = tile * OCO_WAVE;
        long col;
        int iw;
        if (UNIFORM) {
            col = sys0 / nw;
            iw = (int)(sys0 % nw) + lane;
        } else {
            long s = sys0 + lane;
            if (s >= nsys) s = nsys - 1;   /* past the end: duplicate of the last system, no stores */
            col = s / nw;
            iw = (int)(s % nw);
        }
        const bool valid = UNIFORM || sys0 + lane < nsys;
        double w[OCO_WT_STRIDE];
        if (!GREY) {
#pragma unroll
            for (int k = 0; k < OCO_WT_STRIDE; k++) w[k] = a.wtab[iw * OCO_WT_STRIDE + k];
        }
        const long osys = (long)iw * a.ncol + col;
        oco_sys d;
        d.nz = a.nz;
        d.dz = a.dz; d.dz1 = a.dz1; d.dz2 = a.dz2; d.sigma_grey = a.sigma_grey;
        d.cells = a.cells + col * a.nz * OCO_CELL_STRIDE;
        d.rho = a.rho + col * a.nz;
        d.T = a.T + col * a.nz;
        d.U = a.U + osys * (a.nz + 6);
        d.V = a.V + osys * a.nz;
        d.A = a.A + osys * a.nz;
        d.key = col * nw + iw;
        solve_system<GREY, UNIFORM, 1, OCO_BACK_CHUNK, MIX>(d, valid, w, scr, a.fail, M, S);
    }
}

/* FP64 drop-in sweep.  UNIFORM: nw is a multiple of 64, so a wavefront's
 * systems share one column. */
template <bool GREY, bool UNIFORM>
__global__ __launch_bounds__(OCO_WAVE) OCO_SWEEP_ATTR void sweep_fp64(const oco_sweep_args a)
{
    __shared__ oco_mtab smt;
    __shared__ oco_stage stg;
    load_mtab(&smt);
    sweep_tiles<GREY, UNIFORM, false>(a, &smt, &stg);
}

/* Mixed-precision sweep (non-grey): FP32-core opacity, FP64 elimination.
 * Its elementary functions use no tables. */
template <bool UNIFORM>
__global__ __launch_bounds__(OCO_WAVE) OCO_SWEEP_ATTR void sweep_mixed(const oco_sweep_args a)
{
    __shared__ oco_stage stg;
    sweep_tiles<false, UNIFORM, true>(a, nullptr, &stg);
}

/* First-pillar canaries (source/flux_feautrier.f90:198-201): sum(a), sum(b),
 * sum(c), sum(d) of system (ix,iy,iw) = (1,1,1).  Printed diagnostics only
 * (Fortran leaves the summation order of SUM to the compiler), so points,
 * rows and the sums are done in parallel by one workgroup.  `buf` holds
 * 4 + 5*nz doubles: the sums, then alpha, 1/alpha, alpha^2, albedo, S. */
template <bool GREY>
__global__ __launch_bounds__(256) void canary_fp64(const oco_sweep_args a, double* buf)
{
    __shared__ oco_mtab smt;
    __shared__ double red[4][256];
    load_mtab(&smt);
    const int nz = a.nz, tid = threadIdx.x;
    double* pts = buf + 4;
    double w[OCO_WT_STRIDE];
    if (!GREY)
        for (int k = 0; k < OCO_WT_STRIDE; k++) w[k] = a.wtab[k];
    for (int z = tid; z < nz; z += blockDim.x) {
        double rec[CR_NFIELD];
        if (!GREY) load_record<false>(a, 0, z, rec);
        oco_pt p = point_rec<GREY, false>(a, 0, z, rec, w, &smt);
        double* q = pts + 5 * z;
        q[0] = p.alpha; q[1] = p.alpha1; q[2] = p.alpha2; q[3] = p.albedo; q[4] = p.S;
    }
    __syncthreads();
    double sa = 0.0, sb = 0.0, sc = 0.0, sd = 0.0;
    for (int r = tid; r < nz; r += blockDim.x) {
        const double* q = pts + 5 * r;
        oco_pt pc;
        pc.alpha = q[0]; pc.alpha1 = q[1]; pc.alpha2 = q[2]; pc.albedo = q[3]; pc.S = q[4];
        double av, bv, cv, dv;
        if (r == 0) {
            boundary_row(pc, a.dz, a.dz2, bv, dv);
            av = 0.0;
            cv = pc.alpha;
        } else if (r == nz - 1) {
            boundary_row(pc, a.dz, a.dz2, bv, dv);
            av = pc.alpha;
            cv = 0.0;
        } else {
            interior_row(pts[5 * (r - 1)], pc, pts[5 * (r + 1)], a.dz2, av, bv, cv, dv);
        }
        sa += av; sb += bv; sc += cv; sd += dv;
    }
    red[0][tid] = sa; red[1][tid] = sb; red[2][tid] = sc; red[3][tid] = sd;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if (tid < h)
            for (int k = 0; k < 4; k++) red[k][tid] += red[k][tid + h];
        __syncthreads();
    }
    if (tid < 4) buf[tid] = red[tid][0];
}

/* ---- batched uneven columns ----
 * Tiles are (column, block of 64 wavelengths), sorted longest column first
 * on the host; a persistent grid draws them from an atomic counter, so
 * waves that finish short columns take the next tile (dynamic balancing).
 * Every column carries its own nz, grid step and wavelength grid. */
__global__ void batch_reset(unsigned int* counter, unsigned long long* fail)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *counter = 0u;
        *fail = OCO_FAIL_NONE;
    }
}

typedef const __attribute__((address_space(4))) int* oco_cptr_i;
typedef const __attribute__((address_space(4))) long* oco_cptr_l;
typedef const __attribute__((address_space(4))) double* oco_cptr_d;

template <int BC, bool MIX>
__device__ static inline void batch_tiles(const oco_batch_args& a, const oco_mtab* M, int* tile_s, oco_stage* S)
{
    const int lane = threadIdx.x;
    double* scr = a.scratch + (long)blockIdx.x * a.scr_stride + lane;
    oco_cptr_i tiles = (oco_cptr_i)a.tiles;
    oco_cptr_i nzs = (oco_cptr_i)a.nz;
    oco_cptr_i nws = (oco_cptr_i)a.nw;
    oco_cptr_l coff = (oco_cptr_l)a.cell_off;
    oco_cptr_l woff = (oco_cptr_l)a.wt_off;
    oco_cptr_l uoff = (oco_cptr_l)a.u_off;
    oco_cptr_l voff = (oco_cptr_l)a.v_off;
    oco_cptr_d dzs = (oco_cptr_d)a.dz;
    for (;;) {
        if (lane == 0) *tile_s = (int)atomicAdd(a.counter, 1u);
        __syncthreads();
        int t = __builtin_amdgcn_readfirstlane(*tile_s);
        __syncthreads();
        if (t >= a.ntiles) break;
        const int col = tiles[2 * t];
        const int nz = nzs[col];
        const int nw = nws[col];
        const bool valid = tiles[2 * t + 1] + lane < nw;
        /* lanes past the column's last wavelength run a duplicate of it, no stores */
        const int iw = valid ? tiles[2 * t + 1] + lane : nw - 1;
        {
            oco_sys d;
            d.nz = nz;
            d.dz = dzs[3 * col];
            d.dz1 = dzs[3 * col + 1];
            d.dz2 = dzs[3 * col + 2];
            d.sigma_grey = 0.0;
            d.cells = a.cells + coff[col] * OCO_CELL_STRIDE;
            d.rho = a.rho + coff[col];
            d.T = a.T + coff[col];
            d.U = a.U + uoff[col] + (long)iw * (nz + 6);
            d.V = a.V + voff[col] + (long)iw * nz;
            d.A = a.A + voff[col] + (long)iw * nz;
            d.key = woff[col] + iw;
            double w[OCO_WT_STRIDE];
            const double* wr = a.wtab + (woff[col] + iw) * OCO_WT_STRIDE;
#pragma unroll
            for (int k = 0; k < OCO_WT_STRIDE; k++) w[k] = wr[k];
            solve_system<false, true, 1, BC, MIX>(d, valid, w, scr, a.fail, M, S);
        }
    }
}

template <int BC>
__global__ __launch_bounds__(OCO_WAVE) void batch_fp64(const oco_batch_args a)
{
    __shared__ oco_mtab smt;
    __shared__ int tile_s;
    __shared__ oco_stage stg;
    load_mtab(&smt);
    batch_tiles<BC, false>(a, &smt, &tile_s, &stg);
}

template <int BC>
__global__ __launch_bounds__(OCO_WAVE) void batch_mixed(const oco_batch_args a)
{
    __shared__ int tile_s;
    __shared__ oco_stage stg;
    batch_tiles<BC, true>(a, nullptr, &tile_s, &stg);
}

hipError_t oco_launch_batch(const oco_batch_args& a, int mixed, int variant, int nblocks, hipStream_t stream)
{
    hipLaunchKernelGGL(batch_reset, dim3(1), dim3(64), 0, stream, a.counter, a.fail);
    const dim3 g(nblocks), b(OCO_WAVE);
    if (mixed) {
        switch (variant) {
        case 0: hipLaunchKernelGGL((batch_mixed<1>), g, b, 0, stream, a); break;
        case 1: hipLaunchKernelGGL((batch_mixed<2>), g, b, 0, stream, a); break;
        default: hipLaunchKernelGGL((batch_mixed<4>), g, b, 0, stream, a); break;
        }
    } else {
        switch (variant) {
        case 0: hipLaunchKernelGGL((batch_fp64<1>), g, b, 0, stream, a); break;
        case 1: hipLaunchKernelGGL((batch_fp64<2>), g, b, 0, stream, a); break;
        default: hipLaunchKernelGGL((batch_fp64<4>), g, b, 0, stream, a); break;
        }
    }
    return hipGetLastError();
}

int oco_batch_variants(void) { return 3; }

hipError_t oco_launch_cells(const double* rho, const double* T, long ncells, double fully_ionized_T,
                            double* cells, hipStream_t stream)
{
    if (ncells <= 0) return hipSuccess;
    long blocks = (ncells + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(cells_fp64, dim3((unsigned)blocks), dim3(256), 0, stream, rho, T, ncells,
                       fully_ionized_T, cells);
    return hipGetLastError();
}

/* Workgroups of the sweep kernel that fit on the device at once (occupancy x
 * CUs): the grid-stride sweep is launched with exactly that many, so every
 * workgroup is resident from the start and takes an equal share of tiles. */
int oco_sweep_grid(int lgrey, int mixed, int uniform, int device)
{
    const void* k;
    if (lgrey) k = (const void*)sweep_fp64<true, false>;
    else if (mixed) k = uniform ? (const void*)sweep_mixed<true> : (const void*)sweep_mixed<false>;
    else k = uniform ? (const void*)sweep_fp64<false, true> : (const void*)sweep_fp64<false, false>;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, OCO_WAVE, 0) != hipSuccess || per_cu < 1)
        per_cu = 8;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1)
        cus = 256;
    return per_cu * cus;
}

hipError_t oco_launch_sweep(const oco_sweep_args& a, int lgrey, int mixed, int nblocks, hipStream_t stream)
{
    const bool uniform = (a.nw % OCO_WAVE) == 0;
    if (a.canary) {
        if (lgrey)
            hipLaunchKernelGGL((canary_fp64<true>), dim3(1), dim3(256), 0, stream, a, a.canary);
        else
            hipLaunchKernelGGL((canary_fp64<false>), dim3(1), dim3(256), 0, stream, a, a.canary);
    }
    const dim3 g(nblocks), b(OCO_WAVE);
    if (lgrey) {
        /* grey opacity has no per-wavelength transcendentals: one kernel for both modes */
        hipLaunchKernelGGL((sweep_fp64<true, false>), g, b, 0, stream, a);
    } else if (mixed) {
        if (uniform) hipLaunchKernelGGL((sweep_mixed<true>), g, b, 0, stream, a);
        else hipLaunchKernelGGL((sweep_mixed<false>), g, b, 0, stream, a);
    } else if (uniform) {
        hipLaunchKernelGGL((sweep_fp64<false, true>), g, b, 0, stream, a);
    } else {
        hipLaunchKernelGGL((sweep_fp64<false, false>), g, b, 0, stream, a);
    }
    return hipGetLastError();
}
