/*
 * oco_kernels.hip -- the gfx950 kernels of the path.
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
 *   sweep_tiny_fp64  the drop-in for columns of 2 or 3 cells, one thread
 *               per system.
 *   sweep_mixed_wave  the difference-form sweep of OCO_MODE_MIXED (opacity
 *               of oco_mixed.h) and OCO_MODE_ACCURATE (the drop-in's FP64
 *               opacity): an aligned group of L lanes per system.
 *   batch_fp64, batch_mixed_wave  the same per column of a batch of uneven
 *               columns; serve_tiles, serve_wave  the same for the server's
 *               jobs, emergent flux only; athena_slab, opacity_points  input
 *               construction and the opacity query.
 *
 * The FP64 path is compiled with -ffp-contract=off and its elementary
 * functions are oco_math.h's, read from an LDS copy of their tables.
 */
#include <mutex>

#include "oco_internal.h"
#include "oco_physics.h"
#include "oco_mixed.h"

__constant__ oco_mtab oco_mtab_dev = OCO_MTAB_INIT;

#define OCO_WAVE 64
#define OCO_XCDS 8   /* accelerator complexes (each with its own L2) of an MI355X */
#ifndef OCO_BACK_CHUNK
#define OCO_BACK_CHUNK 8   /* back-substitution rows whose scratch loads are batched */
#endif

/* lane of the calling thread in its wavefront */
__device__ static inline int lane_id() { return (int)threadIdx.x & (OCO_WAVE - 1); }

__device__ static inline void load_mtab(oco_mtab* smt)
{
    const double* src = (const double*)&oco_mtab_dev;
    double* dst = (double*)smt;
    for (int i = threadIdx.x; i < (int)(sizeof(oco_mtab) / sizeof(double)); i += blockDim.x) dst[i] = src[i];
    __syncthreads();
}

/* SOA = false: records of OCO_CELL_STRIDE doubles, cell-major (read by the
 * lane-per-system sweeps, a wavefront's 64 lanes sharing one cell);
 * SOA = true: field-major per column, field k of cell z of column c at
 * (c*CR_NFIELD + k)*nz + z (read by the wavefront-per-system sweep, 64 lanes
 * on 64 consecutive cells of one column). */
template <bool SOA>
__global__ __launch_bounds__(256) void cells_fp64(const double* __restrict__ rho, const double* __restrict__ T,
                                                  long ncells, int nz, double fully_ionized_T,
                                                  double* __restrict__ cells)
{
    __shared__ oco_mtab smt;
    load_mtab(&smt);
    if constexpr (SOA) {
        for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < ncells; i += (long)gridDim.x * blockDim.x) {
            double rec[CR_NFIELD];
            oco_cell_record(rho[i], T[i], fully_ionized_T, &smt, rec);
            const long col = i / nz;
            double* out = cells + col * (long)CR_NFIELD * nz + (i - col * nz);
#pragma unroll
            for (int k = 0; k < CR_NFIELD; k++) out[(long)k * nz] = rec[k];
        }
    } else {
        /* the workgroup's 256 records go through LDS, so its stores cover a
         * contiguous 48 KB run (a thread's own 192-byte record written
         * directly puts each store instruction on 64 lines; measured on
         * MI355X, 256 x 256 x 512 cells: 2.78-2.86 ms -> 1.87-1.92 ms).  Field k of thread
         * t's record sits at t S + (k + t) mod S, S = OCO_CELL_STRIDE, so the
         * 64 lanes' writes of one field spread over the LDS banks. */
        constexpr int S = OCO_CELL_STRIDE;
        __shared__ double stg[256 * S];
        const int t = (int)threadIdx.x;
        const long step = (long)gridDim.x * 256;
        for (long base = (long)blockIdx.x * 256; base < ncells; base += step) {   /* workgroup-uniform */
            const long i = base + t;
            if (i < ncells) {
                double rec[CR_NFIELD];
                oco_cell_record(rho[i], T[i], fully_ionized_T, &smt, rec);
#pragma unroll
                for (int k = 0; k < S; k++) stg[t * S + (k + t) % S] = k < CR_NFIELD ? rec[k] : 0.0;
            }
            __syncthreads();
            const int n = (int)(ncells - base < 256 ? ncells - base : 256) * S;
            double* out = cells + base * S;
            for (int e = t; e < n; e += 256) {
                const int c = e / S, k = e - c * S;
                out[e] = stg[c * S + (k + c) % S];
            }
            __syncthreads();
        }
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

/* ---- coalesced output stores ----
 * A lane owns one system, and its outputs U(:), V(:), alpha(:) are contiguous
 * runs, but consecutive lanes' runs lie (nz+6)*ncol doubles apart in the
 * reference's (z, column, wavelength) layout, so a plain per-row store sends
 * 64 lines per instruction.  Rows are staged in LDS instead, R per lane, and
 * every R rows the wavefront writes them out cooperatively: each store
 * instruction then covers 64/R whole runs of R consecutive doubles.
 * Every lane of the wavefront takes part in a flush, so lanes without a
 * system of their own run a duplicate one with their stores switched off. */
/* rows staged per lane: 16, so each lane's runs are 128 bytes (whole cache
 * lines).  The FP64 kernels also hold the elementary-function tables in LDS;
 * with 16 rows their LDS allows 6 wavefronts per CU against 8 at 8 rows, and
 * they still run faster (measured on a 128 x 128 x 512, nw = 128 block:
 * 26.8 ms at 8 rows, 24.7 ms at 16 rows with 8 rows of scratch prefetch). */
#ifndef OCO_STAGE_ROWS
#define OCO_STAGE_ROWS 16
#endif
/* How an FP64 back pass gets 1/alpha: OCO_A1_SCRATCH from the scratch row;
 * OCO_A1_DIRECT by reading alpha from the lane's own output; OCO_A1_STAGED
 * by loading alpha blocks of R rows into the V stage, coalesced (the reverse
 * of the output flush), a block ahead.  The last two recompute 1/alpha as
 * the forward pass computed it (1.0 / alpha, correctly rounded: the same
 * bits) and keep [gam, u] only in scratch. */
#define OCO_A1_SCRATCH 0
#define OCO_A1_DIRECT 1
#define OCO_A1_STAGED 2
template <int R, bool NT = true, int AR = OCO_A1_SCRATCH>
struct oco_stage_t {
    static constexpr int rows = R;
    static constexpr bool emergent = false;
    static constexpr bool nt = NT;   /* outputs by nontemporal stores */
    static constexpr int reload = AR;
    double ub[OCO_WAVE * (R + 1)];   /* U rows; alpha rows during the forward pass */
    double vb[OCO_WAVE * (R + 1)];   /* V rows */
    double* base[3][OCO_WAVE];       /* per lane: U, V, alpha; null = lane stores nothing */
};
/* FP64 block sweep: a lane's alpha is a run of its own output, the 64 lanes'
 * runs ncol*nz doubles apart, so alpha comes back through the stage in
 * coalesced 128-byte runs per lane (read directly, each lane's read lands
 * on a page of its own: 256 x 256 x 512 block 79 ms -> 238 ms). */
#ifndef OCO_BLOCK_A1
#define OCO_BLOCK_A1 OCO_A1_STAGED
#endif
typedef oco_stage_t<OCO_STAGE_ROWS, true, OCO_BLOCK_A1> oco_stage;
/* The FP64 batch kernel stages 16-row runs like the block sweep and stores
 * them nontemporally: a batch's device rows are padded to whole lines
 * (oco_batch_ldu, oco_batch_ldv), so every run is one line (measured on
 * MI355X, 4096 uneven columns: 8-row runs with ordinary stores 7.22-7.28 ms,
 * 16-row 6.93-6.96, 16-row nontemporal 6.67-6.78).  A batched column's
 * alpha is one contiguous (ld, nw) array, so a lane's reads of its own alpha
 * stay within a few pages of its neighbours': the kernel reads it directly
 * (9.98 ms with 1/alpha in scratch, 9.39 ms, before the padding). */
#ifndef OCO_BATCH_A1
#define OCO_BATCH_A1 OCO_A1_DIRECT
#endif
typedef oco_stage_t<16, true, OCO_BATCH_A1> oco_stage_batch;
/* wavefronts per workgroup of the FP64 batch kernel: each takes its own
 * tiles, and they share one LDS copy of the math tables, so two workgroups'
 * stages fit a CU (8 wavefronts, 2 per SIMD) */
#define OCO_BATCH_WG 4
static_assert(OCO_BATCH_WG <= OCO_BATCH_WG_MAX, "scratch slots per launch");

/* Output policy of the serving kernels: only the emergent flux V(1), V(nz)
 * of each system is stored (d.V[0], d.V[1]); U and alpha are not written. */
struct oco_stage_emergent {
    static constexpr int rows = 8;
    static constexpr bool emergent = true;
    static constexpr bool nt = false;
    static constexpr int reload = OCO_A1_SCRATCH;
    double ub[1], vb[1];
    double* base[3][1];
};

/* LDS traffic of one wavefront is performed in issue order, so ordering the
 * wavefront's own accesses needs only a compiler barrier at wavefront scope */
__device__ static inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

template <int R>
__device__ static inline void stage_put(double* buf, int idx, double v)
{
    buf[lane_id() * (R + 1) + (idx & (R - 1))] = v;
}

/* Outputs are written once and not read back by the kernel: nontemporal
 * stores keep them from displacing the scratch and cell records in L2.
 * Tuning knob: OCO_WAVE_NT=false gives the wavefront sweeps ordinary stores
 * (measured on MI355X, 256 x 256 columns, nz = 512, nw = 128: 39.1 ms
 * against 37.35 with nontemporal stores). */
#ifndef OCO_WAVE_NT
#define OCO_WAVE_NT true
#endif
/* Output, scratch and record pointers reach the kernels through structs and
 * LDS tables, where the compiler cannot see that they point to global
 * memory, and then emits flat instructions, which count against the LDS
 * counter as well as the vector memory one: a wait for an LDS result (the
 * math tables, the stages) would also wait for every store in flight.  These
 * casts make them global accesses. */
typedef __attribute__((address_space(1))) double oco_gdouble;
__device__ static inline oco_gdouble* gptr(double* p) { return (oco_gdouble*)p; }
__device__ static inline const oco_gdouble* gptr(const double* p) { return (const oco_gdouble*)p; }

template <bool NT = true>
__device__ static inline void out_store(double* p, double v)
{
    if constexpr (NT) __builtin_nontemporal_store(v, gptr(p));
    else *gptr(p) = v;
}

/* write block b (elements b*R .. b*R+R-1, clipped to [0, len)) of every lane:
 * thread t handles element r = t mod R of lanes t/R + k*(64/R), k = 0..R-1,
 * so R consecutive threads store one lane's run */
template <int R, bool NT = true>
__device__ static inline void stage_flush(const double* buf, double* const* base, int b, int len, int sh = 0)
{
    constexpr int L = OCO_WAVE / R;
    const int r = lane_id() & (R - 1);
    const int s0 = lane_id() / R;
    const double* bp = buf + s0 * (R + 1) + r;
    double* const* pp = base + s0;
    const int idx = b * R + r - sh;
    const bool inb = idx >= 0 && idx < len;
#pragma unroll
    for (int k = 0; k < R; k++) {
        double* p = pp[k * L];
        double v = bp[k * L * (R + 1)];
        if (p != nullptr && inb) out_store<NT>(p + idx, v);
    }
}

/* Row coefficients (source/auxiliary.f90:147-197) -- interior row from the
 * cell and its neighbours' alpha, boundary rows from the cell alone. */
__device__ static inline void interior_row(double am, const oco_pt& pc, double ap, double dz2, double& av,
                                           double& bv, double& cv, double& dv)
{
    double kp = 0.5 * (ap + pc.alpha);
    double km = 0.5 * (pc.alpha + am);
    av = pc.alpha2 / km;
    cv = pc.alpha2 / kp;
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

/* The same rows in difference form (mixed mode): -A x[i-1] + (A + C + H) x[i]
 * - C x[i+1] = R with A, C, H, R >= 0; H is the row's excess over A + C
 * (zeta inside; alpha^2 (dz + zeta) at the boundary rows, whose coupling is
 * alpha).  An interior row is scaled by the product of its interface means
 * k- k+ (A = alpha^2 k+, C = alpha^2 k-, H and R times k- k+): no reciprocal,
 * and the elimination's ratios within the row are unchanged. */
__device__ static inline void rh_interior_row(double am, const oco_pt& pc, double ap, double dz2, double& A,
                                              double& C, double& H, double& R)
{
#pragma clang fp contract(on)
    const double km = 0.5 * (pc.alpha + am), kp = 0.5 * (ap + pc.alpha);
    A = pc.alpha2 * kp;
    C = pc.alpha2 * km;
    H = ((dz2 * (pc.alpha2 * pc.alpha)) * (1.0 - pc.albedo)) * (km * kp);
    R = pc.S * H;
}

__device__ static inline void rh_boundary_row(const oco_pt& pc, double dz, double dz2, double& H, double& R)
{
#pragma clang fp contract(on)
    const double ze = ((pc.alpha * dz2) * (1.0 - pc.albedo)) * 0.25;
    H = (dz + ze) * pc.alpha2;
    R = (pc.S * ze) * pc.alpha2;
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

template <bool GREY, int WS = 1>
__device__ static inline oco_pt point_sys(const oco_sys& d, int z, const double* rec, const double* w,
                                          const oco_mtab* M)
{
    if (GREY) return oco_grey_point(d.rho[z], d.T[z], d.sigma_grey, M);
    return oco_point_s<WS>(rec, w, M);
}

/* Tuning knob: OCO_SWEEP_WAVES_PER_EU=n asks the compiler to fit n waves per
 * SIMD (register budget 512/n VGPRs). */
#ifdef OCO_SWEEP_WAVES_PER_EU
#define OCO_SWEEP_ATTR __attribute__((amdgpu_waves_per_eu(OCO_SWEEP_WAVES_PER_EU)))
#else
#define OCO_SWEEP_ATTR
#endif

/* Elimination scratch of the lane-per-system sweeps, one row per call:
 * [gam, u, 1/alpha] as doubles, each a coalesced line of 64 lanes; with
 * 1/alpha recomputed from a reloaded alpha (AR, OCO_A1_*), [gam, u], 16
 * bytes.  `scr` is the lane's slot pointer (wave slot + lane). */
/* The [gam, u] scratch rows are written once and read once, nz rows later:
 * nontemporal stores and loads keep them from displacing the cell records
 * and the alpha being reloaded in L2.  Tuning knob: OCO_SCRATCH_NT = 0
 * (ordinary accesses), 1 (nontemporal stores only), 2 (both); measured on
 * MI355X, FP64 drop-in, 256 x 256 columns, nz = 512, nw = 128: 74.6, 74.1
 * and 73.7 ms. */
#ifndef OCO_SCRATCH_NT
#define OCO_SCRATCH_NT 2
#endif
template <int AR = OCO_A1_SCRATCH>
__device__ static inline void scr_put(double* scr, int j, double gam, double u, double a1)
{
    if constexpr (AR != OCO_A1_SCRATCH) {
        (void)a1;
        double* sp = scr + (long)j * 2 * OCO_WAVE;
        out_store<(OCO_SCRATCH_NT >= 1)>(sp, gam);
        out_store<(OCO_SCRATCH_NT >= 1)>(sp + OCO_WAVE, u);
    } else {
        double* sp = scr + (long)j * 3 * OCO_WAVE;
        sp[0] = gam;
        sp[OCO_WAVE] = u;
        sp[2 * OCO_WAVE] = a1;
    }
}

/* a1: 1/alpha of row j (OCO_A1_DIRECT: alpha of row j, read from the lane's
 * output A, the caller takes a1_of; OCO_A1_STAGED: not set) */
template <int AR = OCO_A1_SCRATCH>
__device__ static inline void scr_get(const double* scr, int j, double& gam, double& u, double& a1,
                                      const double* A = nullptr)
{
    if constexpr (AR == OCO_A1_DIRECT) {
        const double* sp = scr + (long)j * 2 * OCO_WAVE;
        gam = sp[0];
        u = sp[OCO_WAVE];
        a1 = A[j];
    } else if constexpr (AR == OCO_A1_STAGED) {
        (void)A;
        const double* sp = scr + (long)j * 2 * OCO_WAVE;
        if constexpr (OCO_SCRATCH_NT >= 2) {
            gam = __builtin_nontemporal_load(gptr(sp));
            u = __builtin_nontemporal_load(gptr(sp + OCO_WAVE));
        } else {
            gam = sp[0];
            u = sp[OCO_WAVE];
        }
        a1 = 0.0;
    } else {
        (void)A;
        const double* sp = scr + (long)j * 3 * OCO_WAVE;
        gam = sp[0];
        u = sp[OCO_WAVE];
        a1 = sp[2 * OCO_WAVE];
    }
}

template <int AR>
__device__ static inline double a1_of(double v)
{
    if constexpr (AR == OCO_A1_DIRECT) return 1.0 / v;   /* alpha -> 1/alpha, as oco_point's alpha1 */
    else return v;
}

/* How a lane sweep with this mode and output policy gets 1/alpha in its back
 * pass (OCO_A1_*; alpha is written by the forward pass). */
template <class STAGE>
constexpr int alpha_reload() { return STAGE::emergent ? OCO_A1_SCRATCH : STAGE::reload; }

/* OCO_A1_STAGED: alpha rows [b*R, b*R + R) of every lane into registers
 * (thread t holds element t mod R of lanes t/R + k*64/R, the output flush's
 * pattern: R consecutive threads read one lane's 8R-byte run), then into
 * the V stage's slots; a lane without an alpha output gets 1. */
template <int R>
__device__ static inline void alpha_fetch(double* regs, double* const* base, int b, int len, int sh = 0)
{
    constexpr int L = OCO_WAVE / R;
    const int r = lane_id() & (R - 1), s0 = lane_id() / R, idx = b * R + r - sh;
#pragma unroll
    for (int k = 0; k < R; k++) {
        const double* p = base[s0 + k * L];
        regs[k] = p != nullptr && idx >= 0 && idx < len ? gptr(p)[idx] : 1.0;
    }
}

/* The offset, in doubles within a run of R, of every valid lane's row p
 * when all of them share it (so staged runs can start on whole lines), else
 * 0; q, when given, must share the same offset. */
template <int R>
__device__ static inline int line_shift(const double* p, bool valid, const double* q = nullptr)
{
    const int sl = (int)(((unsigned long)p >> 3) & (R - 1));
    const int ql = q ? (int)(((unsigned long)q >> 3) & (R - 1)) : sl;
    const int s0 = __builtin_amdgcn_readfirstlane(sl);
    return __builtin_amdgcn_ballot_w64(valid && (sl != s0 || ql != s0)) == 0 ? s0 : 0;
}

template <int R>
__device__ static inline void alpha_park(double* buf, const double* regs)
{
    constexpr int L = OCO_WAVE / R;
    const int r = lane_id() & (R - 1), s0 = lane_id() / R;
#pragma unroll
    for (int k = 0; k < R; k++) buf[(s0 + k * L) * (R + 1) + r] = regs[k];
}

/* One (column, wavelength) system per lane, in two halves.  `w` is the
 * lane's wavelength-table row with stride WS; `scr` its scratch column
 * (stride 64 doubles per entry).  The FP64 drop-in's arithmetic.
 *
 * Forward half: points (the opacity of the cells recomputed in registers),
 * rows and the elimination, alpha staged to its output; every row's gam, u
 * and 1/alpha go to scratch.  Returns true when the system hit the
 * reference's zero-pivot stop (FP64; recorded in `fail`). */
template <bool GREY, bool UNIFORM, int WS, class STAGE>
__device__ static inline bool solve_forward(const oco_sys& d, bool valid, const double* w, double* scr,
                                            unsigned long long* fail, const oco_mtab* M, STAGE* S)
{
    const int nz = d.nz;
    const double dz = d.dz, dz2 = d.dz2;
    constexpr int R = STAGE::rows;
    constexpr bool EM = STAGE::emergent;
    const int lane = lane_id();
    /* alpha staged at j + as: runs on whole lines (line_shift) */
    int as = 0;
    if constexpr (!EM) {
        wave_sync();   /* the previous system's flushes have read the buffers */
        S->base[2][lane] = valid ? d.A : nullptr;
        as = line_shift<R>(d.A, valid);
    }

    /* ---- forward elimination (source/auxiliary.f90:25-41) ----
     * Two-deep pipeline: iteration j eliminates row j (which needs the
     * points j-1, j, j+1, all computed earlier) and evaluates point j+2,
     * whose work is independent of the elimination chain. */
    double rec[CR_NFIELD];
    if (!GREY) load_rec<UNIFORM>(d.cells, 0, rec);
    oco_pt pB = point_sys<GREY, WS>(d, 0, rec, w, M);
    if (!GREY) load_rec<UNIFORM>(d.cells, 1, rec);
    oco_pt pC = point_sys<GREY, WS>(d, 1, rec, w, M);
    if (!GREY) load_rec<UNIFORM>(d.cells, 2, rec);

    double bv, dv, av, cv;
    int fail_stage = 0;
    int fail_row = 1;
    double bet = 0.0, u;
    boundary_row(pB, dz, dz2, bv, dv);  /* row 0: a = 0, c = alpha */
    cv = pB.alpha;
    fail_stage = (bv == 0.0) ? 1 : 0;
    bet = bv;
    u = dv / bet;
    scr_put<alpha_reload<STAGE>()>(scr, 0, 0.0, u, pB.alpha1);
    if constexpr (!EM) {
        stage_put<R>(S->ub, as, pB.alpha);
        if (as == R - 1) {   /* a shifted first run that ends at row 0 */
            wave_sync();
            stage_flush<R, STAGE::nt>(S->ub, S->base[2], 0, nz, as);
            wave_sync();
        }
    }
    double aA = pB.alpha;   /* alpha of row j-1 */
    pB = pC;                /* row j = 1 */
    pC = point_sys<GREY, WS>(d, 2, rec, w, M);

    for (int j = 1; j <= nz - 3; j++) {
        if (!GREY) load_rec<UNIFORM>(d.cells, j + 2, rec);
        /* row j and its elimination step */
        double gam;
        double cprev = cv;
        interior_row(aA, pB, pC.alpha, dz2, av, bv, cv, dv);
        gam = cprev / bet;
        bet = bv - av * gam;
        int first = (bet == 0.0) && (fail_stage == 0);
        fail_row = first ? j + 1 : fail_row;
        fail_stage = first ? 2 : fail_stage;
        u = (dv - av * u) / bet;
        scr_put<alpha_reload<STAGE>()>(scr, j, gam, u, pB.alpha1);
        if constexpr (!EM) {
            stage_put<R>(S->ub, j + as, pB.alpha);
            if (((j + as) & (R - 1)) == R - 1) {
                wave_sync();
                stage_flush<R, STAGE::nt>(S->ub, S->base[2], (j + as) / R, nz, as);
                wave_sync();
            }
        }
        /* point j+2 (independent of the chain above) */
        oco_pt pD = point_sys<GREY, WS>(d, j + 2, rec, w, M);
        aA = pB.alpha;
        pB = pC;
        pC = pD;
    }
    {   /* row nz-2 (interior, no new point) and row nz-1 (boundary) */
        int j = nz - 2;
        double gam;
        double cprev = cv;
        interior_row(aA, pB, pC.alpha, dz2, av, bv, cv, dv);
        gam = cprev / bet;
        bet = bv - av * gam;
        int first = (bet == 0.0) && (fail_stage == 0);
        fail_row = first ? j + 1 : fail_row;
        fail_stage = first ? 2 : fail_stage;
        u = (dv - av * u) / bet;
        scr_put<alpha_reload<STAGE>()>(scr, j, gam, u, pB.alpha1);
        if constexpr (!EM) {
            stage_put<R>(S->ub, j + as, pB.alpha);
            if (((j + as) & (R - 1)) == R - 1) {
                wave_sync();
                stage_flush<R, STAGE::nt>(S->ub, S->base[2], (j + as) / R, nz, as);
                wave_sync();
            }
        }
        j = nz - 1;
        {
            double cprev = cv;
            boundary_row(pC, dz, dz2, bv, dv);
            av = pC.alpha;                       /* a(nz) = alpha, c(nz) = 0 */
            gam = cprev / bet;
            bet = bv - av * gam;
            int first = (bet == 0.0) && (fail_stage == 0);
            fail_row = first ? j + 1 : fail_row;
            fail_stage = first ? 2 : fail_stage;
            u = (dv - av * u) / bet;
        }
        scr_put<alpha_reload<STAGE>()>(scr, j, gam, u, pC.alpha1);
        if constexpr (!EM) {
            stage_put<R>(S->ub, j + as, pC.alpha);
            wave_sync();
            stage_flush<R, STAGE::nt>(S->ub, S->base[2], (j + as) / R, nz, as);   /* last (possibly partial) block */
        }
    }
    (void)u;
    /* the reference stops at this system (source/auxiliary.f90:27-39): its
     * U and V are not written */
    if (fail_stage != 0 && valid) record_fail(fail, d.key, fail_stage, fail_row);
    return fail_stage != 0;
}

/* Back half of one system, from the scratch rows of solve_forward:
 * back-substitution fused with the ghosts, the derivative and V; U and V
 * staged to their outputs.  `valid` false: nothing is stored (a lane past
 * the end, or a system that stopped). */
template <int BC, class STAGE>
__device__ static inline void solve_backward(const oco_sys& d, bool valid, const double* scr, STAGE* S)
{
    const int nz = d.nz;
    const double dz1 = d.dz1;
    constexpr int R = STAGE::rows;
    constexpr bool EM = STAGE::emergent;
    const int lane = lane_id();
    if constexpr (!EM) {
        wave_sync();   /* the forward half's flushes have read the buffers */
        S->base[0][lane] = valid ? d.U : nullptr;
        S->base[1][lane] = valid ? d.V : nullptr;
    }
    /* Runs in whole lines: U(e) is staged at e + us and V(z) (with the
     * alpha it replaces) at z + vs, so each flushed run of R doubles starts
     * on an R-double boundary of memory, when every valid lane's row has the
     * same offset in its line (the wavefront's systems are one column's
     * consecutive wavelengths, ncol (nz + 6) or ncol nz doubles apart);
     * otherwise 0.  Measured on MI355X, FP64 drop-in, 256 x 256 columns,
     * nw = 128: nz = 512 (U rows of 518 doubles) 74.2 -> 71.1 ms; nz = 500
     * and 504, whose V and alpha rows are not whole lines either, 74.3 ->
     * 69.0 and 74.0 -> 70.5 ms (nz = 496: 68.2-70.4 ms). */
    int us = 0, vs = 0;
    if constexpr (!EM) {
        us = line_shift<R>(d.U, valid);
        vs = line_shift<R>(d.V, valid, alpha_reload<STAGE>() == OCO_A1_STAGED ? d.A : nullptr);
    }

    /* ---- back-substitution (:43-45) fused with update_ghosts (:49-62),
     *      der (:64-78) and calc_flux (:96-109) ---- */
    const double c60 = 1.0 / 60.0;
    /* window w0..w6 = extended U at e = j .. j+6 (e = z + 3) */
    double gnext, a1q0, Un;              /* gam of the row above; 1/alpha of the V row pending */
    constexpr int AR = alpha_reload<STAGE>();
    constexpr bool ST = AR == OCO_A1_STAGED && !EM;
    /* OCO_A1_STAGED: the V stage's slot of row e holds alpha(e) until V(e)
     * replaces it; apf holds the next block's alpha, in flight */
    double apf[ST ? R : 1];
    if constexpr (ST) {
        const int bt = (nz - 1 + vs) / R;
        alpha_fetch<R>(apf, S->base[2], bt, nz, vs);
        alpha_park<R>(S->vb, apf);
        if (bt >= 1) alpha_fetch<R>(apf, S->base[2], bt - 1, nz, vs);
        wave_sync();
    }
    /* 1/alpha of V row e (a1p: the register pipeline's value) */
    auto a1v = [&](int e, double a1p) -> double {
        if constexpr (ST) return 1.0 / S->vb[lane * (R + 1) + ((e + vs) & (R - 1))];
        else return a1p;
    };
    /* V block b has been flushed: park block b - 1's alpha, fetch b - 2's */
    auto vflushed = [&](int b) {
        if constexpr (ST) {
            if (b >= 1) {
                alpha_park<R>(S->vb, apf);
                wave_sync();
                if (b >= 2) alpha_fetch<R>(apf, S->base[2], b - 2, nz, vs);
            }
        }
    };
    scr_get<AR>(scr, nz - 1, gnext, Un, a1q0, d.A);   /* U(nz) = u(nz) (:43) */
    a1q0 = a1_of<AR>(a1q0);
    double a1q1 = 0.0, a1q2 = 0.0, a1q3 = 0.0;
    double t[4];
    t[0] = Un;
    for (int k = 1; k < 4; k++) {
        int j = nz - 1 - k;
        double gj, uj, a1;
        scr_get<AR>(scr, j, gj, uj, a1, d.A);
        double Uj = uj - gnext * t[k - 1];
        gnext = gj;
        a1q3 = a1q2; a1q2 = a1q1; a1q1 = a1q0; a1q0 = a1_of<AR>(a1);
        t[k] = Uj;
    }
    /* top ghosts (f(n2+i) = 2 f(n2) - f(n2-i)) */
    double g1 = 2.0 * t[0] - t[1], g2 = 2.0 * t[0] - t[2], g3 = 2.0 * t[0] - t[3];
    /* U(e) for e = nz+5 down to nz-1, in descending order: a block is
     * flushed when its lowest element is staged */
    if constexpr (!EM) {
        const double top[7] = {g3, g2, g1, t[0], t[1], t[2], t[3]};
#pragma unroll
        for (int k = 0; k < 7; k++) {
            const int e = nz + 5 - k;
            stage_put<R>(S->ub, e + us, top[k]);
            if (((e + us) & (R - 1)) == 0) {
                wave_sync();
                stage_flush<R, STAGE::nt>(S->ub, S->base[0], (e + us) / R, nz + 6, us);
                wave_sync();
            }
        }
    }
    double w0 = t[3], w1 = t[2], w2 = t[1], w3 = t[0], w4 = g1, w5 = g2, w6 = g3;
    {
        double df = c60 * (((45.0 * (w4 - w2)) - (9.0 * (w5 - w1))) + (w6 - w0));
        const double vtop = ((-1.0 * a1v(nz - 1, a1q3)) * df) * dz1;
        if constexpr (EM) {
            if (valid) *gptr(d.V + 1) = vtop;
        } else {
            stage_put<R>(S->vb, nz - 1 + vs, vtop);
            if (((nz - 1 + vs) & (R - 1)) == 0) {
                wave_sync();
                stage_flush<R, STAGE::nt>(S->vb, S->base[1], (nz - 1 + vs) / R, nz, vs);
                wave_sync();
                vflushed((nz - 1 + vs) / R);
            }
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
            scr_get<AR>(scr, j < 0 ? 0 : j, cg[k], cu[k], ca[k], d.A);
        }
        for (; jc >= 0; jc -= BC) {
#pragma unroll
            for (int k = 0; k < BC; k++) {      /* next chunk, in flight */
                int j = jc - BC - k;
                scr_get<AR>(scr, j < 0 ? 0 : j, ng[k], nu[k], na[k], d.A);
            }
#pragma unroll
            for (int k = 0; k < BC; k++) {
                int j = jc - k;
                if (j >= 0) {
                    double Uj = cu[k] - gnext * w0;
                    gnext = cg[k];
                    a1q3 = a1q2; a1q2 = a1q1; a1q1 = a1q0; a1q0 = a1_of<AR>(ca[k]);
                    const int e = j + 3;   /* U(e) and V(z = e) */
                    w6 = w5; w5 = w4; w4 = w3; w3 = w2; w2 = w1; w1 = w0; w0 = Uj;
                    if constexpr (!EM) {
                        stage_put<R>(S->ub, e + us, Uj);
                        double df = c60 * (((45.0 * (w4 - w2)) - (9.0 * (w5 - w1))) + (w6 - w0));
                        stage_put<R>(S->vb, e + vs, ((-1.0 * a1v(e, a1q3)) * df) * dz1);
                        const bool ue = ((e + us) & (R - 1)) == 0, ve = ((e + vs) & (R - 1)) == 0;
                        if (ue || ve) {
                            wave_sync();
                            if (ue) stage_flush<R, STAGE::nt>(S->ub, S->base[0], (e + us) / R, nz + 6, us);
                            if (ve) stage_flush<R, STAGE::nt>(S->vb, S->base[1], (e + vs) / R, nz, vs);
                            wave_sync();
                            if (ve) vflushed((e + vs) / R);
                        }
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < BC; k++) { cg[k] = ng[k]; cu[k] = nu[k]; ca[k] = na[k]; }
        }
    }
    /* bottom ghosts (f(n1-i) = 2 f(n1) - f(n1+i)); window now e = 3..9 */
    double u0 = w0, G1 = 2.0 * u0 - w1, G2 = 2.0 * u0 - w2, G3 = 2.0 * u0 - w3;
    if constexpr (!EM) {
        const double gv[3] = {G1, G2, G3};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int e = 2 - k;
            stage_put<R>(S->ub, e + us, gv[k]);
            if (e + us >= R && ((e + us) & (R - 1)) == 0) {   /* a shifted block that ends here */
                wave_sync();
                stage_flush<R, STAGE::nt>(S->ub, S->base[0], (e + us) / R, nz + 6, us);
                wave_sync();
            }
        }
    }
    const double gh[3] = {G1, G2, G3};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        w6 = w5; w5 = w4; w4 = w3; w3 = w2; w2 = w1; w1 = w0; w0 = gh[k];
        a1q3 = a1q2; a1q2 = a1q1; a1q1 = a1q0;
        if (!EM || k == 2) {
            double df = c60 * (((45.0 * (w4 - w2)) - (9.0 * (w5 - w1))) + (w6 - w0));
            const double v = ((-1.0 * a1v(2 - k, a1q3)) * df) * dz1;
            if constexpr (EM) {
                if (valid) *gptr(d.V) = v;
            } else {
                const int z = 2 - k;
                stage_put<R>(S->vb, z + vs, v);
                if (z + vs >= R && ((z + vs) & (R - 1)) == 0) {   /* a shifted block that ends here */
                    wave_sync();
                    stage_flush<R, STAGE::nt>(S->vb, S->base[1], (z + vs) / R, nz, vs);
                    wave_sync();
                    vflushed((z + vs) / R);
                }
            }
        }
    }
    if constexpr (!EM) {
        wave_sync();
        stage_flush<R, STAGE::nt>(S->ub, S->base[0], 0, nz + 6, us);
        stage_flush<R, STAGE::nt>(S->vb, S->base[1], 0, nz, vs);
    }
}

/* One (column, wavelength) system, one lane: the two halves in turn. */
template <bool GREY, bool UNIFORM, int WS, int BC, class STAGE>
__device__ static inline void solve_system(const oco_sys& d, bool valid, const double* w, double* scr,
                                           unsigned long long* fail, const oco_mtab* M, STAGE* S)
{
    const bool stopped = solve_forward<GREY, UNIFORM, WS, STAGE>(d, valid, w, scr, fail, M, S);
    /* the back pass reads alpha where the forward pass stored it, through
     * other lanes of this wavefront (the staged flushes): a workgroup-scope
     * fence makes those stores visible to this workgroup's loads */
    if constexpr (alpha_reload<STAGE>()) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    solve_backward<BC, STAGE>(d, valid && !stopped, scr, S);
}

/* One wavefront per workgroup, grid-stride over tiles of 64 systems; the
 * lane's wavelength-table row lives in registers. */
/* The 64 systems of tile `tile` (64 consecutive systems in (column,
 * wavelength) order): this lane's system, its wavelength and whether it
 * exists (lanes past the end run a duplicate of the last system and store
 * nothing). */
template <bool UNIFORM>
__device__ static inline oco_sys tile_system(const oco_sweep_args& a, long tile, int lane, bool& valid, int& iw)
{
    const int nw = a.nw;
    const long nsys = a.ncol * nw;
    const long sys0 = tile * OCO_WAVE;
    long col;
    if (UNIFORM) {
        col = sys0 / nw;
        iw = (int)(sys0 % nw) + lane;
    } else {
        long s = sys0 + lane;
        if (s >= nsys) s = nsys - 1;
        col = s / nw;
        iw = (int)(s % nw);
    }
    valid = UNIFORM || sys0 + lane < nsys;
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
    return d;
}

/* One wavefront per workgroup, grid-stride over tiles of 64 systems; the
 * lane's wavelength-table row lives in registers. */
/* Work queues of the sweeps.  Wavefronts doing the same work progress at
 * rates up to twice apart (they share SIMDs, L2 and memory channels
 * unevenly), so a fixed share per wavefront leaves the grid waiting on its
 * slowest; instead each wavefront takes its next ticket from a counter when
 * it is done with the last, and the grid drains together.  Counters are
 * cleared by tickets_reset before every sweep launch. */
__global__ void tickets_reset(unsigned long long* t)
{
    if (threadIdx.x < OCO_TICKET_QUEUES) t[threadIdx.x * OCO_TICKET_STRIDE] = 0ull;
}

/* The first of the next n items of a queue, the same in every lane of the
 * calling wavefront. */
__device__ __forceinline__ long wave_take(unsigned long long* ctr, unsigned long long n)
{
    unsigned long long v = 0ull;
    if (lane_id() == 0) v = atomicAdd(ctr, n);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return (long)(((unsigned long long)hi << 32) | lo);
}

template <bool GREY, bool UNIFORM, class STAGE>
__device__ static inline void sweep_tiles(const oco_sweep_args& a, const oco_mtab* M, STAGE* S)
{
    const int lane = lane_id();
    const long ntiles = (a.ncol * a.nw + OCO_WAVE - 1) / OCO_WAVE;
    /* one scratch slot per wavefront of the grid */
    const long slot = (long)blockIdx.x * (blockDim.x / OCO_WAVE) + threadIdx.x / OCO_WAVE;
    double* scr = a.scratch + slot * a.nz * 3 * OCO_WAVE + lane;
    /* one queue of tiles in order: the tiles in flight stay neighbours */
    for (long tile = wave_take(a.tickets, 1); tile < ntiles; tile = wave_take(a.tickets, 1)) {
        bool valid;
        int iw;
        const oco_sys d = tile_system<UNIFORM>(a, tile, lane, valid, iw);
        double w[OCO_WT_STRIDE];
        if (!GREY) {
#pragma unroll
            for (int k = 0; k < OCO_WT_STRIDE; k++) w[k] = a.wtab[iw * OCO_WT_STRIDE + k];
        }
        solve_system<GREY, UNIFORM, 1, OCO_BACK_CHUNK, STAGE>(d, valid, w, scr, a.fail, M, S);
    }
}

/* FP64 drop-in sweep.  UNIFORM: nw is a multiple of 64, so a wavefront's
 * systems share one column.  OCO_SWEEP_WG wavefronts per workgroup, each
 * taking its own tiles, share one LDS copy of the math tables: with four,
 * two workgroups (8 wavefronts, the register limit) fit a CU's 160 KB of LDS,
 * where one-wavefront workgroups (24 KB each) fit 6. */
#ifndef OCO_SWEEP_WG
#define OCO_SWEEP_WG 4
#endif
template <bool GREY, bool UNIFORM>
__global__ __launch_bounds__(OCO_WAVE * OCO_SWEEP_WG) OCO_SWEEP_ATTR void sweep_fp64(const oco_sweep_args a)
{
    __shared__ oco_mtab smt;
    __shared__ oco_stage stg[OCO_SWEEP_WG];
    load_mtab(&smt);
    sweep_tiles<GREY, UNIFORM>(a, &smt, &stg[threadIdx.x / OCO_WAVE]);
}

/* First-pillar canaries (source/flux_feautrier.f90:198-201): sum(a), sum(b),
 * sum(c), sum(d) of system (ix,iy,iw) = (1,1,1), from the FP64 opacity in
 * either precision mode (the cell records are recomputed here, so the kernel
 * does not depend on the record layout of the sweep).  Printed diagnostics
 * only (Fortran leaves the summation order of SUM to the compiler), so points,
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
        if (!GREY) oco_cell_record(a.rho[z], a.T[z], a.fully_ionized_T, &smt, rec);
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

/* Columns of 2 or 3 cells (FP64 drop-in): one thread per (column, wavelength)
 * system, grid-stride, the whole system in registers.  The same point,
 * row and elimination code as the lane sweep, in the reference's order:
 * get_tridag_coefficients (source/auxiliary.f90:111-201; nz = 2 is two
 * boundary rows), tridag (:14-47) with its zero-pivot stops, update_ghosts
 * (:49-62) in its interleaved order (with fewer than 4 cells its reads
 * f(n1+i), f(n2-i) reach ghosts the same loop set earlier), der (:64-78)
 * and calc_flux (:96-109).  Outputs are stored directly: a system is a few
 * dozen bytes. */
#define OCO_TINY_NZ 3
template <bool GREY>
__global__ __launch_bounds__(256) void sweep_tiny_fp64(const oco_sweep_args a)
{
    __shared__ oco_mtab smt;
    load_mtab(&smt);
    const int nz = a.nz;
    const long nsys = a.ncol * a.nw;
    for (long s = (long)blockIdx.x * blockDim.x + threadIdx.x; s < nsys; s += (long)gridDim.x * blockDim.x) {
        const long col = s / a.nw;
        const int iw = (int)(s - col * a.nw);
        double w[OCO_WT_STRIDE];
        if (!GREY) {
#pragma unroll
            for (int k = 0; k < OCO_WT_STRIDE; k++) w[k] = a.wtab[(long)iw * OCO_WT_STRIDE + k];
        }
        oco_pt p[OCO_TINY_NZ];
#pragma unroll
        for (int z = 0; z < OCO_TINY_NZ; z++) {
            if (z < nz) {
                double rec[CR_NFIELD];
                if (!GREY) load_record<false>(a, col, z, rec);
                p[z] = point_rec<GREY, false>(a, col, z, rec, w, &smt);
            }
        }
        double av[OCO_TINY_NZ], bv[OCO_TINY_NZ], cv[OCO_TINY_NZ], dv[OCO_TINY_NZ];
        boundary_row(p[0], a.dz, a.dz2, bv[0], dv[0]);
        av[0] = 0.0;
        cv[0] = p[0].alpha;
        if (nz == 3) interior_row(p[0].alpha, p[1], p[2].alpha, a.dz2, av[1], bv[1], cv[1], dv[1]);
        const int n = nz - 1;
        boundary_row(p[n], a.dz, a.dz2, bv[n], dv[n]);
        av[n] = p[n].alpha;
        cv[n] = 0.0;
        double u[OCO_TINY_NZ], gam[OCO_TINY_NZ];
        int stage = 0, row = 0;
        double bet = bv[0];
        if (bet == 0.0) {
            stage = 1;
            row = 1;
        } else {
            u[0] = dv[0] / bet;
            for (int j = 1; j < nz; j++) {
                gam[j] = cv[j - 1] / bet;
                bet = bv[j] - av[j] * gam[j];
                if (bet == 0.0) {
                    stage = 2;
                    row = j + 1;
                    break;
                }
                u[j] = (dv[j] - av[j] * u[j - 1]) / bet;
            }
        }
        const long key = col * a.nw + iw;
        if (stage != 0) {   /* the reference stops at this system: nothing of it is stored */
            record_fail(a.fail, key, stage, row);
            continue;
        }
        for (int j = nz - 2; j >= 0; j--) u[j] = u[j] - gam[j + 1] * u[j + 1];
        double f[OCO_TINY_NZ + 6];
#pragma unroll
        for (int j = 0; j < OCO_TINY_NZ; j++)
            if (j < nz) f[3 + j] = u[j];
        const int n1 = 3, n2 = 3 + nz - 1;
        for (int i = 1; i <= 3; i++) {
            f[n1 - i] = 2.0 * f[n1] - f[n1 + i];
            f[n2 + i] = 2.0 * f[n2] - f[n2 - i];
        }
        const long osys = (long)iw * a.ncol + col;
        double* U = a.U + osys * (nz + 6);
        double* V = a.V + osys * nz;
        double* A = a.A + osys * nz;
        for (int e = 0; e < nz + 6; e++) U[e] = f[e];
        const double c60 = 1.0 / 60.0;
        for (int j = 0; j < nz; j++) {
            const double* g = f + n1 + j;
            const double df = c60 * (((45.0 * (g[1] - g[-1])) - (9.0 * (g[2] - g[-2]))) + (g[3] - g[-3]));
            V[j] = ((-1.0 * p[j].alpha1) * df) * a.dz1;
            A[j] = p[j].alpha;
        }
    }
}

/* ---- mixed precision, one wavefront per system ----
 * The lane-per-system sweep keeps every system's elimination state (gam, u,
 * 1/alpha per row) in an HBM scratch slot: 48 bytes per point of extra
 * traffic.  The mixed mode is not bound to the reference's elimination
 * order, so here a whole wavefront solves one (column, wavelength) system
 * with its state in LDS and registers, in the Rybicki-Hummer difference form
 * (Rybicki & Hummer 1991, A&A 245, 171, appendix A): every row is written
 *     -A x[i-1] + (A + C + H) x[i] - C x[i+1] = R,   A, C, H, R >= 0,
 * with H the row's excess over A + C (zeta; at the boundary rows
 * alpha^2 (dz + ze)).  The elimination never forms A + C + H and subtracts,
 * so it keeps full precision where H << A + C (optically thin cells, where
 * the reference's b = -((a + c) + zeta) rounds zeta away).
 *   1. points, lane z mod 64: opacity from the column's field-major cell
 *      records (coalesced loads); alpha, albedo and S go to LDS, alpha also
 *      to the output;
 *   2. lane l takes m consecutive rows (m odd: chunk strides hit distinct
 *      LDS banks) and eliminates them keeping the previous chunk's last
 *      unknown xL as a parameter: x[i] = D x[i+1] + W xL + Z, with the
 *      deficit E = 1 - D - W carried as its own positive quantity;
 *      products along the chunk give its first row in terms of xL and the
 *      chunk's last unknown;
 *   3. the chunks' last unknowns G satisfy a 64-row system of the same form,
 *      solved across lanes by parallel cyclic reduction in difference form
 *      (H' = H + k1 H[-s] + k2 H[+s]: no subtraction at any step);
 *   4. back-substitution inside each chunk, ghosts, the 6th-order derivative
 *      and V, with every output store coalesced.
 * LDS: alpha [nz], albedo [nz], S -> U [nz + 6]; the chunk's D, W, Z stay in
 * registers (MM rows at most). */
/* Word 3 of a raw buffer descriptor on gfx9-family GPUs (32-bit data
 * format; the layout of the public ISA documentation). */
#define OCO_BUF_DWORD3 0x00020000

/* v of lane (byte address addr / 4) of the wavefront: a rotation when addr
 * is ((lane -+ s) & 63) * 4. */
__device__ static inline double lane_get(double v, int addr)
{
    const int2 w = __builtin_bit_cast(int2, v);
    int2 r;
    r.x = __builtin_amdgcn_ds_bpermute(addr, w.x);
    r.y = __builtin_amdgcn_ds_bpermute(addr, w.y);
    return __builtin_bit_cast(double, r);
}

/* v from the neighbouring lane without an LDS round trip: lane - 1 (0 in
 * lane 0) and lane + 1 (0 in lane 63) by DPP whole-wavefront shifts */
__device__ static inline double lane_up1(double v)
{
    const int2 w = __builtin_bit_cast(int2, v);
    int2 r;
    r.x = __builtin_amdgcn_mov_dpp(w.x, 0x138, 0xf, 0xf, true);   /* wave_shr:1, bound_ctrl */
    r.y = __builtin_amdgcn_mov_dpp(w.y, 0x138, 0xf, 0xf, true);
    return __builtin_bit_cast(double, r);
}
__device__ static inline double lane_dn1(double v)
{
    const int2 w = __builtin_bit_cast(int2, v);
    int2 r;
    r.x = __builtin_amdgcn_mov_dpp(w.x, 0x130, 0xf, 0xf, true);   /* wave_shl:1, bound_ctrl */
    r.y = __builtin_amdgcn_mov_dpp(w.y, 0x130, 0xf, 0xf, true);
    return __builtin_bit_cast(double, r);
}

/* v across the two halves of the wavefront (v_permlane32_swap): up holds
 * lane - 32 in lanes 32..63, dn lane + 32 in lanes 0..31; the other half of
 * each holds the lane's own v */
__device__ static inline void lane_swap32(double v, double& up, double& dn)
{
    const int2 w = __builtin_bit_cast(int2, v);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)w.x, (unsigned)w.x, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)w.y, (unsigned)w.y, false, false);
    up = __builtin_bit_cast(double, int2{(int)lo[0], (int)hi[0]});
    dn = __builtin_bit_cast(double, int2{(int)lo[1], (int)hi[1]});
}

/* rows per lane of the wavefront sweep with L lanes per system for columns
 * of nz cells (odd: the lanes' chunk strides then hit distinct LDS banks) */
__host__ __device__ static inline int wave_rows(int nz, int L = OCO_WAVE)
{
    return ((nz + L - 1) / L) | 1;
}

/* Words of LDS a system of the wavefront sweep takes: alpha [nz], albedo
 * [nz], S -> U with ghosts [nz + 6], and S of the three cells at each
 * boundary [6], kept for the emergent flux after U has replaced S. */
__host__ __device__ static inline long wave_lds_words(int nz) { return 3L * nz + 12; }

/* Point kinds of the difference-form sweeps (OCO_PK_*): the mixed-precision
 * opacity (oco_point_mixed, OCO_MODE_MIXED), the drop-in's FP64 opacity
 * (oco_point, the same alpha, albedo and S bit for bit, OCO_MODE_ACCURATE),
 * and the grey parameters from rho and T (oco_grey_point, both modes). */

/* 1. points of nk <= S systems of one column, S consecutive wavelengths iw0
 * .. iw0 + nk - 1 (source/continuous_opacity.f90:175-201, or :377-390 grey):
 * each cell's record is loaded once per part of the workgroup and evaluated
 * for that part's wavelengths, into system k's LDS region (alpha, albedo,
 * S; regions lstr >= wave_lds_words(nz) doubles apart) and the alpha output
 * of system k.  The K wavefronts split the cells when a column has at least
 * 64 K / 2 of them (one part: thread t takes cells t, t + 64 K, ...);
 * shorter columns need only W < K wavefronts per pass over their cells, so
 * the workgroup splits into P = K / W parts, part p taking wavelengths p, p
 * + P, ...  A wavefront's wavelengths are the same for all its lanes, so its
 * table rows are scalar loads.  The column's field-major records (field k of
 * cell z at cc[k*nz + z]) are read through a buffer descriptor: the field
 * offset is a scalar, the lane offset a vector register, no per-field
 * address arithmetic.  Grey columns read rho at cc[z] and T at cT[z]. */
template <int K, int PK>
__device__ static inline void wave_points(const double* cc, const double* cT, int nz, const double* wtab, int iw0,
                                          int nk, double* A0, long astride, double* lds, long lstr,
                                          double sigma_grey, const oco_mtab* M)
{
    const __amdgpu_buffer_rsrc_t crs =
        __builtin_amdgcn_make_buffer_rsrc((void*)cc, (short)0, CR_NFIELD * nz * (int)sizeof(double), OCO_BUF_DWORD3);
    int W = 1;
    while (W < K && W * OCO_WAVE * 2 <= nz + OCO_WAVE - 1) W <<= 1;   /* wavefronts per pass over the cells */
    const int wid = __builtin_amdgcn_readfirstlane((int)threadIdx.x / OCO_WAVE);
    const int P = K / W, part = wid / W;
    const int t0 = (wid - part * W) * OCO_WAVE + lane_id();
#ifdef OCO_WAVE_POINTS_UNROLL
#pragma unroll OCO_WAVE_POINTS_UNROLL
#endif
    for (int z = t0; z < nz; z += OCO_WAVE * W) {
        double rec[CR_NFIELD];
        if constexpr (PK == OCO_PK_GREY) {
            rec[0] = cc[z];
            rec[1] = cT[z];
        } else {
            const int zo = z * (int)sizeof(double);
#pragma unroll
            for (int k = 0; k < CR_NFIELD; k++)
                rec[k] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(crs, zo,
                                                                                        k * nz * (int)sizeof(double), 0));
        }
        for (int k = part; k < nk; k += P) {
            oco_pt p;
            if constexpr (PK == OCO_PK_GREY) {
                p = oco_grey_point(rec[0], rec[1], sigma_grey, M);
            } else {
                double w[OCO_WT_STRIDE];
                oco_cptr wp = (oco_cptr)(wtab + (long)(iw0 + k) * OCO_WT_STRIDE);
#pragma unroll
                for (int f = 0; f < OCO_WT_STRIDE; f++) w[f] = wp[f];
                if constexpr (PK == OCO_PK_MIXED) p = oco_point_mixed<1>(rec, w);
                else p = oco_point(rec, w, M);
            }
            double* L = lds + k * lstr;
            L[z] = p.alpha;
            L[nz + z] = p.albedo;
            L[2 * nz + z + 3] = p.S;
            out_store<OCO_WAVE_NT>(A0 + k * astride + z, p.alpha);
        }
    }
}

/* A lane's chunk of rows in the wavefront-per-system solve (difference form):
 * x[i] = D x[i+1] + W xL + Z for the chunk's rows, xL the previous chunk's last
 * unknown, the deficit E = 1 - D - W carried positively. */
struct rh_chunk {
    double Dp, Wp, Zp, Ep;   /* the latest row's D, W, Z, E (the parameter row x = xL: -, 1, 0, 0) */
    double pr, sD, sE;       /* the chunk's first row: pr G + (..) xL + sD, deficit sE */
    double DL, WL, ZL, EL;   /* its last row (rh_chunk_close) */
    double mlo;              /* mean alpha of the interface below the current row */
};

__device__ static inline void rh_chunk_init(rh_chunk& c)
{
    c.Dp = 0.0; c.Wp = 1.0; c.Zp = 0.0; c.Ep = 0.0;
    c.pr = 1.0; c.sD = 0.0; c.sE = 0.0;
    c.DL = 0.0; c.WL = 0.0; c.ZL = 0.0; c.EL = 0.0;
    c.mlo = 0.0;
}

/* Row z of the system (source/auxiliary.f90:147-197 in difference form) and
 * its elimination step: al, om, S the cell's point, anext the alpha of cell
 * z + 1 (read for z < nz - 1), first whether z is the chunk's first row.  The
 * first-row summary takes in every row but the chunk's last: each row adds
 * the previous one's, so the test is on the first row (the same for every
 * lane of the wavefront) rather than the last, and the last row's values are
 * the state left behind (rh_chunk_close).  An interior row is scaled by the
 * product of its two interface means m-, m+ (A = alpha^2 m+, C = alpha^2 m-,
 * H and R times m- m+), which takes the two reciprocals out of the row; D, W,
 * Z and E are ratios within one row, so the scale leaves them as they are.
 * The mean below is the one above of the previous row; a chunk starting above
 * row 0 sets c.mlo from its first cell and the cell below before its first
 * row.  D, W, Z: the row's coefficients of x[z+1], xL and 1. */
__device__ static inline void rh_chunk_row(rh_chunk& c, int z, int nz, bool first, double al, double om, double S,
                                           double anext, double dz, double dz2, double& D, double& W, double& Z)
{
#pragma clang fp contract(on)   /* mixed mode: error bounds, not the reference's bits */
    if (!first) {
        c.sD += c.pr * c.Zp;
        c.sE += c.pr * c.Ep;
        c.pr *= c.Dp;
    }
    const double al2 = al * al;
    double Ac, Cc, H, R;
    if (z == 0 || z == nz - 1) {
        oco_pt pc;
        pc.alpha = al; pc.albedo = om; pc.S = S; pc.alpha2 = al2; pc.alpha1 = 0.0;
        rh_boundary_row(pc, dz, dz2, H, R);
        Ac = z == 0 ? 0.0 : al;
        Cc = z == 0 ? al : 0.0;
        if (z == 0) c.mlo = 0.5 * (anext + al);
    } else {
        const double mhi = 0.5 * (anext + al);
        Ac = al2 * mhi;
        Cc = al2 * c.mlo;
        H = ((dz2 * (al2 * al)) * (1.0 - om)) * (c.mlo * mhi);
        R = S * H;
        c.mlo = mhi;
    }
    const double r = mx_rcp((Ac * (c.Wp + c.Ep) + Cc) + H);   /* 1 - D[i-1] = W + E */
    D = Cc * r;
    W = (Ac * c.Wp) * r;
    Z = (R + Ac * c.Zp) * r;
    c.Ep = (Ac * c.Ep + H) * r;
    c.Dp = D; c.Wp = W; c.Zp = Z;
}

/* after the chunk's last row */
__device__ static inline void rh_chunk_close(rh_chunk& c)
{
    c.DL = c.Dp; c.WL = c.Wp; c.ZL = c.Zp; c.EL = c.Ep;
}

/* The reduced system in the chunks' last unknowns, all lanes of the wavefront:
 * -WL G[k-1] + (WL + DL Gn + EL + DL En) G[k] - DL Gn G[k+1] = ZL + DL Dn
 * (Gn, Dn, En: the next chunk's first row), by parallel cyclic reduction in
 * difference form (H' = H + k1 H[-s] + k2 H[+s]: no subtraction).  Neighbours
 * are read by rotation, without masking the lanes whose neighbour does not
 * exist: those lanes' couplings are exactly zero (the first chunk's W, the
 * last row's D, and after each reduction step A = 0 for lanes < s, C = 0 for
 * lanes > 63 - s), and every row's A + C + H is positive and finite, so the
 * wrapped-around values are multiplied by zero.  Empty chunks are identity
 * rows.  Each step, every lane divides its row by its own diagonal
 * A + C + H and shares the quotients: the neighbour above needs a, h, r and
 * the one below c, h, r (k1 A- = A a-, k1 H- = A h-, ...), so one reciprocal
 * and six shuffled doubles per step.  Shifts by one lane are DPP moves and
 * the step across the halves one v_permlane32_swap: where these give a lane
 * 0 or its own value instead of the wrapped-around neighbour's, its
 * coupling to that neighbour is exactly zero as well (before the last step,
 * A = 0 in lanes 0..31 and C = 0 in lanes 32..63).  G: this chunk's last
 * unknown, GL the previous chunk's.  With L < 64 lanes per system the
 * wavefront holds 64 / L systems in aligned groups of L lanes and only the
 * steps below L run; the shifts and rotations still cross the groups, where
 * the couplings are exactly zero for the same reason (a system's first chunk
 * has W = 0, its last D = 0, and an absent system is identity rows). */
template <int L>
__device__ static inline void rh_reduced(const rh_chunk& c, bool rows, double& G, double& GL)
{
#pragma clang fp contract(on)
    const int lane = lane_id();
    const double Gn = lane_dn1(c.pr), Dn = lane_dn1(c.sD), En = lane_dn1(c.sE);
    double Ar = 0.0, Cr = 0.0, Hr = 1.0, Rr = 0.0;
    if (rows) {
        Ar = c.WL;
        Cr = c.DL * Gn;
        Hr = c.EL + c.DL * En;
        Rr = c.ZL + c.DL * Dn;
    }
#pragma unroll
    for (int sft = 1; sft < L; sft <<= 1) {
        const int up = ((lane - sft) & (OCO_WAVE - 1)) * 4, dn = ((lane + sft) & (OCO_WAVE - 1)) * 4;
        const double di = mx_rcp((Ar + Cr) + Hr);
        const double an = Ar * di, cn = Cr * di, hn = Hr * di, rn = Rr * di;
        double am, hm, rm, cp, hp, rp;
        if (sft == 1) {
            am = lane_up1(an); hm = lane_up1(hn); rm = lane_up1(rn);
            cp = lane_dn1(cn); hp = lane_dn1(hn); rp = lane_dn1(rn);
        } else if (L == OCO_WAVE && sft == OCO_WAVE / 2) {
            double x;
            lane_swap32(an, am, x);
            lane_swap32(cn, x, cp);
            lane_swap32(hn, hm, hp);
            lane_swap32(rn, rm, rp);
        } else {
            am = lane_get(an, up); hm = lane_get(hn, up); rm = lane_get(rn, up);
            cp = lane_get(cn, dn); hp = lane_get(hn, dn); rp = lane_get(rn, dn);
        }
        Hr = (Hr + Ar * hm) + Cr * hp;
        Rr = (Rr + Ar * rm) + Cr * rp;
        Ar = Ar * am;
        Cr = Cr * cp;
    }
    G = Rr * mx_rcp((Ar + Cr) + Hr);
    GL = lane_up1(G);
}

/* x[z] from x[z+1] and the previous chunk's last unknown */
__device__ static inline double rh_back(double D, double W, double Z, double x, double GL)
{
#pragma clang fp contract(on)
    return (D * x + W * GL) + Z;
}

/* 6th-order derivative at e from U(e-3 .. e+3) (source/auxiliary.f90:64-78) */
__device__ static inline double der7(double um3, double um2, double um1, double up1, double up2, double up3)
{
#pragma clang fp contract(on)
    const double c60 = 1.0 / 60.0;
    return c60 * (((45.0 * (up1 - um1)) - (9.0 * (up2 - um2))) + (up3 - um3));
}

/* calc_flux (source/auxiliary.f90:96-109) of a cell from its alpha and der */
__device__ static inline double flux_v(double al, double df, double dz1)
{
    return ((-1.0 * mx_rcp(al)) * df) * dz1;
}

/* ghost f(n -+ i) = 2 f(n) - f(n +- i) (source/auxiliary.f90:49-62) */
__device__ static inline double ghost(double fn, double fi)
{
#pragma clang fp contract(on)
    return 2.0 * fn - fi;
}

/* The derivative der of the difference-form solves at a boundary cell,
 * counted inward, without differences of U.  With the ghosts of
 * source/auxiliary.f90:49-62, der (:64-78) at a boundary cell is
 *     der = (37 d0 - 8 d1 + d2) / 30,   dk = U(k + 1) - U(k),
 * cells k = 0, 1, 2, 3 counted inward from the boundary (at the top, der is
 * minus this).  On an optically thin cell the U agree to ~13 digits, so dk
 * formed from them keeps few: the emergent flux of an FP64 U rounded to the
 * last bit is off by up to 3e-3 on a 1e-15 g/cm^3, 2000 K floor.  The rows
 * give each difference without that subtraction, from U and S, which are
 * accurate to full precision:
 *   boundary row (:187-197):  d0 = a0 (dz U0 + ze (U0 - S0)),
 *                             ze = a0 dz^2 (1 - w0) / 4;
 *   interior row (:159-165), c (U(k+1) - U(k)) - a (U(k) - U(k-1)) = zeta (U(k) - S(k)):
 *     dk = kap(k+1/2) (d(k-1) / kap(k-1/2) + dz^2 ak (1 - wk) (Uk - Sk)),
 * kap the interface means of alpha.  In a thick cell (alpha dz > 1), U is
 * close to S and Uk - Sk is the smaller, less accurate difference, so there
 * dk = U(k+1) - U(k) directly.  Both are the same der in exact arithmetic;
 * measured on C4-profile columns, V from this is within 1e-11 of the exact
 * solution where der(U) in FP64 is off by 3e-3. */
__device__ static inline double rh_boundary_der(double u0, double u1, double u2, double u3, double a0, double a1,
                                                double a2, double a3, double w0, double w1, double w2, double s0,
                                                double s1, double s2, double dz, double dz2)
{
#pragma clang fp contract(on)
    const double ze = ((a0 * dz2) * (1.0 - w0)) * 0.25;
    const double k01 = 0.5 * (a0 + a1), k12 = 0.5 * (a1 + a2), k23 = 0.5 * (a2 + a3);
    const double d0 = a0 * dz <= 1.0 ? a0 * (dz * u0 + ze * (u0 - s0)) : u1 - u0;
    const double d1 = a1 * dz <= 1.0 ? k12 * (d0 * mx_rcp(k01) + ((dz2 * a1) * (1.0 - w1)) * (u1 - s1)) : u2 - u1;
    const double d2 = a2 * dz <= 1.0 ? k23 * (d1 * mx_rcp(k12) + ((dz2 * a2) * (1.0 - w2)) * (u2 - s2)) : u3 - u2;
    return ((37.0 * d0 - 8.0 * d1) + d2) * (1.0 / 30.0);
}

/* 2.-4. one system on an aligned group of L lanes, from its points in `lds`
 * (written by wave_points; the caller has synchronised the workgroup since).
 * Every lane of the wavefront calls this; a group without a system
 * (valid false) takes empty chunks, which makes its reduced rows identity
 * rows, and stores nothing. */
template <int MM, int L>
__device__ static inline void solve_system_wave(const oco_sys& d, double* lds, bool valid)
{
    const int nz = d.nz, ll = lane_id() & (L - 1);
    const double dz = d.dz, dz1 = d.dz1, dz2 = d.dz2;
    double* LA = lds;              /* alpha */
    double* LW = lds + nz;         /* albedo */
    double* LS = lds + 2 * nz;     /* extended, e = z + 3: S, then U with ghosts */
    double* LB = lds + 3 * nz + 6; /* S of cells 0, 1, 2, nz - 1, nz - 2, nz - 3 */
    const int m = wave_rows(nz, L);
    const int z0 = valid && ll * m < nz ? ll * m : nz;
    const int z1 = z0 + m < nz ? z0 + m : nz;
    /* boundary lanes: lane 0 of the group the bottom (cells 0, 1, 2, ...
     * upward), lane 1 the top */
    const int bz = ll == 0 ? 0 : nz - 1, bs = ll == 0 ? 1 : -1;
    if (valid && ll < 6) LB[ll] = LS[3 + (ll < 3 ? ll : nz + 2 - ll)];

    /* 2. rows (source/auxiliary.f90:147-197) in difference form and the
     * chunk's forward elimination */
    double Dr[MM], Wr[MM], Zr[MM];
    rh_chunk c;
    rh_chunk_init(c);
    if (z0 > 0 && z0 < z1) c.mlo = 0.5 * (LA[z0] + LA[z0 - 1]);
#pragma unroll
    for (int j = 0; j < MM; j++) {
        const int z = z0 + j;
        Dr[j] = 0.0; Wr[j] = 0.0; Zr[j] = 0.0;
        if (z < z1)
            rh_chunk_row(c, z, nz, j == 0, LA[z], LW[z], LS[z + 3], z < nz - 1 ? LA[z + 1] : 0.0, dz, dz2,
                         Dr[j], Wr[j], Zr[j]);
    }
    rh_chunk_close(c);
    /* 3. the reduced system */
    double G, GL;
    rh_reduced<L>(c, z1 > z0, G, GL);
    /* 4. back-substitution inside the chunk */
    double x = G;
#pragma unroll
    for (int j = MM - 1; j >= 0; j--) {
        const int z = z0 + j;
        if (z < z1) {
            if (z < z1 - 1) x = rh_back(Dr[j], Wr[j], Zr[j], x, GL);
            LS[z + 3] = x;
        }
    }
    wave_sync();
    /* ghosts (source/auxiliary.f90:49-62) */
    if (valid) {
        if (ll < 3) {
            const int i = ll + 1;   /* f(n2+i) = 2 f(n2) - f(n2-i) */
            LS[nz + 2 + i] = ghost(LS[nz + 2], LS[nz + 2 - i]);
        } else if (ll < 6) {
            const int i = ll - 2;   /* f(n1-i) = 2 f(n1) - f(n1+i) */
            LS[3 - i] = ghost(LS[3], LS[3 + i]);
        }
    }
    wave_sync();
    if (!valid) return;
    /* der and calc_flux (:64-78, :96-109): the emergent V(1), V(nz) from
     * rh_boundary_der, every other V from der of U */
    if (ll < 2) {
        const double* u = LS + 3 + bz;
        const double* al = LA + bz;
        const double* om = LW + bz;
        const double* sb = LB + 3 * ll;
        const double df = rh_boundary_der(u[0], u[bs], u[2 * bs], u[3 * bs], al[0], al[bs], al[2 * bs], al[3 * bs],
                                          om[0], om[bs], om[2 * bs], sb[0], sb[1], sb[2], dz, dz2);
        out_store<OCO_WAVE_NT>(d.V + bz, flux_v(al[0], ll == 0 ? df : -df, dz1));
    }
    /* a lane's cells z = ll + 1 + k L, k < MM (nz - 2 <= MM L): straight-line
     * code, every LDS read from an index clamped into the system, so all of
     * them are issued before the first result is needed */
#pragma unroll
    for (int k = 0; k < MM; k++) {
        const int z = ll + 1 + k * L;
        const double* u = LS + (z < nz - 1 ? z : nz - 2) + 3;
        const double v = flux_v(u[-3 - 2 * nz], der7(u[-3], u[-2], u[-1], u[1], u[2], u[3]), dz1);
        if (z < nz - 1) out_store<OCO_WAVE_NT>(d.V + z, v);
    }
    /* U with ghosts, nz + 6 doubles: a system's U starts anywhere in a
     * 128-byte line (nz + 6 is not a multiple of 16), so with 32 or 64 lanes
     * the lanes are shifted by the start's offset `off` in the line, e = k L
     * + ll - off, and every store but the first and last covers whole lines
     * (measured on MI355X, 256 x 256 columns, nz = 512, nw = 128: 37.35 ->
     * 36.7 ms).  k <= MM: nz + 6 + 15 <= (MM + 1) L for L >= 32.  With 8 or 16
     * lanes, e = ll + k L (nz + 6 <= (MM + 1) L). */
    if constexpr (L >= 32) {
        const int off = (int)(((unsigned long)d.U >> 3) & 15);
#pragma unroll
        for (int k = 0; k <= MM; k++) {
            const int e = k * L + ll - off;
            const double u = LS[e < 0 ? 0 : (e < nz + 6 ? e : nz + 5)];
            if (e >= 0 && e < nz + 6) out_store<OCO_WAVE_NT>(d.U + e, u);
        }
        return;
    }
#pragma unroll
    for (int k = 0; k <= MM; k++) {
        const int e = ll + k * L;
        const double u = LS[e < nz + 6 ? e : nz + 5];
        if (e < nz + 6) out_store<OCO_WAVE_NT>(d.U + e, u);
    }
}

/* The difference-form sweep of the block (OCO_MODE_MIXED, OCO_MODE_ACCURATE):
 * L lanes per system, so a wavefront holds 64 / L systems, and K wavefronts
 * per workgroup on S = K 64 / L consecutive wavelengths of one column. */
/* Tuning knob: OCO_WAVE_WAVES_PER_EU=n asks the compiler to fit n waves of
 * the wavefront sweep per SIMD. */
#ifdef OCO_WAVE_WAVES_PER_EU
#define OCO_WAVE_ATTR __attribute__((amdgpu_waves_per_eu(OCO_WAVE_WAVES_PER_EU)))
#else
#define OCO_WAVE_ATTR
#endif

/* Systems a workgroup takes per ticket of the wavefront sweep: a ticket is
 * 32 / S groups of S consecutive wavelengths of one column (at least one).
 * Measured on MI355X with workgroups of 4 wavefronts of one system each,
 * 256 x 256 columns, nz = 512, nw = 128: 8 systems 36.6-36.7 ms, 16
 * 35.7-35.8, 32 35.4-35.5; on 64 x 64 columns of 256 cells, 1.39 and 1.40 ms
 * for 8 and 32. */
#ifndef OCO_WAVE_TICKET
#define OCO_WAVE_TICKET 32
#endif

template <int MM, int K, int L, int PK>
__global__ __launch_bounds__(OCO_WAVE * K) OCO_WAVE_ATTR void sweep_mixed_wave(const oco_sweep_args a)
{
    extern __shared__ double lds[];   /* S systems' regions, then the FP64 math tables (not in mixed mode) */
    __shared__ long ticket;
    constexpr int S = K * (OCO_WAVE / L);   /* systems per workgroup */
    const int nw = a.nw, nz = a.nz;
    const long lstr = wave_lds_words(nz);
    oco_mtab* smt = PK == OCO_PK_MIXED ? nullptr : (oco_mtab*)(lds + S * lstr);
    if constexpr (PK != OCO_PK_MIXED) load_mtab(smt);
    const int ngrp = (nw + S - 1) / S;   /* wavelength groups per column */
    const int wid = (int)threadIdx.x / OCO_WAVE;
    const int g = wid * (OCO_WAVE / L) + lane_id() / L;   /* this lane group's system in the workgroup's S */
    constexpr long TG = OCO_WAVE_TICKET / S > 0 ? OCO_WAVE_TICKET / S : 1;   /* groups per ticket */
    const long astride = a.ncol * (long)nz;   /* alpha output: one wavelength to the next */
    /* XCD-aware work queues: workgroups are dispatched to the 8 XCDs
     * round-robin, so workgroup b runs on XCD b mod 8.  With a grid that is a
     * multiple of 8, queue x holds the columns x, x + 8, x + 16, ..., each
     * column's wavelength groups in turn, and the workgroups of XCD x draw
     * tickets of TG groups from it: a column's cell records (22 fields x nz
     * doubles) are brought into one L2 and reused there, and each record a
     * thread loads serves the wavelengths of its group.  A workgroup whose
     * queue is empty goes on to the next XCD's, so the grid drains together.
     * Any other grid: one queue in (column, wavelength) order.  Every
     * workgroup ends once it has found all queues empty (the ticket is read
     * by every thread after a barrier, so the exit is workgroup-uniform). */
    const int nq = gridDim.x % OCO_XCDS == 0 ? OCO_XCDS : 1;
    const int home = (int)(blockIdx.x % nq);
    for (int v = 0; v < nq; v++) {
        const int q = (home + v) % nq;
        const long ngroups = (a.ncol - q + nq - 1) / nq * ngrp;   /* columns of queue q x groups */
        unsigned long long* ctr = a.tickets + q * OCO_TICKET_STRIDE;
        for (;;) {
            __syncthreads();   /* the last group's systems are done with LDS, and every thread has read `ticket` */
            if (threadIdx.x == 0) ticket = (long)atomicAdd(ctr, (unsigned long long)TG);
            __syncthreads();
            const long t0 = ticket;
            if (t0 >= ngroups) break;
            const long t1 = t0 + TG < ngroups ? t0 + TG : ngroups;
            for (long t = t0; t < t1; t++) {
                const long col = (t / ngrp) * nq + q;
                const int iw0 = (int)(t % ngrp) * S;
                const int nk = nw - iw0 < S ? nw - iw0 : S;
                if (t > t0) __syncthreads();   /* the previous group's systems are done with LDS */
                const double* cc = PK == OCO_PK_GREY ? a.rho + col * (long)nz : a.cells + col * (long)CR_NFIELD * nz;
                wave_points<K, PK>(cc, a.T + col * (long)nz, nz, a.wtab, iw0, nk,
                                   a.A + ((long)iw0 * a.ncol + col) * nz, astride, lds, lstr, a.sigma_grey, smt);
                __syncthreads();
                const bool valid = g < nk;
                const int iw = iw0 + (valid ? g : 0);
                const long osys = (long)iw * a.ncol + col;
                oco_sys d;
                d.nz = nz;
                d.dz = a.dz; d.dz1 = a.dz1; d.dz2 = a.dz2; d.sigma_grey = a.sigma_grey;
                d.cells = nullptr; d.rho = nullptr; d.T = nullptr;
                d.U = a.U + osys * (nz + 6);
                d.V = a.V + osys * nz;
                d.A = a.A + osys * nz;
                d.key = col * nw + iw;
                solve_system_wave<MM, L>(d, lds + g * lstr, valid);
            }
        }
    }
}

/* LDS of a workgroup of the block sweep: S = k 64 / lanes systems' regions,
 * plus the FP64 math tables unless the points are mixed-precision */
static size_t wave_lds_bytes(int nz, int k, int lanes = OCO_WAVE, int pk = OCO_PK_MIXED)
{
    return (size_t)k * (OCO_WAVE / lanes) * wave_lds_words(nz) * sizeof(double) +
           (pk == OCO_PK_MIXED ? 0 : sizeof(oco_mtab));
}

/* Rows per lane: at most 17 with 64 lanes per system (columns up to 1088
 * cells), 9 with fewer lanes (a longer column takes more lanes). */
#define OCO_WAVE_MAX_ROWS 17
#define OCO_WAVE_MAX_ROWS_SUB 9
static int wave_max_rows(int lanes) { return lanes == OCO_WAVE ? OCO_WAVE_MAX_ROWS : OCO_WAVE_MAX_ROWS_SUB; }

/* Wavefronts per workgroup of the wavefront sweep: each cell record a thread
 * loads serves the workgroup's wavelengths.  Measured on MI355X (256 x 256
 * columns, nz = 512, nw = 128), see oco_wave_group. */
#ifndef OCO_WAVE_GROUP
#define OCO_WAVE_GROUP 4
#endif

static int lds_limit(int device)
{
    int lim = 0;
    if (hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess || lim <= 0)
        lim = 65536;
    return lim;
}

static bool lanes_ok(int lanes) { return lanes == 8 || lanes == 16 || lanes == 32 || lanes == 64; }

/* Lanes per system for columns of nz cells: `want` (8, 16, 32, 64), or for
 * 0 the fewest lanes whose chunks are at most 9 rows (64 above 576 cells). */
int oco_wave_lanes(int nz, int want)
{
    if (lanes_ok(want)) return want;
    for (int l = 8; l < OCO_WAVE; l <<= 1)
        if (wave_rows(nz, l) <= OCO_WAVE_MAX_ROWS_SUB) return l;
    return OCO_WAVE;
}

/* 1 when the difference-form sweep with `lanes` lanes per system takes
 * columns of nz cells: rows per lane within its limit and one wavefront's
 * LDS within a workgroup's */
int oco_wave_fits(int nz, int lanes, int pk, int device)
{
    return nz >= 4 && lanes_ok(lanes) && wave_rows(nz, lanes) <= wave_max_rows(lanes) &&
           wave_lds_bytes(nz, 1, lanes, pk) <= (size_t)lds_limit(device);
}

/* Wavefronts per workgroup for columns of nz cells: `want` (1, 2 or 4; 0 for
 * OCO_WAVE_GROUP), halved until the workgroup's LDS fits. */
int oco_wave_group(int nz, int lanes, int pk, int want, int device)
{
    int k = want == 1 || want == 2 || want == 4 ? want : OCO_WAVE_GROUP;
    while (k > 1 && wave_lds_bytes(nz, k, lanes, pk) > (size_t)lds_limit(device)) k /= 2;
    return k;
}

/* The kernel of a variant: (rows, lanes) as instantiated, k wavefronts per
 * workgroup, point kind pk (grey: one wavefront per workgroup). */
template <int K, int PK>
static const void* wave_kernel_kp(int nz, int lanes)
{
    const int m = wave_rows(nz, lanes);
    switch (lanes) {
    case 8: return (const void*)sweep_mixed_wave<9, K, 8, PK>;
    case 16: return (const void*)sweep_mixed_wave<9, K, 16, PK>;
    case 32: return (const void*)sweep_mixed_wave<9, K, 32, PK>;
    default:
        if (m <= 5) return (const void*)sweep_mixed_wave<5, K, 64, PK>;
        if (m <= 9) return (const void*)sweep_mixed_wave<9, K, 64, PK>;
        return (const void*)sweep_mixed_wave<17, K, 64, PK>;
    }
}

static const void* wave_kernel(int nz, int k, int lanes, int pk)
{
    if (pk == OCO_PK_GREY) return wave_kernel_kp<1, OCO_PK_GREY>(nz, lanes);
    if (pk == OCO_PK_FP64)
        return k == 4 ? wave_kernel_kp<4, OCO_PK_FP64>(nz, lanes)
                      : k == 2 ? wave_kernel_kp<2, OCO_PK_FP64>(nz, lanes) : wave_kernel_kp<1, OCO_PK_FP64>(nz, lanes);
    return k == 4 ? wave_kernel_kp<4, OCO_PK_MIXED>(nz, lanes)
                  : k == 2 ? wave_kernel_kp<2, OCO_PK_MIXED>(nz, lanes) : wave_kernel_kp<1, OCO_PK_MIXED>(nz, lanes);
}

/* Resident workgroups of the wavefront sweep with k wavefronts each. */
int oco_wave_grid(int nz, int k, int lanes, int pk, int device)
{
    if (pk == OCO_PK_GREY) k = 1;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, wave_kernel(nz, k, lanes, pk), OCO_WAVE * k,
                                                     wave_lds_bytes(nz, k, lanes, pk)) != hipSuccess || per_cu < 1)
        per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1)
        cus = 256;
    const int g = per_cu * cus;
    return g >= OCO_XCDS ? g / OCO_XCDS * OCO_XCDS : g;   /* a multiple of 8: the XCD-aware order */
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
        if (fail) *fail = OCO_FAIL_NONE;
    }
}

/* fill n failure keys with OCO_FAIL_NONE (serving: a slot's key at admission) */
__global__ void fail_reset(unsigned long long* fail, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) fail[i] = OCO_FAIL_NONE;
}

/* The first non-finite value of an output array (oco_launch_finite_scan):
 * grid-stride over the array; a lane that finds one forms its key in the
 * reference's loop order and takes the minimum. */
__global__ __launch_bounds__(256) void finite_scan(const double* __restrict__ p, long n, int inner, int which,
                                                   int nyloc, int nxloc, int iw0, int nw, int nz,
                                                   unsigned long long* key)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double v = __builtin_nontemporal_load(gptr(p + i));
        if (__builtin_isfinite(v)) continue;
        const long e = i % inner;
        long t = i / inner;
        const long iy = t % nyloc;
        t /= nyloc;
        const long ix = t % nxloc;
        const long iw = iw0 + t / nxloc;
        const unsigned long long sys = (unsigned long long)((ix * nyloc + iy) * nw + iw);
        atomicMin(key, (sys * 3ull + (unsigned long long)which) * (unsigned long long)(nz + 6) + (unsigned long long)e);
    }
}

hipError_t oco_launch_finite_scan(const double* p, long n, int inner, int which, int nyloc, int nxloc, int iw0, int nw,
                                  int nz, unsigned long long* key, hipStream_t stream)
{
    (void)hipGetLastError();
    if (n <= 0) return hipSuccess;
    long blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(finite_scan, dim3((unsigned)blocks), dim3(256), 0, stream, p, n, inner, which, nyloc, nxloc,
                       iw0, nw, nz, key);
    return hipGetLastError();
}

hipError_t oco_launch_fail_reset(unsigned long long* fail, int n, hipStream_t stream)
{
    (void)hipGetLastError();   /* a pending not-ready status from an event query is not this launch's */
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fail_reset, dim3((n + 63) / 64), dim3(64), 0, stream, fail, n);
    return hipGetLastError();
}

typedef const __attribute__((address_space(4))) int* oco_cptr_i;
typedef const __attribute__((address_space(4))) long* oco_cptr_l;
typedef const __attribute__((address_space(4))) double* oco_cptr_d;

template <int BC, class STAGE>
__device__ static inline void batch_tiles(const oco_batch_args& a, const oco_mtab* M, STAGE* S)
{
    const int lane = lane_id();
    /* one scratch slot per wavefront of the grid */
    const long slot = (long)blockIdx.x * (blockDim.x / OCO_WAVE) + threadIdx.x / OCO_WAVE;
    double* scr = a.scratch + slot * a.scr_stride + lane;
    oco_cptr_i tiles = (oco_cptr_i)a.tiles;
    oco_cptr_i nzs = (oco_cptr_i)a.nz;
    oco_cptr_i nws = (oco_cptr_i)a.nw;
    oco_cptr_l coff = (oco_cptr_l)a.cell_off;
    oco_cptr_l woff = (oco_cptr_l)a.wt_off;
    oco_cptr_l uoff = (oco_cptr_l)a.u_off;
    oco_cptr_l voff = (oco_cptr_l)a.v_off;
    oco_cptr_d dzs = (oco_cptr_d)a.dz;
    for (;;) {
        unsigned tv = 0u;   /* this wavefront's next tile */
        if (lane == 0) tv = atomicAdd(a.counter, 1u);
        const int t = (int)__builtin_amdgcn_readfirstlane(tv);
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
            d.U = a.U + uoff[col] + (long)iw * oco_batch_ldu(nz);
            d.V = a.V + voff[col] + (long)iw * oco_batch_ldv(nz);
            d.A = a.A + voff[col] + (long)iw * oco_batch_ldv(nz);
            d.key = woff[col] + iw;
            double w[OCO_WT_STRIDE];
            const double* wr = a.wtab + (woff[col] + iw) * OCO_WT_STRIDE;
#pragma unroll
            for (int k = 0; k < OCO_WT_STRIDE; k++) w[k] = gptr(wr)[k];
            solve_system<false, true, 1, BC, STAGE>(d, valid, w, scr, a.fail, M, S);
        }
    }
}

template <int BC>
__global__ __launch_bounds__(OCO_WAVE * OCO_BATCH_WG) void batch_fp64(const oco_batch_args a)
{
    __shared__ oco_mtab smt;
    __shared__ oco_stage_batch stg[OCO_BATCH_WG];
    load_mtab(&smt);
    batch_tiles<BC>(a, &smt, &stg[threadIdx.x / OCO_WAVE]);
}


/* ---- serving: tiles of many jobs in one persistent launch ----
 * Each tile is (job slot, column, first of 64 wavelengths); the job's
 * descriptor gives its cell records, wavelength table and emergent-flux
 * output.  A launch mixes tiles of every job admitted so far, so jobs that
 * arrive while others run join at the next launch (continuous batching);
 * each system is solved by one lane exactly as in a launch of its own. */
template <int BC>
__global__ __launch_bounds__(OCO_WAVE) void serve_tiles(const oco_serve_args a)
{
    __shared__ oco_mtab smt;
    __shared__ int tile_s;
    __shared__ oco_stage_emergent none;
    load_mtab(&smt);
    const int lane = lane_id();
    double* scr = a.scratch + (long)blockIdx.x * a.scr_stride + lane;
    oco_cptr_i tiles = (oco_cptr_i)a.tiles;
    for (;;) {
        if (lane == 0) tile_s = (int)atomicAdd(a.counter, 1u);
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(tile_s);
        __syncthreads();
        if (t >= a.ntiles) break;
        const int slot = tiles[3 * t];
        const long col = tiles[3 * t + 1];
        const oco_serve_job jb = a.jobs[slot];
        const bool valid = tiles[3 * t + 2] + lane < jb.nw;
        const int iw = valid ? tiles[3 * t + 2] + lane : jb.nw - 1;
        oco_sys d;
        d.nz = jb.nz;
        d.dz = jb.dz; d.dz1 = jb.dz1; d.dz2 = jb.dz2;
        d.sigma_grey = 0.0;
        d.cells = jb.cells + col * jb.nz * OCO_CELL_STRIDE;
        d.rho = nullptr; d.T = nullptr;
        d.U = nullptr; d.A = nullptr;
        d.V = jb.em + ((long)iw * jb.ncol + col) * 2;
        d.key = col * jb.nw + iw;
        double w[OCO_WT_STRIDE];
        const double* wr = jb.wtab + (long)iw * OCO_WT_STRIDE;
#pragma unroll
        for (int k = 0; k < OCO_WT_STRIDE; k++) w[k] = gptr(wr)[k];
        solve_system<false, true, 1, BC, oco_stage_emergent>(d, valid, w, scr, a.fail + slot, &smt, &none);
    }
}

/* ---- serving, mixed mode: one wavefront per system, emergent flux only ----
 * The same partitioned difference-form solve as solve_system_wave (its rows,
 * reduced system, back-substitution and derivative are the shared functions
 * above, so a job's V(1), V(nz) are those of the block wavefront sweep, bit
 * for bit), restructured for outputs at the two boundaries only:
 *   - each lane evaluates the points of its own chunk of rows, loading its
 *     records from the chunk-major layout (one 512-byte line per field
 *     across the wavefront) and evaluating each for K wavelengths, and
 *     eliminates each row as soon as the next cell's alpha is known; a
 *     system's state is a few registers whatever nz is (no LDS copy of the
 *     column, no per-row registers), so any nz up to the server's max_nz is
 *     taken;
 *   - a chunk's first row needs the alpha of the cell below it (the previous
 *     lane's last cell): each lane evaluates that cell's alpha itself; its
 *     last row needs the next lane's first alpha, read across lanes;
 *   - only the chunks holding rows 0..3 and nz-4..nz-1 keep their rows'
 *     D, W, Z and alpha (LDS, the boundary window) and back-substitute;
 *     two lanes then form V(1) and V(nz).
 * A tile is (job, column, 64 wavelengths), taken by a workgroup of
 * OCO_SERVE_WG wavefronts, each solving every OCO_SERVE_WG-th group of K of
 * its wavelengths: the wavefronts stream the same column's records at the
 * same time, so they come from L2 once per workgroup rather than from HBM
 * once per group (one wavefront per tile: capacity 7.0e10 points/s on the
 * serving workload). */
#ifndef OCO_SERVE_WG
#define OCO_SERVE_WG 4
#endif
/* Wavelengths each record load serves (K) and wavefronts per SIMD: measured
 * on MI355X, serving workload (capacity burst of 12 jobs): K = 4 at 2 per
 * SIMD (256 VGPRs, 128 bytes of spills per lane) 1.05e11 points/s; K = 2 at
 * 2 per SIMD 9.3e10, at 3 per SIMD 8.9-9.0e10; K = 1 at 2 per SIMD 5.0e10;
 * K = 6 at 2 per SIMD 5.7e10 (spills); K = 8 at 1 per SIMD 6.7e10; K = 4 with
 * 2-wavefront workgroups 1.0e11. */
#ifndef OCO_SERVE_K
#define OCO_SERVE_K 4
#endif
/* Doubles per row of the serving kernel's boundary window: D (then x), W,
 * Z, alpha, albedo, S */
#define OCO_SERVE_SLOT 6
/* Rows per lane of the serving solve: the wavefront sweep's, at least
 * OCO_SERVE_MIN_ROWS (short columns then leave lanes empty: longer sequential
 * chunks and a smaller reduced system) */
#ifndef OCO_SERVE_MIN_ROWS
#define OCO_SERVE_MIN_ROWS 5
#endif
__host__ __device__ static inline int serve_rows(int nz)
{
    const int m = wave_rows(nz);
    return m < OCO_SERVE_MIN_ROWS ? OCO_SERVE_MIN_ROWS : m;
}

int oco_serve_win_rows(int max_nz) { return 2 * (serve_rows(max_nz) + 3); }

size_t oco_cells_chunked_doubles(long ncol, int nz) { return (size_t)ncol * serve_rows(nz) * CR_NFIELD * OCO_WAVE; }

/* Thread i takes slot (column, row in chunk j, lane l), lane fastest, so a
 * wavefront's field stores are one 512-byte line each (cell-order threads
 * would store 64 lines 11 KB apart); padding slots (l*m + j >= nz) are
 * skipped. */
__global__ __launch_bounds__(256) void cells_chunked_fp64(const double* __restrict__ rho, const double* __restrict__ T,
                                                          long ncol, int nz, double fully_ionized_T,
                                                          double* __restrict__ cells)
{
    __shared__ oco_mtab smt;
    load_mtab(&smt);
    const int m = serve_rows(nz);
    const long per_col = (long)m * OCO_WAVE, nslots = ncol * per_col;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nslots; i += (long)gridDim.x * blockDim.x) {
        const long col = i / per_col;
        const int r = (int)(i - col * per_col), j = r / OCO_WAVE, l = r - j * OCO_WAVE, z = l * m + j;
        if (z >= nz) continue;
        const long c = col * nz + z;
        double rec[CR_NFIELD];
        oco_cell_record(rho[c], T[c], fully_ionized_T, &smt, rec);
        double* out = cells + col * per_col * CR_NFIELD + (long)j * CR_NFIELD * OCO_WAVE + l;
#pragma unroll
        for (int k = 0; k < CR_NFIELD; k++) out[k * OCO_WAVE] = rec[k];
    }
}

hipError_t oco_launch_cells_chunked(const double* rho, const double* T, long ncells, int nz, double fully_ionized_T,
                                    double* cells, hipStream_t stream)
{
    (void)hipGetLastError();
    if (ncells <= 0 || nz <= 0) return hipSuccess;
    const long ncol = ncells / nz, nslots = ncol * serve_rows(nz) * OCO_WAVE;
    long blocks = (nslots + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(cells_chunked_fp64, dim3((unsigned)blocks), dim3(256), 0, stream, rho, T, ncol, nz,
                       fully_ionized_T, cells);
    return hipGetLastError();
}

/* Waves of the serving kernel per SIMD the compiler is asked to fit (with
 * OCO_SERVE_K, above). */
#ifndef OCO_SERVE_WAVES_PER_EU
#define OCO_SERVE_WAVES_PER_EU 2
#endif
#if OCO_SERVE_WAVES_PER_EU > 0
#define OCO_SERVE_ATTR __attribute__((amdgpu_waves_per_eu(OCO_SERVE_WAVES_PER_EU)))
#else
#define OCO_SERVE_ATTR
#endif
template <int K>
__global__ __launch_bounds__(OCO_WAVE * OCO_SERVE_WG) OCO_SERVE_ATTR void serve_wave(const oco_serve_args a)
{
    extern __shared__ double win[];   /* per wavefront: K systems x win_rows rows x (D or x, W, Z, alpha, albedo, S) */
    __shared__ int tile_s;
    /* the wavefront's index, wave-uniform (its wavelengths' table rows are then scalar loads) */
    const int lane = lane_id(), wid = __builtin_amdgcn_readfirstlane((int)threadIdx.x / OCO_WAVE);
    const int WR = a.win_rows;
    double* lds = win + (long)wid * K * WR * OCO_SERVE_SLOT;
    oco_cptr_i tiles = (oco_cptr_i)a.tiles;
    for (;;) {
        __syncthreads();   /* every wavefront has read the last tile_s */
        if (threadIdx.x == 0) tile_s = (int)atomicAdd(a.counter, 1u);
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(tile_s);
        if (t >= a.ntiles) break;
        const int slot = tiles[3 * t];
        const long col = tiles[3 * t + 1];
        const int wt0 = tiles[3 * t + 2];
        const oco_serve_job jb = a.jobs[slot];
        const int nz = jb.nz, m = serve_rows(nz);
        const double dz = jb.dz, dz1 = jb.dz1, dz2 = jb.dz2;
        const int z0 = lane * m < nz ? lane * m : nz;
        const int z1 = z0 + m < nz ? z0 + m : nz;
        const double* cc = jb.cells + col * (long)m * CR_NFIELD * OCO_WAVE;
        /* boundary window: rows [0, nB) and [t0, nz), whole chunks */
        const int nB = nz < (3 / m + 1) * m ? nz : (3 / m + 1) * m;
        const int t0 = ((nz - 4) / m) * m > nB ? ((nz - 4) / m) * m : nB;
        const bool win = z0 < z1 && (z0 < nB || z1 > t0);
        auto wslot = [&](int k, int z) -> double* {
            return lds + ((long)k * WR + (z < nB ? z : nB + (z - t0))) * OCO_SERVE_SLOT;
        };
        const int wend = wt0 + OCO_WAVE < jb.nw ? wt0 + OCO_WAVE : jb.nw;
        for (int iw0 = wt0 + wid * K; iw0 < wend; iw0 += K * OCO_SERVE_WG) {
            const int nk = wend - iw0 < K ? wend - iw0 : K;
            rh_chunk st[K];
            double ap[K], op[K], Sp[K], a0[K], anb[K];
#pragma unroll
            for (int k = 0; k < K; k++) {
                rh_chunk_init(st[k]);
                ap[k] = op[k] = Sp[k] = a0[k] = anb[k] = 0.0;
            }
            double rec[CR_NFIELD];
            if (z0 > 0 && z0 < z1) {   /* the cell below the chunk: the previous lane's last */
#pragma unroll
                for (int f = 0; f < CR_NFIELD; f++) rec[f] = gptr(cc)[(long)((m - 1) * CR_NFIELD + f) * OCO_WAVE + lane - 1];
#pragma unroll
                for (int k = 0; k < K; k++) {
                    if (k < nk) {
                        double w[OCO_WT_STRIDE];
                        oco_cptr wp = (oco_cptr)(jb.wtab + (long)(iw0 + k) * OCO_WT_STRIDE);
#pragma unroll
                        for (int f = 0; f < OCO_WT_STRIDE; f++) w[f] = wp[f];
                        anb[k] = oco_point_mixed<1>(rec, w).alpha;
                    }
                }
            }
            for (int j = 0; j < m; j++) {
                const int z = z0 + j;
                if (z < z1) {
#pragma unroll
                    for (int f = 0; f < CR_NFIELD; f++) rec[f] = gptr(cc)[(long)(j * CR_NFIELD + f) * OCO_WAVE + lane];
#pragma unroll
                    for (int k = 0; k < K; k++) {
                        if (k < nk) {
                            double w[OCO_WT_STRIDE];
                            oco_cptr wp = (oco_cptr)(jb.wtab + (long)(iw0 + k) * OCO_WT_STRIDE);
#pragma unroll
                            for (int f = 0; f < OCO_WT_STRIDE; f++) w[f] = wp[f];
                            const oco_pt p = oco_point_mixed<1>(rec, w);
                            if (j == 0) {
                                a0[k] = p.alpha;
                                if (z0 > 0) st[k].mlo = 0.5 * (p.alpha + anb[k]);
                            } else {   /* row z - 1, now that alpha(z) is known */
                                double D, W, Z;
                                rh_chunk_row(st[k], z - 1, nz, j == 1, ap[k], op[k], Sp[k], p.alpha, dz, dz2, D, W, Z);
                                if (win) {
                                    double* q = wslot(k, z - 1);
                                    q[0] = D; q[1] = W; q[2] = Z; q[3] = ap[k]; q[4] = op[k]; q[5] = Sp[k];
                                }
                            }
                            ap[k] = p.alpha; op[k] = p.albedo; Sp[k] = p.S;
                        }
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < K; k++) {
                if (k < nk) {
                    /* alpha of cell z1, the next chunk's first (read only where z1 < nz,
                     * so never in lane 63, whose shifted-in value is 0) */
                    const double an = lane_dn1(a0[k]);
                    if (z0 < z1) {   /* the chunk's last row */
                        double D, W, Z;
                        rh_chunk_row(st[k], z1 - 1, nz, z1 - 1 == z0, ap[k], op[k], Sp[k], z1 < nz ? an : 0.0, dz,
                                     dz2, D, W, Z);
                        rh_chunk_close(st[k]);
                        if (win) {
                            double* q = wslot(k, z1 - 1);
                            q[0] = D; q[1] = W; q[2] = Z; q[3] = ap[k]; q[4] = op[k]; q[5] = Sp[k];
                        }
                    }
                    double G, GL;
                    rh_reduced<OCO_WAVE>(st[k], z0 < z1, G, GL);
                    if (win) {   /* back-substitution in the window's chunks; x replaces D */
                        double x = G;
                        wslot(k, z1 - 1)[0] = x;
                        for (int z = z1 - 2; z >= z0; z--) {
                            double* q = wslot(k, z);
                            x = rh_back(q[0], q[1], q[2], x, GL);
                            q[0] = x;
                        }
                    }
                }
            }
            wave_sync();
            if (lane < 2) {
#pragma unroll
                for (int k = 0; k < K; k++) {
                    if (k < nk) {
                        /* V(1) in lane 0, V(nz) in lane 1: the cells counted inward */
                        const int bz = lane == 0 ? 0 : nz - 1, bs = lane == 0 ? 1 : -1;
                        const double* q0 = wslot(k, bz);
                        const double* q1 = wslot(k, bz + bs);
                        const double* q2 = wslot(k, bz + 2 * bs);
                        const double* q3 = wslot(k, bz + 3 * bs);
                        const double df = rh_boundary_der(q0[0], q1[0], q2[0], q3[0], q0[3], q1[3], q2[3], q3[3],
                                                          q0[4], q1[4], q2[4], q0[5], q1[5], q2[5], dz, dz2);
                        *gptr(jb.em + ((long)(iw0 + k) * jb.ncol + col) * 2 + lane) = flux_v(q0[3], lane == 0 ? df : -df, dz1);
                    }
                }
            }
            wave_sync();   /* the window is read before the next group writes it */
        }
    }
}

static size_t serve_wave_lds(int win_rows)
{
    return (size_t)OCO_SERVE_WG * OCO_SERVE_K * win_rows * OCO_SERVE_SLOT * sizeof(double);
}

hipError_t oco_launch_serve(const oco_serve_args& a, int mixed, int nblocks, hipStream_t stream)
{
    (void)hipGetLastError();   /* a pending not-ready status from an event query is not this launch's */
    hipLaunchKernelGGL(batch_reset, dim3(1), dim3(64), 0, stream, a.counter, (unsigned long long*)nullptr);
    if (mixed)
        hipLaunchKernelGGL((serve_wave<OCO_SERVE_K>), dim3(nblocks), dim3(OCO_WAVE * OCO_SERVE_WG),
                           serve_wave_lds(a.win_rows), stream, a);
    else
        hipLaunchKernelGGL((serve_tiles<2>), dim3(nblocks), dim3(OCO_WAVE), 0, stream, a);
    return hipGetLastError();
}

int oco_serve_grid(int mixed, int max_nz, int device)
{
    int per_cu = 0, cus = 0;
    const void* k = mixed ? (const void*)serve_wave<OCO_SERVE_K> : (const void*)serve_tiles<2>;
    const size_t sh = mixed ? serve_wave_lds(oco_serve_win_rows(max_nz)) : 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, mixed ? OCO_WAVE * OCO_SERVE_WG : OCO_WAVE, sh) !=
            hipSuccess || per_cu < 1)
        per_cu = mixed ? 1 : 4;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1) cus = 256;
    return per_cu * cus;
}

/* variants 0, 1, 2: back chunks of 2, 4 and 8 scratch rows (measured on MI355X, 4096 uneven
 * columns: chunks of 1 row were the slowest on every grid, 8 rows 1.5% faster than 4) */
hipError_t oco_launch_batch(const oco_batch_args& a, int variant, int nblocks, hipStream_t stream)
{
    (void)hipGetLastError();   /* a pending not-ready status from an event query is not this launch's */
    hipLaunchKernelGGL(batch_reset, dim3(1), dim3(64), 0, stream, a.counter, a.fail);
    const dim3 g((nblocks + OCO_BATCH_WG - 1) / OCO_BATCH_WG), b(OCO_WAVE * OCO_BATCH_WG);
    switch (variant) {
    case 0: hipLaunchKernelGGL((batch_fp64<2>), g, b, 0, stream, a); break;
    case 1: hipLaunchKernelGGL((batch_fp64<4>), g, b, 0, stream, a); break;
    default: hipLaunchKernelGGL((batch_fp64<8>), g, b, 0, stream, a); break;
    }
    return hipGetLastError();
}

int oco_batch_variants(void) { return 3; }

/* Cell records of a batch, field-major per column (the wavefront sweep's
 * layout with each column's own nz): one workgroup per column. */
__global__ __launch_bounds__(256) void cells_cols_fp64(const double* __restrict__ rho, const double* __restrict__ T,
                                                       const int* __restrict__ nzs, const long* __restrict__ coff,
                                                       double fully_ionized_T, double* __restrict__ cells)
{
    __shared__ oco_mtab smt;
    load_mtab(&smt);
    const int col = blockIdx.x, nz = nzs[col];
    const long c0 = coff[col];
    double* out = cells + c0 * CR_NFIELD;
    for (int z = threadIdx.x; z < nz; z += blockDim.x) {
        double rec[CR_NFIELD];
        oco_cell_record(rho[c0 + z], T[c0 + z], fully_ionized_T, &smt, rec);
#pragma unroll
        for (int k = 0; k < CR_NFIELD; k++) out[(long)k * nz + z] = rec[k];
    }
}

/* Batched mixed mode: the block sweep's wave_points and solve_system_wave
 * (L lanes per system, K wavefronts per workgroup on S = K 64 / L
 * wavelengths of one column), with per-column shapes, so a column gives the
 * same bits as in a block of its own with the same L.  Items (column, first
 * of S wavelengths) are drawn from per-XCD queues of whole columns, longest
 * first; a workgroup whose queue is empty goes on to the next XCD's.  Each
 * system's LDS region is sized for the longest column of the launch. */
template <int MM, int K, int L>
__global__ __launch_bounds__(OCO_WAVE * K) void batch_mixed_wave(const oco_batch_args a)
{
    extern __shared__ double lds[];
    __shared__ int ticket;
    constexpr int S = K * (OCO_WAVE / L);
    const int wid = (int)threadIdx.x / OCO_WAVE;
    const int g = wid * (OCO_WAVE / L) + lane_id() / L;
    oco_cptr_i items = (oco_cptr_i)a.witems;
    oco_cptr_i qoff = (oco_cptr_i)a.wq_off;
    oco_cptr_i nzs = (oco_cptr_i)a.nz;
    oco_cptr_i nws = (oco_cptr_i)a.nw;
    oco_cptr_l coff = (oco_cptr_l)a.cell_off;
    oco_cptr_l woff = (oco_cptr_l)a.wt_off;
    oco_cptr_l uoff = (oco_cptr_l)a.u_off;
    oco_cptr_l voff = (oco_cptr_l)a.v_off;
    oco_cptr_d dzs = (oco_cptr_d)a.dz;
    const long lstr = wave_lds_words(a.max_nz);
    const int nq = gridDim.x % OCO_XCDS == 0 ? OCO_XCDS : 1;
    const int home = (int)(blockIdx.x % nq);
    for (int v = 0; v < nq; v++) {
        const int q = (home + v) % nq;
        const int i0 = nq == 1 ? 0 : qoff[q];
        const int i1 = nq == 1 ? qoff[OCO_TICKET_QUEUES] : qoff[q + 1];
        unsigned long long* ctr = a.wtickets + q * OCO_TICKET_STRIDE;
        for (;;) {
            __syncthreads();   /* the last item's systems are done with LDS, and every thread has read `ticket` */
            if (threadIdx.x == 0) ticket = (int)atomicAdd(ctr, 1ull);
            __syncthreads();
            const int t = i0 + ticket;
            if (t >= i1) break;
            const int col = items[2 * t], iw0 = items[2 * t + 1];
            const int nz = nzs[col], nw = nws[col];
            const int nk = nw - iw0 < S ? nw - iw0 : S;
            const long c0 = coff[col], v0 = voff[col];
            wave_points<K, OCO_PK_MIXED>(a.cells_soa + c0 * CR_NFIELD, nullptr, nz, a.wtab + woff[col] * OCO_WT_STRIDE,
                                         iw0, nk, a.A + v0 + (long)iw0 * oco_batch_ldv(nz), oco_batch_ldv(nz), lds, lstr,
                                         0.0, nullptr);
            __syncthreads();
            const bool valid = g < nk;
            const int iw = iw0 + (valid ? g : 0);
            oco_sys d;
            d.nz = nz;
            d.dz = dzs[3 * col]; d.dz1 = dzs[3 * col + 1]; d.dz2 = dzs[3 * col + 2]; d.sigma_grey = 0.0;
            d.cells = nullptr; d.rho = nullptr; d.T = nullptr;
            d.U = a.U + uoff[col] + (long)iw * oco_batch_ldu(nz);
            d.V = a.V + v0 + (long)iw * oco_batch_ldv(nz);
            d.A = a.A + v0 + (long)iw * oco_batch_ldv(nz);
            d.key = woff[col] + iw;
            solve_system_wave<MM, L>(d, lds + g * lstr, valid);
        }
    }
}

template <int K>
static const void* batch_wave_kernel_k(int max_nz, int lanes)
{
    switch (lanes) {
    case 8: return (const void*)batch_mixed_wave<9, K, 8>;
    case 16: return (const void*)batch_mixed_wave<9, K, 16>;
    case 32: return (const void*)batch_mixed_wave<9, K, 32>;
    default:
        return wave_rows(max_nz) <= 9 ? (const void*)batch_mixed_wave<9, K, 64>
                                      : (const void*)batch_mixed_wave<17, K, 64>;
    }
}

static const void* batch_wave_kernel(int max_nz, int k, int lanes)
{
    return k == 4 ? batch_wave_kernel_k<4>(max_nz, lanes)
                  : k == 2 ? batch_wave_kernel_k<2>(max_nz, lanes) : batch_wave_kernel_k<1>(max_nz, lanes);
}

int oco_batch_wave_grid(int max_nz, int k, int lanes, int device)
{
    if (!oco_wave_fits(max_nz, lanes, OCO_PK_MIXED, device)) return 0;
    if (wave_lds_bytes(max_nz, k, lanes) > (size_t)lds_limit(device)) return 0;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_wave_kernel(max_nz, k, lanes), OCO_WAVE * k,
                                                     wave_lds_bytes(max_nz, k, lanes)) != hipSuccess || per_cu < 1)
        per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1)
        cus = 256;
    const int g = per_cu * cus;
    return g >= OCO_XCDS ? g / OCO_XCDS * OCO_XCDS : g;
}

hipError_t oco_launch_batch_wave_records(const oco_batch_args& a, int ncol, double fully_ionized_T,
                                         hipStream_t stream)
{
    (void)hipGetLastError();   /* a pending not-ready status from an event query is not this launch's */
    if (!a.cells_soa || ncol <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(batch_reset, dim3(1), dim3(64), 0, stream, a.counter, a.fail);
    hipLaunchKernelGGL(cells_cols_fp64, dim3(ncol), dim3(256), 0, stream, a.rho, a.T, a.nz, a.cell_off,
                       fully_ionized_T, (double*)a.cells_soa);
    return hipGetLastError();
}

hipError_t oco_launch_batch_wave(const oco_batch_args& a, int k, int lanes, int nblocks, hipStream_t stream)
{
    (void)hipGetLastError();
    if (!a.wtickets || !a.cells_soa || nblocks <= 0 || !lanes_ok(lanes)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(tickets_reset, dim3(1), dim3(64), 0, stream, a.wtickets);
    const void* kern = batch_wave_kernel(a.max_nz, k, lanes);
    void* args[] = {(void*)&a};
    (void)hipLaunchKernel(kern, dim3(nblocks), dim3(OCO_WAVE * k), args, wave_lds_bytes(a.max_nz, k, lanes), stream);
    return hipGetLastError();
}

/* Athena input construction: one z slab (one id<iproc> file) of raw code-unit
 * fields into the block's rho, T -- the arithmetic of
 * source/read_athena.f90:163-205 (kinetic energy with the momentum-underflow
 * guard of :213-228, internal energy per mass), :232-258 (density unit and
 * floor, temperature unit) and the temperature rule of
 * source/disk.f90:96-104 (below switchTemp: the Gaussian profile or the
 * floor), in the reference's operation order.
 *
 * The file holds x fastest, the block z fastest, so the kernel is a
 * transpose with arithmetic in it.  A workgroup takes a tile of ATH_TX x
 * cells by ATH_TK z cells of one y row: it reads the five fields along x
 * (coalesced), computes rho and T per cell in registers, parks the two
 * results in LDS (rows padded by one double, so the column-wise reads hit
 * distinct banks) and writes them along z (coalesced).  Every cell is read
 * once and written once: 56 bytes per cell. */
#ifndef OCO_ATH_TX
#define OCO_ATH_TX 32
#endif
#ifndef OCO_ATH_TK
#define OCO_ATH_TK 64
#endif
constexpr int ATH_TX = OCO_ATH_TX, ATH_TK = OCO_ATH_TK, ATH_THREADS = 256;
static_assert(ATH_THREADS % ATH_TX == 0 && ATH_THREADS % ATH_TK == 0, "athena tile must divide the workgroup");

__global__ __launch_bounds__(ATH_THREADS) void athena_slab(const oco_athena_args a)
{
    __shared__ double s_rho[ATH_TK][ATH_TX + 1];
    __shared__ double s_T[ATH_TK][ATH_TX + 1];
    const double eps = 0x1.0p-52;   /* epsilon(0.0) under the reference's 8-byte reals */
    const int ntx = (a.nxloc + ATH_TX - 1) / ATH_TX, ntk = (a.nzl + ATH_TK - 1) / ATH_TK;
    const long ntiles = (long)ntx * ntk * a.nyloc;
    const long plane = (long)a.nyloc * a.nxloc;
    for (long t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int tx = (int)(t % ntx);
        const long r = t / ntx;
        const int tk = (int)(r % ntk);
        const int iy = (int)(r / ntk);
        const int ix0 = tx * ATH_TX, k0 = tk * ATH_TK;
        {   /* read phase: lanes along x; every load of the thread's rows is
               issued before the arithmetic (the divides) of the first */
            constexpr int STEP = ATH_THREADS / ATH_TX, ROWS = ATH_TK / STEP;
            const int lx = threadIdx.x % ATH_TX, ix = ix0 + lx, kk0 = threadIdx.x / ATH_TX;
            double f[ROWS][5];
            bool ok[ROWS];
#pragma unroll
            for (int j = 0; j < ROWS; j++) {
                const int k = k0 + kk0 + j * STEP;
                ok[j] = ix < a.nxloc && k < a.nzl;
                const long src = ok[j] ? (long)k * plane + (long)iy * a.nxloc + ix : 0;
                f[j][0] = ok[j] ? a.rho[src] : 1.0;
                f[j][1] = ok[j] ? a.mx[src] : 0.0;
                f[j][2] = ok[j] ? a.my[src] : 0.0;
                f[j][3] = ok[j] ? a.mz[src] : 0.0;
                f[j][4] = ok[j] ? a.e[src] : 0.0;
            }
#pragma unroll
            for (int j = 0; j < ROWS; j++) {
                if (!ok[j]) continue;
                const int kk = kk0 + j * STEP;
                const double rho = f[j][0];
                double e = 0.0;
#pragma unroll
                for (int c = 1; c <= 3; c++) {
                    const double p = f[j][c];
                    if (!a.avoid_momentum_underflow || __builtin_fabs(p) >= eps) e = e + (0.5 * (p * p)) / rho;
                }
                e = (f[j][4] - e) / rho;
                double rc = rho * a.unit_density;
                if (rc < a.rho_floor) rc = a.rho_floor;
                double T = (e * a.gamma1) * a.unit_temperature;
                if (T < a.switch_temp) {
                    const double tg = a.tgauss[a.iz0 + k0 + kk];
                    T = tg > a.floor_temp ? tg : a.floor_temp;
                }
                s_rho[kk][lx] = rc;
                s_T[kk][lx] = T;
            }
        }
        __syncthreads();
        {   /* write phase: lanes along z */
            const int wk = threadIdx.x % ATH_TK, k = k0 + wk;
            for (int xx = threadIdx.x / ATH_TK; xx < ATH_TX; xx += ATH_THREADS / ATH_TK) {
                const int ix = ix0 + xx;
                if (ix >= a.nxloc || k >= a.nzl) continue;
                const long dst = ((long)ix * a.nyloc + iy) * a.nz + a.iz0 + k;
                a.rho3d[dst] = s_rho[wk][xx];
                a.T3d[dst] = s_T[wk][xx];
            }
        }
        __syncthreads();
    }
}

hipError_t oco_launch_athena(const oco_athena_args& a, hipStream_t stream)
{
    (void)hipGetLastError();
    const long ntiles = (long)((a.nxloc + ATH_TX - 1) / ATH_TX) * ((a.nzl + ATH_TK - 1) / ATH_TK) * a.nyloc;
    if (ntiles <= 0) return hipSuccess;
    const long blocks = ntiles < 65536 ? ntiles : 65536;
    hipLaunchKernelGGL(athena_slab, dim3((unsigned)blocks), dim3(ATH_THREADS), 0, stream, a);
    return hipGetLastError();
}

/* Opacity query: alpha, albedo, S of (point, wavelength) pairs, one lane
 * each, wavelength fastest in the outputs (coalesced stores).  Non-grey:
 * the point's cell record (cells_fp64) and the wavelength's table row, the
 * sweeps' per-point code (oco_point / oco_point_mixed); grey: the grey
 * parameters of the point's rho, T. */
template <bool MIX, bool GREY>
__global__ __launch_bounds__(256) void opacity_points(const oco_points_args a)
{
    __shared__ oco_mtab smt;
    if (!MIX || GREY) load_mtab(&smt);
    const long n = (long)a.npts * a.nw;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const long pt = i / a.nw;
        const int iw = (int)(i - pt * a.nw);
        oco_pt p;
        if (GREY) {
            p = oco_grey_point(a.rho[pt], a.T[pt], a.sigma_grey, &smt);
        } else {
            double c[CR_NFIELD], w[OCO_WT_STRIDE];
#pragma unroll
            for (int k = 0; k < CR_NFIELD; k++) c[k] = a.cells[pt * OCO_CELL_STRIDE + k];
#pragma unroll
            for (int k = 0; k < OCO_WT_STRIDE; k++) w[k] = a.wtab[(long)iw * OCO_WT_STRIDE + k];
            if (MIX) p = oco_point_mixed<1>(c, w);
            else p = oco_point(c, w, &smt);
        }
        a.alpha[i] = p.alpha;
        if (a.albedo) a.albedo[i] = p.albedo;
        if (a.S) a.S[i] = p.S;
    }
}

hipError_t oco_launch_points(const oco_points_args& a, int lgrey, int mixed, hipStream_t stream)
{
    (void)hipGetLastError();
    const long n = (long)a.npts * a.nw;
    if (n <= 0) return hipSuccess;
    long blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    const dim3 g((unsigned)blocks), b(256);
    if (lgrey) hipLaunchKernelGGL((opacity_points<false, true>), g, b, 0, stream, a);
    else if (mixed) hipLaunchKernelGGL((opacity_points<true, false>), g, b, 0, stream, a);
    else hipLaunchKernelGGL((opacity_points<false, false>), g, b, 0, stream, a);
    return hipGetLastError();
}

hipError_t oco_launch_cells(const double* rho, const double* T, long ncells, double fully_ionized_T,
                            double* cells, int soa_nz, hipStream_t stream)
{
    (void)hipGetLastError();
    if (ncells <= 0) return hipSuccess;
    long blocks = (ncells + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (soa_nz > 0)
        hipLaunchKernelGGL(cells_fp64<true>, dim3((unsigned)blocks), dim3(256), 0, stream, rho, T, ncells, soa_nz,
                           fully_ionized_T, cells);
    else
        hipLaunchKernelGGL(cells_fp64<false>, dim3((unsigned)blocks), dim3(256), 0, stream, rho, T, ncells, 0,
                           fully_ionized_T, cells);
    return hipGetLastError();
}

/* Workgroups of the sweep kernel that fit on the device at once (occupancy x
 * CUs): the grid-stride sweep is launched with exactly that many, so every
 * workgroup is resident from the start and takes an equal share of tiles. */
/* wavefronts per workgroup of the lane-per-system sweep */
int oco_sweep_wg(void) { return OCO_SWEEP_WG; }

/* resident workgroups of the lane-per-system sweep (the FP64 drop-in) */
int oco_sweep_grid(int lgrey, int uniform, int device)
{
    const void* k;
    if (lgrey) k = (const void*)sweep_fp64<true, false>;
    else k = uniform ? (const void*)sweep_fp64<false, true> : (const void*)sweep_fp64<false, false>;
    const int wg = OCO_SWEEP_WG;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, OCO_WAVE * wg, 0) != hipSuccess || per_cu < 1)
        per_cu = 8 / wg > 0 ? 8 / wg : 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1)
        cus = 256;
    return per_cu * cus;
}

static void launch_wave(const oco_sweep_args& a, int k, int lanes, int pk, int nblocks, hipStream_t stream)
{
    if (pk == OCO_PK_GREY) k = 1;
    const void* kern = wave_kernel(a.nz, k, lanes, pk);
    void* args[] = {(void*)&a};
    (void)hipLaunchKernel(kern, dim3(nblocks), dim3(OCO_WAVE * k), args, wave_lds_bytes(a.nz, k, lanes, pk), stream);
}

hipError_t oco_launch_sweep(const oco_sweep_args& a, int lgrey, int wave, int lanes, int pk, int nblocks,
                            hipStream_t stream)
{
    (void)hipGetLastError();   /* a pending not-ready status from an event query is not this launch's */
    const bool uniform = (a.nw % OCO_WAVE) == 0;
    if (a.canary) {
        if (lgrey)
            hipLaunchKernelGGL((canary_fp64<true>), dim3(1), dim3(256), 0, stream, a, a.canary);
        else
            hipLaunchKernelGGL((canary_fp64<false>), dim3(1), dim3(256), 0, stream, a, a.canary);
    }
    if (!a.tickets) return hipErrorInvalidValue;
    if (a.nz <= OCO_TINY_NZ) {   /* 2 or 3 cells: the FP64 drop-in only */
        if (wave || a.nz < 2) return hipErrorInvalidValue;
        const long nsys = a.ncol * a.nw;
        const long nb = (nsys + 255) / 256 < 4096 ? (nsys + 255) / 256 : 4096;
        if (lgrey)
            hipLaunchKernelGGL((sweep_tiny_fp64<true>), dim3((unsigned)nb), dim3(256), 0, stream, a);
        else
            hipLaunchKernelGGL((sweep_tiny_fp64<false>), dim3((unsigned)nb), dim3(256), 0, stream, a);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(tickets_reset, dim3(1), dim3(64), 0, stream, a.tickets);
    const dim3 g(nblocks), bf(OCO_WAVE * OCO_SWEEP_WG);
    if (wave) {
        launch_wave(a, wave, lanes, pk, nblocks, stream);
    } else if (lgrey) {
        hipLaunchKernelGGL((sweep_fp64<true, false>), g, bf, 0, stream, a);
    } else if (uniform) {
        hipLaunchKernelGGL((sweep_fp64<false, true>), g, bf, 0, stream, a);
    } else {
        hipLaunchKernelGGL((sweep_fp64<false, false>), g, bf, 0, stream, a);
    }
    return hipGetLastError();
}

/* ---- the allocation-free block solve (oco_solve_block_device_noalloc) ----
 * Its scratch -- the work-queue tickets, the failure key, the canary sums,
 * the first pillar's min/max and the field-major cell records -- lives in
 * one file-scope device array, `scratch`, of OCO_STATIC_SCRATCH_BYTES: every
 * GPU the library runs on holds its own copy, at its own address, until the
 * process ends.  Host code never passes the array's name to a kernel (on
 * the host it is a stand-in, not a device address): oco_static_scratch
 * looks up each device's address once, with that device current, and the
 * kernels take that pointer like any other buffer. */
static __device__ __align__(256) unsigned char scratch[OCO_STATIC_SCRATCH_BYTES];

#define OCO_MAX_DEVICES 64
static void* scratch_dev[OCO_MAX_DEVICES];
static std::mutex scratch_mu;

int oco_static_scratch(int device, void** ptr, size_t* bytes)
{
    if (device < 0 || device >= OCO_MAX_DEVICES) return OCO_E_ARG;
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_dev[device]) {
        int cur = -1;
        if (hipGetDevice(&cur) != hipSuccess || cur != device) return OCO_E_HIP;   /* the device must be current */
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return OCO_E_HIP;
        }
        scratch_dev[device] = p;
    }
    *ptr = scratch_dev[device];
    *bytes = OCO_STATIC_SCRATCH_BYTES;
    return OCO_OK;
}

/* min/max of the first pillar's U (interior), V and alpha
 * (source/flux_feautrier.f90:222-224) into out[6]: minU, maxU, minV, maxV,
 * minA, maxA; one thread (nz is at most a few thousand). */
__global__ void pillar_minmax(const double* U, const double* V, const double* A, int nz, double* out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double mu = U[3], xu = U[3], mv = V[0], xv = V[0], ma = A[0], xa = A[0];
    for (int i = 0; i < nz; i++) {
        mu = fmin(mu, U[3 + i]); xu = fmax(xu, U[3 + i]);
        mv = fmin(mv, V[i]); xv = fmax(xv, V[i]);
        ma = fmin(ma, A[i]); xa = fmax(xa, A[i]);
    }
    out[0] = mu; out[1] = xu; out[2] = mv; out[3] = xv; out[4] = ma; out[5] = xa;
}

hipError_t oco_launch_pillar_minmax(const double* U, const double* V, const double* A, int nz, double* out,
                                    hipStream_t stream)
{
    (void)hipGetLastError();
    hipLaunchKernelGGL(pillar_minmax, dim3(1), dim3(64), 0, stream, U, V, A, nz, out);
    return hipGetLastError();
}
