/*
 * oco_api.cpp -- the C ABI (include/ocotillo.h): context, resident
 * wavelength tables, device buffers and the block solve.
 */
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ocotillo.h"
#include "oco_internal.h"
#include "oco_physics.h"

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
};

}  // namespace

struct oco_ctx {
    int device = 0;
    size_t budget = 0;      /* 0: unlimited */
    size_t used = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    int nw = 0;
    std::vector<double> waves;
    double fully_ionized_T = 2e4;
    int lgrey = 0;
    double sigma_grey = 0.4;
    int wave_lanes = 0;   /* lanes per system of the difference-form sweep: 8, 16, 32, 64, 0 = by nz */
    int wave_group = 0;   /* wavefronts per workgroup of the difference-form sweep: 1, 2, 4, 0 = default */
    int checks = 0;       /* OCO_CHECK_* flags of the block solves */
    DevBuf wtab, rho, T, cells, scratch, U, V, A, fail, canary, raw, tg, tickets, fkey;
    /* block outputs to files: copy stream, per-slot events, pinned staging */
    hipStream_t copy_stream = nullptr;
    hipEvent_t ev_sweep[2] = {nullptr, nullptr}, ev_copy[2] = {nullptr, nullptr};
    void* pinned = nullptr;
    size_t pinned_n = 0;
    /* this device's static scratch (oco_solve_block_device_noalloc) */
    void* sscr = nullptr;
    size_t sscr_n = 0;
};

static int hip_status(hipError_t e) { return e == hipSuccess ? OCO_OK : OCO_E_HIP; }

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            fprintf(stderr, "ocotillo: %s failed: %s\n", #x, hipGetErrorString(e_));        \
            return OCO_E_HIP;                                                              \
        }                                                                                  \
    } while (0)

static void release(oco_ctx* c, DevBuf& b)
{
    if (b.p) {
        (void)hipFree(b.p);
        c->used -= b.n;
    }
    b.p = nullptr;
    b.n = 0;
}

static int ensure(oco_ctx* c, DevBuf& b, size_t bytes)
{
    if (bytes == 0) bytes = 8;
    if (b.n >= bytes) return OCO_OK;
    release(c, b);
    if (c->budget && c->used + bytes > c->budget) return OCO_E_OOM;
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return OCO_E_OOM;
    }
    b.p = p;
    b.n = bytes;
    c->used += bytes;
    return OCO_OK;
}

/* Fortran x**n for a run-time integer n, as the reference's compiler
 * evaluates wa**(i-1) and lgwave**(j-1) (source/continuous_opacity.f90:108,136). */
static double powi_rt(double x, int n)
{
    double r = 1.0, b = x;
    unsigned m = (unsigned)n;
    while (m) {
        if (m & 1u) r = r * b;
        m >>= 1;
        if (m) b = b * b;
    }
    return r;
}

/* gaunt (source/continuous_opacity.f90:343-375) */
static double gaunt(int n, double wave_cm)
{
    double lam_n = (double)(n * n) * OCO_HCR1;
    if (wave_cm <= lam_n) {
        double q = (wave_cm / lam_n) - 1.0;
        double eps2 = q * q;
        return 1.0 - (0.17285714285714285 * (1.0 - eps2)) * std::pow((double)n * (1.0 + eps2), -OCO_TWO_THIRDS);
    }
    return 0.0;
}

/* pre_calc_opacity_quantities (source/continuous_opacity.f90:64-141) plus the
 * wavelength-only factors of the per-point formulas.  Runs once per grid on
 * the host, with the host libm exactly as the reference does. */
void oco_build_wtab(const std::vector<double>& waves, std::vector<double>& wt)
{
    static const double a_coeff[7] = {+1.99654, -1.18267e-6, +2.62423e-7, -4.40524e-11,
                                      +3.23992e-15, -1.39568e-19, +2.78701e-24};
    static const double b_coeff[3][5] = {{-2.276300, -1.685000, +0.766610, -0.053356, +0.000000},
                                         {+15.28270, -9.284600, +1.993810, -0.142631, +0.000000},
                                         {-197.789, +190.266, -67.9775, +10.6913, -0.62515}};
    const double sfc = (2.0 * OCO_H_PLANCK) * (OCO_C_LIGHT * OCO_C_LIGHT);
    const size_t nw = waves.size();
    wt.assign(nw * OCO_WT_STRIDE, 0.0);
    for (size_t i = 0; i < nw; i++) {
        double* w = &wt[i * OCO_WT_STRIDE];
        double wa = waves[i];
        double wcm = wa * 1e-8;
        double wcm1 = 1.0 / wcm;
        double nu = OCO_C_LIGHT * wcm1;
        double chi = 1.2398e4 / wa;
        w[WT_CHI] = chi;
        w[WT_WA3] = (wa * wa) * wa;
        for (int n = 1; n <= OCO_MDIM; n++) {
            double n1 = 1.0 / (double)n;
            w[WT_G1 + n - 1] = gaunt(n, wcm) * ((n1 * n1) * n1);
        }
        double s = 0.0;
        for (int k = 1; k <= 7; k++) s = s + (1e-17 * a_coeff[k - 1]) * powi_rt(wa, k - 1);
        w[WT_SCOEFF] = s;
        w[WT_HMBF_ON] = (wa < 16000.0) ? 1.0 : 0.0;
        double lgw = std::log10(wa);
        for (int r = 0; r < 3; r++) {
            double f = 0.0;
            for (int j = 1; j <= 5; j++) f = f + powi_rt(lgw, j - 1) * b_coeff[r][j - 1];
            w[WT_F1 + r] = f;
        }
        w[WT_GFF] = 0.3456 * std::pow(wa * OCO_RANGSTROM, -OCO_ONE_THIRD);
        w[WT_CHI1] = 1.0 / chi;
        w[WT_NU9] = nu * 1e-9;
        w[WT_L3NU] = 3.0 * oco_log(nu);
        w[WT_HNK] = (OCO_H_PLANCK * nu) * OCO_K1_CGS;
        w[WT_WCM1] = wcm1;
        w[WT_SFW5] = sfc * ((((wcm1 * wcm1) * wcm1) * wcm1) * wcm1);
        w[WT_HBFW] = OCO_AHBF * w[WT_WA3];
        w[WT_HFFW] = OCO_ALPHA0 * w[WT_WA3];
        w[WT_F1M26] = w[WT_F1] - 26.0;
        w[WT_GBP] = oco_exp(5.960) * std::pow(w[WT_NU9], -(OCO_SQRT3 * OCO_PI1));
        w[WT_NU3] = oco_exp(-w[WT_L3NU]);
        w[WT_CHIL] = chi * OCO_LN10_HI;
    }
}

int oco_ctx_device(const oco_ctx* c) { return c->device; }
double oco_ctx_fully_ionized_T(const oco_ctx* c) { return c->fully_ionized_T; }

extern "C" {

const char* oco_version(void) { return "ocotillo-mi355x 0.1 (gfx950, FP64 drop-in)"; }

const char* oco_strerror(int s)
{
    switch (s) {
    case OCO_OK: return "ok";
    case OCO_E_PIVOT0_STAGE1: return "ERROR tridiag stage 1: bet=b(1)) = 0";
    case OCO_E_PIVOT0_STAGE2: return "ERROR tridiag stage 2: bet = b(j)-a(j)*gam(j) = 0";
    case OCO_E_ARG: return "invalid argument";
    case OCO_E_OOM: return "device memory exhausted or HBM budget exceeded";
    case OCO_E_HIP: return "HIP runtime error";
    case OCO_E_STATE: return "context not ready (wavelengths not set)";
    case OCO_E_NODEVICE: return "no HIP device";
    case OCO_E_BUSY: return "context busy";
    case OCO_E_IO: return "output file could not be written";
    case OCO_E_NONFINITE: return "an output is NaN or infinite (OCO_CHECK_FINITE)";
    default: return "unknown status";
    }
}

int oco_create(int device, size_t hbm_budget, oco_ctx** out)
{
    if (!out) return OCO_E_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return OCO_E_NODEVICE;
    }
    if (device < 0 || device >= n) return OCO_E_ARG;
    HIPCHK(hipSetDevice(device));
    oco_ctx* c = new (std::nothrow) oco_ctx;
    if (!c) return OCO_E_OOM;
    c->device = device;
    c->budget = hbm_budget;
    /* OCO_WAVE_LANES=8|16|32|64: lanes per system of the difference-form
     * sweep (results within the same bounds; default: by column length) */
    if (const char* e = getenv("OCO_WAVE_LANES")) c->wave_lanes = atoi(e);
    /* OCO_WAVE_GROUP=1|2|4: wavefronts per workgroup of the difference-form
     * sweep (the same bits; default: the library's measured choice) */
    if (const char* e = getenv("OCO_WAVE_GROUP")) c->wave_group = atoi(e);
    /* OCO_CHECK_FINITE=1: oco_set_checks(ctx, OCO_CHECK_FINITE) from the start */
    if (const char* e = getenv("OCO_CHECK_FINITE")) c->checks = atoi(e) ? OCO_CHECK_FINITE : 0;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return OCO_E_HIP;
    }
    for (auto& e : c->ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            oco_destroy(c);
            return OCO_E_HIP;
        }
    }
    /* the allocation-free solve's scratch on this device, looked up here
     * with the device current (once per device for the process) */
    if (oco_static_scratch(device, &c->sscr, &c->sscr_n) != OCO_OK) {
        oco_destroy(c);
        return OCO_E_HIP;
    }
    *out = c;
    return OCO_OK;
}

int oco_destroy(oco_ctx* c)
{
    if (!c) return OCO_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (DevBuf* b : {&c->wtab, &c->rho, &c->T, &c->cells, &c->scratch, &c->U, &c->V, &c->A, &c->fail,
                      &c->canary, &c->raw, &c->tg, &c->tickets, &c->fkey})
        release(c, *b);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (int i = 0; i < 2; i++) {
        if (c->ev_sweep[i]) (void)hipEventDestroy(c->ev_sweep[i]);
        if (c->ev_copy[i]) (void)hipEventDestroy(c->ev_copy[i]);
    }
    if (c->copy_stream) {
        (void)hipStreamSynchronize(c->copy_stream);
        (void)hipStreamDestroy(c->copy_stream);
    }
    if (c->pinned) (void)hipHostFree(c->pinned);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return OCO_OK;
}

int oco_set_wavelengths(oco_ctx* c, int nw, const double* waves)
{
    if (!c || nw <= 0 || !waves) return OCO_E_ARG;
    for (int i = 0; i < nw; i++)
        if (!(waves[i] > 0.0) || !std::isfinite(waves[i])) return OCO_E_ARG;
    HIPCHK(hipSetDevice(c->device));
    std::vector<double> wv(waves, waves + nw), wt;
    oco_build_wtab(wv, wt);
    /* until the new table is on the device the context holds no grid, so a
     * failure below leaves later solves returning OCO_E_STATE (ensure frees
     * the old table before it allocates the new one) */
    c->nw = 0;
    c->waves.clear();
    int st = ensure(c, c->wtab, wt.size() * sizeof(double));
    if (st) return st;
    hipError_t e = hipMemcpyAsync(c->wtab.p, wt.data(), wt.size() * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        fprintf(stderr, "ocotillo: oco_set_wavelengths: %s\n", hipGetErrorString(e));
        return OCO_E_HIP;
    }
    c->waves = std::move(wv);
    c->nw = nw;
    return OCO_OK;
}

int oco_set_checks(oco_ctx* c, int flags)
{
    if (!c || (flags & ~OCO_CHECK_FINITE)) return OCO_E_ARG;
    c->checks = flags;
    return OCO_OK;
}

int oco_set_physics(oco_ctx* c, double fully_ionized_T, int lgrey, double sigma_grey)
{
    if (!c) return OCO_E_ARG;
    c->fully_ionized_T = fully_ionized_T;
    c->lgrey = lgrey ? 1 : 0;
    c->sigma_grey = sigma_grey;
    return OCO_OK;
}

}  /* extern "C" */

/* Kernel and grid for sweeping nw_run wavelengths of ncol columns. */
struct SweepPlan {
    int wave = 0, nblocks = 0;   /* wave: wavefronts per workgroup of the difference-form sweep, 0 for the drop-in */
    int lanes = 64, pk = 0;      /* difference-form sweep: lanes per system, point kind */
    int wg = 1;                  /* drop-in sweep: wavefronts per workgroup (each with its own scratch slot) */
    int ok = 1;                  /* 0: the mode cannot take columns of nz cells */
};

static bool mode_ok(int mode) { return mode == OCO_MODE_FP64 || mode == OCO_MODE_MIXED || mode == OCO_MODE_ACCURATE; }

static SweepPlan sweep_plan(const oco_ctx* c, int nz, long ncol, int nw_run, int mode)
{
    SweepPlan p;
    const long nsys = ncol * nw_run;
    const long ntiles = (nsys + 63) / 64;
    if (mode != OCO_MODE_FP64) {
        /* difference form: grey points from rho and T in both modes */
        p.pk = c->lgrey ? OCO_PK_GREY : mode == OCO_MODE_MIXED ? OCO_PK_MIXED : OCO_PK_FP64;
        p.lanes = oco_wave_lanes(nz, c->wave_lanes);
        if (!oco_wave_fits(nz, p.lanes, p.pk, c->device)) {
            p.ok = 0;
            return p;
        }
        p.wave = p.pk == OCO_PK_GREY ? 1 : oco_wave_group(nz, p.lanes, p.pk, c->wave_group, c->device);
        const int S = p.wave * (64 / p.lanes);
        const long ngroups = ncol * ((nw_run + S - 1) / S);
        const int grid = oco_wave_grid(nz, p.wave, p.lanes, p.pk, c->device);
        p.nblocks = (int)(ngroups < grid ? ngroups : grid);
    } else {
        const int grid = oco_sweep_grid(c->lgrey, (nw_run % 64) == 0, c->device);
        p.wg = oco_sweep_wg();
        const long nb = (ntiles + p.wg - 1) / p.wg;
        p.nblocks = (int)(nb < grid ? nb : grid);
    }
    return p;
}

/* Cell records, scratch, pivot key and canary buffers for a plan. */
static int sweep_buffers(oco_ctx* c, const SweepPlan& p, int nz, long ncells)
{
    int st;
    if (!c->lgrey && (st = ensure(c, c->cells, (size_t)ncells * OCO_CELL_STRIDE * sizeof(double)))) return st;
    if (!p.wave) {
        /* one scratch slot per wavefront */
        const size_t slots = (size_t)p.nblocks * p.wg;
        if ((st = ensure(c, c->scratch, slots * nz * 3 * 64 * sizeof(double)))) return st;
    }
    if ((st = ensure(c, c->tickets, OCO_TICKET_BYTES))) return st;
    if ((st = ensure(c, c->fail, sizeof(unsigned long long)))) return st;
    if ((st = ensure(c, c->canary, (4 + 5 * (size_t)nz) * sizeof(double)))) return st;
    return OCO_OK;
}

static oco_sweep_args sweep_args(const oco_ctx* c, int nz, long ncol, double dz, double dz1, double dz2,
                                 const double* d_rho, const double* d_T)
{
    oco_sweep_args a;
    a.nz = nz;
    a.nw = c->nw;
    a.ncol = ncol;
    a.dz = dz;
    a.dz1 = dz1;
    a.dz2 = dz2;
    a.sigma_grey = c->sigma_grey;
    a.fully_ionized_T = c->fully_ionized_T;
    a.rho = d_rho;
    a.T = d_T;
    a.cells = (const double*)c->cells.p;
    a.wtab = (const double*)c->wtab.p;
    a.U = a.V = a.A = nullptr;
    a.scratch = (double*)c->scratch.p;
    a.fail = (unsigned long long*)c->fail.p;
    a.canary = (double*)c->canary.p;
    a.tickets = (unsigned long long*)c->tickets.p;
    return a;
}

/* First zero pivot of a sweep over wavelengths [iw0, iw0+nw_run) as a key
 * in the reference's loop order over the whole block (column-major, then
 * wavelength, then stage and row), or OCO_FAIL_NONE. */
static unsigned long long global_key(unsigned long long key, int iw0, int nw_run, int nw)
{
    if (key == OCO_FAIL_NONE) return key;
    const unsigned long long sys = key >> 32;
    const unsigned long long gsys = (sys / nw_run) * nw + iw0 + sys % nw_run;
    return (gsys << 32) | (key & 0xffffffffULL);
}

static void decode_key(unsigned long long key, int nw, int nyloc, oco_diag* d)
{
    const long sys = (long)(key >> 32);
    d->status = (int)((key >> 28) & 0xf);
    d->fail_j = (int)(key & 0x0fffffffULL);
    const long col = sys / nw;
    d->fail_iw = (int)(sys % nw) + 1;
    d->fail_ix = (int)(col / nyloc) + 1;
    d->fail_iy = (int)(col % nyloc) + 1;
}

/* min/max of the first pillar's U (interior), V, absorp
 * (source/flux_feautrier.f90:222-224). */
static void pillar_minmax(const double* U, const double* V, const double* A, int nz, oco_diag* d)
{
    const double* Uh = U + 3;
    d->minU = d->maxU = Uh[0];
    d->minV = d->maxV = V[0];
    d->minA = d->maxA = A[0];
    for (int i = 0; i < nz; i++) {
        d->minU = fmin(d->minU, Uh[i]); d->maxU = fmax(d->maxU, Uh[i]);
        d->minV = fmin(d->minV, V[i]); d->maxV = fmax(d->maxV, V[i]);
        d->minA = fmin(d->minA, A[i]); d->maxA = fmax(d->maxA, A[i]);
    }
}

/* OCO_CHECK_FINITE: enqueue the non-finite scan of one chunk of a block's
 * outputs (wavelengths iw0 .. iw0 + nw_run - 1 of nw) into *key. */
static hipError_t finite_scan(int nz, int nyloc, int nxloc, int iw0, int nw_run, int nw, const double* U,
                              const double* V, const double* A, unsigned long long* key, hipStream_t s)
{
    const long ncol = (long)nxloc * nyloc;
    hipError_t e = oco_launch_finite_scan(A, ncol * nz * nw_run, nz, 0, nyloc, nxloc, iw0, nw, nz, key, s);
    if (e == hipSuccess)
        e = oco_launch_finite_scan(U, ncol * (nz + 6) * nw_run, nz + 6, 1, nyloc, nxloc, iw0, nw, nz, key, s);
    if (e == hipSuccess) e = oco_launch_finite_scan(V, ncol * nz * nw_run, nz, 2, nyloc, nxloc, iw0, nw, nz, key, s);
    return e;
}

/* the first non-finite output (finite_scan's key) as OCO_E_NONFINITE and its
 * system and cell */
static void decode_finite(unsigned long long key, int nz, int nw, int nyloc, oco_diag* d)
{
    const unsigned long long mz = (unsigned long long)(nz + 6);
    const long sys = (long)(key / (3 * mz));
    const int which = (int)((key / mz) % 3), e = (int)(key % mz);
    const long col = sys / nw;
    d->status = OCO_E_NONFINITE;
    d->fail_iw = (int)(sys % nw) + 1;
    d->fail_ix = (int)(col / nyloc) + 1;
    d->fail_iy = (int)(col % nyloc) + 1;
    d->fail_j = which == 1 ? e - 2 : e + 1;
}

extern "C" {

int oco_solve_block_device(oco_ctx* c, int nz, int nyloc, int nxloc, double dz, double dz1, double dz2,
                           const double* d_rho, const double* d_T, double* d_U, double* d_V, double* d_A,
                           int mode, oco_diag* diag)
{
    if (!c) return OCO_E_ARG;
    if (nz < 2 || nyloc < 1 || nxloc < 1 || !d_rho || !d_T || !d_U || !d_V || !d_A) return OCO_E_ARG;
    if (!mode_ok(mode)) return OCO_E_ARG;
    if (c->nw <= 0) return OCO_E_STATE;
    if (c->lgrey && c->nw != 1) return OCO_E_ARG;   /* source/flux_feautrier.f90:84-87 */
    HIPCHK(hipSetDevice(c->device));
    const long ncol = (long)nxloc * nyloc;
    const long ncells = ncol * nz;
    const SweepPlan p = sweep_plan(c, nz, ncol, c->nw, mode);
    if (!p.ok) return OCO_E_ARG;
    int st;
    if ((st = sweep_buffers(c, p, nz, ncells))) return st;
    const bool check = (c->checks & OCO_CHECK_FINITE) != 0;
    if (check && (st = ensure(c, c->fkey, sizeof(unsigned long long)))) return st;

    const unsigned long long none = OCO_FAIL_NONE;
    HIPCHK(hipMemcpyAsync(c->fail.p, &none, sizeof(none), hipMemcpyHostToDevice, c->stream));
    if (check) HIPCHK(oco_launch_fail_reset((unsigned long long*)c->fkey.p, 1, c->stream));

    oco_sweep_args a = sweep_args(c, nz, ncol, dz, dz1, dz2, d_rho, d_T);
    a.U = d_U;
    a.V = d_V;
    a.A = d_A;

    HIPCHK(hipEventRecord(c->ev[0], c->stream));
    if (!c->lgrey)
        HIPCHK(oco_launch_cells(d_rho, d_T, ncells, c->fully_ionized_T, (double*)c->cells.p, p.wave ? nz : 0,
                                c->stream));
    HIPCHK(hipEventRecord(c->ev[1], c->stream));
    HIPCHK(oco_launch_sweep(a, c->lgrey, p.wave, p.lanes, p.pk, p.nblocks, c->stream));
    HIPCHK(hipEventRecord(c->ev[2], c->stream));
    unsigned long long fkey = none;
    if (check) {
        HIPCHK(finite_scan(nz, nyloc, nxloc, 0, c->nw, c->nw, d_U, d_V, d_A, (unsigned long long*)c->fkey.p,
                           c->stream));
        HIPCHK(hipMemcpyAsync(&fkey, c->fkey.p, sizeof(fkey), hipMemcpyDeviceToHost, c->stream));
    }

    unsigned long long key = none;
    double can[4] = {0, 0, 0, 0};
    HIPCHK(hipMemcpyAsync(&key, c->fail.p, sizeof(key), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(can, c->canary.p, sizeof(can), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));

    oco_diag d;
    memset(&d, 0, sizeof(d));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev[0], c->ev[1]) == hipSuccess) d.cells_ms = ms;
    if (hipEventElapsedTime(&ms, c->ev[1], c->ev[2]) == hipSuccess) d.sweep_ms = ms;
    if (hipEventElapsedTime(&ms, c->ev[0], c->ev[2]) == hipSuccess) d.total_ms = ms;
    d.sum_a = can[0];
    d.sum_b = can[1];
    d.sum_c = can[2];
    d.sum_d = can[3];
    if (key != none) {
        decode_key(key, c->nw, nyloc, &d);
    } else {
        /* first-pillar canaries: system (ix,iy,iw) = (1,1,1) is the first of each array */
        std::vector<double> h((size_t)(nz + 6) + 2 * (size_t)nz);
        HIPCHK(hipMemcpy(h.data(), d_U, (nz + 6) * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(h.data() + nz + 6, d_V, nz * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(h.data() + 2 * nz + 6, d_A, nz * sizeof(double), hipMemcpyDeviceToHost));
        pillar_minmax(h.data(), h.data() + nz + 6, h.data() + 2 * nz + 6, nz, &d);
        if (fkey != none) decode_finite(fkey, nz, c->nw, nyloc, &d);
    }
    if (diag) *diag = d;
    return d.status;
}

/* Byte layout of the allocation-free solve in the static scratch: tickets,
 * failure key, canary sums, first-pillar min/max, then the field-major cell
 * records (none for grey columns), each 256-byte aligned; the total. */
struct StaticLayout {
    size_t tickets, fail, fkey, canary, minmax, cells, total;
};

static size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

static StaticLayout static_layout(int nz, long ncells, int lgrey)
{
    StaticLayout L;
    L.tickets = 0;
    L.fail = L.tickets + align256(OCO_TICKET_BYTES);
    L.fkey = L.fail + align256(sizeof(unsigned long long));
    L.canary = L.fkey + align256(sizeof(unsigned long long));
    L.minmax = L.canary + align256((4 + 5 * (size_t)nz) * sizeof(double));
    L.cells = L.minmax + align256(6 * sizeof(double));
    L.total = L.cells + (lgrey ? 0 : (size_t)ncells * CR_NFIELD * sizeof(double));
    return L;
}

int oco_solve_block_device_noalloc(oco_ctx* c, int nz, int nyloc, int nxloc, double dz, double dz1, double dz2,
                                   const double* d_rho, const double* d_T, double* d_U, double* d_V, double* d_A,
                                   int mode, oco_diag* diag)
{
    if (!c) return OCO_E_ARG;
    if (nz < 4 || nyloc < 1 || nxloc < 1 || !d_rho || !d_T || !d_U || !d_V || !d_A) return OCO_E_ARG;
    /* the difference-form sweep needs no elimination scratch; the drop-in's
     * lane sweep does (16 B per resident lane and cell) */
    if (mode != OCO_MODE_MIXED && mode != OCO_MODE_ACCURATE) return OCO_E_ARG;
    if (c->nw <= 0) return OCO_E_STATE;
    if (c->lgrey && c->nw != 1) return OCO_E_ARG;   /* source/flux_feautrier.f90:84-87 */
    if (!c->sscr) return OCO_E_STATE;
    const long ncol = (long)nxloc * nyloc;
    const long ncells = ncol * nz;
    const StaticLayout L = static_layout(nz, ncells, c->lgrey);
    if (L.total > c->sscr_n) {
        fprintf(stderr, "ocotillo: oco_solve_block_device_noalloc: %zu bytes of scratch needed, the static scratch "
                        "holds %zu (at most %ld cells)\n",
                L.total, c->sscr_n, (long)((c->sscr_n - L.cells) / (CR_NFIELD * sizeof(double))));
        return OCO_E_ARG;
    }
    HIPCHK(hipSetDevice(c->device));
    const SweepPlan p = sweep_plan(c, nz, ncol, c->nw, mode);
    if (!p.ok) return OCO_E_ARG;
    char* base = (char*)c->sscr;
    unsigned long long* fail = (unsigned long long*)(base + L.fail);
    double* canary = (double*)(base + L.canary);
    double* minmax = (double*)(base + L.minmax);
    double* cells = (double*)(base + L.cells);

    oco_sweep_args a = sweep_args(c, nz, ncol, dz, dz1, dz2, d_rho, d_T);
    a.U = d_U;
    a.V = d_V;
    a.A = d_A;
    a.cells = cells;
    a.scratch = nullptr;
    a.fail = fail;
    a.canary = canary;
    a.tickets = (unsigned long long*)(base + L.tickets);

    HIPCHK(hipEventRecord(c->ev[0], c->stream));
    HIPCHK(oco_launch_fail_reset(fail, 1, c->stream));   /* the last call's key stays otherwise */
    if (!c->lgrey) HIPCHK(oco_launch_cells(d_rho, d_T, ncells, c->fully_ionized_T, cells, nz, c->stream));
    HIPCHK(hipEventRecord(c->ev[1], c->stream));
    HIPCHK(oco_launch_sweep(a, c->lgrey, p.wave, p.lanes, p.pk, p.nblocks, c->stream));   /* resets the tickets */
    HIPCHK(hipEventRecord(c->ev[2], c->stream));
    HIPCHK(oco_launch_pillar_minmax(d_U, d_V, d_A, nz, minmax, c->stream));
    unsigned long long fkey = OCO_FAIL_NONE;
    if (c->checks & OCO_CHECK_FINITE) {
        unsigned long long* dk = (unsigned long long*)(base + L.fkey);
        HIPCHK(oco_launch_fail_reset(dk, 1, c->stream));
        HIPCHK(finite_scan(nz, nyloc, nxloc, 0, c->nw, c->nw, d_U, d_V, d_A, dk, c->stream));
        HIPCHK(hipMemcpyAsync(&fkey, dk, sizeof(fkey), hipMemcpyDeviceToHost, c->stream));
    }

    unsigned long long key = OCO_FAIL_NONE;
    double can[4] = {0, 0, 0, 0}, mm[6] = {0, 0, 0, 0, 0, 0};
    HIPCHK(hipMemcpyAsync(&key, fail, sizeof(key), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(can, canary, sizeof(can), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(mm, minmax, sizeof(mm), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));

    oco_diag d;
    memset(&d, 0, sizeof(d));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev[0], c->ev[1]) == hipSuccess) d.cells_ms = ms;
    if (hipEventElapsedTime(&ms, c->ev[1], c->ev[2]) == hipSuccess) d.sweep_ms = ms;
    if (hipEventElapsedTime(&ms, c->ev[0], c->ev[2]) == hipSuccess) d.total_ms = ms;
    d.sum_a = can[0];
    d.sum_b = can[1];
    d.sum_c = can[2];
    d.sum_d = can[3];
    if (key != OCO_FAIL_NONE) {
        decode_key(key, c->nw, nyloc, &d);
    } else {
        d.minU = mm[0]; d.maxU = mm[1];
        d.minV = mm[2]; d.maxV = mm[3];
        d.minA = mm[4]; d.maxA = mm[5];
        if (fkey != OCO_FAIL_NONE) decode_finite(fkey, nz, c->nw, nyloc, &d);
    }
    if (diag) *diag = d;
    return d.status;
}

int oco_solve_block(oco_ctx* c, int nz, int nyloc, int nxloc, double dz, double dz1, double dz2,
                    const double* rho3d, const double* T3d, double* U, double* V, double* absorp, int mode,
                    oco_diag* diag)
{
    if (!c || !rho3d || !T3d || !U || !V || !absorp) return OCO_E_ARG;
    if (nz < 2 || nyloc < 1 || nxloc < 1) return OCO_E_ARG;
    if (c->nw <= 0) return OCO_E_STATE;
    HIPCHK(hipSetDevice(c->device));
    const size_t ncells = (size_t)nz * nyloc * nxloc;
    const size_t nout = ncells * c->nw;
    const size_t nU = (size_t)(nz + 6) * nyloc * nxloc * c->nw;
    int st;
    if ((st = ensure(c, c->rho, ncells * sizeof(double)))) return st;
    if ((st = ensure(c, c->T, ncells * sizeof(double)))) return st;
    if ((st = ensure(c, c->U, nU * sizeof(double)))) return st;
    if ((st = ensure(c, c->V, nout * sizeof(double)))) return st;
    if ((st = ensure(c, c->A, nout * sizeof(double)))) return st;
    HIPCHK(hipMemcpyAsync(c->rho.p, rho3d, ncells * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->T.p, T3d, ncells * sizeof(double), hipMemcpyHostToDevice, c->stream));
    st = oco_solve_block_device(c, nz, nyloc, nxloc, dz, dz1, dz2, (const double*)c->rho.p,
                                (const double*)c->T.p, (double*)c->U.p, (double*)c->V.p, (double*)c->A.p,
                                mode, diag);
    if (st == OCO_E_ARG || st == OCO_E_HIP || st == OCO_E_OOM || st == OCO_E_STATE) return st;
    /* on a zero pivot the reference stops with its outputs unwritten; the
     * arrays are still returned so a caller can inspect the solved systems */
    HIPCHK(hipMemcpyAsync(U, c->U.p, nU * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(V, c->V.p, nout * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(absorp, c->A.p, nout * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return st;
}

/* ---------------------------------------------------------------------- */
/* Opacity query                                                           */
/* ---------------------------------------------------------------------- */

int oco_opacity_points(oco_ctx* c, int npts, const double* rho, const double* T, int mode, double* alpha,
                       double* albedo, double* S)
{
    if (!c || npts < 1 || !rho || !T || !alpha) return OCO_E_ARG;
    if (!mode_ok(mode)) return OCO_E_ARG;   /* OCO_MODE_ACCURATE: the FP64 opacity */
    if (c->nw <= 0) return OCO_E_STATE;
    for (int i = 0; i < npts; i++)
        if (!(rho[i] > 0.0) || !(T[i] > 0.0)) return OCO_E_ARG;
    HIPCHK(hipSetDevice(c->device));
    const size_t n = (size_t)npts * c->nw;
    int st;
    if ((st = ensure(c, c->rho, (size_t)npts * sizeof(double)))) return st;
    if ((st = ensure(c, c->T, (size_t)npts * sizeof(double)))) return st;
    if (!c->lgrey && (st = ensure(c, c->cells, (size_t)npts * OCO_CELL_STRIDE * sizeof(double)))) return st;
    if ((st = ensure(c, c->U, n * sizeof(double)))) return st;
    if (albedo && (st = ensure(c, c->V, n * sizeof(double)))) return st;
    if (S && (st = ensure(c, c->A, n * sizeof(double)))) return st;
    HIPCHK(hipMemcpyAsync(c->rho.p, rho, (size_t)npts * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->T.p, T, (size_t)npts * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (!c->lgrey)
        HIPCHK(oco_launch_cells((const double*)c->rho.p, (const double*)c->T.p, npts, c->fully_ionized_T,
                                (double*)c->cells.p, 0, c->stream));
    oco_points_args a;
    a.npts = npts;
    a.nw = c->nw;
    a.rho = (const double*)c->rho.p;
    a.T = (const double*)c->T.p;
    a.cells = (const double*)c->cells.p;
    a.wtab = (const double*)c->wtab.p;
    a.sigma_grey = c->sigma_grey;
    a.alpha = (double*)c->U.p;
    a.albedo = albedo ? (double*)c->V.p : nullptr;
    a.S = S ? (double*)c->A.p : nullptr;
    HIPCHK(oco_launch_points(a, c->lgrey, mode == OCO_MODE_MIXED, c->stream));
    HIPCHK(hipMemcpyAsync(alpha, a.alpha, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (albedo) HIPCHK(hipMemcpyAsync(albedo, a.albedo, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (S) HIPCHK(hipMemcpyAsync(S, a.S, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return OCO_OK;
}

/* ---------------------------------------------------------------------- */
/* Block outputs to files                                                  */
/* ---------------------------------------------------------------------- */

}  /* extern "C" */

static int ensure_copy_stream(oco_ctx* c)
{
    if (c->copy_stream) return OCO_OK;
    if (hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking) != hipSuccess) return OCO_E_HIP;
    for (int i = 0; i < 2; i++) {
        if (hipEventCreateWithFlags(&c->ev_sweep[i], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&c->ev_copy[i], hipEventDisableTiming) != hipSuccess)
            return OCO_E_HIP;
    }
    return OCO_OK;
}

static int ensure_pinned(oco_ctx* c, size_t bytes)
{
    if (c->pinned_n >= bytes) return OCO_OK;
    if (c->pinned) (void)hipHostFree(c->pinned);
    c->pinned = nullptr;
    c->pinned_n = 0;
    if (hipHostMalloc(&c->pinned, bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        c->pinned = nullptr;
        return OCO_E_OOM;
    }
    c->pinned_n = bytes;
    return OCO_OK;
}

namespace {

/* Hand-off between the launching thread and the file writer: chunk indices
 * whose copies are enqueued, in order; -1 ends.  `written` counts chunks the
 * writer is done with (written, or skipped after an error). */
struct ChunkQueue {
    std::mutex m;
    std::condition_variable cv;
    std::deque<int> q;
    int written = 0;
    int status = OCO_OK;

    void push(int k)
    {
        {
            std::lock_guard<std::mutex> lk(m);
            q.push_back(k);
        }
        cv.notify_all();
    }
};

}  // namespace

extern "C" {

int oco_solve_block_files(oco_ctx* c, int nz, int nyloc, int nxloc, double dz, double dz1, double dz2,
                          const double* rho3d, const double* T3d, int mode, const char* dir, const char* snapshot,
                          size_t stage_bytes, double* U_col, double* A_col, oco_diag* diag)
{
    const auto t_start = std::chrono::steady_clock::now();
    if (!c || !rho3d || !T3d || !dir || !snapshot) return OCO_E_ARG;
    if (nz < 2 || nyloc < 1 || nxloc < 1) return OCO_E_ARG;
    if (!mode_ok(mode)) return OCO_E_ARG;
    if (c->nw <= 0) return OCO_E_STATE;
    if (c->lgrey && c->nw != 1) return OCO_E_ARG;   /* source/flux_feautrier.f90:84-87 */
    HIPCHK(hipSetDevice(c->device));
    const int nw = c->nw, mz = nz + 6;
    const long ncol = (long)nxloc * nyloc;
    const size_t ncells = (size_t)ncol * nz;
    const size_t bU = (size_t)ncol * mz * sizeof(double), bV = (size_t)ncol * nz * sizeof(double);
    const size_t per_w = bU + 2 * bV;
    if (stage_bytes == 0) stage_bytes = (size_t)512 << 20;
    const int nwc = (int)std::min<size_t>((size_t)nw, std::max<size_t>(1, stage_bytes / per_w));
    const int nch = (nw + nwc - 1) / nwc;
    const int nlast = nw - (nch - 1) * nwc;
    const SweepPlan p_full = sweep_plan(c, nz, ncol, nwc, mode);
    const SweepPlan p_last = sweep_plan(c, nz, ncol, nlast, mode);
    if (!p_full.ok || !p_last.ok) return OCO_E_ARG;
    SweepPlan p_max = p_full;
    p_max.nblocks = std::max(p_full.nblocks, p_last.nblocks);
    const size_t slot_bytes = (size_t)nwc * per_w;
    int st;
    if ((st = sweep_buffers(c, p_max, nz, (long)ncells))) return st;
    const bool check = (c->checks & OCO_CHECK_FINITE) != 0;
    if (check && (st = ensure(c, c->fkey, sizeof(unsigned long long)))) return st;
    if ((st = ensure(c, c->rho, ncells * sizeof(double)))) return st;
    if ((st = ensure(c, c->T, ncells * sizeof(double)))) return st;
    if ((st = ensure(c, c->U, 2 * (size_t)nwc * bU))) return st;
    if ((st = ensure(c, c->V, 2 * (size_t)nwc * bV))) return st;
    if ((st = ensure(c, c->A, 2 * (size_t)nwc * bV))) return st;
    /* pinned: two chunk slots [U][absorp][V], then per-chunk pivot keys, the
     * canary sums, the "no pivot" value and the non-finite key */
    if ((st = ensure_pinned(c, 2 * slot_bytes + ((size_t)nch + 6) * sizeof(double)))) return st;
    if ((st = ensure_copy_stream(c))) return st;
    char* pin = (char*)c->pinned;
    unsigned long long* keys = (unsigned long long*)(pin + 2 * slot_bytes);
    double* can = (double*)(keys + nch);
    unsigned long long* none = (unsigned long long*)(can + 4);
    unsigned long long* fkey = none + 1;
    *none = OCO_FAIL_NONE;
    *fkey = OCO_FAIL_NONE;

    oco_record_file* F[3] = {nullptr, nullptr, nullptr};
    auto close_all = [&](int keep) {
        int r = OCO_OK;
        for (auto& f : F) {
            const int q = oco_record_close(f, keep);
            if (r == OCO_OK) r = q;
            f = nullptr;
        }
        return r;
    };
    {
        const char* names[3] = {"mean_intensity_", "absorption_coefficients_", "flux_"};
        const unsigned long long sizes[3] = {nw * (unsigned long long)bU, nw * (unsigned long long)bV,
                                             nw * (unsigned long long)bV};
        for (int i = 0; i < 3; i++) {
            const std::string path = std::string(dir) + "/" + names[i] + snapshot + ".bin";
            if ((st = oco_record_open(path.c_str(), sizes[i], &F[i]))) {
                close_all(0);
                return st;
            }
        }
    }

    hipError_t he = hipMemcpyAsync(c->rho.p, rho3d, ncells * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (he == hipSuccess) he = hipMemcpyAsync(c->T.p, T3d, ncells * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (he == hipSuccess) he = hipEventRecord(c->ev[0], c->stream);
    if (he == hipSuccess && !c->lgrey)
        he = oco_launch_cells((const double*)c->rho.p, (const double*)c->T.p, (long)ncells, c->fully_ionized_T,
                              (double*)c->cells.p, p_full.wave ? nz : 0, c->stream);
    if (he == hipSuccess) he = hipEventRecord(c->ev[1], c->stream);
    if (he == hipSuccess && check) he = oco_launch_fail_reset((unsigned long long*)c->fkey.p, 1, c->stream);
    if (he != hipSuccess) {
        fprintf(stderr, "ocotillo: block outputs: %s\n", hipGetErrorString(he));
        close_all(0);
        return OCO_E_HIP;
    }

    oco_diag d;
    memset(&d, 0, sizeof(d));
    ChunkQueue w;
    /* writer: waits for each chunk's copy, writes its byte ranges of the three
     * records, picks out the printed column and the first pillar */
    std::thread writer([&]() {
        (void)hipSetDevice(c->device);
        for (;;) {
            int k;
            {
                std::unique_lock<std::mutex> lk(w.m);
                w.cv.wait(lk, [&] { return !w.q.empty(); });
                k = w.q.front();
                w.q.pop_front();
            }
            if (k < 0) break;
            const int slot = k & 1, iw0 = k * nwc, n = std::min(nwc, nw - iw0);
            int r = hipEventSynchronize(c->ev_copy[slot]) == hipSuccess ? OCO_OK : OCO_E_HIP;
            bool skip;
            {
                std::lock_guard<std::mutex> lk(w.m);
                skip = w.status != OCO_OK;
            }
            const char* sb = pin + slot * slot_bytes;
            const double* sU = (const double*)sb;
            const double* sA = (const double*)(sb + (size_t)nwc * bU);
            const double* sV = (const double*)(sb + (size_t)nwc * (bU + bV));
            if (r == OCO_OK && !skip) {
                r = oco_record_write(F[0], (unsigned long long)iw0 * bU, sU, (size_t)n * bU);
                if (r == OCO_OK) r = oco_record_write(F[1], (unsigned long long)iw0 * bV, sA, (size_t)n * bV);
                if (r == OCO_OK) r = oco_record_write(F[2], (unsigned long long)iw0 * bV, sV, (size_t)n * bV);
            }
            if (r == OCO_OK && !skip) {
                for (int j = 0; j < n; j++) {
                    if (U_col)
                        memcpy(U_col + (size_t)(iw0 + j) * mz, sU + ((size_t)j * ncol + ncol - 1) * mz,
                               (size_t)mz * sizeof(double));
                    if (A_col)
                        memcpy(A_col + (size_t)(iw0 + j) * nz, sA + ((size_t)j * ncol + ncol - 1) * nz,
                               (size_t)nz * sizeof(double));
                }
                if (k == 0) pillar_minmax(sU, sV, sA, nz, &d);
            }
            {
                std::lock_guard<std::mutex> lk(w.m);
                if (r != OCO_OK && w.status == OCO_OK) w.status = r;
                w.written = k + 1;
            }
            w.cv.notify_all();
        }
    });

    for (int k = 0; k < nch; k++) {
        const int slot = k & 1, iw0 = k * nwc, n = std::min(nwc, nw - iw0);
        if (k >= 2) {   /* slot reuse: the writer is done with chunk k-2 */
            std::unique_lock<std::mutex> lk(w.m);
            w.cv.wait(lk, [&] { return w.written >= k - 1; });
            if (w.status != OCO_OK) break;
        }
        oco_sweep_args a = sweep_args(c, nz, ncol, dz, dz1, dz2, (const double*)c->rho.p, (const double*)c->T.p);
        a.nw = n;
        a.wtab = (const double*)c->wtab.p + (size_t)iw0 * OCO_WT_STRIDE;
        a.U = (double*)c->U.p + (size_t)slot * nwc * ncol * mz;
        a.A = (double*)c->A.p + (size_t)slot * nwc * ncol * nz;
        a.V = (double*)c->V.p + (size_t)slot * nwc * ncol * nz;
        a.canary = k == 0 ? (double*)c->canary.p : nullptr;
        const SweepPlan& pk = n == nwc ? p_full : p_last;
        char* sb = pin + slot * slot_bytes;
        he = hipMemcpyAsync(c->fail.p, none, sizeof(*none), hipMemcpyHostToDevice, c->stream);
        if (he == hipSuccess) he = oco_launch_sweep(a, c->lgrey, pk.wave, pk.lanes, pk.pk, pk.nblocks, c->stream);
        if (he == hipSuccess && check)
            he = finite_scan(nz, nyloc, nxloc, iw0, n, nw, a.U, a.V, a.A, (unsigned long long*)c->fkey.p, c->stream);
        if (he == hipSuccess && check && k == nch - 1)
            he = hipMemcpyAsync(fkey, c->fkey.p, sizeof(*fkey), hipMemcpyDeviceToHost, c->stream);
        if (he == hipSuccess)
            he = hipMemcpyAsync(keys + k, c->fail.p, sizeof(*keys), hipMemcpyDeviceToHost, c->stream);
        if (he == hipSuccess && k == 0)
            he = hipMemcpyAsync(can, c->canary.p, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (he == hipSuccess) he = hipEventRecord(c->ev_sweep[slot], c->stream);
        if (he == hipSuccess && k == nch - 1) he = hipEventRecord(c->ev[2], c->stream);
        if (he == hipSuccess) he = hipStreamWaitEvent(c->copy_stream, c->ev_sweep[slot], 0);
        if (he == hipSuccess) he = hipMemcpyAsync(sb, a.U, (size_t)n * bU, hipMemcpyDeviceToHost, c->copy_stream);
        if (he == hipSuccess)
            he = hipMemcpyAsync(sb + (size_t)nwc * bU, a.A, (size_t)n * bV, hipMemcpyDeviceToHost, c->copy_stream);
        if (he == hipSuccess)
            he = hipMemcpyAsync(sb + (size_t)nwc * (bU + bV), a.V, (size_t)n * bV, hipMemcpyDeviceToHost,
                                c->copy_stream);
        if (he == hipSuccess) he = hipEventRecord(c->ev_copy[slot], c->copy_stream);
        if (he != hipSuccess) {
            fprintf(stderr, "ocotillo: block outputs: %s\n", hipGetErrorString(he));
            st = OCO_E_HIP;
            break;
        }
        w.push(k);
    }
    w.push(-1);
    writer.join();
    if (hipStreamSynchronize(c->stream) != hipSuccess && st == OCO_OK) st = OCO_E_HIP;
    if (hipStreamSynchronize(c->copy_stream) != hipSuccess && st == OCO_OK) st = OCO_E_HIP;
    if (st == OCO_OK) st = w.status;

    unsigned long long best = OCO_FAIL_NONE;
    if (st == OCO_OK) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->ev[0], c->ev[1]) == hipSuccess) d.cells_ms = ms;
        if (hipEventElapsedTime(&ms, c->ev[1], c->ev[2]) == hipSuccess) d.sweep_ms = ms;
        d.sum_a = can[0];
        d.sum_b = can[1];
        d.sum_c = can[2];
        d.sum_d = can[3];
        for (int k = 0; k < nch; k++)
            best = std::min(best, global_key(keys[k], k * nwc, std::min(nwc, nw - k * nwc), nw));
        if (best != OCO_FAIL_NONE) {
            decode_key(best, nw, nyloc, &d);
            d.minU = d.maxU = d.minV = d.maxV = d.minA = d.maxA = 0.0;
        } else if (*fkey != OCO_FAIL_NONE) {
            /* the reference's trapping build stops before output_binary */
            decode_finite(*fkey, nz, nw, nyloc, &d);
            best = *fkey;
        }
    }
    const int keep = st == OCO_OK && best == OCO_FAIL_NONE;
    const int cst = close_all(keep);
    if (st == OCO_OK && keep) st = cst;
    if (st == OCO_OK && !keep) st = d.status;
    d.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    if (diag) *diag = d;
    return st;
}

/* ---------------------------------------------------------------------- */
/* Athena input construction                                              */
/* ---------------------------------------------------------------------- */

static int athena_check(int nz, int nyloc, int nxloc, int nzl, int iz0, const oco_athena_units* u)
{
    if (nz < 1 || nyloc < 1 || nxloc < 1 || nzl < 1 || iz0 < 0 || iz0 + nzl > nz || !u) return OCO_E_ARG;
    return OCO_OK;
}

static oco_athena_args athena_args(int nz, int nyloc, int nxloc, int nzl, int iz0, const oco_athena_units* u)
{
    oco_athena_args a;
    a.nz = nz; a.nyloc = nyloc; a.nxloc = nxloc; a.nzl = nzl; a.iz0 = iz0;
    a.unit_density = u->unit_density;
    a.unit_temperature = u->unit_temperature;
    a.gamma1 = u->gamma1;
    a.rho_floor = u->rho_floor;
    a.switch_temp = u->switch_temp;
    a.floor_temp = u->floor_temp;
    a.avoid_momentum_underflow = u->avoid_momentum_underflow;
    return a;
}

int oco_athena_slab_device(oco_ctx* c, int nz, int nyloc, int nxloc, int nzl, int iz0, const double* rho_raw,
                           const double* mx, const double* my, const double* mz, const double* e,
                           const double* tgauss, const oco_athena_units* u, double* rho3d, double* T3d)
{
    if (!c || !rho_raw || !mx || !my || !mz || !e || !tgauss || !rho3d || !T3d) return OCO_E_ARG;
    int st = athena_check(nz, nyloc, nxloc, nzl, iz0, u);
    if (st) return st;
    HIPCHK(hipSetDevice(c->device));
    oco_athena_args a = athena_args(nz, nyloc, nxloc, nzl, iz0, u);
    a.rho = rho_raw; a.mx = mx; a.my = my; a.mz = mz; a.e = e;
    a.tgauss = tgauss;
    a.rho3d = rho3d;
    a.T3d = T3d;
    HIPCHK(oco_launch_athena(a, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return OCO_OK;
}

int oco_athena_slab(oco_ctx* c, int nz, int nyloc, int nxloc, int nzl, int iz0, const double* rho_raw,
                    const double* mx, const double* my, const double* mz, const double* e, const double* tgauss,
                    const oco_athena_units* u, double* rho3d, double* T3d)
{
    if (!c || !rho_raw || !mx || !my || !mz || !e || !tgauss || !rho3d || !T3d) return OCO_E_ARG;
    int st = athena_check(nz, nyloc, nxloc, nzl, iz0, u);
    if (st) return st;
    HIPCHK(hipSetDevice(c->device));
    const size_t ns = (size_t)nxloc * nyloc * nzl, ncell = (size_t)nxloc * nyloc * nz;
    if ((st = ensure(c, c->raw, 5 * ns * sizeof(double)))) return st;
    if ((st = ensure(c, c->tg, (size_t)nz * sizeof(double)))) return st;
    if ((st = ensure(c, c->rho, ncell * sizeof(double)))) return st;
    if ((st = ensure(c, c->T, ncell * sizeof(double)))) return st;
    double* d = (double*)c->raw.p;
    const double* src[5] = {rho_raw, mx, my, mz, e};
    for (int f = 0; f < 5; f++)
        HIPCHK(hipMemcpyAsync(d + f * ns, src[f], ns * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->tg.p, tgauss, (size_t)nz * sizeof(double), hipMemcpyHostToDevice, c->stream));
    oco_athena_args a = athena_args(nz, nyloc, nxloc, nzl, iz0, u);
    a.rho = d; a.mx = d + ns; a.my = d + 2 * ns; a.mz = d + 3 * ns; a.e = d + 4 * ns;
    a.tgauss = (const double*)c->tg.p;
    a.rho3d = (double*)c->rho.p;
    a.T3d = (double*)c->T.p;
    HIPCHK(oco_launch_athena(a, c->stream));
    /* the slab's z range of every column back into the caller's arrays */
    const size_t pitch = (size_t)nz * sizeof(double), width = (size_t)nzl * sizeof(double);
    HIPCHK(hipMemcpy2DAsync(rho3d + iz0, pitch, (const double*)c->rho.p + iz0, pitch, width,
                            (size_t)nxloc * nyloc, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpy2DAsync(T3d + iz0, pitch, (const double*)c->T.p + iz0, pitch, width, (size_t)nxloc * nyloc,
                            hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return OCO_OK;
}

/* ---------------------------------------------------------------------- */
/* Batched uneven columns                                                  */
/* ---------------------------------------------------------------------- */

/* column classes of the batched difference-form kernel (batch_class); each
 * class's workgroup size (one of BATCH_WAVE_K's three) is tuned on its own */
#define OCO_BATCH_CLASSES 6

struct oco_batch {
    oco_ctx* ctx = nullptr;
    int ncol = 0, max_nz = 0, ntiles = 0;
    long ncells = 0, nsys = 0, nU = 0, nVA = 0;   /* nU, nVA: the packed (host) layout's totals */
    long dU = 0, dVA = 0;                         /* the device buffers' totals (padded rows) */
    std::vector<int> nz, nw;
    std::vector<long> woff, uoff, voff;           /* device offsets of each column's U and V, absorp */
    DevBuf ints, longs, dz, wtab, rho, T, cells, U, V, A, scratch, counter, fail;
    /* mixed mode, wavefront per system (when every column fits it): item
     * lists for workgroups of 1, 2 and 4 wavefronts, their queue tickets and
     * the field-major records */
    DevBuf witems, wtickets, cells_soa;
    bool wave_ok = false;
    /* OCO_BATCH_CLASSES classes of columns for that kernel (batch_class),
     * each launched for its own longest column, and each class's tuned index
     * into BATCH_WAVE_K */
    int cls_max_nz[OCO_BATCH_CLASSES] = {0};
    long wio[OCO_BATCH_CLASSES][3] = {{0}}, wqo[OCO_BATCH_CLASSES][3] = {{0}};
    int cls_k[OCO_BATCH_CLASSES] = {0};
    /* per precision mode: autotuned kernel variant and grid, captured graph
     * and the context's fully_ionized_T it was captured with (FP64 variants
     * 0..2: lane per system with 2, 4, 8 back-substitution rows per chunk;
     * mixed: the classes' indices into BATCH_WAVE_K as base-3 digits, class 0
     * the most significant) */
    int nblocks[2] = {0, 0}, variant[2] = {1, 0}, tuned[2] = {0, 0};
    float tune_ms[2][32] = {{0}};
    int tune_n[2] = {0, 0};
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    double graph_fiT[2] = {0.0, 0.0};
};

static void batch_release(oco_batch* b)
{
    oco_ctx* c = b->ctx;
    for (DevBuf* d : {&b->ints, &b->longs, &b->dz, &b->wtab, &b->rho, &b->T, &b->cells, &b->U, &b->V, &b->A,
                      &b->scratch, &b->counter, &b->fail, &b->witems, &b->wtickets, &b->cells_soa})
        release(c, *d);
    for (int m = 0; m < 2; m++) {
        if (b->exec[m]) (void)hipGraphExecDestroy(b->exec[m]);
        if (b->graph[m]) (void)hipGraphDestroy(b->graph[m]);
        b->exec[m] = nullptr;
        b->graph[m] = nullptr;
    }
}

static oco_batch_args batch_args(oco_batch* b)
{
    oco_batch_args a;
    const int ncol = b->ncol;
    a.ntiles = b->ntiles;
    a.nz = (const int*)b->ints.p;
    a.nw = a.nz + ncol;
    a.tiles = a.nw + ncol;
    a.cell_off = (const long*)b->longs.p;
    a.wt_off = a.cell_off + ncol;
    a.u_off = a.wt_off + ncol;
    a.v_off = a.u_off + ncol;
    a.dz = (const double*)b->dz.p;
    a.cells = (const double*)b->cells.p;
    a.rho = (const double*)b->rho.p;
    a.T = (const double*)b->T.p;
    a.wtab = (const double*)b->wtab.p;
    a.U = (double*)b->U.p;
    a.V = (double*)b->V.p;
    a.A = (double*)b->A.p;
    a.scratch = (double*)b->scratch.p;
    a.scr_stride = (long)b->max_nz * 3 * 64;
    a.counter = (unsigned int*)b->counter.p;
    a.fail = (unsigned long long*)b->fail.p;
    a.witems = nullptr;
    a.wq_off = nullptr;
    a.wtickets = (unsigned long long*)b->wtickets.p;
    a.cells_soa = (const double*)b->cells_soa.p;
    a.max_nz = b->max_nz;
    return a;
}

static const int BATCH_WAVE_K[3] = {1, 2, 4};
/* The column classes of the batched difference-form kernel, each launched
 * with LDS for its longest column and its own lanes per system: up to 72,
 * 144 and 288 cells on 8, 16 and 32 lanes (at most 9 rows per lane, as the
 * block sweep chooses), then 64 lanes up to 512, 768 and 1088 cells.  The LDS
 * of a class's longest column sets how many wavefronts a CU holds, so the
 * 64-lane columns are split three ways (at 512 cells, 12 wavefronts, 3 per
 * SIMD, where the sweep stops gaining; without the 768 class, columns of
 * 513-768 cells would run at 6). */
static const int BATCH_CLASS_NZ[OCO_BATCH_CLASSES] = {72, 144, 288, 512, 768, 1088};
static const int BATCH_CLASS_LANES[OCO_BATCH_CLASSES] = {8, 16, 32, 64, 64, 64};
static int batch_class(int nz)
{
    int cl = 0;
    while (cl < OCO_BATCH_CLASSES - 1 && nz > BATCH_CLASS_NZ[cl]) cl++;
    return cl;
}

/* enqueue the sweep of class cl with workgroups of BATCH_WAVE_K[i] */
static hipError_t batch_class_enqueue(oco_batch* b, int cl, int i)
{
    oco_ctx* c = b->ctx;
    oco_batch_args a = batch_args(b);
    const int K = BATCH_WAVE_K[i], L = BATCH_CLASS_LANES[cl];
    a.max_nz = b->cls_max_nz[cl];
    a.witems = (const int*)b->witems.p + b->wio[cl][i];
    a.wq_off = (const int*)b->witems.p + b->wqo[cl][i];
    a.wtickets = (unsigned long long*)b->wtickets.p + cl * OCO_TICKET_QUEUES * OCO_TICKET_STRIDE;
    return oco_launch_batch_wave(a, K, L, oco_batch_wave_grid(a.max_nz, K, L, c->device), c->stream);
}

/* enqueue one full batch solve with the given kernel variant / grid (mixed:
 * each class with its tuned workgroup size) */
static hipError_t batch_enqueue(oco_batch* b, int mixed, int variant, int nblocks)
{
    oco_ctx* c = b->ctx;
    if (mixed) {
        (void)variant;
        (void)nblocks;   /* each class takes its own resident grid */
        hipError_t e = oco_launch_batch_wave_records(batch_args(b), b->ncol, c->fully_ionized_T, c->stream);
        for (int cl = 0; cl < OCO_BATCH_CLASSES && e == hipSuccess; cl++)
            if (b->cls_max_nz[cl] > 0) e = batch_class_enqueue(b, cl, b->cls_k[cl]);
        return e;
    }
    hipError_t e = oco_launch_cells((const double*)b->rho.p, (const double*)b->T.p, b->ncells, c->fully_ionized_T,
                                    (double*)b->cells.p, 0, c->stream);
    if (e != hipSuccess) return e;
    return oco_launch_batch(batch_args(b), variant, nblocks, c->stream);
}

int oco_batch_create(oco_ctx* c, int ncol, const int* nz, const int* nw, const double* waves, const double* dz,
                     oco_batch** out)
{
    if (!c || !out || ncol <= 0 || !nz || !nw || !waves || !dz) return OCO_E_ARG;
    if (c->lgrey) return OCO_E_ARG;
    *out = nullptr;
    oco_batch* b = new (std::nothrow) oco_batch;
    if (!b) return OCO_E_OOM;
    b->ctx = c;
    b->ncol = ncol;
    b->nz.assign(nz, nz + ncol);
    b->nw.assign(nw, nw + ncol);
    std::vector<long> coff(ncol), woff(ncol), uoff(ncol), voff(ncol);
    std::vector<double> dz3(3 * (size_t)ncol);
    long nc = 0, ns = 0, nU = 0, nVA = 0, pU = 0, pVA = 0;
    for (int i = 0; i < ncol; i++) {
        if (nz[i] < 4 || nw[i] < 1 || !(dz[i] > 0.0)) {
            delete b;
            return OCO_E_ARG;
        }
        coff[i] = nc;
        woff[i] = ns;
        uoff[i] = nU;
        voff[i] = nVA;
        nc += nz[i];
        ns += nw[i];
        nU += (long)oco_batch_ldu(nz[i]) * nw[i];
        nVA += (long)oco_batch_ldv(nz[i]) * nw[i];
        pU += (long)(nz[i] + 6) * nw[i];
        pVA += (long)nz[i] * nw[i];
        b->max_nz = std::max(b->max_nz, nz[i]);
        dz3[3 * i] = dz[i];
        dz3[3 * i + 1] = 1.0 / dz[i];
        dz3[3 * i + 2] = dz[i] * dz[i];
    }
    for (long i = 0; i < ns; i++)
        if (!(waves[i] > 0.0) || !std::isfinite(waves[i])) {
            delete b;
            return OCO_E_ARG;
        }
    b->ncells = nc;
    b->nsys = ns;
    b->nU = pU;
    b->nVA = pVA;
    b->dU = nU;
    b->dVA = nVA;
    b->woff = woff;
    b->uoff = uoff;
    b->voff = voff;
    /* tiles of 64 wavelengths, longest columns first (LPT order) */
    std::vector<int> order(ncol);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return nz[x] > nz[y]; });
    std::vector<int> tiles;
    for (int i : order)
        for (int w0 = 0; w0 < nw[i]; w0 += 64) {
            tiles.push_back(i);
            tiles.push_back(w0);
        }
    b->ntiles = (int)(tiles.size() / 2);
    /* wavelength tables, one grid per column */
    std::vector<double> wt((size_t)ns * OCO_WT_STRIDE), one;
    for (int i = 0; i < ncol; i++) {
        std::vector<double> g(waves + woff[i], waves + woff[i] + nw[i]);
        oco_build_wtab(g, one);
        std::copy(one.begin(), one.end(), wt.begin() + woff[i] * OCO_WT_STRIDE);
    }
    /* mixed mode, difference form: for workgroups of K wavefronts, items
     * (column, first of the S = K 64 / L wavelengths of a workgroup); column
     * p of the longest-first order goes to queue p mod 8, so each XCD's L2
     * holds its own columns' records */
    std::vector<int> witems;
    b->wave_ok = b->max_nz <= BATCH_CLASS_NZ[OCO_BATCH_CLASSES - 1] &&
                 oco_batch_wave_grid(b->max_nz, 1, BATCH_CLASS_LANES[batch_class(b->max_nz)], c->device) > 0;
    if (b->wave_ok) {
        for (int i : order) {
            const int cl = batch_class(nz[i]);
            b->cls_max_nz[cl] = std::max(b->cls_max_nz[cl], nz[i]);
        }
        for (int cl = 0; cl < OCO_BATCH_CLASSES; cl++) {
            for (int i = 0; i < 3; i++) {
                const int K = BATCH_WAVE_K[i] * (64 / BATCH_CLASS_LANES[cl]);   /* wavelengths per item */
                std::vector<int> q[OCO_TICKET_QUEUES];
                int p = 0;   /* position among the class's columns, longest first */
                for (int col : order) {
                    if (batch_class(nz[col]) != cl) continue;
                    for (int w0 = 0; w0 < nw[col]; w0 += K) {
                        q[p % OCO_TICKET_QUEUES].push_back(col);
                        q[p % OCO_TICKET_QUEUES].push_back(w0);
                    }
                    p++;
                }
                b->wio[cl][i] = (long)witems.size();
                int n = 0;
                std::vector<int> qoff(OCO_TICKET_QUEUES + 1);
                for (int k = 0; k < OCO_TICKET_QUEUES; k++) {
                    qoff[k] = n;
                    n += (int)(q[k].size() / 2);
                    witems.insert(witems.end(), q[k].begin(), q[k].end());
                }
                qoff[OCO_TICKET_QUEUES] = n;
                b->wqo[cl][i] = (long)witems.size();
                witems.insert(witems.end(), qoff.begin(), qoff.end());
            }
        }
    }
    int dev_cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess && prop.multiProcessorCount > 0)
        dev_cus = prop.multiProcessorCount;
    b->nblocks[0] = b->nblocks[1] = std::min(b->ntiles, dev_cus * 4 * 3);
    const int max_blocks = dev_cus * 4 * 4;
    int st = OCO_OK;
    std::vector<int> ints;
    ints.insert(ints.end(), nz, nz + ncol);
    ints.insert(ints.end(), nw, nw + ncol);
    ints.insert(ints.end(), tiles.begin(), tiles.end());
    std::vector<long> longs;
    for (auto* v : {&coff, &woff, &uoff, &voff}) longs.insert(longs.end(), v->begin(), v->end());
    if (!st) st = ensure(c, b->ints, ints.size() * sizeof(int));
    if (!st) st = ensure(c, b->longs, longs.size() * sizeof(long));
    if (!st) st = ensure(c, b->dz, dz3.size() * sizeof(double));
    if (!st) st = ensure(c, b->wtab, wt.size() * sizeof(double));
    if (!st) st = ensure(c, b->rho, nc * sizeof(double));
    if (!st) st = ensure(c, b->T, nc * sizeof(double));
    if (!st) st = ensure(c, b->cells, (size_t)nc * OCO_CELL_STRIDE * sizeof(double));
    if (!st) st = ensure(c, b->U, nU * sizeof(double));
    if (!st) st = ensure(c, b->V, nVA * sizeof(double));
    if (!st) st = ensure(c, b->A, nVA * sizeof(double));
    /* a scratch slot per wavefront: the FP64 kernel's grid rounds the tuned
     * count up to whole workgroups of OCO_BATCH_WG_MAX at most */
    if (!st) st = ensure(c, b->scratch, (size_t)(std::min(b->ntiles, max_blocks) + OCO_BATCH_WG_MAX) * b->max_nz * 3 * 64 *
                                            sizeof(double));
    if (!st) st = ensure(c, b->counter, sizeof(unsigned int));
    if (!st) st = ensure(c, b->fail, sizeof(unsigned long long));
    if (!st && b->wave_ok) {
        st = ensure(c, b->witems, witems.size() * sizeof(int));
        if (!st) st = ensure(c, b->wtickets, OCO_BATCH_CLASSES * OCO_TICKET_BYTES);
        if (!st) st = ensure(c, b->cells_soa, (size_t)nc * CR_NFIELD * sizeof(double));
    }
    if (!st) {
        hipError_t e = hipMemcpyAsync(b->ints.p, ints.data(), ints.size() * sizeof(int), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess && b->wave_ok)
            e = hipMemcpyAsync(b->witems.p, witems.data(), witems.size() * sizeof(int), hipMemcpyHostToDevice,
                               c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(b->longs.p, longs.data(), longs.size() * sizeof(long), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(b->dz.p, dz3.data(), dz3.size() * sizeof(double), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(b->wtab.p, wt.data(), wt.size() * sizeof(double), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        st = hip_status(e);
    }
    if (st) {
        batch_release(b);
        delete b;
        return st;
    }
    *out = b;
    return OCO_OK;
}

int oco_batch_destroy(oco_batch* b)
{
    if (!b) return OCO_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    batch_release(b);
    delete b;
    return OCO_OK;
}

int oco_batch_layout(oco_batch* b, size_t* ncells, size_t* nsys, size_t* nU, size_t* nVA)
{
    if (!b) return OCO_E_ARG;
    if (ncells) *ncells = (size_t)b->ncells;
    if (nsys) *nsys = (size_t)b->nsys;
    if (nU) *nU = (size_t)b->nU;
    if (nVA) *nVA = (size_t)b->nVA;
    return OCO_OK;
}

int oco_batch_column(oco_batch* b, int col, size_t* u_off, int* ldu, size_t* va_off, int* ldva)
{
    if (!b || col < 0 || col >= b->ncol) return OCO_E_ARG;
    if (u_off) *u_off = (size_t)b->uoff[col];
    if (ldu) *ldu = oco_batch_ldu(b->nz[col]);
    if (va_off) *va_off = (size_t)b->voff[col];
    if (ldva) *ldva = oco_batch_ldv(b->nz[col]);
    return OCO_OK;
}

int oco_batch_buffers(oco_batch* b, double** rho, double** T, double** U, double** V, double** absorp)
{
    if (!b) return OCO_E_ARG;
    if (rho) *rho = (double*)b->rho.p;
    if (T) *T = (double*)b->T.p;
    if (U) *U = (double*)b->U.p;
    if (V) *V = (double*)b->V.p;
    if (absorp) *absorp = (double*)b->A.p;
    return OCO_OK;
}

int oco_batch_tune_times(oco_batch* b, int mode, float* ms, int cap)
{
    if (!b || (mode != OCO_MODE_FP64 && mode != OCO_MODE_MIXED) || cap < 0 || (cap > 0 && !ms)) return -OCO_E_ARG;
    const int n = b->tune_n[mode];
    for (int i = 0; i < n && i < cap; i++) ms[i] = b->tune_ms[mode][i];
    return n;
}

int oco_batch_tuning(oco_batch* b, int mode, int* variant, int* nblocks, int* tuned)
{
    if (!b || (mode != OCO_MODE_FP64 && mode != OCO_MODE_MIXED)) return OCO_E_ARG;
    if (variant) *variant = b->variant[mode];
    if (nblocks) *nblocks = b->nblocks[mode];
    if (tuned) *tuned = b->tuned[mode];
    return OCO_OK;
}

/* First run in a mode: time every kernel variant x grid size three times
 * and keep the one with the best median; then capture the whole solve
 * (reset, cells, sweep) as one hipGraph.  The kernels take fully_ionized_T
 * by value, so a graph captured under another value of the context's is
 * captured again. */
static int batch_prepare(oco_batch* b, int m)
{
    oco_ctx* c = b->ctx;
    if (m == OCO_MODE_MIXED && !b->wave_ok) return OCO_E_ARG;
    if (!b->tuned[m]) {
        int dev_cus = 256;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, c->device) == hipSuccess && prop.multiProcessorCount > 0)
            dev_cus = prop.multiProcessorCount;
        int k = 0;
        /* the median of three timings of a launch sequence */
        auto time3 = [&](auto&& enqueue, float& ms) -> int {
            float t3[3];
            for (int r = 0; r < 3; r++) {
                HIPCHK(hipEventRecord(c->ev[0], c->stream));
                HIPCHK(enqueue());
                HIPCHK(hipEventRecord(c->ev[1], c->stream));
                HIPCHK(hipEventSynchronize(c->ev[1]));
                t3[r] = 0.f;
                HIPCHK(hipEventElapsedTime(&t3[r], c->ev[0], c->ev[1]));
            }
            std::sort(t3, t3 + 3);
            ms = t3[1];
            if (k < 32) b->tune_ms[m][k++] = ms;
            b->tune_n[m] = k;
            return OCO_OK;
        };
        if (m == OCO_MODE_MIXED) {
            /* each class on its own: its launch with each workgroup size
             * (the records once, before) */
            HIPCHK(oco_launch_batch_wave_records(batch_args(b), b->ncol, c->fully_ionized_T, c->stream));
            int code = 0;
            for (int cl = 0; cl < OCO_BATCH_CLASSES; cl++) {
                float best = 1e30f;
                b->cls_k[cl] = 0;
                if (b->cls_max_nz[cl] > 0) {
                    for (int i = 0; i < 3; i++) {
                        const int L = BATCH_CLASS_LANES[cl];
                        if (oco_batch_wave_grid(b->cls_max_nz[cl], BATCH_WAVE_K[i], L, c->device) <= 0) continue;
                        float ms = 0.f;
                        int st = time3([&] { return batch_class_enqueue(b, cl, i); }, ms);
                        if (st) return st;
                        if (ms < best) {
                            best = ms;
                            b->cls_k[cl] = i;
                        }
                    }
                }
                code = 3 * code + b->cls_k[cl];
            }
            b->variant[m] = code;
            b->nblocks[m] = 0;
        } else {
            float best = 1e30f;
            for (int v = 0; v < oco_batch_variants(); v++) {
                for (int wps = 2; wps <= 4; wps++) {
                    const int nb = std::min(b->ntiles, dev_cus * 4 * wps);
                    float ms = 0.f;
                    int st = time3([&] { return batch_enqueue(b, m, v, nb); }, ms);
                    if (st) return st;
                    if (ms < best) {
                        best = ms;
                        b->variant[m] = v;
                        b->nblocks[m] = nb;
                    }
                }
            }
        }
        b->tuned[m] = 1;
    }
    if (b->exec[m] && b->graph_fiT[m] != c->fully_ionized_T) {
        (void)hipGraphExecDestroy(b->exec[m]);
        (void)hipGraphDestroy(b->graph[m]);
        b->exec[m] = nullptr;
        b->graph[m] = nullptr;
    }
    if (!b->exec[m]) {
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = batch_enqueue(b, m, b->variant[m], b->nblocks[m]);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(c->stream, &g);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return OCO_E_HIP;
        }
        b->graph[m] = g;
        HIPCHK(hipGraphInstantiate(&b->exec[m], g, nullptr, nullptr, 0));
        b->graph_fiT[m] = c->fully_ionized_T;
    }
    return OCO_OK;
}

int oco_batch_run(oco_batch* b, int mode, oco_diag* diag)
{
    if (!b) return OCO_E_ARG;
    if (mode != OCO_MODE_FP64 && mode != OCO_MODE_MIXED) return OCO_E_ARG;
    oco_ctx* c = b->ctx;
    HIPCHK(hipSetDevice(c->device));
    int st = batch_prepare(b, mode);
    if (st) return st;
    HIPCHK(hipEventRecord(c->ev[0], c->stream));
    HIPCHK(hipGraphLaunch(b->exec[mode], c->stream));
    HIPCHK(hipEventRecord(c->ev[2], c->stream));
    unsigned long long key = OCO_FAIL_NONE;
    HIPCHK(hipMemcpyAsync(&key, b->fail.p, sizeof(key), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    oco_diag d;
    memset(&d, 0, sizeof(d));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev[0], c->ev[2]) == hipSuccess) d.total_ms = d.sweep_ms = ms;
    if (key != OCO_FAIL_NONE) {
        long sys = (long)(key >> 32);
        d.status = (int)((key >> 28) & 0xf);
        d.fail_j = (int)(key & 0x0fffffffULL);
        int col = (int)(std::upper_bound(b->woff.begin(), b->woff.end(), sys) - b->woff.begin()) - 1;
        d.fail_ix = col + 1;
        d.fail_iy = 1;
        d.fail_iw = (int)(sys - b->woff[col]) + 1;
    }
    if (diag) *diag = d;
    return d.status;
}

int oco_batch_run_host(oco_batch* b, const double* rho, const double* T, double* U, double* V, double* absorp,
                       int mode, oco_diag* diag)
{
    if (!b || !rho || !T || !U || !V || !absorp) return OCO_E_ARG;
    oco_ctx* c = b->ctx;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(b->rho.p, rho, b->ncells * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(b->T.p, T, b->ncells * sizeof(double), hipMemcpyHostToDevice, c->stream));
    int st = oco_batch_run(b, mode, diag);
    if (st == OCO_E_ARG || st == OCO_E_HIP || st == OCO_E_OOM) return st;
    /* the device rows are padded (oco_batch_ldu, oco_batch_ldv); the host
     * arrays get the packed layout, one pitched copy per column and array */
    long pu = 0, pv = 0;
    for (int i = 0; i < b->ncol; i++) {
        const size_t nz = (size_t)b->nz[i], nw = (size_t)b->nw[i];
        HIPCHK(hipMemcpy2DAsync(U + pu, (nz + 6) * sizeof(double), (const double*)b->U.p + b->uoff[i],
                                (size_t)oco_batch_ldu((int)nz) * sizeof(double), (nz + 6) * sizeof(double), nw,
                                hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpy2DAsync(V + pv, nz * sizeof(double), (const double*)b->V.p + b->voff[i],
                                (size_t)oco_batch_ldv((int)nz) * sizeof(double), nz * sizeof(double), nw,
                                hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpy2DAsync(absorp + pv, nz * sizeof(double), (const double*)b->A.p + b->voff[i],
                                (size_t)oco_batch_ldv((int)nz) * sizeof(double), nz * sizeof(double), nw,
                                hipMemcpyDeviceToHost, c->stream));
        pu += (long)(nz + 6) * nw;
        pv += (long)nz * nw;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return st;
}

}  /* extern "C" */
