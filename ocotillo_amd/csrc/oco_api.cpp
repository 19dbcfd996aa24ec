/*
 * oco_api.cpp -- the C ABI (include/ocotillo.h): context, resident
 * wavelength tables, device buffers and the block solve.
 */
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
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
    DevBuf wtab, rho, T, cells, scratch, U, V, A, fail, canary;
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
static void build_wtab(const std::vector<double>& waves, std::vector<double>& wt)
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
    }
}

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
    *out = c;
    return OCO_OK;
}

int oco_destroy(oco_ctx* c)
{
    if (!c) return OCO_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (DevBuf* b : {&c->wtab, &c->rho, &c->T, &c->cells, &c->scratch, &c->U, &c->V, &c->A, &c->fail,
                      &c->canary})
        release(c, *b);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
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
    c->waves.assign(waves, waves + nw);
    std::vector<double> wt;
    build_wtab(c->waves, wt);
    int st = ensure(c, c->wtab, wt.size() * sizeof(double));
    if (st) return st;
    HIPCHK(hipMemcpyAsync(c->wtab.p, wt.data(), wt.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->nw = nw;
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

int oco_solve_block_device(oco_ctx* c, int nz, int nyloc, int nxloc, double dz, double dz1, double dz2,
                           const double* d_rho, const double* d_T, double* d_U, double* d_V, double* d_A,
                           int mode, oco_diag* diag)
{
    if (!c) return OCO_E_ARG;
    if (nz < 4 || nyloc < 1 || nxloc < 1 || !d_rho || !d_T || !d_U || !d_V || !d_A) return OCO_E_ARG;
    if (mode != OCO_MODE_FP64) return OCO_E_ARG;
    if (c->nw <= 0) return OCO_E_STATE;
    if (c->lgrey && c->nw != 1) return OCO_E_ARG;   /* source/flux_feautrier.f90:84-87 */
    HIPCHK(hipSetDevice(c->device));
    const long ncol = (long)nxloc * nyloc;
    const long ncells = ncol * nz;
    const long nsys = ncol * c->nw;
    const long ntiles = (nsys + 63) / 64;
    int nblocks = (int)(ntiles < oco_sweep_max_blocks() ? ntiles : oco_sweep_max_blocks());
    int st;
    if (!c->lgrey && (st = ensure(c, c->cells, (size_t)ncells * OCO_CELL_STRIDE * sizeof(double)))) return st;
    /* one scratch slot per wavefront; multi-wavefront workgroups round up */
    const size_t slots = ((size_t)nblocks + 7) / 8 * 8;
    if ((st = ensure(c, c->scratch, slots * nz * 3 * 64 * sizeof(double)))) return st;
    if ((st = ensure(c, c->fail, sizeof(unsigned long long)))) return st;
    if ((st = ensure(c, c->canary, (4 + 5 * (size_t)nz) * sizeof(double)))) return st;

    const unsigned long long none = OCO_FAIL_NONE;
    HIPCHK(hipMemcpyAsync(c->fail.p, &none, sizeof(none), hipMemcpyHostToDevice, c->stream));

    oco_sweep_args a;
    a.nz = nz;
    a.nw = c->nw;
    a.ncol = ncol;
    a.dz = dz;
    a.dz1 = dz1;
    a.dz2 = dz2;
    a.sigma_grey = c->sigma_grey;
    a.rho = d_rho;
    a.T = d_T;
    a.cells = (const double*)c->cells.p;
    a.wtab = (const double*)c->wtab.p;
    a.U = d_U;
    a.V = d_V;
    a.A = d_A;
    a.scratch = (double*)c->scratch.p;
    a.fail = (unsigned long long*)c->fail.p;
    a.canary = (double*)c->canary.p;

    HIPCHK(hipEventRecord(c->ev[0], c->stream));
    if (!c->lgrey) HIPCHK(oco_launch_cells(d_rho, d_T, ncells, c->fully_ionized_T, (double*)c->cells.p, c->stream));
    HIPCHK(hipEventRecord(c->ev[1], c->stream));
    HIPCHK(oco_launch_sweep(a, c->lgrey, nblocks, c->stream));
    HIPCHK(hipEventRecord(c->ev[2], c->stream));

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
        long sys = (long)(key >> 32);
        d.status = (int)((key >> 28) & 0xf);
        d.fail_j = (int)(key & 0x0fffffffULL);
        long col = sys / c->nw;
        d.fail_iw = (int)(sys % c->nw) + 1;
        d.fail_ix = (int)(col / nyloc) + 1;
        d.fail_iy = (int)(col % nyloc) + 1;
    } else {
        /* first-pillar canaries: system (ix,iy,iw) = (1,1,1) is the first of each array */
        std::vector<double> h((size_t)(nz + 6) + 2 * (size_t)nz);
        HIPCHK(hipMemcpy(h.data(), d_U, (nz + 6) * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(h.data() + nz + 6, d_V, nz * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(h.data() + 2 * nz + 6, d_A, nz * sizeof(double), hipMemcpyDeviceToHost));
        const double* Uh = h.data() + 3;
        const double* Vh = h.data() + nz + 6;
        const double* Ah = h.data() + 2 * nz + 6;
        d.minU = d.maxU = Uh[0];
        d.minV = d.maxV = Vh[0];
        d.minA = d.maxA = Ah[0];
        for (int i = 0; i < nz; i++) {
            d.minU = fmin(d.minU, Uh[i]); d.maxU = fmax(d.maxU, Uh[i]);
            d.minV = fmin(d.minV, Vh[i]); d.maxV = fmax(d.maxV, Vh[i]);
            d.minA = fmin(d.minA, Ah[i]); d.maxA = fmax(d.maxA, Ah[i]);
        }
    }
    if (diag) *diag = d;
    return d.status;
}

int oco_solve_block(oco_ctx* c, int nz, int nyloc, int nxloc, double dz, double dz1, double dz2,
                    const double* rho3d, const double* T3d, double* U, double* V, double* absorp, int mode,
                    oco_diag* diag)
{
    if (!c || !rho3d || !T3d || !U || !V || !absorp) return OCO_E_ARG;
    if (nz < 4 || nyloc < 1 || nxloc < 1) return OCO_E_ARG;
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

}  /* extern "C" */
