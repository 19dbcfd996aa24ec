/*
 * abi_check.c -- drives every entry point of include/ocotillo.h through the
 * library built with host AddressSanitizer and UndefinedBehaviorSanitizer
 * (ocotillo_amd/libocotillo_san.so): test infrastructure only.
 *
 *   abi_check cpu   argument checks, status texts, the record writer and,
 *                   without a GPU, the no-device path (no kernel runs)
 *   abi_check gpu   the same, then small solves on device 0: a block in the
 *                   three precision modes (host and device buffers, the
 *                   allocation-free entry point), the opacity query, the
 *                   streamed block outputs, an Athena slab, a batch in both
 *                   modes, a server in both modes with a job that can never
 *                   fit, a budget too small for a second wavelength grid
 *
 * A sanitizer finding aborts the program (-fno-sanitize-recover); any
 * unexpected status prints the step and exits 1.  The program ends by _exit
 * after its summary line (see the end of main).  Every step prints its
 * name first, so a failure names the case.
 */
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "../../include/ocotillo.h"

static int failures = 0;

#define STEP(name) (printf("step: %s\n", name), fflush(stdout))
#define EXPECT(cond, what)                                                  \
    do {                                                                    \
        if (!(cond)) {                                                      \
            printf("FAILED: %s (%s:%d)\n", what, __FILE__, __LINE__);       \
            fflush(stdout);                                                 \
            failures++;                                                     \
        }                                                                   \
    } while (0)

static int all_finite(const double* x, size_t n)
{
    for (size_t i = 0; i < n; i++)
        if (!isfinite(x[i])) return 0;
    return 1;
}

/* seeded disk-like column profiles: rho (nz, ny, nx), T (nz, ny, nx) */
static void profiles(int nz, int ny, int nx, double* rho, double* T, unsigned seed)
{
    srand(seed);
    for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++) {
            const double r0 = 1e-10 * (1.0 + 9.0 * rand() / (double)RAND_MAX);
            const double t0 = 2e4 + 8e4 * rand() / (double)RAND_MAX;
            for (int z = 0; z < nz; z++) {
                const double x = (2.0 * z / (nz - 1) - 1.0) * 3.0;
                const size_t i = ((size_t)ix * ny + iy) * nz + z;
                rho[i] = fmax(r0 * exp(-0.5 * x * x), 1e-15);
                T[i] = fmax(t0 * exp(-0.25 * x * x), 2000.0);
            }
        }
}

static void waves_log(int nw, double* w)
{
    for (int i = 0; i < nw; i++) w[i] = 100.0 * pow(1000.0, nw > 1 ? i / (double)(nw - 1) : 0.0);
}

static void cpu_checks(void)
{
    STEP("status texts and version");
    for (int s = -1; s <= 12; s++) EXPECT(oco_strerror(s) != NULL && strlen(oco_strerror(s)) > 0, "oco_strerror");
    EXPECT(oco_version() != NULL, "oco_version");

    STEP("null and out-of-range arguments");
    oco_ctx* c = NULL;
    double x = 1.0;
    oco_diag d;
    EXPECT(oco_create(-1, 0, &c) == OCO_E_ARG || oco_create(-1, 0, &c) == OCO_E_NODEVICE, "create(-1)");
    EXPECT(oco_create(0, 0, NULL) == OCO_E_ARG, "create(out NULL)");
    EXPECT(oco_destroy(NULL) == OCO_OK || oco_destroy(NULL) == OCO_E_ARG, "destroy(NULL)");
    EXPECT(oco_set_wavelengths(NULL, 1, &x) == OCO_E_ARG, "set_wavelengths(NULL)");
    EXPECT(oco_set_physics(NULL, 2e4, 0, 0.4) == OCO_E_ARG, "set_physics(NULL)");
    EXPECT(oco_set_checks(NULL, OCO_CHECK_FINITE) == OCO_E_ARG, "set_checks(NULL)");
    EXPECT(oco_solve_block(NULL, 4, 1, 1, 1.0, 1.0, 1.0, &x, &x, &x, &x, &x, OCO_MODE_FP64, &d) == OCO_E_ARG,
           "solve_block(NULL)");
    EXPECT(oco_solve_block_device(NULL, 4, 1, 1, 1.0, 1.0, 1.0, &x, &x, &x, &x, &x, OCO_MODE_FP64, &d) ==
               OCO_E_ARG, "solve_block_device(NULL)");
    EXPECT(oco_opacity_points(NULL, 1, &x, &x, OCO_MODE_FP64, &x, NULL, NULL) == OCO_E_ARG, "opacity(NULL)");
    EXPECT(oco_batch_create(NULL, 1, NULL, NULL, NULL, NULL, NULL) == OCO_E_ARG, "batch_create(NULL)");
    EXPECT(oco_batch_destroy(NULL) == OCO_OK || oco_batch_destroy(NULL) == OCO_E_ARG, "batch_destroy(NULL)");
    EXPECT(oco_server_create(NULL, 1 << 30, 64, OCO_MODE_MIXED, NULL) == OCO_E_ARG, "server_create(NULL)");
    EXPECT(oco_server_submit(NULL, NULL, NULL) == OCO_E_ARG, "server_submit(NULL)");
    EXPECT(oco_write_record(NULL, &x, 8) != OCO_OK, "write_record(NULL)");

    STEP("record writer");
    char path[] = "/tmp/oco_abi_check_XXXXXX";
    const int fd = mkstemp(path);
    EXPECT(fd >= 0, "mkstemp");
    if (fd >= 0) {
        close(fd);
        double a[37];
        for (int i = 0; i < 37; i++) a[i] = 0.5 * i - 3.0;
        EXPECT(oco_write_record(path, a, sizeof(a)) == OCO_OK, "write_record");
        FILE* f = fopen(path, "rb");
        unsigned char buf[4 + sizeof(a) + 4];
        const size_t got = f ? fread(buf, 1, sizeof(buf), f) : 0;
        if (f) fclose(f);
        int m0 = 0, m1 = 0;
        memcpy(&m0, buf, 4);
        memcpy(&m1, buf + 4 + sizeof(a), 4);
        EXPECT(got == sizeof(buf) && m0 == (int)sizeof(a) && m1 == (int)sizeof(a) &&
                   memcmp(buf + 4, a, sizeof(a)) == 0, "record layout");
        unlink(path);
    }
}

static void gpu_checks(void)
{
    oco_ctx* c = NULL;
    oco_diag d;
    STEP("context");
    int st = oco_create(0, 0, &c);
    EXPECT(st == OCO_OK, "oco_create");
    if (st != OCO_OK) return;
    enum { NX = 3, NY = 2, NZ = 40, NW = 20 };
    double waves[NW];
    waves_log(NW, waves);
    static double rho[NX * NY * NZ], T[NX * NY * NZ];
    static double U[(NZ + 6) * NY * NX * NW], V[NZ * NY * NX * NW], A[NZ * NY * NX * NW];
    profiles(NZ, NY, NX, rho, T, 7);
    const double dz = 1e13, dz1 = 1.0 / dz, dz2 = dz * dz;
    EXPECT(oco_solve_block(c, NZ, NY, NX, dz, dz1, dz2, rho, T, U, V, A, OCO_MODE_FP64, &d) == OCO_E_STATE,
           "solve before the wavelengths are set");
    EXPECT(oco_set_wavelengths(c, NW, waves) == OCO_OK, "set_wavelengths");
    EXPECT(oco_set_physics(c, 2e4, 0, 0.4) == OCO_OK, "set_physics");
    EXPECT(oco_solve_block(c, 1, NY, NX, dz, dz1, dz2, rho, T, U, V, A, OCO_MODE_FP64, &d) == OCO_E_ARG,
           "nz = 1");
    EXPECT(oco_solve_block(c, 3, NY, NX, dz, dz1, dz2, rho, T, U, V, A, OCO_MODE_MIXED, &d) == OCO_E_ARG,
           "mixed nz = 3");

    const int modes[3] = {OCO_MODE_FP64, OCO_MODE_MIXED, OCO_MODE_ACCURATE};
    for (int m = 0; m < 3; m++) {
        STEP(m == 0 ? "block FP64" : m == 1 ? "block MIXED" : "block ACCURATE");
        memset(U, 0, sizeof(U));
        st = oco_solve_block(c, NZ, NY, NX, dz, dz1, dz2, rho, T, U, V, A, modes[m], &d);
        EXPECT(st == OCO_OK && d.status == OCO_OK, "solve_block");
        EXPECT(all_finite(U, sizeof(U) / 8) && all_finite(V, sizeof(V) / 8) && all_finite(A, sizeof(A) / 8),
               "finite outputs");
    }

    STEP("finite check");
    EXPECT(oco_set_checks(c, 2) == OCO_E_ARG, "unknown check flag");
    EXPECT(oco_set_checks(c, OCO_CHECK_FINITE) == OCO_OK, "set_checks");
    EXPECT(oco_solve_block(c, NZ, NY, NX, dz, dz1, dz2, rho, T, U, V, A, OCO_MODE_FP64, &d) == OCO_OK,
           "clean block under the check");
    {
        const double keep = rho[NZ + 3];
        rho[NZ + 3] = NAN;   /* column (ix, iy) = (1, 2), cell 4 */
        st = oco_solve_block(c, NZ, NY, NX, dz, dz1, dz2, rho, T, U, V, A, OCO_MODE_MIXED, &d);
        EXPECT(st == OCO_E_NONFINITE && d.fail_ix == 1 && d.fail_iy == 2 && d.fail_iw == 1 && d.fail_j == 4,
               "NaN located");
        rho[NZ + 3] = keep;
    }
    EXPECT(oco_set_checks(c, 0) == OCO_OK, "checks off");

    STEP("device buffers and the allocation-free solve");
    double *d_rho = NULL, *d_T = NULL, *d_U = NULL, *d_V = NULL, *d_A = NULL;
    EXPECT(hipMalloc((void**)&d_rho, sizeof(rho)) == hipSuccess && hipMalloc((void**)&d_T, sizeof(T)) == hipSuccess &&
               hipMalloc((void**)&d_U, sizeof(U)) == hipSuccess && hipMalloc((void**)&d_V, sizeof(V)) == hipSuccess &&
               hipMalloc((void**)&d_A, sizeof(A)) == hipSuccess, "hipMalloc");
    EXPECT(hipMemcpy(d_rho, rho, sizeof(rho), hipMemcpyHostToDevice) == hipSuccess &&
               hipMemcpy(d_T, T, sizeof(T), hipMemcpyHostToDevice) == hipSuccess, "upload");
    EXPECT(oco_solve_block_device(c, NZ, NY, NX, dz, dz1, dz2, d_rho, d_T, d_U, d_V, d_A, OCO_MODE_MIXED, &d) ==
               OCO_OK, "solve_block_device");
    EXPECT(oco_solve_block_device_noalloc(c, NZ, NY, NX, dz, dz1, dz2, d_rho, d_T, d_U, d_V, d_A, OCO_MODE_MIXED,
                                          &d) == OCO_OK, "noalloc mixed");
    EXPECT(oco_solve_block_device_noalloc(c, NZ, NY, NX, dz, dz1, dz2, d_rho, d_T, d_U, d_V, d_A, OCO_MODE_FP64,
                                          &d) == OCO_E_ARG, "noalloc FP64 refused");
    (void)hipFree(d_rho);
    (void)hipFree(d_T);
    (void)hipFree(d_U);
    (void)hipFree(d_V);
    (void)hipFree(d_A);

    STEP("opacity query");
    {
        static double al[NW * 5], om[NW * 5], S[NW * 5];
        const double pr[5] = {1e-12, 1e-10, 1e-9, 1e-8, 1e-15}, pt[5] = {3000.0, 8000.0, 2e4, 6e4, 2000.0};
        EXPECT(oco_opacity_points(c, 5, pr, pt, OCO_MODE_FP64, al, om, S) == OCO_OK, "opacity FP64");
        EXPECT(oco_opacity_points(c, 5, pr, pt, OCO_MODE_MIXED, al, NULL, NULL) == OCO_OK, "opacity mixed");
        EXPECT(all_finite(al, NW * 5), "finite alpha");
    }

    STEP("streamed block outputs");
    {
        char dir[] = "/tmp/oco_abi_dir_XXXXXX";
        if (mkdtemp(dir)) {
            static double ucol[(NZ + 6) * NW], acol[NZ * NW];
            EXPECT(oco_solve_block_files(c, NZ, NY, NX, dz, dz1, dz2, rho, T, OCO_MODE_FP64, dir, "0000", 4096, ucol,
                                         acol, &d) == OCO_OK, "solve_block_files");
            const char* names[3] = {"mean_intensity", "absorption_coefficients", "flux"};
            for (int i = 0; i < 3; i++) {
                char p[512];
                snprintf(p, sizeof(p), "%s/%s_0000.bin", dir, names[i]);
                EXPECT(access(p, F_OK) == 0, "output file");
                unlink(p);
            }
            rmdir(dir);
        }
    }

    STEP("Athena slab");
    {
        enum { NZL = NZ / 2 };
        static double raw[5][NX * NY * NZL];
        double tg[NZ];
        for (int k = 0; k < NX * NY * NZL; k++) {
            raw[0][k] = 1.0 + 0.01 * k;
            raw[1][k] = raw[2][k] = raw[3][k] = 0.001 * k;
            raw[4][k] = 2.0 + 0.02 * k;
        }
        for (int z = 0; z < NZ; z++) tg[z] = 3000.0;
        const oco_athena_units u = {1e-9, 1e4, 0.4, 1e-15, 1.0, 2000.0, 1};
        static double r3[NX * NY * NZ], t3[NX * NY * NZ];
        EXPECT(oco_athena_slab(c, NZ, NY, NX, NZL, NZL, raw[0], raw[1], raw[2], raw[3], raw[4], tg, &u, r3, t3) ==
                   OCO_OK, "athena_slab");
        EXPECT(oco_athena_slab(c, NZ, NY, NX, NZL, NZ, raw[0], raw[1], raw[2], raw[3], raw[4], tg, &u, r3, t3) ==
                   OCO_E_ARG, "athena_slab past the block");
    }

    STEP("batch, both modes");
    {
        const int ncol = 3, nzs[3] = {40, 17, 90}, nws[3] = {5, 20, 64};
        double bw[89], bdz[3] = {1e13, 2e13, 5e12};
        waves_log(5, bw);
        waves_log(20, bw + 5);
        waves_log(64, bw + 25);
        static double br[147], bT[147];
        for (int c0 = 0, i = 0; i < ncol; c0 += nzs[i], i++) profiles(nzs[i], 1, 1, br + c0, bT + c0, 11 + i);
        oco_batch* b = NULL;
        EXPECT(oco_batch_create(c, ncol, nzs, nws, bw, bdz, &b) == OCO_OK, "batch_create");
        if (b) {
            size_t nc = 0, ns = 0, nU = 0, nVA = 0;
            EXPECT(oco_batch_layout(b, &nc, &ns, &nU, &nVA) == OCO_OK && nc == 147, "batch_layout");
            double* bU = (double*)malloc(nU * 8);
            double* bV = (double*)malloc(nVA * 8);
            double* bA = (double*)malloc(nVA * 8);
            for (int m = 0; m < 2; m++)
                for (int rep = 0; rep < 2; rep++)
                    EXPECT(oco_batch_run_host(b, br, bT, bU, bV, bA, m ? OCO_MODE_MIXED : OCO_MODE_FP64, &d) == OCO_OK,
                           "batch_run_host");
            EXPECT(all_finite(bU, nU) && all_finite(bV, nVA), "batch finite");
            size_t uo = 0, vo = 0;
            int lu = 0, lv = 0;
            EXPECT(oco_batch_column(b, 2, &uo, &lu, &vo, &lv) == OCO_OK && lu == 96 && lv == 96 &&
                       uo == (size_t)(48 * 5 + 32 * 20) && vo == (size_t)(48 * 5 + 32 * 20),
                   "batch_column offsets and padded rows");
            EXPECT(oco_batch_column(b, 3, &uo, &lu, &vo, &lv) == OCO_E_ARG, "batch_column past the end");
            float ms[32];
            EXPECT(oco_batch_tune_times(b, OCO_MODE_MIXED, ms, 32) > 0, "tune times");
            free(bU);
            free(bV);
            free(bA);
            EXPECT(oco_batch_destroy(b) == OCO_OK, "batch_destroy");
        }
    }

    for (int m = 0; m < 2; m++) {
        STEP(m ? "server MIXED" : "server FP64");
        oco_server* s = NULL;
        EXPECT(oco_server_create(c, (size_t)64 << 20, 128, m ? OCO_MODE_MIXED : OCO_MODE_FP64, &s) == OCO_OK,
               "server_create");
        if (!s) continue;
        static double em[3][2 * NY * NX * NW];
        long long tk[3];
        oco_job j = {NX, NY, NZ, NW, dz, rho, T, waves, NULL};
        for (int k = 0; k < 3; k++) {
            j.emergent = em[k];
            EXPECT(oco_server_submit(s, &j, &tk[k]) == OCO_OK, "server_submit");
        }
        oco_job big = j;
        big.nx = 40000;   /* rho/T are not read: the job is refused at submit */
        long long tb = 0;
        EXPECT(oco_server_submit(s, &big, &tb) == OCO_E_OOM, "job that never fits");
        for (int k = 0; k < 3; k++) {
            oco_job_result r;
            EXPECT(oco_server_wait(s, tk[k], &r) == OCO_OK && r.status == OCO_OK, "server_wait");
        }
        EXPECT(memcmp(em[0], em[1], sizeof(em[0])) == 0 && memcmp(em[0], em[2], sizeof(em[0])) == 0,
               "the same job, the same bits");
        oco_serve_stats ss;
        EXPECT(oco_server_stats(s, &ss) == OCO_OK && ss.completed == 3, "server_stats");
        EXPECT(oco_server_destroy(s) == OCO_OK, "server_destroy");
    }
    EXPECT(oco_destroy(c) == OCO_OK, "oco_destroy");

    STEP("budget too small for a second grid");
    {
        oco_ctx* c2 = NULL;
        EXPECT(oco_create(0, 64 << 10, &c2) == OCO_OK, "oco_create with a budget");
        if (c2) {
            static double wbig[4096];
            waves_log(4096, wbig);
            EXPECT(oco_set_wavelengths(c2, 4, waves) == OCO_OK, "small grid");
            st = oco_set_wavelengths(c2, 4096, wbig);
            EXPECT(st == OCO_E_OOM, "larger grid over the budget");
            EXPECT(oco_solve_block(c2, NZ, NY, NX, dz, dz1, dz2, rho, T, U, V, A, OCO_MODE_FP64, &d) == OCO_E_STATE,
                   "solve after the failed grid");
            EXPECT(oco_destroy(c2) == OCO_OK, "oco_destroy");
        }
    }
}

int main(int argc, char** argv)
{
    const int gpu = argc > 1 && strcmp(argv[1], "gpu") == 0;
    cpu_checks();
    if (gpu) {
        gpu_checks();
    } else {
        STEP("no device");
        oco_ctx* c = NULL;
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) EXPECT(oco_create(0, 0, &c) == OCO_E_NODEVICE, "no device");
    }
    printf("abi_check %s: %d failure(s)\n", gpu ? "gpu" : "cpu", failures);
    fflush(stdout);
    /* Leave without the runtimes' exit-time teardown: under AddressSanitizer
     * the HIP runtime's own destructors can free device memory the runtime
     * kept after the sanitizer has marked the device runtime unloaded, which
     * trips the sanitizer's device-allocator check (sanitizer_allocator_device.h)
     * whether or not this program and the library freed everything they
     * allocated (every context, plan and server above is destroyed). */
    _exit(failures ? 1 : 0);
}
