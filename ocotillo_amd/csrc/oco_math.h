/*
 * oco_math.h -- FP64 elementary functions that give the same bits on the
 * host (x86-64, gcc/clang with -ffp-contract=off and hardware FMA) and on
 * gfx950.  Used by the HIP kernels; also usable from plain C.
 *
 * The reference evaluates exp, log, log10, sqrt and real**real through the
 * host libm.  A GPU libm never matches a CPU libm bit for bit, and several
 * quantities on this path amplify a one-ulp change (the Saha root
 * x = (sqrt(C^2+4C)-C)/2 near full ionisation, and the Feautrier rows of
 * optically thin cells, where b = -(a+c+zeta) with zeta << a+c).  These
 * functions are therefore written once, with every rounding spelled out, so
 * that a CPU build using them reproduces the GPU's results exactly.
 * Accuracy is about 0.5 ulp plus 2^-60 relative (exp, log, log10, exp10) and
 * about 2^-62 relative (pow); NaN/inf/zero inputs follow C99 Annex F for the
 * cases the solver can produce.  Results are nearly always the correctly
 * rounded ones, as glibc's are, which keeps the two libms in agreement on
 * the inputs where the reference's formulas amplify a one-ulp difference.
 * The common path has no branches (range and special cases are clamps and
 * selects), so a wavefront never diverges here.
 *
 * Methods (standard, see e.g. Tang, ACM TOMS 15(2) 1989 and Muller,
 * "Elementary Functions", ch. 6):
 *   exp:   x = k ln2/128 + r, |r| <= ln2/256;  exp(x) = 2^(k/128) (1 + p(r)),
 *          with 2^(i/128) tabulated as hi*(1+lo) and p a degree-6 Taylor
 *          polynomial.  Two-part arguments (hi + lo) are accepted.
 *   log:   x = 2^k z; r = z*invc - 1 is exact (invc has <= 9 bits);
 *          log x = k ln2 - log(invc) + log1p(r), carried as a double-double
 *          (about 2^-68 absolute) so pow and log10 can reuse it.  The cells
 *          touching 1 have invc = 1, so no separate path is needed near 1.
 *   pow:   exp(y log x) with y*log x formed exactly by fma.
 *   exp10: exp(y ln10) with y*ln10 formed exactly by fma.
 *   sqrt:  IEEE square root (correctly rounded on both targets).
 */
#ifndef OCO_MATH_H
#define OCO_MATH_H

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define OCO_HD __host__ __device__
#else
#define OCO_HD
#endif

#include "oco_math_tables.h"

typedef struct {
    double exp_thi[128];
    double exp_tlo[128];
    double log_invc[128];
    double log_lhi[128];
    double log_llo[128];
} oco_mtab;

#define OCO_MTAB_INIT                                                                    \
    {                                                                                    \
        OCO_EXP_THI_INIT, OCO_EXP_TLO_INIT, OCO_LOG_INVC_INIT, OCO_LOG_LHI_INIT, OCO_LOG_LLO_INIT \
    }

#define OCO_INVLN2N    0x1.71547652b82fep+7    /* RN(128/ln2) */
#define OCO_LN2N_HI    0x1.62e42fefc0000p-8    /* ln2/128, 35 bits: k*HI exact, |k|<2^18 */
#define OCO_LN2N_LO   -0x1.c610ca86c3899p-44
#define OCO_LN2_HI     0x1.62e42fefa3800p-1    /* ln2, 42 bits: k*HI exact, |k|<2^11 */
#define OCO_LN2_LO     0x1.ef35793c76730p-45
#define OCO_LN10_HI    0x1.26bb1bbb55516p+1
#define OCO_LN10_LO   -0x1.f48ad494ea3e9p-53
#define OCO_IVLN10_HI  0x1.bcb7b1526e50ep-2
#define OCO_IVLN10_LO  0x1.95355baaafad3p-57
#define OCO_LOG_OFF    0x3fe6a00000000000ULL   /* bits of 0x1.6ap-1 */
#define OCO_SHIFT      0x1.8p52

OCO_HD static inline uint64_t oco_asu64(double d)
{
    uint64_t u;
    __builtin_memcpy(&u, &d, sizeof(u));
    return u;
}

OCO_HD static inline double oco_asf64(uint64_t u)
{
    double d;
    __builtin_memcpy(&d, &u, sizeof(d));
    return d;
}

OCO_HD static inline double oco_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }

#define OCO_INF (0x1p1023 * 2.0)

OCO_HD static inline double oco_sel(int c, double a, double b) { return c ? a : b; }

/* exp(hi + lo), |lo| <= ulp(hi).  No branch on the common path: the
 * argument is clamped to +-2000 (beyond which the result is inf or 0 anyway),
 * the table scale's exponent to the normal range, and ldexp applies the rest
 * (exact, or one rounding into the subnormal range, or overflow to inf);
 * only NaN needs a select. */
OCO_HD static inline double oco_exp_dd(double hi, double lo, const oco_mtab* T)
{
    double hc = __builtin_fmin(__builtin_fmax(hi, -2000.0), 2000.0);
    double kd = hc * OCO_INVLN2N + OCO_SHIFT;
    kd = kd - OCO_SHIFT;
    int ki = (int)kd;
    double r = (hc - kd * OCO_LN2N_HI) - kd * OCO_LN2N_LO;
    r = r + lo;
    int idx = ki & 127;
    int e = ki >> 7;
    double r2 = r * r;
    double p = oco_fma(r2, oco_fma(r, 0x1.5555555555555p-3, 0.5), r);
    p = oco_fma(r2 * r2, oco_fma(r2, 0x1.6c16c16c16c17p-10, oco_fma(r, 0x1.1111111111111p-7, 0x1.5555555555555p-5)), p);
    double tmp = T->exp_tlo[idx] + p;
    int ec = e < -1000 ? -1000 : (e > 1023 ? 1023 : e);
    double scale = oco_asf64(oco_asu64(T->exp_thi[idx]) + ((uint64_t)(int64_t)ec << 52));
    double y = __builtin_ldexp(oco_fma(scale, tmp, scale), e - ec);
    return oco_sel(hi != hi, hi + hi, y);
}

OCO_HD static inline double oco_exp_t(double x, const oco_mtab* T) { return oco_exp_dd(x, 0.0, T); }

/* log(x) as a double-double *hi + *lo for finite x > 0 (other inputs give
 * values that the callers replace).  The two table cells touching 1 have
 * invc = 1 and log c = 0, so near 1 the result is log1p(x - 1) with x - 1
 * exact. */
OCO_HD static inline void oco_log_dd(double x, const oco_mtab* T, double* hi, double* lo)
{
    uint64_t ix = oco_asu64(x);
    int sub = ix < 0x0010000000000000ULL;    /* subnormal: scale by 2^52 */
    ix = sub ? oco_asu64(x * 0x1p52) : ix;
    uint64_t tmp = ix - OCO_LOG_OFF;
    int i = (int)((tmp >> 45) & 127);
    int k = (int)((int64_t)tmp >> 52) - (sub ? 52 : 0);
    double z = oco_asf64(ix - (tmp & (0xfffULL << 52)));
    double r = oco_fma(z, T->log_invc[i], -1.0);  /* exact */
    double kd = (double)k;
    double t1 = kd * OCO_LN2_HI;               /* exact */
    double lc = T->log_lhi[i];
    double s = t1 + lc;
    double bb = s - t1;
    double e1 = (t1 - (s - bb)) + (lc - bb);
    double h = s + r;
    bb = h - s;
    double e2 = (s - (h - bb)) + (r - bb);
    double r2 = r * r;
    double r2e = oco_fma(r, r, -r2);
    /* log1p(r) = r - r^2/2 + r^3 q(r), q = sum_{n=3..11} (-1)^(n+1) r^(n-3) / n */
    double q = (1.0 / 11);
    q = -(1.0 / 10) + r * q;
    q = (1.0 / 9) + r * q;
    q = -(1.0 / 8) + r * q;
    q = (1.0 / 7) + r * q;
    q = -(1.0 / 6) + r * q;
    q = (1.0 / 5) + r * q;
    q = -(1.0 / 4) + r * q;
    q = (1.0 / 3) + r * q;
    double p = (r * r2) * q;
    double l = -0.5 * r2 + ((p - 0.5 * r2e) + (((kd * OCO_LN2_LO + T->log_llo[i]) + e1) + e2));
    double hs = h + l;
    *lo = l - (hs - h);
    *hi = hs;
}

#define OCO_QNAN oco_asf64(0x7ff8000000000000ULL)

/* result for x outside (0, inf): log(0) = -inf, log(inf) = inf, else NaN */
OCO_HD static inline double oco_log_special(double x)
{
    return oco_sel(x == 0.0, -OCO_INF, oco_sel(x == OCO_INF, x, OCO_QNAN));
}

OCO_HD static inline double oco_log_t(double x, const oco_mtab* T)
{
    double h, l;
    oco_log_dd(x, T, &h, &l);   /* log(1) = 0 exactly: invc = 1 on its cell */
    return oco_sel(x > 0.0 && x < OCO_INF, h, oco_log_special(x));
}

OCO_HD static inline double oco_log10_t(double x, const oco_mtab* T)
{
    double h, l;
    oco_log_dd(x, T, &h, &l);
    double p = h * OCO_IVLN10_HI;
    double pe = oco_fma(h, OCO_IVLN10_HI, -p);
    double y = p + (pe + (h * OCO_IVLN10_LO + l * OCO_IVLN10_HI));
    return oco_sel(x > 0.0 && x < OCO_INF, y, oco_log_special(x));
}

/* 10**y */
OCO_HD static inline double oco_exp10_t(double y, const oco_mtab* T)
{
    double hi = y * OCO_LN10_HI;
    double lo = oco_fma(y, OCO_LN10_HI, -hi) + y * OCO_LN10_LO;
    return oco_exp_dd(hi, lo, T);
}

/* x**y for x >= 0 (the solver never raises a negative base). */
OCO_HD static inline double oco_pow_t(double x, double y, const oco_mtab* T)
{
    double h, l;
    oco_log_dd(x, T, &h, &l);
    double yh = y * h;
    double yl = oco_fma(y, h, -yh) + y * l;
    double r = oco_exp_dd(yh, yl, T);
    int finite_pos = x > 0.0 && x < OCO_INF;
    r = oco_sel(finite_pos, r, oco_sel(x == 0.0, oco_sel(y > 0.0, 0.0, OCO_INF),
                                       oco_sel(x == OCO_INF, oco_sel(y > 0.0, OCO_INF, 0.0), OCO_QNAN)));
    r = oco_sel(x != x || y != y, x + y, r);
    return oco_sel(y == 0.0 || x == 1.0, 1.0, r);
}

OCO_HD static inline double oco_sqrt(double x) { return __builtin_sqrt(x); }

#if !defined(__HIP_DEVICE_COMPILE__)
/* Host-side table instance and table-free wrappers (host code only). */
static const oco_mtab oco_mtab_host = OCO_MTAB_INIT;
static inline double oco_exp(double x) { return oco_exp_t(x, &oco_mtab_host); }
static inline double oco_log(double x) { return oco_log_t(x, &oco_mtab_host); }
static inline double oco_log10(double x) { return oco_log10_t(x, &oco_mtab_host); }
static inline double oco_exp10(double y) { return oco_exp10_t(y, &oco_mtab_host); }
static inline double oco_pow(double x, double y) { return oco_pow_t(x, y, &oco_mtab_host); }
#endif

#endif
