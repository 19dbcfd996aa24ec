/*
 * oco_mixed.h -- the mixed-precision opacity of OCO_MODE_MIXED (device only).
 *
 * On MI355X with packed FP32 disabled, FP32 VALU arithmetic issues at the
 * same rate as FP64, so products and sums stay FP64 (which also keeps the
 * 1e-50..1e-10 range of alpha^2 and S representable).  What is cheaper in
 * FP32 is the transcendental core: exp and log use the hardware FP32
 * v_exp_f32 / v_log_f32 on an FP64-reduced argument, with the FP64 exponent
 * put back by ldexp ("exponent scaling"), and divides use v_rcp_f64 with one
 * Newton step.  Relative error: about 1.2e-7 per exp (the FP32 core), about 1e-7
 * absolute per log, ~1e-16 per divide.  The cancelling forms need no
 * transcendental at all: exp(d) - 1 is one FP64 series on the argument
 * reduced by multiples of ln 2, scaled back by 2^n, which keeps full
 * relative accuracy at long wavelengths and high temperature, and the
 * 1 - exp(-y) of stimulated emission (1 - 10^(-x)) and of the Gaunt factor
 * follow from it, their y being the same h nu / kT.
 *
 * The Saha state, number densities and every other temperature-only factor
 * come from the FP64 cell kernel, unchanged.
 */
#ifndef OCO_MIXED_H
#define OCO_MIXED_H

#include "oco_physics.h"

#define OCO_LOG2E    0x1.71547652b82fep+0
#define OCO_LOG2_10  0x1.a934f0979a371p+1
#define OCO_LN2      0x1.62e42fefa39efp-1
#define OCO_LN10     0x1.26bb1bbb55516p+1

/* 2^t for |t| < 2^31: t = n + f, |f| <= 1/2, 2^f in FP32, 2^n by ldexp
 * (which gives 0 or inf beyond the double range).  Callers bound their
 * arguments: every exponent of the opacity is bounded by the cell's
 * temperature-only factors. */
__device__ static inline double mx_exp2(double t)
{
    double n = __builtin_rint(t);
    float f = (float)(t - n);
    double m = (double)__builtin_amdgcn_exp2f(f);
    return __builtin_amdgcn_ldexp(m, (int)n);
}

__device__ static inline double mx_exp(double x) { return mx_exp2(x * OCO_LOG2E); }
__device__ static inline double mx_exp10(double x) { return mx_exp2(x * OCO_LOG2_10); }

/* ln x for finite x > 0: x = m 2^e, m in [1/2, 1), log2 m in FP32.  The error
 * is absolute (about 1e-7): callers use the result as an exponent or add it
 * to terms of order one. */
__device__ static inline double mx_log(double x)
{
    double m = __builtin_amdgcn_frexp_mant(x);
    int e = __builtin_amdgcn_frexp_exp(x);
    float l = __builtin_amdgcn_logf((float)m);
    return ((double)e + (double)l) * OCO_LN2;
}

/* 1 - exp(-y) = y (1 - y/2 + y^2/6 - ...) for |y| < 1/2: the Taylor series to
 * degree 9 (truncation below 1e-9 relative at 1/2, 2e-11 on the reduced
 * range |y| <= ln2/2 where the callers use it), Horner with FMA */
__device__ static inline double mx_series_omexp(double y)
{
    double q = (1.0 / 362880);
    q = __builtin_fma(y, q, -(1.0 / 40320));
    q = __builtin_fma(y, q, (1.0 / 5040));
    q = __builtin_fma(y, q, -(1.0 / 720));
    q = __builtin_fma(y, q, (1.0 / 120));
    q = __builtin_fma(y, q, -(1.0 / 24));
    q = __builtin_fma(y, q, (1.0 / 6));
    q = __builtin_fma(y, q, -0.5);
    q = __builtin_fma(y, q, 1.0);
    return y * q;
}

/* exp(d) - 1, 0 <= d <= 708: d = n ln2 + r, |r| <= ln2/2, gives
 * exp(d) - 1 = 2^n (exp(r) - 1) + (2^n - 1), with exp(r) - 1 = -(1 - exp(r))
 * from the series; for n = 0 the series alone */
__device__ static inline double mx_expm1(double d)
{
    const double n = __builtin_rint(d * OCO_LOG2E);
    const double r = __builtin_fma(-n, OCO_LN2, d);
    const double t = __builtin_amdgcn_ldexp(1.0, (int)n);
    return __builtin_fma(t, -mx_series_omexp(-r), t - 1.0);
}

/* 1/y and x/y: v_rcp_f64 refined by one Newton step */
__device__ static inline double mx_rcp(double y)
{
    double r = __builtin_amdgcn_rcp(y);
    return __builtin_fma(r, __builtin_fma(-y, r, 1.0), r);
}
__device__ static inline double mx_div(double x, double y) { return x * mx_rcp(y); }

/* 1 - exp(-y) for y close to d, from E = exp(d) - 1 and r1 = 1 / (1 + E):
 * 1 - exp(-y) = 1 - exp(-d) exp(-delta) = (E + (1 - exp(-delta))) r1 with
 * delta = y - d, the bracket by its series to third order.  The opacity's
 * exponents h nu / kT of stimulated emission (chi theta ln10, from the
 * rounded 1.2398e4 eV A and 5040 K) and of the bremsstrahlung Gaunt factor
 * (h nu k1 / T) are the Planck term's d times 1 + 1e-5 and 1 + 1e-16; the
 * truncation, delta^4 / 24 against E, is then below 1e-19.  When d is
 * clamped to 308, delta may be large, but the result is 1 in double as long
 * as the series (y clamped to 1000) stays far below E = exp(308). */
__device__ static inline double mx_omexp_near(double y, double d, double E, double r1)
{
#pragma clang fp contract(on)
    const double dl = __builtin_fmin(y, 1000.0) - d;
    const double q = dl * __builtin_fma(dl, __builtin_fma(dl, 1.0 / 6.0, -0.5), 1.0);
    return (E + q) * r1;
}

/* the opacity, albedo and source function of one (cell, wavelength), mixed
 * precision; same formulas as oco_point_s (source/continuous_opacity.f90) */
template <int WS>
__device__ static inline oco_pt oco_point_mixed(const double* c, const double* w)
{
    /* products feeding sums may be fused (one rounding instead of two): the
     * mixed mode is held to error bounds, not to the reference's bits */
#pragma clang fp contract(on)
    oco_pt p;
    double theta = c[CR_THETA], fn = c[CR_FN];
    /* the Planck term's exp(d) - 1, and 1/E, 1/(1 + E) from one reciprocal
     * of E (1 + E) (below 1e268 for d <= 308) */
    const double damp = __builtin_fmin(c[CR_DT1] * w[WT_WCM1 * WS], OCO_LOG_OVF);
    const double E = mx_expm1(damp);
    const double E1 = 1.0 + E;
    const double rP = mx_rcp(E * E1);
    const double rE = E1 * rP, r1 = E * rP;
    /* The same sums, regrouped: stim*fn is common to the three
     * stimulated-emission terms, A_Hbf wa^3 and alpha0 wa^3 come from the
     * table, and the neutral albedo is e_scatter over the sum before the
     * factor 1/mp (which cancels). */
    const double sf = mx_omexp_near(w[WT_CHIL * WS] * theta, damp, E, r1) * fn;
    /* Unsold tail + sum_n g(n) n^-3 exp(-chi_n/kT), from the cell's term */
    double sm = c[CR_UNSOLD] + w[WT_G1 * WS];
    sm = sm + w[WT_G2 * WS] * c[CR_E2];
    sm = sm + w[WT_G3 * WS] * c[CR_E3];
    sm = sm + w[WT_G4 * WS] * c[CR_E4];
    sm = sm + w[WT_G5 * WS] * c[CR_E5];
    sm = sm + w[WT_G6 * WS] * c[CR_E6];
    /* H- bound-free below 16000 A: the switch selects the wavelength's
     * coefficient (0 above), not the term, so no per-point select */
    const double sco = w[WT_HMBF_ON * WS] != 0.0 ? w[WT_SCOEFF * WS] : 0.0;
    const double kbf = w[WT_HBFW * WS] * sm + c[CR_HMBF] * sco;   /* (H bf + H- bf) / (stim fn) */
    double flam = (w[WT_F1M26 * WS] + w[WT_F2 * WS] * c[CR_LGT]) + w[WT_F3 * WS] * c[CR_LGT2];
    const double s1 = sf * kbf + (c[CR_PHMFF] * mx_exp10(flam) + c[CR_ES]);   /* kappa mp before free-free */
    const double rho = c[CR_RHO];
    double alpha, kappa;
    if (c[CR_NEUTRAL] != 0.0) {
        /* kappa mp += H ff; alpha = (s1 + s2) rho / mp (the double count of
         * source/continuous_opacity.f90:194); albedo = e_scatter / s2 */
        double g_ff = 1.0 + w[WT_GFF * WS] * ((c[CR_U0] * w[WT_CHI1 * WS]) + 0.5);
        const double s2 = s1 + ((sf * w[WT_HFFW * WS]) * g_ff) * c[CR_U1];
        alpha = (s1 + s2) * (rho * OCO_MP1);
        kappa = s2;
    } else {
        /* bremsstrahlung: exp(5.960 - sqrt3/pi log(inner)) = WT_GBP CR_MQ and
         * exp(lnQ) = CR_MR WT_NU3 (1 - exp(-h nu / kT)), so one logarithm is
         * left where the reference's form has three logarithms and two
         * exponentials */
        const double gB = mx_log(w[WT_GBP * WS] * c[CR_MQ] + OCO_EXP1);
        const double Q = (c[CR_MR] * w[WT_NU3 * WS]) * mx_omexp_near(w[WT_HNK * WS] * c[CR_U2], damp, E, r1);
        const double kRho = (OCO_C_BREMS * Q) * gB;
        alpha = s1 * (rho * OCO_MP1) + kRho;
        kappa = s1 + alpha * (c[CR_U3] * (1.0 / OCO_MP1));   /* kappa mp */
    }
    p.albedo = mx_div(c[CR_ES], kappa);
    p.alpha = alpha;
    p.alpha1 = mx_rcp(alpha);
    p.alpha2 = alpha * alpha;
    p.S = w[WT_SFW5 * WS] * rE;
    return p;
}

#endif
