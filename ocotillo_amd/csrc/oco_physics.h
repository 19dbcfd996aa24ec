/*
 * oco_physics.h -- FP64 physics of the per-column, per-wavelength path,
 * written for the GPU kernels (host-callable too, for the wavelength tables).
 *
 * Evaluation order follows the reference expression by expression (Fortran
 * left-to-right, unary minus after multiplication) and the code is compiled
 * with -ffp-contract=off, so with the elementary functions of oco_math.h the
 * result of every operation is fixed.  Sub-expressions that depend only on
 * the cell (T, rho) are evaluated once per cell and stored in the cell
 * record; each is exactly the value the reference recomputes per wavelength,
 * so hoisting changes no bit.
 */
#ifndef OCO_PHYSICS_H
#define OCO_PHYSICS_H

#include "oco_math.h"

/* source/common.f90:23-81 as the reference's compiler folds them (checked
 * against the compiled reference by the golden tests). */
#define OCO_SIGMA_SB     0x1.dbaa5d09e7374p-15   /* 5.6703744191844314e-05 */
#define OCO_LOG_OVF      0x1.3400000000000p+8    /* floor(log10(huge)) = 308 */
#define OCO_MP           0x1.02d3617bf22f4p-79   /* 1.67262192369e-24 */
#define OCO_MP1          0x1.fa690885f1ef2p+78   /* 1/mp */
#define OCO_H_PLANCK     0x1.067c559cb516cp-87   /* 6.62607015e-27 */
#define OCO_PI1          0x1.45f306dc9c883p-2    /* 1/pi */
#define OCO_SQRT3        0x1.bb67ae8584caap+0
#define OCO_EXP1         0x1.5bf0a8b145769p+1    /* e */
#define OCO_K_CGS        0x1.3e5b21cc8f74fp-53   /* 1.380649e-16 */
#define OCO_K1_CGS       0x1.9bb717cba28f0p+52   /* 1/k */
#define OCO_K1_EV        0x1.6aa4251ce947ap+13   /* 1/k_eV */
#define OCO_HION_EV      0x1.b333333333333p+3    /* 13.6 */
#define OCO_C_LIGHT      0x1.beb9bf3a00000p+34   /* 2.99792458e10 */
#define OCO_RYD          0x1.7f7ccd45b2805p-36   /* 2.1798741e-11 */
#define OCO_RYD1         0x1.55ca1c19b9fa9p+35   /* 1/Ryd */
#define OCO_LOG10E       0x1.bcb7b1526e50ep-2
#define OCO_IEV          0x1.b369160866805p+3    /* h c Rcm 6.2419e11 */
#define OCO_IEV1         0x1.2d07e731cc2b6p-4
#define OCO_C_BREMS      0x1.60213dead66d7p+28   /* 369234910.67735046 */
#define OCO_HCR1         0x1.31c52e6096baep-17   /* h c / Ryd */
#define OCO_ONE_THIRD    0x1.5555555555555p-2
#define OCO_TWO_THIRDS   0x1.5555555555555p-1
/* source/continuous_opacity.f90:20-25 */
#define OCO_SWITCH_ION   1e-2
#define OCO_AHBF         1.0449e-26
#define OCO_RANGSTROM    1.0968e-3
#define OCO_ALPHA0       1.0443e-26
#define OCO_MDIM         6

/* Per-cell record (everything the per-wavelength stage reads from a cell). */
enum {
    CR_THETA = 0,   /* theta = 5040/T */
    CR_FN,          /* neutral fraction nHI/(nHI+nHII) */
    CR_E2, CR_E3, CR_E4, CR_E5, CR_E6,   /* exp(-chi_n/kT), n=2..6 (n=1 is exactly 1) */
    CR_UNSOLD,      /* Unsold tail */
    CR_HMBF,        /* H- bound-free factor */
    CR_LGT, CR_LGT2,/* log10(theta), its square */
    CR_PHMFF,       /* ((fn*ne)*k)*T */
    CR_ES,          /* Thomson e_scatter */
    CR_RHO,
    CR_NEUTRAL,     /* 1.0 if (1-x) > 0.01 else 0.0 */
    CR_U0, CR_U1, CR_U2, CR_U3,  /* neutral: theta1*log10e, energyfactor; ionised: (T*1e-4)^-1.5,
                                    log(ne n) + 2 log 1 - 0.5 log T, T1, rho1 */
    CR_ESM,         /* e_scatter * mp1 */
    CR_DT1,         /* damping_factor_const * T1 */
    CR_MQ, CR_MR,   /* mixed mode only; ionised: U0^-(sqrt3/pi) = (T*1e-4)^(1.5 sqrt3/pi), exp(U1) =
                       ne n / sqrt(T); neutral: 0 */
    CR_NFIELD
};
#define OCO_CELL_STRIDE 24   /* doubles per record (padded to 192 B) */

/* Per-wavelength table (host-built once per grid).  Built with the
 * reference's expressions (source/continuous_opacity.f90:64-141) plus the
 * wavelength-only sub-expressions of the per-point formulas. */
enum {
    WT_CHI = 0, WT_WA3, WT_G1, WT_G2, WT_G3, WT_G4, WT_G5, WT_G6,  /* g(iw,n)*n1(n)**3 */
    WT_SCOEFF, WT_HMBF_ON,   /* s_coeff, 1.0 if wa < 16000 */
    WT_F1, WT_F2, WT_F3, WT_GFF, WT_CHI1,
    WT_NU9,                  /* nu*1e-9 */
    WT_L3NU,                 /* 3*log(nu) */
    WT_HNK,                  /* (h*nu)*k1 */
    WT_WCM1, WT_SFW5,        /* 1/wcm, source_function_const*w1cm5 */
    WT_HBFW, WT_HFFW,        /* mixed mode only: A_Hbf*wa^3, alpha0*wa^3, */
    WT_F1M26,                /* f_coeff(1) - 26 */
    WT_GBP,                  /* mixed mode only: exp(5.960) (nu*1e-9)^-(sqrt3/pi) */
    WT_NU3,                  /* mixed mode only: nu^-3 = exp(-3 log nu) */
    WT_CHIL,                 /* mixed mode only: chi ln10 */
    WT_NFIELD
};
#define OCO_WT_STRIDE 26

/* wavelength_independent_pillars (source/gas_state.f90:30-110 and
 * source/continuous_opacity.f90:143-163, 233-244, 262-296): fills rec. */
OCO_HD static inline void oco_cell_record(double rho, double T, double fully_ionized_T,
                                          const oco_mtab* M, double* rec)
{
    double rho1 = 1.0 / rho;
    double T1 = 1.0 / T;
    double constants = 2414683039571967.0 * oco_pow_t(T, 1.5, M);
    double exparg = (T1 * OCO_HION_EV) * OCO_K1_EV;
    double niine = constants * oco_exp_t(-exparg, M);
    double c_ion = (OCO_MP * rho1) * niine;
    double x = 0.5 * (oco_sqrt(c_ion * c_ion + 4.0 * c_ion) - c_ion);
    if (T > fully_ionized_T) x = 1.0;
    double n = rho * OCO_MP1;
    double ninv = rho1 * OCO_MP;
    double nHII = x * n;
    double nHI = n - nHII;
    double ne = nHII;
    double fn = (nHI != 0.0) ? nHI / (nHI + nHII) : 0.0;
    double Pe = (ne * OCO_K_CGS) * T;
    double es = (0.6648e-24 * nHII) * ninv;
    double theta = 5040.0 * T1;
    double theta1 = 1.0 / theta;
    double lgt = oco_log10_t(theta, M);
    double hmbf = ((4.158e-10 * Pe) * oco_pow_t(theta, 2.5, M)) * oco_exp10_t(0.754 * theta, M);
    /* T-only factors of get_kappa_H_bf (:234-242) */
    double ktemp = OCO_K_CGS * T;
    double ktemp1 = OCO_K1_CGS * T1;
    rec[CR_THETA] = theta;
    rec[CR_FN] = fn;
#pragma unroll
    for (int k = 2; k <= OCO_MDIM; k++) {
        double n1 = 1.0 / (double)k;
        double chi_n = OCO_RYD * (1.0 - 1.0 * (n1 * n1));
        rec[CR_E2 + k - 2] = oco_exp_t(-(chi_n * ktemp1), M);
    }
    double n16 = 1.0 / (double)OCO_MDIM;
    double chi_m = OCO_RYD * (1.0 - n16 * n16);
    rec[CR_UNSOLD] = ((0.5 * ktemp) * OCO_RYD1) * (oco_exp_t(-(chi_m * ktemp1), M) - oco_exp_t(-(OCO_RYD * ktemp1), M));
    rec[CR_HMBF] = hmbf;
    rec[CR_LGT] = lgt;
    rec[CR_LGT2] = lgt * lgt;
    rec[CR_PHMFF] = ((fn * ne) * OCO_K_CGS) * T;
    rec[CR_ES] = es;
    rec[CR_RHO] = rho;
    int neutral = (1.0 - x) > OCO_SWITCH_ION;
    rec[CR_NEUTRAL] = neutral ? 1.0 : 0.0;
    if (neutral) {
        rec[CR_U0] = theta1 * OCO_LOG10E;
        rec[CR_U1] = (((0.5 * OCO_LOG10E) * theta1) * OCO_IEV1) * oco_exp10_t(-(theta * OCO_IEV), M);
        rec[CR_U2] = 0.0;
        rec[CR_U3] = 0.0;
    } else {
        rec[CR_U0] = oco_pow_t(T * 1e-4, -1.5, M);
        rec[CR_U1] = (oco_log_t(ne * (nHI + nHII), M) + 2.0 * oco_log_t(1.0, M)) - 0.5 * oco_log_t(T, M);
        rec[CR_U2] = T1;
        rec[CR_U3] = rho1;
    }
    rec[CR_ESM] = es * OCO_MP1;
    /* damping_factor_const = (h*c)*k1 (:73) */
    rec[CR_DT1] = ((OCO_H_PLANCK * OCO_C_LIGHT) * OCO_K1_CGS) * T1;
    /* the cell factors of the mixed-precision bremsstrahlung term: the
     * reference's exp(5.960 - sqrt3/pi log(nu 1e-9 U0)) and exp(U1 - 3 log nu)
     * (source/continuous_opacity.f90:281-300) split into a cell factor here
     * and a wavelength factor in the table (WT_GBP, WT_NU3) */
    rec[CR_MQ] = neutral ? 0.0 : oco_exp_t(-(OCO_SQRT3 * OCO_PI1) * oco_log_t(rec[CR_U0], M), M);
    rec[CR_MR] = neutral ? 0.0 : oco_exp_t(rec[CR_U1], M);
}

typedef struct {
    double alpha, albedo, alpha1, alpha2, S;
} oco_pt;

/* calc_opacity_and_albedo for one (cell, wavelength):
 * source/continuous_opacity.f90:175-201, 203-341, 392-409.
 * w[k * WS] is field k of this wavelength's table row (WS = 1 for a row in
 * registers, 64 for a [field][lane] table in LDS). */
template <int WS>
OCO_HD static inline oco_pt oco_point_s(const double* c, const double* w, const oco_mtab* M)
{
    oco_pt p;
    double theta = c[CR_THETA], fn = c[CR_FN];
    double stim = 1.0 - oco_exp10_t(-(w[WT_CHI * WS] * theta), M);
    double sm = w[WT_G1 * WS];   /* n = 1: g*1*exp(-0) */
    sm = sm + w[WT_G2 * WS] * c[CR_E2];
    sm = sm + w[WT_G3 * WS] * c[CR_E3];
    sm = sm + w[WT_G4 * WS] * c[CR_E4];
    sm = sm + w[WT_G5 * WS] * c[CR_E5];
    sm = sm + w[WT_G6 * WS] * c[CR_E6];
    double kHbf = (((OCO_AHBF * stim) * fn) * w[WT_WA3 * WS]) * (c[CR_UNSOLD] + sm);
    double kHmbf = (w[WT_HMBF_ON * WS] != 0.0) ? ((c[CR_HMBF] * stim) * fn) * w[WT_SCOEFF * WS] : 0.0;
    double flam = ((w[WT_F1 * WS] + w[WT_F2 * WS] * c[CR_LGT]) + w[WT_F3 * WS] * c[CR_LGT2]) - 26.0;
    double kHmff = c[CR_PHMFF] * oco_exp10_t(flam, M);
    double kappa = (((kHbf + kHmbf) + kHmff) + c[CR_ES]) * OCO_MP1;
    double rho = c[CR_RHO];
    double alpha = kappa * rho;
    if (c[CR_NEUTRAL] != 0.0) {
        double g_ff = 1.0 + w[WT_GFF * WS] * ((c[CR_U0] * w[WT_CHI1 * WS]) + 0.5);
        double kHff = ((((stim * fn) * OCO_ALPHA0) * w[WT_WA3 * WS]) * g_ff) * c[CR_U1];
        kappa = kappa + kHff * OCO_MP1;
        alpha = alpha + kappa * rho;
    } else {
        double inner = w[WT_NU9 * WS] * c[CR_U0];
        double gB = oco_log_t(oco_exp_t(5.960 - (OCO_SQRT3 * OCO_PI1) * oco_log_t(inner, M), M) + OCO_EXP1, M);
        double lnQ = (c[CR_U1] - w[WT_L3NU * WS]) + oco_log_t(1.0 - oco_exp_t(-(w[WT_HNK * WS] * c[CR_U2]), M), M);
        double kRho = (OCO_C_BREMS * oco_exp_t(lnQ, M)) * gB;
        alpha = alpha + kRho;
        kappa = kappa + alpha * c[CR_U3];
    }
    p.albedo = c[CR_ESM] / kappa;
    p.alpha = alpha;
    p.alpha1 = 1.0 / alpha;
    p.alpha2 = alpha * alpha;
    double damp = c[CR_DT1] * w[WT_WCM1 * WS];
    if (damp > OCO_LOG_OVF) damp = OCO_LOG_OVF;
    p.S = w[WT_SFW5 * WS] / (oco_exp_t(damp, M) - 1.0);
    return p;
}

OCO_HD static inline oco_pt oco_point(const double* c, const double* w, const oco_mtab* M)
{
    return oco_point_s<1>(c, w, M);
}

/* grey_parameters (source/continuous_opacity.f90:377-390) from rho, T. */
OCO_HD static inline oco_pt oco_grey_point(double rho, double T, double sigma_grey, const oco_mtab* M)
{
    oco_pt p;
    p.S = OCO_SIGMA_SB * (((T * T) * T) * T);
    double ag = (3.68e22 * oco_pow_t(T, -3.5, M)) * rho;
    p.alpha = (ag + sigma_grey) * rho;
    p.albedo = sigma_grey / (ag + sigma_grey);
    p.alpha1 = 1.0 / p.alpha;
    p.alpha2 = p.alpha * p.alpha;
    return p;
}

#endif
