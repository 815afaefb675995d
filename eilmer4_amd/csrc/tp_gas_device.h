// tp_gas_device.h -- thermally perfect gas mixture on the device: NASA-9 (CEA)
// curves with linear blending, the bracketed Newton temperature solve with the
// reference's side effect on u, pressure and frozen sound speed.
//
// Follows gas/thermo/cea_thermo_curves.d:48-169, therm_perf_gas_mix_eos.d:63-153,
// nm/newton.d:40-107, perf_gas_mix_eos.d:43-94 and therm_perf_gas.d:177-196,
// 345-385.  Templated on the arithmetic mode M of gasdyn_device.h (division
// rule); these functions are always evaluated in FP64.  ModeExact evaluates
// every species' curve as the reference does; the fast modes, when all species
// share their breakpoints and blend ranges, evaluate the mixture as one curve
// whose coefficients are the Y_s R_s weighted sums of the species' (the energy
// and heat capacity are linear in the coefficients), with the reciprocal of T
// shared by all terms.
#pragma once
#include "gasdyn_device.h"

namespace e4 {

constexpr int TP_NSP_MAX = 8;
constexpr int TP_NSEG_MAX = 4;

struct CeaCurve {
    double R;
    int nseg;
    double Tb[TP_NSEG_MAX + 1];
    double Tbl[TP_NSEG_MAX];
    double rTbl[TP_NSEG_MAX];                     // 1 / Tbl (fast modes' blend weights)
    double c[TP_NSEG_MAX][9];
};

// ln x of a positive, normal, finite x for the fast modes: x = m 2^e with m
// in [1/sqrt 2, sqrt 2), ln m = 2 atanh z with z = (m - 1) / (m + 1), so
// |z| < 0.1716, and 2 atanh z = 2 z (1 + z^2/3 + z^4/5 + ... + z^18/19), the
// first dropped term below 2.2e-17 relative; e ln 2 added in two parts.
// Within a few ulp, in about a quarter of the instructions of the correctly
// rounded log (the energy curve evaluates one per temperature).
__device__ __forceinline__ double log_fast(double x)
{
    int e;
    double m = frexp(x, &e);
    if (m < 0.70710678118654752440) {
        m += m;
        e -= 1;
    }
    const double z = dv_seed(m - 1.0, m + 1.0, rcp_seed(m + 1.0));
    const double w = z * z;
    double p = 1.0 / 19.0;
    p = fma(p, w, 1.0 / 17.0);
    p = fma(p, w, 1.0 / 15.0);
    p = fma(p, w, 1.0 / 13.0);
    p = fma(p, w, 1.0 / 11.0);
    p = fma(p, w, 1.0 / 9.0);
    p = fma(p, w, 1.0 / 7.0);
    p = fma(p, w, 1.0 / 5.0);
    p = fma(p, w, 1.0 / 3.0);
    p = fma(p, w, 1.0);
    const double de = (double)e;
    // ln 2 = 6.93147180559945286227e-01 + 2.31904681384629955842e-17
    return fma(de, 6.93147180559945286227e-01, fma(de, 2.31904681384629955842e-17, 2.0 * z * p));
}

struct TPGas {
    int nsp;
    int shared;                                   // all species share nseg, Tb, Tbl
    double R[TP_NSP_MAX];
    CeaCurve curve[TP_NSP_MAX];
    double RC[TP_NSEG_MAX][8][TP_NSP_MAX];        // R_s a_{s,k} of segment q: RC[q][k][s]
};

template <int NSP>
struct TPState {
    double rho, p, T, u, a;
    double Y[NSP];
};

// cea_thermo_curves.d:136-169
__device__ __forceinline__ void cea_coefficients(const CeaCurve& cv, double T, double a[9])
{
    const int nb = cv.nseg + 1;
    int seg = -1;
    double wB = 0.0;
    if (T < (cv.Tb[1] - 0.5 * cv.Tbl[0])) {
        seg = 0;
    } else if (T > (cv.Tb[nb - 2] + 0.5 * cv.Tbl[nb - 3])) {
        seg = cv.nseg - 1;
    } else {
        for (int i = 1; i < nb - 1; ++i) {
            double T_blend_low = cv.Tb[i] - 0.5 * cv.Tbl[i - 1];
            double T_blend_high = cv.Tb[i] + 0.5 * cv.Tbl[i - 1];
            if (T >= T_blend_low && T <= T_blend_high) {
                wB = (1. / cv.Tbl[i - 1]) * (T - T_blend_low);
                double wA = 1.0 - wB;
#pragma unroll
                for (int j = 0; j < 9; ++j) a[j] = wA * cv.c[i - 1][j] + wB * cv.c[i][j];
                return;
            }
            if (T > T_blend_high && T < (cv.Tb[i + 1] - 0.5 * cv.Tbl[i])) { seg = i; break; }
        }
        if (seg < 0) seg = 0;   // unreachable for T inside the curve range
    }
#pragma unroll
    for (int j = 0; j < 9; ++j) a[j] = cv.c[seg][j];
}

// cea_thermo_curves.d:48-62
template <class M>
__device__ __forceinline__ double cea_Cp(const CeaCurve& cv, double T)
{
    using MD = typename std::conditional<M::kFast, ModeFast64, ModeExact>::type;
    if (T < cv.Tb[0]) T = cv.Tb[0];
    if (T > cv.Tb[cv.nseg]) T = cv.Tb[cv.nseg];
    double a[9];
    cea_coefficients(cv, T, a);
    double Cp_on_R = dv<MD>(a[0], T * T) + dv<MD>(a[1], T) + a[2] + a[3] * T;
    Cp_on_R += a[4] * T * T + a[5] * T * T * T + a[6] * T * T * T * T;
    return cv.R * Cp_on_R;
}

// cea_thermo_curves.d:69-90 inside the curve range
template <class M>
__device__ __forceinline__ double cea_h_inside(const CeaCurve& cv, double T)
{
    using MD = typename std::conditional<M::kFast, ModeFast64, ModeExact>::type;
    double a[9];
    cea_coefficients(cv, T, a);
    double h_on_RT = dv<MD>(-a[0], T * T) + dv<MD>(a[1], T) * log(T) + a[2] + dv<MD>(a[3] * T, 2.0);
    h_on_RT += dv<MD>(a[4] * T * T, 3.0) + dv<MD>(a[5] * T * T * T, 4.0) + dv<MD>(a[6] * T * T * T * T, 5.0) +
               dv<MD>(a[7], T);
    return cv.R * T * h_on_RT;
}

// cea_thermo_curves.d:63-91 (constant-Cp extension outside the range)
template <class M>
__device__ __forceinline__ double cea_h(const CeaCurve& cv, double T)
{
    if (T < cv.Tb[0]) {
        double T_low = cv.Tb[0];
        return cea_h_inside<M>(cv, T_low) - cea_Cp<M>(cv, T_low) * (T_low - T);
    }
    if (T > cv.Tb[cv.nseg]) {
        double T_high = cv.Tb[cv.nseg];
        return cea_h_inside<M>(cv, T_high) + cea_Cp<M>(cv, T_high) * (T - T_high);
    }
    return cea_h_inside<M>(cv, T);
}

template <int NSP>
__device__ __forceinline__ double tp_Rmix(const TPGas& g, const double* Y)
{
    double Rmix = 0.0;
#pragma unroll
    for (int s = 0; s < NSP; ++s) Rmix += Y[s] * g.R[s];
    return Rmix;
}

// Mixture coefficients M_k = sum_s Y_s R_s a_{s,k} at T (fast modes, shared
// breakpoints): the segment choice and the linear blend of cea_coefficients
// (cea_thermo_curves.d:136-169), applied to the mixture.
template <int NSP>
__device__ __forceinline__ void tp_mix_coeffs(const TPGas& g, const double* Y, double T, double Mk[8])
{
    const CeaCurve& cv = g.curve[0];
    const int nb = cv.nseg + 1;
    int s0 = -1, s1 = -1;
    double wB = 0.0;
    if (T < (cv.Tb[1] - 0.5 * cv.Tbl[0])) {
        s0 = 0;
    } else if (T > (cv.Tb[nb - 2] + 0.5 * cv.Tbl[nb - 3])) {
        s0 = cv.nseg - 1;
    } else {
        for (int i = 1; i < nb - 1; ++i) {
            const double lo = cv.Tb[i] - 0.5 * cv.Tbl[i - 1], hi = cv.Tb[i] + 0.5 * cv.Tbl[i - 1];
            if (T >= lo && T <= hi) { s0 = i - 1; s1 = i; wB = cv.rTbl[i - 1] * (T - lo); break; }
            if (T > hi && T < (cv.Tb[i + 1] - 0.5 * cv.Tbl[i])) { s0 = i; break; }
        }
        if (s0 < 0) s0 = 0;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        double m = 0.0;
#pragma unroll
        for (int s = 0; s < NSP; ++s) m += Y[s] * g.RC[s0][k][s];
        Mk[k] = m;
    }
    if (s1 >= 0) {
        const double wA = 1.0 - wB;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            double m = 0.0;
#pragma unroll
            for (int s = 0; s < NSP; ++s) m += Y[s] * g.RC[s1][k][s];
            Mk[k] = wA * Mk[k] + wB * m;
        }
    }
}

// Mixture enthalpy of the mixture coefficients M at T (r = 1/T, lnT = log T)
__device__ __forceinline__ double tp_mix_h(const double* M, double T, double r, double lnT)
{
    return -M[0] * r + M[1] * lnT + M[7] +
           T * (M[2] + T * (M[3] * 0.5 + T * (M[4] * (1.0 / 3.0) + T * (M[5] * 0.25 + T * (M[6] * 0.2)))));
}

// Mixture energy u = sum_s Y_s (h_s - R_s T) and Cv = sum_s Y_s (Cp_s - R_s)
// at T (fast modes, shared breakpoints); outside the curve range the
// constant-Cp extension of cea_thermo_curves.d:63-91.
template <int NSP>
__device__ __forceinline__ void tp_mix_u_cv(const TPGas& g, const double* Y, double Rmix, double T, double* u,
                                            double* cvmix)
{
    const CeaCurve& c0 = g.curve[0];
    const double Tlo = c0.Tb[0], Thi = c0.Tb[c0.nseg];
    const double Tc = fmin(fmax(T, Tlo), Thi);
    double Mk[8];
    tp_mix_coeffs<NSP>(g, Y, Tc, Mk);
    const double r = dv_seed(1.0, Tc, rcp_seed(Tc));
    const double lnT = log_fast(Tc);
    const double T2 = Tc * Tc, T3 = T2 * Tc, T4 = T2 * T2;
    double h = tp_mix_h(Mk, Tc, r, lnT);
    const double cp = Mk[0] * r * r + Mk[1] * r + Mk[2] + Mk[3] * Tc + Mk[4] * T2 + Mk[5] * T3 + Mk[6] * T4;
    if (T < Tlo) h -= cp * (Tlo - T);
    else if (T > Thi) h += cp * (T - Thi);
    *u = h - Rmix * T;
    *cvmix = cp - Rmix;
}

// The fast modes' curves: three segments (two blend ranges), every species on
// the same breakpoints; other mixtures run the reference-form kernels.
constexpr int TP_FAST_NSEG = 3;

// The mixture's coefficient sums of every segment, M[q][k] = sum_s Y_s R_s
// a_{s,q,k}: formed once per temperature solve (the composition is fixed
// during the solve), so an evaluation at any T only weights the three sets.
struct TPMix {
    double M[TP_FAST_NSEG][8];
};

template <int NSP>
__device__ __forceinline__ void tp_mix_all(const TPGas& g, const double* Y, TPMix& X)
{
#pragma unroll
    for (int q = 0; q < TP_FAST_NSEG; ++q)
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            double m = 0.0;
#pragma unroll
            for (int s = 0; s < NSP; ++s) m += Y[s] * g.RC[q][k][s];
            X.M[q][k] = m;
        }
}

// A value that is the same in every lane of the wave (here: a field of the
// gas model, whose address is uniform), pinned to a scalar register: without
// this the compiler may turn a per-lane choice between two such fields into
// a per-lane load, a memory round trip inside every curve evaluation.
__device__ __forceinline__ double uniform_f64(double x)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const unsigned int lo = (unsigned int)__builtin_amdgcn_readfirstlane((int)(unsigned int)b);
    const unsigned int hi = (unsigned int)__builtin_amdgcn_readfirstlane((int)(unsigned int)(b >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Weights of the three segments at Tc and their T derivatives: one segment,
// or the linear blend of two inside a blend range (cea_thermo_curves.d:136-169,
// the same comparisons).
__device__ __forceinline__ void tp_seg_weights(const CeaCurve& cv, double Tc, double w[3], double dw[3])
{
    const double lo1 = uniform_f64(cv.Tb[1] - 0.5 * cv.Tbl[0]), hi1 = uniform_f64(cv.Tb[1] + 0.5 * cv.Tbl[0]);
    const double lo2 = uniform_f64(cv.Tb[2] - 0.5 * cv.Tbl[1]), hi2 = uniform_f64(cv.Tb[2] + 0.5 * cv.Tbl[1]);
    const double r1 = uniform_f64(cv.rTbl[0]), r2 = uniform_f64(cv.rTbl[1]);
    const bool in1 = Tc >= lo1 && Tc <= hi1, in2 = Tc >= lo2 && Tc <= hi2;
    const double wb = in1 ? r1 * (Tc - lo1) : r2 * (Tc - lo2);
    const double rb = in1 ? r1 : r2;
    w[0] = (Tc < lo1) ? 1.0 : (in1 ? 1.0 - wb : 0.0);
    w[1] = in1 ? wb : (in2 ? 1.0 - wb : ((Tc > hi1 && Tc < lo2) ? 1.0 : 0.0));
    w[2] = in2 ? wb : ((Tc > hi2) ? 1.0 : 0.0);
    dw[0] = in1 ? -rb : 0.0;
    dw[1] = in1 ? rb : (in2 ? -rb : 0.0);
    dw[2] = in2 ? rb : 0.0;
}

// Which smooth piece of the fast modes' curve T lies on: below the curve
// range (0), segment 0 (1), the first blend range (2), segment 1 (3), the
// second blend range (4), segment 2 (5), above the range (6) -- the same
// comparisons as tp_seg_weights.  On one piece the energy is one smooth
// expression of T (a segment's polynomial and log terms, their linear blend,
// or the constant-Cp extension).
__device__ __forceinline__ int tp_curve_piece(const CeaCurve& cv, double T)
{
    if (T < uniform_f64(cv.Tb[0])) return 0;
    if (T > uniform_f64(cv.Tb[TP_FAST_NSEG])) return 6;
    const double lo1 = uniform_f64(cv.Tb[1] - 0.5 * cv.Tbl[0]), hi1 = uniform_f64(cv.Tb[1] + 0.5 * cv.Tbl[0]);
    const double lo2 = uniform_f64(cv.Tb[2] - 0.5 * cv.Tbl[1]), hi2 = uniform_f64(cv.Tb[2] + 0.5 * cv.Tbl[1]);
    if (T < lo1) return 1;
    if (T <= hi1) return 2;
    if (T < lo2) return 3;
    if (T <= hi2) return 4;
    return 5;
}

// u, du/dT and d2u/dT2 of the mixture at T from its segment sums.  SLOPE:
// du/dT is the exact derivative of the blended curve, which inside a blend
// adds (dw_B/dT)(h_B - h_A) to the blended Cv; without it du/dT is the blended
// Cv (the reference's dzdT, therm_perf_gas_mix_eos.d:116-123) and d2u is not
// formed.  Outside the curve range: the constant-Cp extension
// (cea_thermo_curves.d:63-91).
template <bool SLOPE>
__device__ __forceinline__ void tp_mix_eval(const TPMix& X, const CeaCurve& c0, double Rmix, double T, double* u,
                                            double* du, double* d2u)
{
    const double Tlo = uniform_f64(c0.Tb[0]), Thi = uniform_f64(c0.Tb[TP_FAST_NSEG]);
    const double Tc = fmin(fmax(T, Tlo), Thi);
    double w[3], dw[3];
    tp_seg_weights(c0, Tc, w, dw);
    double Mk[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) Mk[k] = w[0] * X.M[0][k] + w[1] * X.M[1][k] + w[2] * X.M[2][k];
    const double r = dv_seed(1.0, Tc, rcp_seed(Tc));
    const double lnT = log_fast(Tc);
    const double T2 = Tc * Tc, T3 = T2 * Tc, T4 = T2 * T2;
    double h = tp_mix_h(Mk, Tc, r, lnT);
    const double cp = Mk[0] * r * r + Mk[1] * r + Mk[2] + Mk[3] * Tc + Mk[4] * T2 + Mk[5] * T3 + Mk[6] * T4;
    double dh = cp, d2h = 0.0;
    if constexpr (SLOPE) {
        d2h = -2.0 * Mk[0] * r * r * r - Mk[1] * r * r + Mk[3] + 2.0 * Mk[4] * Tc + 3.0 * Mk[5] * T2 +
              4.0 * Mk[6] * T3;
        if (dw[0] != 0.0 || dw[1] != 0.0 || dw[2] != 0.0) {
            double dM[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) dM[k] = dw[0] * X.M[0][k] + dw[1] * X.M[1][k] + dw[2] * X.M[2][k];
            dh += tp_mix_h(dM, Tc, r, lnT);
            d2h += 2.0 * (dM[0] * r * r + dM[1] * r + dM[2] + dM[3] * Tc + dM[4] * T2 + dM[5] * T3 + dM[6] * T4);
        }
    }
    if (T < Tlo) { h -= cp * (Tlo - T); dh = cp; d2h = 0.0; }
    else if (T > Thi) { h += cp * (T - Thi); dh = cp; d2h = 0.0; }
    *u = h - Rmix * T;
    *du = dh - Rmix;
    *d2u = d2h;
}

// tp_mix_eval from the mass fractions instead of tp_mix_all's sums: each
// evaluation forms the species sums of the segments that some lane of the
// wave weights (a wave-uniform test), so that a solve keeps NSP mass
// fractions in registers instead of all the segments' sums.  Per lane the
// arithmetic is that of tp_mix_eval: a segment no lane weights would add
// w * M = 0.
template <bool SLOPE, int NSP>
__device__ __forceinline__ void tp_mix_eval_y(const TPGas& g, const double* Y, double Rmix, double T, double* u,
                                              double* du, double* d2u)
{
    const CeaCurve& c0 = g.curve[0];
    const double Tlo = uniform_f64(c0.Tb[0]), Thi = uniform_f64(c0.Tb[TP_FAST_NSEG]);
    const double Tc = fmin(fmax(T, Tlo), Thi);
    double w[3], dw[3];
    tp_seg_weights(c0, Tc, w, dw);
    double Mk[8], dM[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) { Mk[k] = 0.0; dM[k] = 0.0; }
#pragma unroll
    for (int q = 0; q < TP_FAST_NSEG; ++q) {
        if (__ballot(w[q] != 0.0 || (SLOPE && dw[q] != 0.0)) == 0) continue;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            double m = 0.0;
#pragma unroll
            for (int s = 0; s < NSP; ++s) m += Y[s] * g.RC[q][k][s];
            Mk[k] += w[q] * m;
            if constexpr (SLOPE) dM[k] += dw[q] * m;
        }
    }
    const double r = dv_seed(1.0, Tc, rcp_seed(Tc));
    const double lnT = log_fast(Tc);
    const double T2 = Tc * Tc, T3 = T2 * Tc, T4 = T2 * T2;
    double h = tp_mix_h(Mk, Tc, r, lnT);
    const double cp = Mk[0] * r * r + Mk[1] * r + Mk[2] + Mk[3] * Tc + Mk[4] * T2 + Mk[5] * T3 + Mk[6] * T4;
    double dh = cp, d2h = 0.0;
    if constexpr (SLOPE) {
        d2h = -2.0 * Mk[0] * r * r * r - Mk[1] * r * r + Mk[3] + 2.0 * Mk[4] * Tc + 3.0 * Mk[5] * T2 +
              4.0 * Mk[6] * T3;
        if (dw[0] != 0.0 || dw[1] != 0.0 || dw[2] != 0.0) {
            dh += tp_mix_h(dM, Tc, r, lnT);
            d2h += 2.0 * (dM[0] * r * r + dM[1] * r + dM[2] + dM[3] * Tc + dM[4] * T2 + dM[5] * T3 + dM[6] * T4);
        }
    }
    if (T < Tlo) { h -= cp * (Tlo - T); dh = cp; d2h = 0.0; }
    else if (T > Thi) { h += cp * (T - Thi); dh = cp; d2h = 0.0; }
    *u = h - Rmix * T;
    *du = dh - Rmix;
    *d2u = d2h;
}

// The composition as the fast solves see it through tp_mix_eval: TPMix, the
// segment sums formed once per solve, or TPMixY, the mass fractions with the
// sums formed at each evaluation (tp_mix_eval_y: fewer registers held through
// the solve, more arithmetic per evaluation).
template <int NSP>
struct TPMixY {
    const TPGas* g;
    double Y[NSP];
};

template <int NSP>
__device__ __forceinline__ TPMixY<NSP> tp_mix_y(const TPGas& g, const double* Y)
{
    TPMixY<NSP> X;
    X.g = &g;
#pragma unroll
    for (int s = 0; s < NSP; ++s) X.Y[s] = Y[s];
    return X;
}

template <bool SLOPE, int NSP>
__device__ __forceinline__ void tp_mix_eval(const TPMixY<NSP>& X, const CeaCurve&, double Rmix, double T, double* u,
                                            double* du, double* d2u)
{
    tp_mix_eval_y<SLOPE, NSP>(*X.g, X.Y, Rmix, T, u, du, d2u);
}

// therm_perf_gas_mix_eos.d:63-69 (mass_average, gas_model.d:330-338)
template <class M, int NSP>
__device__ __forceinline__ double tp_energy(const TPGas& g, const double* Y, double T)
{
    if constexpr (M::kFast) {
        double u, cv;
        tp_mix_u_cv<NSP>(g, Y, tp_Rmix<NSP>(g, Y), T, &u, &cv);
        return u;
    } else {
        double result = 0.0;
#pragma unroll
        for (int s = 0; s < NSP; ++s) {
            double e = cea_h<M>(g.curve[s], T) - g.R[s] * T;
            result += Y[s] * e;
        }
        return result;
    }
}

// -(sum Y_s (Cp_s - R_s)), the dzdT closure of therm_perf_gas_mix_eos.d:116-123
template <class M, int NSP>
__device__ __forceinline__ double tp_dzdT(const TPGas& g, const double* Y, double T)
{
    if constexpr (M::kFast) {
        double u, cv;
        tp_mix_u_cv<NSP>(g, Y, tp_Rmix<NSP>(g, Y), T, &u, &cv);
        return -1.0 * cv;
    } else {
        double result = 0.0;
#pragma unroll
        for (int s = 0; s < NSP; ++s) {
            double Cv = cea_Cp<M>(g.curve[s], T) - g.R[s];
            result += Y[s] * Cv;
        }
        return -1.0 * result;
    }
}

// Mixture energy at T and, when want_d, the dzdT closure's value
// -(sum Y_s Cv_s) (therm_perf_gas_mix_eos.d:104-123); the fast modes take
// both from one mixture evaluation.
template <class M, int NSP>
__device__ __forceinline__ void tp_u_dzdT(const TPGas& g, const double* Y, double T, bool want_d, double* u,
                                          double* dz)
{
    if constexpr (M::kFast) {
        double cv;
        tp_mix_u_cv<NSP>(g, Y, tp_Rmix<NSP>(g, Y), T, u, &cv);
        *dz = -1.0 * cv;
    } else {
        *u = tp_energy<M, NSP>(g, Y, T);
        *dz = want_d ? tp_dzdT<M, NSP>(g, Y, T) : 0.0;
    }
}

// nm/newton.d:40-107 with the zeroFn/dzdT closures; Q.T and Q.u receive the
// closures' side effects.  Returns 0 and the root, or -1 when it would throw.
// All energy evaluations (the two bracket ends, then the iterates from x0)
// go through one call site in one loop, which keeps the inlined code to one
// copy of the energy evaluation.
template <class M, int NSP, class MX>
__device__ __forceinline__ int tp_newton(const TPGas& g, const MX& X, double Rmix, TPState<NSP>& Q, double e_tgt,
                                         double x0, double xMin, double xMax, double tol, double* root, int* fevals)
{
    using MD = typename std::conditional<M::kFast, ModeFast64, ModeExact>::type;
    const int MAXIT = 30;
    double xL = xMin, xH = xMax, fL = 0.0;
    double rts = x0;
    double dxold = (xMax - xMin);
    double dx = dxold;
#pragma nounroll
    for (int n = 0;; ++n) {
        // evaluation n: 0 the lower bracket end, 1 the upper, 2+ the iterates
        const double T = (n == 0) ? xMin : (n == 1) ? xMax : rts;
        double u, df0;
        if constexpr (M::kFast) {
            double du, d2u;
            tp_mix_eval<false>(X, g.curve[0], Rmix, T, &u, &du, &d2u);
            df0 = -1.0 * du;
        } else {
            tp_u_dzdT<M, NSP>(g, Q.Y, T, n >= 2, &u, &df0);
        }
        Q.T = T;
        Q.u = u;
        (*fevals)++;
        const double f0 = e_tgt - u;
        if (n == 0) { fL = f0; continue; }
        if (n == 1) {
            const double fH = f0;
            if ((fL > 0.0 && fH > 0.0) || (fL < 0.0 && fH < 0.0)) return -1;
            if (fL == 0.0) { *root = xMin; return 0; }
            if (fH == 0.0) { *root = xMax; return 0; }
            if (fL < 0.0) { xL = xMin; xH = xMax; } else { xH = xMin; xL = xMax; }
            continue;
        }
        const int j = n - 2;   // Newton steps taken so far
        if (j > 0) { if (f0 < 0.0) xL = rts; else xH = rts; }
        if (j == MAXIT) return -1;
        if ((((rts - xH) * df0 - f0) * ((rts - xL) * df0 - f0) > 0.0) || (fabs(2.0 * f0) > fabs(dxold * df0))) {
            dxold = dx;
            dx = 0.5 * (xH - xL);
            rts = xL + dx;
            if (xL == rts) { *root = rts; return 0; }
        } else {
            dxold = dx;
            dx = dv<MD>(f0, df0);
            double tmp = rts;
            rts -= dx;
            if (tmp == rts) { *root = rts; return 0; }
        }
        if (fabs(dx) < tol) { *root = rts; return 0; }
    }
}

// therm_perf_gas_mix_eos.d:70-153: guess = Q.T, bracket [max(T-500, 20), +1000],
// then [20, 1e5]; tolerance 1e-6 K.  X and Rmix: the composition's segment
// sums and gas constant (fast modes; unused by ModeExact).
template <class M, int NSP, class MX>
__device__ __forceinline__ int tp_update_temperature_x(const TPGas& g, const MX& X, double Rmix, TPState<NSP>& Q,
                                                       int* fevals)
{
    const double Tsave = Q.T;
    const double e_tgt = Q.u;
    const double delT = 1000.0;
    const double T1 = fmax(Q.T - 0.5 * delT, 20.0);
#pragma nounroll
    for (int attempt = 0; attempt < 2; ++attempt) {
        const double lo = attempt ? 20.0 : T1, hi = attempt ? 100000.0 : T1 + delT;
        double root;
        if (tp_newton<M, NSP>(g, X, Rmix, Q, e_tgt, Tsave, lo, hi, 1.0e-6, &root, fevals) == 0) {
            Q.T = root;
            return 0;
        }
    }
    Q.T = Tsave;
    if constexpr (M::kFast) {
        double du, d2u;
        tp_mix_eval<false>(X, g.curve[0], Rmix, Tsave, &Q.u, &du, &d2u);
    } else {
        Q.u = tp_energy<M, NSP>(g, Q.Y, Tsave);
    }
    return -1;
}

template <class M, int NSP>
__device__ __forceinline__ int tp_update_temperature(const TPGas& g, TPState<NSP>& Q, int* fevals)
{
    if constexpr (M::kFast) {
        // the composition with its segment sums formed at each evaluation
        // (fewer registers than holding all segments' sums through the solve)
        const TPMixY<NSP> X = tp_mix_y<NSP>(g, Q.Y);
        return tp_update_temperature_x<M, NSP>(g, X, tp_Rmix<NSP>(g, Q.Y), Q, fevals);
    } else {
        TPMix X;   // unused by the reference form
        return tp_update_temperature_x<M, NSP>(g, X, 0.0, Q, fevals);
    }
}

// Frozen sound speed a = sqrt(Cp/Cv R T) (therm_perf_gas.d:345-385) from the
// composition's segment sums (fast modes).
template <class M, class MX>
__device__ __forceinline__ double tp_sound_speed_x(const TPGas& g, const MX& X, double Rmix, double T)
{
    using MD = typename std::conditional<M::kFast, ModeFast64, ModeExact>::type;
    double u, cv, d2u;
    tp_mix_eval<false>(X, g.curve[0], Rmix, T, &u, &cv, &d2u);
    return sqrt(dv<MD>(cv + Rmix, cv) * (Rmix * T));
}

// therm_perf_gas.d:187-191 with perf_gas_mix_eos.d:43-49
template <class M, int NSP>
__device__ __forceinline__ int tp_update_thermo_from_rhou(const TPGas& g, TPState<NSP>& Q, int* fevals)
{
    if (tp_update_temperature<M, NSP>(g, Q, fevals) != 0) return -1;
    double Rmix = tp_Rmix<NSP>(g, Q.Y);
    Q.p = Q.rho * Rmix * Q.T;
    return 0;
}

// therm_perf_gas.d:177-181 (low-T fallback path of decode)
template <class M, int NSP>
__device__ __forceinline__ void tp_update_thermo_from_rhoT(const TPGas& g, TPState<NSP>& Q)
{
    Q.u = tp_energy<M, NSP>(g, Q.Y, Q.T);
    double Rmix = tp_Rmix<NSP>(g, Q.Y);
    Q.p = Q.rho * Rmix * Q.T;
}

// therm_perf_gas.d:345-385: a = sqrt(Cp/Cv * R T), frozen
template <class M, int NSP>
__device__ __forceinline__ double tp_sound_speed(const TPGas& g, const TPState<NSP>& Q)
{
    using MD = typename std::conditional<M::kFast, ModeFast64, ModeExact>::type;
    if constexpr (M::kFast) {
        const double R = tp_Rmix<NSP>(g, Q.Y);
        double u, cv;
        tp_mix_u_cv<NSP>(g, Q.Y, R, Q.T, &u, &cv);
        return sqrt(dv<MD>(cv + R, cv) * (R * Q.T));
    } else {
        double Cp = 0.0, Cv = 0.0;
#pragma unroll
        for (int s = 0; s < NSP; ++s) Cp += Q.Y[s] * cea_Cp<M>(g.curve[s], Q.T);
#pragma unroll
        for (int s = 0; s < NSP; ++s) Cv += Q.Y[s] * (cea_Cp<M>(g.curve[s], Q.T) - g.R[s]);
        double R = tp_Rmix<NSP>(g, Q.Y);
        return sqrt(dv<MD>(Cp, Cv) * (R * Q.T));
    }
}

// gas/gas_model.d:296-328 (tolerance 0, assert_error_tolerance 0.1)
// (in FP64, or in FP32 for the mixed form's face reconstruction)
template <class M, int NSP, class R>
__device__ __forceinline__ int scale_mass_fractions(R* Y)
{
    using MD = typename std::conditional<std::is_same<R, float>::value, ModeMixed,
                                         typename std::conditional<M::kFast, ModeFast64, ModeExact>::type>::type;
    R sum = R(0.0);
#pragma unroll
    for (int s = 0; s < NSP; ++s) {
        Y[s] = Y[s] >= R(0.0) ? Y[s] : R(0.0);
        sum += Y[s];
    }
    if (ab(sum - R(1.0)) > R(0.1)) return -1;
    if (ab(sum - R(1.0)) > R(0.0)) {
#pragma unroll
        for (int s = 0; s < NSP; ++s) Y[s] = dv<MD>(Y[s], sum);
    }
    return 0;
}

} // namespace e4
