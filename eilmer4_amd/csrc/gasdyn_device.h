// gasdyn_device.h -- device functions of the explicit gas-dynamic update for
// ideal gas: decode, MUSCL l2r2 reconstruction with van Albada limiting and
// extrema clipping, face EOS, AUSMDV and HLLC.  Every expression keeps the
// reference's operand order and association so that, built without FMA
// contraction, the results are the reference arithmetic's roundings.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace e4 {

enum { FLUX_AUSMDV = 0, FLUX_HLLC = 9 };

// Cell state used by the stencil: density, velocity (global x,y,z or face
// frame n,t1,t2 depending on context), internal energy u and sound speed a.
// Pressure and temperature are not stored: for an ideal gas they are the
// functions T = u/Cv, p = rho R T of (rho, u) that update_thermo_from_rhou
// computes (gas/ideal_gas.d:98-106).
struct Prim {
    double rho, v0, v1, v2, u, a;
};

// Weights set by l2r2_prepare (eilmer/onedinterp.d:118-133) for one direction.
struct L2R2W {
    double aL0, aR0, lenL0, lenR0;
    double two_over_lenL0_plus_lenL1, two_over_lenR0_plus_lenL0, two_over_lenR1_plus_lenR0;
    double two_lenL0_plus_lenL1, two_lenR0_plus_lenR1;
};

// Ideal-gas constants (gas/ideal_gas.d:65-68); gammaR = gamma*Rgas as the
// sound-speed product gamma*Rgas*T is associated (ideal_gas.d:143).
struct IdealGas {
    double Rgas, gamma, Cv, Cvinv, gammaR;
};

// limiters.d:30-38
__device__ __forceinline__ double clip_to_limits(double q, double A, double B)
{
    double lower_limit = (A <= B) ? A : B;
    double upper_limit = (A > B) ? A : B;
    double qclipped = (q > lower_limit) ? q : lower_limit;
    return (qclipped <= upper_limit) ? qclipped : upper_limit;
}

// onedinterp.d:136-162
__device__ __forceinline__ void interp_l2r2_scalar(const L2R2W& w, bool apply_limiter, bool extrema_clipping,
                                                   double qL1, double qL0, double qR0, double qR1,
                                                   double& qL, double& qR)
{
    const double eps = 1.0e-12;   // onedinterp.d:22
    double delLminus = (qL0 - qL1) * w.two_over_lenL0_plus_lenL1;
    double del = (qR0 - qL0) * w.two_over_lenR0_plus_lenL0;
    double delRplus = (qR1 - qR0) * w.two_over_lenR1_plus_lenR0;
    double sL = 1.0;
    double sR = 1.0;
    if (apply_limiter) {
        sL = (delLminus * del + fabs(delLminus * del) + eps) / (delLminus * delLminus + del * del + eps);
        sR = (del * delRplus + fabs(del * delRplus) + eps) / (del * del + delRplus * delRplus + eps);
    }
    qL = qL0 + sL * w.aL0 * (del * w.two_lenL0_plus_lenL1 + delLminus * w.lenR0);
    qR = qR0 - sR * w.aR0 * (delRplus * w.lenL0 + del * w.two_lenR0_plus_lenR1);
    if (extrema_clipping) {
        qL = clip_to_limits(qL, qL0, qR0);
        qR = clip_to_limits(qR, qL0, qR0);
    }
}

// decode_conserved for an ideal gas (eilmer/fvcell.d:590-651, 680, 703 with
// gas/ideal_gas.d:98-106, 137-144).  No validity checks: the stage kernel
// checks every state it writes.
__device__ __forceinline__ Prim decode_ideal(double U0, double U1, double U2, double U3, double U4,
                                             const IdealGas& g)
{
    Prim P;
    double rho = U0;
    P.rho = rho;
    double dinv = 1.0 / rho;
    P.v0 = U1 * dinv;
    P.v1 = U2 * dinv;
    P.v2 = U3 * dinv;
    double u = U4 * dinv;
    double ke = 0.5 * (P.v0 * P.v0 + P.v1 * P.v1 + P.v2 * P.v2);
    u -= ke;
    P.u = u;
    double T = u * g.Cvinv;
    P.a = sqrt(g.gammaR * T);
    return P;
}

// A face-side state after reconstruction and the face EOS.
struct FaceState {
    double rho, p, u, a, vn, vt1, vt2;
};

// interp_l2r2 (onedinterp.d:435-585) for an ideal gas with thermo_interpolator
// rhou, the EOS mixins of onedinterp.d:1039-1068 (fall back to the cell state
// when update_thermo_from_rhou rejects the face values) and the copy of the
// cell-centre sound speed (sfluidblock.d:2199-2208).  Inputs are in the face
// frame.  With order 1 the face states are the adjacent cells'.
__device__ __forceinline__ void reconstruct_ideal(const Prim& L1, const Prim& L0, const Prim& R0, const Prim& R1,
                                                  const L2R2W& w, const IdealGas& g, bool high_order,
                                                  bool apply_limiter, bool extrema_clipping,
                                                  FaceState& Lf, FaceState& Rf)
{
    Lf.a = L0.a;
    Rf.a = R0.a;
    bool okL = false, okR = false;
    if (high_order) {
        interp_l2r2_scalar(w, apply_limiter, extrema_clipping, L1.v0, L0.v0, R0.v0, R1.v0, Lf.vn, Rf.vn);
        interp_l2r2_scalar(w, apply_limiter, extrema_clipping, L1.v1, L0.v1, R0.v1, R1.v1, Lf.vt1, Rf.vt1);
        interp_l2r2_scalar(w, apply_limiter, extrema_clipping, L1.v2, L0.v2, R0.v2, R1.v2, Lf.vt2, Rf.vt2);
        interp_l2r2_scalar(w, apply_limiter, extrema_clipping, L1.rho, L0.rho, R0.rho, R1.rho, Lf.rho, Rf.rho);
        interp_l2r2_scalar(w, apply_limiter, extrema_clipping, L1.u, L0.u, R0.u, R1.u, Lf.u, Rf.u);
        okL = !(Lf.u <= 0.0 || Lf.rho <= 0.0);
        okR = !(Rf.u <= 0.0 || Rf.rho <= 0.0);
    }
    if (!okL) { Lf.rho = L0.rho; Lf.u = L0.u; Lf.vn = L0.v0; Lf.vt1 = L0.v1; Lf.vt2 = L0.v2; }
    if (!okR) { Rf.rho = R0.rho; Rf.u = R0.u; Rf.vn = R0.v0; Rf.vt1 = R0.v1; Rf.vt2 = R0.v2; }
    double TL = Lf.u * g.Cvinv;
    Lf.p = Lf.rho * g.Rgas * TL;
    double TR = Rf.u * g.Cvinv;
    Rf.p = Rf.rho * g.Rgas * TR;
}

// Flux in the face frame: mass, normal momentum, two tangential momenta, energy.
struct Flux5 {
    double mass, mn, mt1, mt2, E;
};

// ausmdv (eilmer/fluxcalc.d:376-529); eL = gmodel.internal_energy = Q.u.
__device__ __forceinline__ Flux5 ausmdv(const FaceState& L, const FaceState& R)
{
    double rL = L.rho;
    double pL = L.p;
    double pLrL = pL / rL;
    double uL = L.vn;
    double vL = L.vt1;
    double wL = L.vt2;
    double eL = L.u;
    double aL = L.a;
    double keL = 0.5 * (uL * uL + vL * vL + wL * wL);
    double HL = eL + pLrL + keL;
    double rR = R.rho;
    double pR = R.p;
    double pRrR = pR / rR;
    double uR = R.vn;
    double vR = R.vt1;
    double wR = R.vt2;
    double eR = R.u;
    double aR = R.a;
    double keR = 0.5 * (uR * uR + vR * vR + wR * wR);
    double HR = eR + pRrR + keR;
    double alphaL = 2.0 * pLrL / (pLrL + pRrR);
    double alphaR = 2.0 * pRrR / (pLrL + pRrR);
    double am = fmax(aL, aR);
    double ML = uL / am;
    double MR = uR / am;
    double pLplus, uLplus;
    double duL = 0.5 * (uL + fabs(uL));
    if (fabs(ML) <= 1.0) {
        pLplus = pL * (ML + 1.0) * (ML + 1.0) * (2.0 - ML) * 0.25;
        uLplus = alphaL * ((uL + am) * (uL + am) / (4.0 * am) - duL) + duL;
    } else {
        pLplus = pL * duL / uL;
        uLplus = duL;
    }
    double pRminus, uRminus;
    double duR = 0.5 * (uR - fabs(uR));
    if (fabs(MR) <= 1.0) {
        pRminus = pR * (MR - 1.0) * (MR - 1.0) * (2.0 + MR) * 0.25;
        uRminus = alphaR * (-(uR - am) * (uR - am) / (4.0 * am) - duR) + duR;
    } else {
        pRminus = pR * duR / uR;
        uRminus = duR;
    }
    double ru_half = uLplus * rL + uRminus * rR;
    double p_half = pLplus + pRminus;
    double dp = pL - pR;
    const double K_SWITCH = 10.0;
    dp = K_SWITCH * fabs(dp) / fmin(pL, pR);
    double s = 0.5 * fmin(1.0, dp);
    double ru2_AUSMV = uLplus * rL * uL + uRminus * rR * uR;
    double ru2_AUSMD = 0.5 * (ru_half * (uL + uR) - fabs(ru_half) * (uR - uL));
    double ru2_half = (0.5 + s) * ru2_AUSMV + (0.5 - s) * ru2_AUSMD;
    Flux5 F;
    F.mass = ru_half;
    F.mn = ru2_half + p_half;
    if (ru_half >= 0.0) {
        F.mt1 = ru_half * vL;
        F.mt2 = ru_half * wL;
        F.E = ru_half * HL;
    } else {
        F.mt1 = ru_half * vR;
        F.mt2 = ru_half * wR;
        F.E = ru_half * HR;
    }
    const double C_EFIX = 0.125;
    bool caseA = ((uL - aL) < 0.0) && ((uR - aR) > 0.0);
    bool caseB = ((uL + aL) < 0.0) && ((uR + aR) > 0.0);
    double d_ua = 0.0;
    if (caseA && !caseB) d_ua = C_EFIX * ((uR - aR) - (uL - aL));
    if (caseB && !caseA) d_ua = C_EFIX * ((uR + aR) - (uL + aL));
    if (d_ua != 0.0) {
        F.mass -= d_ua * (rR - rL);
        F.mn -= d_ua * (rR * uR - rL * uL);
        F.mt1 -= d_ua * (rR * vR - rL * vL);
        F.mt2 -= d_ua * (rR * wR - rL * wL);
        F.E -= d_ua * (rR * HR - rL * HL);
    }
    return F;
}

// HLLC (Toro 2009, section 10.4: star states eq. 10.39, contact speed
// eq. 10.37, Davis wave-speed bounds eq. 10.48).  The reference snapshot has
// no HLLC; this is the same formula set as the CPU oracle's.
__device__ __forceinline__ Flux5 hllc(const FaceState& L, const FaceState& R)
{
    double rL = L.rho, pL = L.p, uL = L.vn, vL = L.vt1, wL = L.vt2, aL = L.a;
    double EL = rL * (L.u + 0.5 * (uL * uL + vL * vL + wL * wL));
    double rR = R.rho, pR = R.p, uR = R.vn, vR = R.vt1, wR = R.vt2, aR = R.a;
    double ER = rR * (R.u + 0.5 * (uR * uR + vR * vR + wR * wR));
    double SL = fmin(uL - aL, uR - aR);
    double SR = fmax(uL + aL, uR + aR);
    double mL = rL * (SL - uL);
    double mR = rR * (SR - uR);
    double Sstar = (pR - pL + mL * uL - mR * uR) / (mL - mR);
    Flux5 F;
    if (0.0 <= SL) {
        double mf = rL * uL;
        F.mass = mf; F.mn = mf * uL + pL; F.mt1 = mf * vL; F.mt2 = mf * wL; F.E = uL * (EL + pL);
        return F;
    }
    if (SR <= 0.0) {
        double mf = rR * uR;
        F.mass = mf; F.mn = mf * uR + pR; F.mt1 = mf * vR; F.mt2 = mf * wR; F.E = uR * (ER + pR);
        return F;
    }
    const bool left = (0.0 <= Sstar);
    const double rK = left ? rL : rR, uK = left ? uL : uR, vK = left ? vL : vR, wK = left ? wL : wR;
    const double pK = left ? pL : pR, EK = left ? EL : ER, SK = left ? SL : SR;
    double mf = rK * uK;
    double FK0 = mf, FK1 = mf * uK + pK, FK2 = mf * vK, FK3 = mf * wK, FK4 = uK * (EK + pK);
    double coef = rK * (SK - uK) / (SK - Sstar);
    double Us0 = coef;
    double Us1 = coef * Sstar;
    double Us2 = coef * vK;
    double Us3 = coef * wK;
    double Us4 = coef * (EK / rK + (Sstar - uK) * (Sstar + pK / (rK * (SK - uK))));
    F.mass = FK0 + SK * (Us0 - rK);
    F.mn = FK1 + SK * (Us1 - rK * uK);
    F.mt1 = FK2 + SK * (Us2 - rK * vK);
    F.mt2 = FK3 + SK * (Us3 - rK * wK);
    F.E = FK4 + SK * (Us4 - EK);
    return F;
}

template <int FLUX>
__device__ __forceinline__ Flux5 flux_calc(const FaceState& L, const FaceState& R)
{
    if constexpr (FLUX == FLUX_HLLC) return hllc(L, R);
    else return ausmdv(L, R);
}

// Reconstruct and evaluate the flux at one face; stencil states in the face frame.
template <int FLUX>
__device__ __forceinline__ Flux5 face_flux_ideal(const Prim& L1, const Prim& L0, const Prim& R0, const Prim& R1,
                                                 const L2R2W& w, const IdealGas& g, bool high_order,
                                                 bool apply_limiter, bool extrema_clipping)
{
    FaceState Lf, Rf;
    reconstruct_ideal(L1, L0, R0, R1, w, g, high_order, apply_limiter, extrema_clipping, Lf, Rf);
    return flux_calc<FLUX>(Lf, Rf);
}

} // namespace e4
