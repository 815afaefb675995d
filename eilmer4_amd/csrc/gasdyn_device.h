// gasdyn_device.h -- device functions of the explicit gas-dynamic update for
// ideal gas: decode, MUSCL l2r2 reconstruction with van Albada limiting and
// extrema clipping, face EOS, AUSMDV and HLLC.
//
// Every function is templated on an arithmetic mode M:
//   ModeExact  FP64; IEEE division; the translation unit is compiled without
//              FMA contraction, and every expression keeps the reference's
//              operand order, so results are the reference arithmetic's
//              roundings (bit-identical to the CPU oracle).
//   ModeFast64 FP64; FMA contraction; a/b evaluated as a*rcp(b) with the
//              reciprocal refined by one Newton step and the quotient by one
//              residual correction (within ~1 ulp); reciprocals of repeated
//              denominators are shared; p/rho of a face state is Rgas T.
//   ModeMixed  FP32 reconstruction and flux (same division scheme in FP32);
//              decode, accumulation and update stay FP64 in the stage kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>

namespace e4 {

// Flux calculators (FluxCalculator order, eilmer/fvcore.d:184-194; hllc
// appended).  FLUX_SHOCK is the kernels' instantiation for the calculators
// that read the shock indicator: adaptive_hanel_ausmdv, and hanel as its
// always-Hanel case.
// FLUX_OTHER is the instantiation of the calculators without a shock
// indicator other than ausmdv and hllc: ausm_plus_up and roe, chosen at run
// time.  FLUX_SHOCK also serves efm and adaptive_efm_ausmdv (EFM as the
// near-shock calculator).
enum { FLUX_AUSMDV = 0, FLUX_HANEL = 1, FLUX_EFM = 2, FLUX_AUSM_PLUS_UP = 3, FLUX_ADAPTIVE_EFM_AUSMDV = 4,
       FLUX_ADAPTIVE_HANEL_AUSMDV = 5, FLUX_ROE = 8, FLUX_HLLC = 9, FLUX_SHOCK = 5, FLUX_OTHER = 3 };

struct ModeExact { using R = double; static constexpr bool kFast = false; };
struct ModeFast64 { using R = double; static constexpr bool kFast = true; };
struct ModeMixed { using R = float; static constexpr bool kFast = true; };

// Reciprocal seeds for the fast division: v_rcp_f64 is good to about 2^-23
// relative, one Newton step takes it to about 2^-46; v_rcp_f32 is good to
// about one ulp and is used as is.
__device__ __forceinline__ double rcp_seed(double b)
{
    double r = __builtin_amdgcn_rcp(b);
    double e = fma(-b, r, 1.0);
    return fma(e, r, r);
}

__device__ __forceinline__ float rcp_seed(float b) { return __builtin_amdgcn_rcpf(b); }

// a/b from a reciprocal seed rb of b: q = a rb, then one correction with the
// exact residual a - b q; the result is within about one ulp of a/b.
template <class R>
__device__ __forceinline__ R dv_seed(R a, R b, R rb)
{
    R q = a * rb;
    R res = fma(-b, q, a);
    return fma(res, rb, q);
}

template <class M>
__device__ __forceinline__ typename M::R dv(typename M::R a, typename M::R b)
{
    if constexpr (!M::kFast) return a / b;
    else return dv_seed(a, b, rcp_seed(b));
}

template <class R> __device__ __forceinline__ R mx(R a, R b) { return fmax(a, b); }
template <class R> __device__ __forceinline__ R mn(R a, R b) { return fmin(a, b); }
template <class R> __device__ __forceinline__ R ab(R a) { return fabs(a); }

// Cell state used by the stencil: density, velocity (global x,y,z or face
// frame n,t1,t2 depending on context), internal energy u and sound speed a.
// Pressure and temperature are not stored: for an ideal gas they are the
// functions T = u/Cv, p = rho R T of (rho, u) that update_thermo_from_rhou
// computes (gas/ideal_gas.d:98-106).
template <class R>
struct PrimT {
    R rho, v0, v1, v2, u, a;
};
using Prim = PrimT<double>;

// Weights set by l2r2_prepare (eilmer/onedinterp.d:118-133) for one direction.
template <class R>
struct L2R2WT {
    R aL0, aR0, lenL0, lenR0;
    R two_over_lenL0_plus_lenL1, two_over_lenR0_plus_lenL0, two_over_lenR1_plus_lenR0;
    R two_lenL0_plus_lenL1, two_lenR0_plus_lenR1;
};
using L2R2W = L2R2WT<double>;

// Ideal-gas constants (gas/ideal_gas.d:65-68); gammaR = gamma*Rgas as the
// sound-speed product gamma*Rgas*T is associated (ideal_gas.d:143).
// CpCv is gamma(Q) = Cp/Cv of the gas model (gas_model.d:162), with Cp as
// ideal_gas.d:68 forms it.
template <class R>
struct IdealGasT {
    R Rgas, gamma, Cv, Cvinv, gammaR, CpCv;
};
using IdealGas = IdealGasT<double>;

template <class R>
__device__ __forceinline__ L2R2WT<R> cast_w(const L2R2W& w)
{
    return L2R2WT<R>{R(w.aL0), R(w.aR0), R(w.lenL0), R(w.lenR0), R(w.two_over_lenL0_plus_lenL1),
                     R(w.two_over_lenR0_plus_lenL0), R(w.two_over_lenR1_plus_lenR0), R(w.two_lenL0_plus_lenL1),
                     R(w.two_lenR0_plus_lenR1)};
}

// The weights of l2r2_prepare on a uniform grid: with every cell length equal
// to len, the three 2/(len+len) are one number c, aL0 == aR0 == a,
// lenL0 == lenR0 == len and 2 lenL0 + lenL1 == 2 lenR0 + lenR1 == len3.
template <class R>
struct L2R2UT {
    R c, a, len, len3;
};
using L2R2U = L2R2UT<double>;

template <class R>
__device__ __forceinline__ L2R2UT<R> cast_wu(const L2R2U& w)
{
    return L2R2UT<R>{R(w.c), R(w.a), R(w.len), R(w.len3)};
}

template <class R>
__device__ __forceinline__ IdealGasT<R> cast_gas(const IdealGas& g)
{
    return IdealGasT<R>{R(g.Rgas), R(g.gamma), R(g.Cv), R(g.Cvinv), R(g.gammaR), R(g.CpCv)};
}

// limiters.d:30-38
template <class R>
__device__ __forceinline__ R clip_to_limits(R q, R A, R B)
{
    R lower_limit = (A <= B) ? A : B;
    R upper_limit = (A > B) ? A : B;
    R qclipped = (q > lower_limit) ? q : lower_limit;
    return (qclipped <= upper_limit) ? qclipped : upper_limit;
}

// clip_to_limits in the mode's form: the reference's comparisons for
// ModeExact; max/min (FP64) or the median of three (FP32) for the fast modes
// (the same value for ordered, non-NaN operands up to the sign of a zero).
template <class M>
__device__ __forceinline__ typename M::R clip_limits(typename M::R q, typename M::R A, typename M::R B)
{
    if constexpr (!M::kFast) return clip_to_limits(q, A, B);
    else if constexpr (std::is_same<typename M::R, float>::value) return __builtin_amdgcn_fmed3f(q, A, B);
    else return mn(mx(q, mn(A, B)), mx(A, B));
}

// onedinterp.d:136-162
template <class M>
__device__ __forceinline__ void interp_l2r2_scalar(const L2R2WT<typename M::R>& w, bool apply_limiter,
                                                   bool extrema_clipping, typename M::R qL1, typename M::R qL0,
                                                   typename M::R qR0, typename M::R qR1, typename M::R& qL,
                                                   typename M::R& qR)
{
    using R = typename M::R;
    const R eps = R(1.0e-12);   // onedinterp.d:22
    R delLminus = (qL0 - qL1) * w.two_over_lenL0_plus_lenL1;
    R del = (qR0 - qL0) * w.two_over_lenR0_plus_lenL0;
    R delRplus = (qR1 - qR0) * w.two_over_lenR1_plus_lenR0;
    R sL = R(1.0);
    R sR = R(1.0);
    if (apply_limiter) {
        sL = dv<M>(delLminus * del + ab(delLminus * del) + eps, delLminus * delLminus + del * del + eps);
        sR = dv<M>(del * delRplus + ab(del * delRplus) + eps, del * del + delRplus * delRplus + eps);
    }
    qL = qL0 + sL * w.aL0 * (del * w.two_lenL0_plus_lenL1 + delLminus * w.lenR0);
    qR = qR0 - sR * w.aR0 * (delRplus * w.lenL0 + del * w.two_lenR0_plus_lenR1);
    if (extrema_clipping) {
        qL = clip_to_limits(qL, qL0, qR0);
        qR = clip_to_limits(qR, qL0, qR0);
    }
}

// decode_conserved for an ideal gas (eilmer/fvcell.d:590-651, 680, 703 with
// gas/ideal_gas.d:98-106, 137-144), always FP64.  The reciprocal of rho and
// the square root follow the mode's division rule.
template <class M>
__device__ __forceinline__ Prim decode_ideal(double U0, double U1, double U2, double U3, double U4,
                                             const IdealGas& g)
{
    using MD = typename std::conditional<M::kFast, ModeFast64, ModeExact>::type;
    Prim P;
    double rho = U0;
    P.rho = rho;
    double dinv = dv<MD>(1.0, rho);
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

// A face-side state after reconstruction and the face EOS; RT = Rgas T, the
// ratio p/rho of the ideal gas, is what the fast modes use for p/rho; T is
// the face temperature u/Cv (used by efm).
template <class R>
struct FaceStateT {
    R rho, p, u, a, vn, vt1, vt2, RT, T;
};

// Flux in the face frame: mass, normal momentum, two tangential momenta, energy.
template <class R>
struct Flux5T {
    R mass, mn, mt1, mt2, E;
};

// ausmdv (eilmer/fluxcalc.d:376-529); eL = gmodel.internal_energy = Q.u.
template <class M>
__device__ __forceinline__ Flux5T<typename M::R> ausmdv(const FaceStateT<typename M::R>& L,
                                                        const FaceStateT<typename M::R>& Rs)
{
    using R = typename M::R;
    R rL = L.rho;
    R pL = L.p;
    R pLrL = M::kFast ? L.RT : dv<M>(pL, rL);
    R uL = L.vn;
    R vL = L.vt1;
    R wL = L.vt2;
    R eL = L.u;
    R aL = L.a;
    R keL = R(0.5) * (uL * uL + vL * vL + wL * wL);
    R HL = eL + pLrL + keL;
    R rR = Rs.rho;
    R pR = Rs.p;
    R pRrR = M::kFast ? Rs.RT : dv<M>(pR, rR);
    R uR = Rs.vn;
    R vR = Rs.vt1;
    R wR = Rs.vt2;
    R eR = Rs.u;
    R aR = Rs.a;
    R keR = R(0.5) * (uR * uR + vR * vR + wR * wR);
    R HR = eR + pRrR + keR;
    R alphaL = dv<M>(R(2.0) * pLrL, pLrL + pRrR);
    R alphaR = dv<M>(R(2.0) * pRrR, pLrL + pRrR);
    R am = mx(aL, aR);
    R ML, MR, r4am = R(0);
    if constexpr (M::kFast) {
        // one reciprocal seed serves uL/am, uR/am and x/(4 am) (4 is exact)
        R ram = rcp_seed(am);
        ML = dv_seed(uL, am, ram);
        MR = dv_seed(uR, am, ram);
        r4am = R(0.25) * ram;
    } else {
        ML = dv<M>(uL, am);
        MR = dv<M>(uR, am);
    }
    auto over4am = [&](R x) { return M::kFast ? dv_seed(x, R(4.0) * am, r4am) : dv<M>(x, R(4.0) * am); };
    R pLplus, uLplus;
    R duL = R(0.5) * (uL + ab(uL));
    if (ab(ML) <= R(1.0)) {
        pLplus = pL * (ML + R(1.0)) * (ML + R(1.0)) * (R(2.0) - ML) * R(0.25);
        uLplus = alphaL * (over4am((uL + am) * (uL + am)) - duL) + duL;
    } else {
        pLplus = dv<M>(pL * duL, uL);
        uLplus = duL;
    }
    R pRminus, uRminus;
    R duR = R(0.5) * (uR - ab(uR));
    if (ab(MR) <= R(1.0)) {
        pRminus = pR * (MR - R(1.0)) * (MR - R(1.0)) * (R(2.0) + MR) * R(0.25);
        uRminus = alphaR * (over4am(-(uR - am) * (uR - am)) - duR) + duR;
    } else {
        pRminus = dv<M>(pR * duR, uR);
        uRminus = duR;
    }
    R ru_half = uLplus * rL + uRminus * rR;
    R p_half = pLplus + pRminus;
    R dp = pL - pR;
    const R K_SWITCH = R(10.0);
    dp = dv<M>(K_SWITCH * ab(dp), mn(pL, pR));
    R s = R(0.5) * mn(R(1.0), dp);
    R ru2_AUSMV = uLplus * rL * uL + uRminus * rR * uR;
    R ru2_AUSMD = R(0.5) * (ru_half * (uL + uR) - ab(ru_half) * (uR - uL));
    R ru2_half = (R(0.5) + s) * ru2_AUSMV + (R(0.5) - s) * ru2_AUSMD;
    Flux5T<R> F;
    F.mass = ru_half;
    F.mn = ru2_half + p_half;
    if (ru_half >= R(0.0)) {
        F.mt1 = ru_half * vL;
        F.mt2 = ru_half * wL;
        F.E = ru_half * HL;
    } else {
        F.mt1 = ru_half * vR;
        F.mt2 = ru_half * wR;
        F.E = ru_half * HR;
    }
    const R C_EFIX = R(0.125);
    bool caseA = ((uL - aL) < R(0.0)) && ((uR - aR) > R(0.0));
    bool caseB = ((uL + aL) < R(0.0)) && ((uR + aR) > R(0.0));
    R d_ua = R(0.0);
    if (caseA && !caseB) d_ua = C_EFIX * ((uR - aR) - (uL - aL));
    if (caseB && !caseA) d_ua = C_EFIX * ((uR + aR) - (uL + aL));
    if (d_ua != R(0.0)) {
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
template <class M>
__device__ __forceinline__ Flux5T<typename M::R> hllc(const FaceStateT<typename M::R>& L,
                                                      const FaceStateT<typename M::R>& Rs)
{
    using R = typename M::R;
    R rL = L.rho, pL = L.p, uL = L.vn, vL = L.vt1, wL = L.vt2, aL = L.a;
    R EL = rL * (L.u + R(0.5) * (uL * uL + vL * vL + wL * wL));
    R rR = Rs.rho, pR = Rs.p, uR = Rs.vn, vR = Rs.vt1, wR = Rs.vt2, aR = Rs.a;
    R ER = rR * (Rs.u + R(0.5) * (uR * uR + vR * vR + wR * wR));
    R SL = mn(uL - aL, uR - aR);
    R SR = mx(uL + aL, uR + aR);
    R mL = rL * (SL - uL);
    R mR = rR * (SR - uR);
    R Sstar = dv<M>(pR - pL + mL * uL - mR * uR, mL - mR);
    Flux5T<R> F;
    if (R(0.0) <= SL) {
        R mf = rL * uL;
        F.mass = mf; F.mn = mf * uL + pL; F.mt1 = mf * vL; F.mt2 = mf * wL; F.E = uL * (EL + pL);
        return F;
    }
    if (SR <= R(0.0)) {
        R mf = rR * uR;
        F.mass = mf; F.mn = mf * uR + pR; F.mt1 = mf * vR; F.mt2 = mf * wR; F.E = uR * (ER + pR);
        return F;
    }
    const bool left = (R(0.0) <= Sstar);
    const R rK = left ? rL : rR, uK = left ? uL : uR, vK = left ? vL : vR, wK = left ? wL : wR;
    const R pK = left ? pL : pR, EK = left ? EL : ER, SK = left ? SL : SR;
    R mf = rK * uK;
    R FK0 = mf, FK1 = mf * uK + pK, FK2 = mf * vK, FK3 = mf * wK, FK4 = uK * (EK + pK);
    R coef = dv<M>(rK * (SK - uK), SK - Sstar);
    R Us0 = coef;
    R Us1 = coef * Sstar;
    R Us2 = coef * vK;
    R Us3 = coef * wK;
    R Us4 = coef * (dv<M>(EK, rK) + (Sstar - uK) * (Sstar + dv<M>(pK, rK * (SK - uK))));
    F.mass = FK0 + SK * (Us0 - rK);
    F.mn = FK1 + SK * (Us1 - rK * uK);
    F.mt1 = FK2 + SK * (Us2 - rK * vK);
    F.mt2 = FK3 + SK * (Us3 - rK * wK);
    F.E = FK4 + SK * (Us4 - EK);
    return F;
}

// fluxcalc.d:532-623: Hanel's flux vector splitting (van Leer's split
// velocities and pressures), in the reference's operation order.
template <class M>
__device__ __forceinline__ Flux5T<typename M::R> hanel(const FaceStateT<typename M::R>& L,
                                                       const FaceStateT<typename M::R>& Rs)
{
    using R = typename M::R;
    const R rL = L.rho, pL = L.p;
    const R pLrL = dv<M>(pL, rL);
    const R uL = L.vn, vL = L.vt1, wL = L.vt2;
    const R eL = L.u, aL = L.a;
    const R keL = R(0.5) * (uL * uL + vL * vL + wL * wL);
    const R HL = eL + pLrL + keL;
    const R rR = Rs.rho, pR = Rs.p;
    const R pRrR = dv<M>(pR, rR);
    const R uR = Rs.vn, vR = Rs.vt1, wR = Rs.vt2;
    const R eR = Rs.u, aR = Rs.a;
    const R keR = R(0.5) * (uR * uR + vR * vR + wR * wR);
    const R HR = eR + pRrR + keR;
    R pLplus, uLplus;
    if (ab(uL) <= aL) {
        uLplus = dv<M>(R(1.0), R(4.0) * aL) * (uL + aL) * (uL + aL);
        pLplus = pL * uLplus * (dv<M>(R(1.0), aL) * (R(2.0) - dv<M>(uL, aL)));
    } else {
        uLplus = R(0.5) * (uL + ab(uL));
        pLplus = pL * uLplus * dv<M>(R(1.0), uL);
    }
    R pRminus, uRminus;
    if (ab(uR) <= aR) {
        uRminus = dv<M>(R(-1.0), R(4.0) * aR) * (uR - aR) * (uR - aR);
        pRminus = pR * uRminus * (dv<M>(R(1.0), aR) * (R(-2.0) - dv<M>(uR, aR)));
    } else {
        uRminus = R(0.5) * (uR - ab(uR));
        pRminus = pR * uRminus * dv<M>(R(1.0), uR);
    }
    const R p_half = pLplus + pRminus;
    Flux5T<R> F;
    F.mass = uLplus * rL + uRminus * rR;
    F.mn = uLplus * rL * uL + uRminus * rR * uR + p_half;
    F.mt1 = uLplus * rL * vL + uRminus * rR * vR;
    F.mt2 = uLplus * rL * wL + uRminus * rR * wR;
    F.E = uLplus * rL * HL + uRminus * rR * HR;
    return F;
}

// Whether adaptive_hanel_ausmdv (fluxcalc.d:822-836) takes Hanel at a face:
// either side is a shock point and the tangential velocity jump, scaled by the
// mean sound speed, is within the shear tolerance.
template <class M>
__device__ __forceinline__ bool hanel_at_face(const FaceStateT<typename M::R>& L, const FaceStateT<typename M::R>& Rs,
                                              bool shock, typename M::R shear_tol)
{
    using R = typename M::R;
    const R sound_speed = R(0.5) * (L.a + Rs.a);
    const R shear_y = dv<M>(ab(L.vt1 - Rs.vt1), sound_speed);
    const R shear_z = dv<M>(ab(L.vt2 - Rs.vt2), sound_speed);
    const bool shear_is_small = mx(shear_y, shear_z) <= shear_tol;
    return shock && shear_is_small;
}

template <class R> __device__ __forceinline__ R sq_rt(R x) { return sqrt(x); }
template <> __device__ __forceinline__ float sq_rt<float>(float x) { return sqrtf(x); }
template <class R> __device__ __forceinline__ R ex_p(R x) { return exp(x); }
template <> __device__ __forceinline__ float ex_p<float>(float x) { return expf(x); }

// fluxcalc.d:792-820: exp(-x^2) and erf(x) by the reference's polynomial
// (Abramowitz and Stegun 7.1.26), clamped beyond |x| = 5.
template <class M>
__device__ __forceinline__ void exxef(typename M::R sn, typename M::R& exx, typename M::R& ef)
{
    using R = typename M::R;
    const R P = R(0.327591100), A1 = R(0.254829592), A2 = R(-0.284496736), A3 = R(1.421413741);
    const R A4 = R(-1.453152027), A5 = R(1.061405429);
    // a zero argument counts as +0: the polynomial's erf(0) is 1e-9, so
    // copysign makes the sign of a zero matter; the reference's velocities
    // reach the flux through frame rotations that give +0 where the device's
    // mirrored ghost states carry -0
    sn = sn + R(0.0);
    R ef1;
    if (ab(sn) > R(5.0)) {
        exx = R(0.138879e-10);
        ef1 = R(1.0);
    } else {
        const R snsq = sn * sn;
        exx = ex_p<R>(-snsq);
        const R y = dv<M>(R(1.0), R(1.0) + P * ab(sn));
        ef1 = R(1.0) - y * (A1 + y * (A2 + y * (A3 + y * (A4 + A5 * y)))) * exx;
    }
    ef = copysign(ef1, sn);
}

// fluxcalc.d:626-790 (efmflx): Macrossan and Pullin's equilibrium flux method
// with the gas "constants" of the two sides weighted by sqrt(rho); Cv of an
// ideal gas is the constant Cv, R of each side p/(rho T), and the right side's
// RT is p/rho where the left one's is R T, as the reference forms them.
template <class M>
__device__ __forceinline__ Flux5T<typename M::R> efm(const FaceStateT<typename M::R>& L,
                                                     const FaceStateT<typename M::R>& Rs,
                                                     const IdealGasT<typename M::R>& g)
{
    using R = typename M::R;
    const R dtwspi = R(0.282094792);
    const R rhoL = L.rho, presL = L.p, eL = L.u, tL = L.T;
    const R hL = eL + dv<M>(presL, rhoL);
    const R vnL = L.vn, vpL = L.vt1, vqL = L.vt2;
    const R rhoR = Rs.rho, presR = Rs.p, eR = Rs.u, tR = Rs.T;
    const R hR = eR + dv<M>(presR, rhoR);
    const R vnR = Rs.vn, vpR = Rs.vt1, vqR = Rs.vt2;
    const R cvL = g.Cv, cvR = g.Cv;
    const R RgasL = dv<M>(presL, rhoL * tL);
    const R RgasR = dv<M>(presR, rhoR * tR);
    const R rLsqrt = sq_rt<R>(rhoL), rRsqrt = sq_rt<R>(rhoR);
    const R alpha = dv<M>(rLsqrt, rLsqrt + rRsqrt);
    const R cv = alpha * cvL + (R(1.0) - alpha) * cvR;
    const R Rgas = alpha * RgasL + (R(1.0) - alpha) * RgasR;
    const R cp = cv + Rgas;
    const R gam = dv<M>(cp, cv);
    const R con = dv<M>(R(0.5) * (gam + R(1.0)), gam - R(1.0));
    const R rtL = Rgas * tL;
    const R cmpL = sq_rt<R>(R(2.0) * rtL);
    const R hvsqL = R(0.5) * (vnL * vnL + vpL * vpL + vqL * vqL);
    const R snL = dv<M>(vnL, R(1.0) * cmpL);
    R exL, efL;
    exxef<M>(snL, exL, efL);
    const R wL = R(0.5) * (R(1.0) + efL);
    const R dL = exL * dtwspi;
    const R rtR = dv<M>(presR, rhoR);
    const R cmpR = sq_rt<R>(R(2.0) * rtR);
    const R hvsqR = R(0.5) * (vnR * vnR + vpR * vpR + vqR * vqR);
    const R snR = dv<M>(vnR, R(1.0) * cmpR);
    R exR, efR;
    exxef<M>(snR, exR, efR);
    const R wR = R(0.5) * (R(1.0) - efR);
    const R dR = -exR * dtwspi;
    const R fmsL = (wL * rhoL * vnL) + (dL * cmpL * rhoL);
    const R fmsR = (wR * rhoR * vnR) + (dR * cmpR * rhoR);
    Flux5T<R> F;
    F.mass = fmsL + fmsR;
    F.mn = fmsL * vnL + fmsR * vnR + wL * presL + wR * presR;
    F.mt1 = fmsL * vpL + fmsR * vpR;
    F.mt2 = fmsL * vqL + fmsR * vqR;
    F.E = (wL * rhoL * vnL) * (hvsqL + hL) + (wR * rhoR * vnR) * (hvsqR + hR) +
          (dL * cmpL * rhoL) * (hvsqL + con * rtL) + (dR * cmpR * rhoR) * (hvsqR + con * rtR);
    return F;
}

// fluxcalc.d:860-1037: Liou's AUSM+-up (2006) with the mean sound speed at
// the face and the reference's M_inf floor on M0^2.  The split functions
// (M1, M2, M4, P5) are those of the reference's nested helpers.
template <class R> __device__ __forceinline__ R aup_M1p(R M) { return R(0.5) * (M + ab(M)); }
template <class R> __device__ __forceinline__ R aup_M1m(R M) { return R(0.5) * (M - ab(M)); }
template <class R> __device__ __forceinline__ R aup_M2p(R M) { return R(0.25) * (M + R(1.0)) * (M + R(1.0)); }
template <class R> __device__ __forceinline__ R aup_M2m(R M) { return R(-0.25) * (M - R(1.0)) * (M - R(1.0)); }

template <class M>
__device__ __forceinline__ Flux5T<typename M::R> ausm_plus_up(const FaceStateT<typename M::R>& L,
                                                              const FaceStateT<typename M::R>& Rs,
                                                              typename M::R M_inf)
{
    using R = typename M::R;
    const R rL = L.rho, pL = L.p, uL = L.vn, vL = L.vt1, wL = L.vt2, eL = L.u, aL = L.a;
    const R keL = R(0.5) * (uL * uL + vL * vL + wL * wL);
    const R HL = eL + dv<M>(pL, rL) + keL;
    const R rR = Rs.rho, pR = Rs.p, uR = Rs.vn, vR = Rs.vt1, wR = Rs.vt2, eR = Rs.u, aR = Rs.a;
    const R keR = R(0.5) * (uR * uR + vR * vR + wR * wR);
    const R HR = eR + dv<M>(pR, rR) + keR;
    const R a_half = R(0.5) * (aR + aL);
    const R ML = dv<M>(uL, a_half);
    const R MR = dv<M>(uR, a_half);
    const R MbarSq = dv<M>(uL * uL + uR * uR, R(2.0) * a_half * a_half);
    const R M0Sq = mn(R(1.0), mx(MbarSq, M_inf));
    const R sM0 = sq_rt<R>(M0Sq);
    const R fa = sM0 * (R(2.0) - sM0);
    const R alpha = R(0.1875) * (R(-4.0) + R(5.0) * fa * fa);
    const R beta = R(0.125);
    R M4plus_ML, P5plus_ML, M4minus_MR, P5minus_MR;
    if (ab(ML) >= R(1.0)) {
        M4plus_ML = aup_M1p(ML);
        P5plus_ML = dv<M>(R(1.0), ML) * aup_M1p(ML);
    } else {
        const R M2p = aup_M2p(ML), M2m = aup_M2m(ML);
        M4plus_ML = M2p * (R(1.0) - R(16.0) * beta * M2m);
        P5plus_ML = M2p * ((R(2.0) - ML) - R(16.0) * alpha * ML * M2m);
    }
    if (ab(MR) >= R(1.0)) {
        M4minus_MR = aup_M1m(MR);
        P5minus_MR = dv<M>(R(1.0), MR) * aup_M1m(MR);
    } else {
        const R M2p = aup_M2p(MR), M2m = aup_M2m(MR);
        M4minus_MR = M2m * (R(1.0) + R(16.0) * beta * M2p);
        P5minus_MR = M2m * ((R(-2.0) - MR) + R(16.0) * alpha * MR * M2p);
    }
    const R KP = R(0.25), KU = R(0.75), SIGMA = R(1.0);
    const R r_half = R(0.5) * (rL + rR);
    const R Mp = dv<M>(dv<M>(-KP, fa) * mx((R(1.0) - SIGMA * MbarSq), R(0.0)) * (pR - pL), r_half * a_half * a_half);
    const R Pu = -KU * P5plus_ML * P5minus_MR * (rL + rR) * fa * a_half * (uR - uL);
    const R M_half = M4plus_ML + M4minus_MR + Mp;
    R ru_half = a_half * M_half;
    if (M_half > R(0.0)) ru_half *= rL;
    else ru_half *= rR;
    const R p_half = P5plus_ML * pL + P5minus_MR * pR + Pu;
    Flux5T<R> F;
    F.mass = ru_half;
    if (ru_half >= R(0.0)) {
        F.mn = ru_half * uL + p_half; F.mt1 = ru_half * vL; F.mt2 = ru_half * wL; F.E = ru_half * HL;
    } else {
        F.mn = ru_half * uR + p_half; F.mt1 = ru_half * vR; F.mt2 = ru_half * wR; F.E = ru_half * HR;
    }
    return F;
}

// fluxcalc.d:1202-1371: Roe-Pike (Toro 2009, section 11.3) with the
// Kermani-Plett entropy fix on the acoustic eigenvalues; gamma of each side is
// the gas model's Cp/Cv (for an ideal gas one constant).  Each flux is the
// central part less the five wave terms in the reference's order (eigenvector
// entries 0 and 1 included, as the reference sums them).
template <class M>
__device__ __forceinline__ Flux5T<typename M::R> roe(const FaceStateT<typename M::R>& L,
                                                     const FaceStateT<typename M::R>& Rs, typename M::R gL,
                                                     typename M::R gR)
{
    using R = typename M::R;
    const R rL = L.rho, pL = L.p, uL = L.vn, vL = L.vt1, wL = L.vt2, eL = L.u, aL = L.a;
    const R pLrL = dv<M>(pL, rL);
    const R keL = R(0.5) * (uL * uL + vL * vL + wL * wL);
    const R HL = eL + pLrL + keL;
    const R rR = Rs.rho, pR = Rs.p, uR = Rs.vn, vR = Rs.vt1, wR = Rs.vt2, eR = Rs.u, aR = Rs.a;
    const R pRrR = dv<M>(pR, rR);
    const R keR = R(0.5) * (uR * uR + vR * vR + wR * wR);
    const R HR = eR + pRrR + keR;
    const R ghat = (gR + gL) / R(2.0);
    const R rhat = sq_rt<R>(rL * rR);
    const R sL = sq_rt<R>(rL), sR = sq_rt<R>(rR);
    const R den = sL + sR;
    const R uhat = dv<M>(sL * uL + sR * uR, den);
    const R vhat = dv<M>(sL * vL + sR * vR, den);
    const R what = dv<M>(sL * wL + sR * wR, den);
    const R Hhat = dv<M>(sL * HL + sR * HR, den);
    const R ke_hat = R(0.5) * (uhat * uhat + vhat * vhat + what * what);
    const R ahat2 = (ghat - R(1.0)) * (Hhat - ke_hat);
    const R ahat = sq_rt<R>(ahat2);
    R lam[5] = {uhat - ahat, uhat, uhat, uhat, uhat + ahat};
    const R lambdaL = uL - aL, lambdaR = uR - aR;
    const R eps = R(1.0e-12);
    const R EPS = R(4.0) * mx(R(0.0), (lambdaR - lambdaL));
    if (ab(lam[0]) < EPS) lam[0] = dv<M>(lam[0] * lam[0] + EPS * EPS, R(2.0) * EPS + eps);
    if (ab(lam[4]) < EPS) lam[4] = dv<M>(lam[4] * lam[4] + EPS * EPS, R(2.0) * EPS + eps);
    const R K[5][5] = {{R(1.0), uhat - ahat, vhat, what, Hhat - uhat * ahat},
                       {R(1.0), uhat, vhat, what, ke_hat},
                       {R(0.0), R(0.0), R(1.0), R(0.0), vhat},
                       {R(0.0), R(0.0), R(0.0), R(1.0), what},
                       {R(1.0), uhat + ahat, vhat, what, Hhat + uhat * ahat}};
    const R inv2a2 = dv<M>(R(1.0), R(2.0) * ahat2);
    R alpha[5];
    alpha[0] = inv2a2 * ((pR - pL) - rhat * ahat * (uR - uL));
    alpha[1] = (rR - rL) - dv<M>(pR - pL, ahat2);
    alpha[2] = rhat * (vR - vL);
    alpha[3] = rhat * (wR - wL);
    alpha[4] = inv2a2 * ((pR - pL) + rhat * ahat * (uR - uL));
    R c[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) c[i] = R(0.5) * alpha[i] * ab(lam[i]);
    R Fq[5];
    Fq[0] = R(0.5) * (rL * uL + rR * uR);
    Fq[1] = R(0.5) * ((pL + rL * uL * uL) + (pR + rR * uR * uR));
    Fq[2] = R(0.5) * (rL * uL * vL + rR * uR * vR);
    Fq[3] = R(0.5) * (rL * uL * wL + rR * uR * wR);
    Fq[4] = R(0.5) * ((rL * eL + rL * (uL * uL + vL * vL + wL * wL) / R(2.0) + pL) * uL +
                      (rR * eR + rR * (uR * uR + vR * vR + wR * wR) / R(2.0) + pR) * uR);
#pragma unroll
    for (int q = 0; q < 5; ++q) {
#pragma unroll
        for (int i = 0; i < 5; ++i) Fq[q] -= c[i] * K[i][q];
    }
    Flux5T<R> F;
    F.mass = Fq[0]; F.mn = Fq[1]; F.mt1 = Fq[2]; F.mt2 = Fq[3]; F.E = Fq[4];
    return F;
}

// Options of the flux calculators: the shear tolerance of the adaptive ones,
// whether every face takes the near-shock calculator (hanel, efm), whether
// that calculator is efm rather than hanel, whether FLUX_OTHER is roe rather
// than ausm_plus_up, and ausm_plus_up's M_inf.
template <class R>
struct FluxOpt {
    R shear_tol;
    bool always_shock_flux;
    bool efm;
    bool roe;
    R M_inf;
};

template <class M, int FLUX>
__device__ __forceinline__ Flux5T<typename M::R> flux_calc(const FaceStateT<typename M::R>& L,
                                                           const FaceStateT<typename M::R>& R,
                                                           const IdealGasT<typename M::R>& g,
                                                           bool shock = false,
                                                           FluxOpt<typename M::R> opt = {})
{
    if constexpr (FLUX == FLUX_HLLC) return hllc<M>(L, R);
    else if constexpr (FLUX == FLUX_SHOCK) {
        // adaptive_hanel_ausmdv / adaptive_efm_ausmdv (fluxcalc.d:802-836), or
        // hanel / efm at every face
        if (opt.always_shock_flux || hanel_at_face<M>(L, R, shock, opt.shear_tol))
            return opt.efm ? efm<M>(L, R, g) : hanel<M>(L, R);
        return ausmdv<M>(L, R);
    } else if constexpr (FLUX == FLUX_OTHER) {
        if (opt.roe) return roe<M>(L, R, g.CpCv, g.CpCv);
        return ausm_plus_up<M>(L, R, opt.M_inf);
    } else return ausmdv<M>(L, R);
}

// compute_flux_at_left_wall (left) / compute_flux_at_right_wall with the wall
// at rest: rho, p, a, vn the gas state beside the wall (vn along the face
// normal), gam its Cp/Cv.  Returns the wall pressure; the flux is that
// pressure in the normal momentum and zero elsewhere.  FP64 in every mode.
__device__ inline double wall_pressure(double rho, double p, double a, double vn, double gam, bool left)
{
    const double vstar = 0.0;
    double tmp;
    if (left) {
        const double Ubar = vn - 2.0 * a / (gam - 1.0);
        tmp = (vstar - Ubar) * (gam - 1.0) / (2.0 * sqrt(gam)) * sqrt(rho / pow(p, 1.0 / gam));
    } else {
        const double Ubar = vn + 2.0 * a / (gam - 1.0);
        tmp = (Ubar - vstar) * (gam - 1.0) / (2.0 * sqrt(gam)) * sqrt(rho / pow(p, 1.0 / gam));
    }
    const double ptiny = 0.1;   // FlowStateLimits.min_pressure (fvcore.d:26)
    double pstar = (tmp > 0.0) ? pow(tmp, 2.0 * gam / (gam - 1.0)) : ptiny;
    if (pstar > 1.1 * p) {
        // the wave is a shock: secant steps on the shock relation, at most 10
        auto f = [&](double ps) {
            const double xi = ps / p;
            const double M1sq = 1.0 + (gam + 1.0) / 2.0 / gam * (xi - 1.0);
            const double u1 = sqrt(M1sq) * a;
            const double u2 = u1 * ((gam - 1.0) * M1sq + 2.0) / ((gam + 1.0) * M1sq);
            return left ? vstar - u1 + u2 - vn : vstar + u1 - u2 - vn;
        };
        int count = 0;
        double incr;
        do {
            const double f0 = f(pstar);
            const double dp = 0.001 * pstar;
            const double f1 = f(pstar + dp);
            incr = -f0 * dp / (f1 - f0);
            pstar += incr;
            count += 1;
        } while (fabs(incr) / pstar > 0.01 && count < 10);
    }
    return fmin(pstar, p * 10.0);
}

} // namespace e4
