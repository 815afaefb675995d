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
enum { FLUX_AUSMDV = 0, FLUX_HANEL = 1, FLUX_ADAPTIVE_HANEL_AUSMDV = 5, FLUX_HLLC = 9, FLUX_SHOCK = 5 };

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
template <class R>
struct IdealGasT {
    R Rgas, gamma, Cv, Cvinv, gammaR;
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
    return IdealGasT<R>{R(g.Rgas), R(g.gamma), R(g.Cv), R(g.Cvinv), R(g.gammaR)};
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
// ratio p/rho of the ideal gas, is what the fast modes use for p/rho.
template <class R>
struct FaceStateT {
    R rho, p, u, a, vn, vt1, vt2, RT;
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

// Options of the shock-indicator calculators: the shear tolerance, and whether
// every face takes Hanel (the hanel calculator itself).
template <class R>
struct ShockOpt {
    R shear_tol;
    bool always_hanel;
};

template <class M, int FLUX>
__device__ __forceinline__ Flux5T<typename M::R> flux_calc(const FaceStateT<typename M::R>& L,
                                                           const FaceStateT<typename M::R>& R,
                                                           bool shock = false,
                                                           ShockOpt<typename M::R> opt = {})
{
    if constexpr (FLUX == FLUX_HLLC) return hllc<M>(L, R);
    else if constexpr (FLUX == FLUX_SHOCK) {
        if (opt.always_hanel || hanel_at_face<M>(L, R, shock, opt.shear_tol)) return hanel<M>(L, R);
        return ausmdv<M>(L, R);
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
