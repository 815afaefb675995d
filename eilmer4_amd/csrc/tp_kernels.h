// tp_kernels.h -- the explicit update for a thermally perfect gas mixture.
//
// The face EOS is a Newton solve on each side of every face (the rhou thermo
// interpolator, onedinterp.d:534-551 with the mixin of 1039-1068), so this path
// is compute-bound; it keeps a per-cell primitive array (rho, vel, u, T, p, a,
// Y[NSP]; SoA, ping-ponged between stages) because u and T are not functions
// of U alone here: T is the previous stage's Newton result (the next solve's
// guess) and u is the energy of the last Newton iterate
// (therm_perf_gas_mix_eos.d:104-114).
//
//   tp_face_kernel<M,NSP,FLUX,DIR>  one thread per face of direction DIR:
//       l2r2 reconstruction of velocity (face frame), mass fractions (then
//       scale_mass_fractions), rho and u; face Newton; flux; global frame.
//   tp_update_kernel<M,NSP,GUESS>   one thread per cell: time_derivatives,
//       stage update, decode_conserved with Newton (guess: the cell's T),
//       low-T fallback, check_data; writes U and the primitive state.
#pragma once
#include "stage_args.h"
#include "tp_gas_device.h"

namespace e4 {

// Modes of the thermally perfect path: ModeExact and ModeFast64 as for the
// ideal gas; ModeMixed keeps the FP64 arithmetic of ModeFast64 and replaces
// the bracketed Newton solves by the polish of learned_eos.h (Newton from the
// cell's temperature, accepted by its energy residual, else the reference's
// bracketed solve): the face EOS starts from the face's cell, the decode from
// the stored temperature (or the learned guess when one is installed).
template <class M> struct TPArith { using type = M; static constexpr bool polish = false; };
template <> struct TPArith<ModeMixed> { using type = ModeFast64; static constexpr bool polish = true; };
// The arithmetic of the face kernel's reconstruction and flux: that of the
// mode, except that the mixed form reconstructs and evaluates the flux in FP32
// (the face temperature solve stays in FP64, TPArith).
template <class M> struct TPFlux { using type = typename TPArith<M>::type; };
template <> struct TPFlux<ModeMixed> { using type = ModeMixed; };

// Element type of a mixture block's face-flux arrays: the mixed form computes
// its fluxes in FP32 and keeps them as floats (packed in the first half of each
// array), which halves the flux traffic between the face and update kernels.
template <class M>
using TPFluxStore = typename std::conditional<std::is_same<M, ModeMixed>::value, float, double>::type;

// update_thermo_from_rhou of a face state (therm_perf_gas.d:187-191); with
// polish, Newton from Q.T first (*fallbacks counts the solves it hands on).
template <class M, int NSP, bool POLISH>
__device__ __forceinline__ int tp_face_eos(const TPGas& g, TPState<NSP>& Q, int* fevals, int* fallbacks)
{
    if constexpr (M::kFast) {
        // the composition for the polish and the bracketed solve, its
        // segment sums formed at each evaluation: fewer registers in the face
        // kernels (three waves per SIMD)
        const TPMixY<NSP> X = tp_mix_y<NSP>(g, Q.Y);
        const double Rmix = tp_Rmix<NSP>(g, Q.Y);
        int rc = -1;
        if constexpr (POLISH) {
            rc = tp_polish_x<M, NSP>(g, X, Rmix, Q, Q.u, Q.T, fevals);
            if (rc != 0) (*fallbacks)++;
        }
        if (rc != 0) rc = tp_update_temperature_x<M, NSP>(g, X, Rmix, Q, fevals);
        if (rc == 0) Q.p = Q.rho * Rmix * Q.T;
        return rc;
    } else {
        if constexpr (POLISH) {
            if (tp_polish<M, NSP>(g, Q, Q.u, Q.T, fevals) == 0) {
                Q.p = Q.rho * tp_Rmix<NSP>(g, Q.Y) * Q.T;
                return 0;
            }
            (*fallbacks)++;
        }
        return tp_update_thermo_from_rhou<M, NSP>(g, Q, fevals);
    }
}

// Mixed form, the temperatures of the two states of a face: Halley steps on
// the mixture curve from each side's cell temperature, at most three
// evaluations per side.  A step s with |s| <= 1e-3 T that stays on the
// smooth piece of the curve it started from ends a side's solve: Halley's
// cubic convergence then leaves the new temperature within about
// |s|^3 / T^2 <= 1e-9 T of the root, far inside the FP32 rounding of the
// flux the face state feeds; a step that lands on another piece is followed
// by one from there.  The two sides' evaluations sit side by side in one
// loop, so each one's chain of dependent FP64 operations fills the other's
// latency.  okL/okR tell which sides settled (TL, TR then hold their
// temperatures); a face with a side that did not goes to the retry kernel,
// whose solve is the polish and bracketed Newton of tp_face_eos.
template <int NSP>
__device__ __forceinline__ void tp_face_halley_pair(const TPGas& g, const double* YL, const double* YR, double RL,
                                                    double RR, double eL, double eR, double& TL, double& TR,
                                                    bool& okL, bool& okR, int* fevals)
{
    const CeaCurve& c0 = g.curve[0];
    bool liveL = TL >= 20.0 && TL <= 100000.0, liveR = TR >= 20.0 && TR <= 100000.0;
    okL = false;
    okR = false;
    double tL = TL, tR = TR;
    auto step = [&](double t, double u, double du, double d2u, double e, double& out, bool& live, bool& ok) {
        if (!live) return t;
        (*fevals)++;
        const double f = u - e;
        const double s = dv<ModeFast64>(2.0 * f * du, 2.0 * du * du - f * d2u);
        const double tn = t - s;
        if (tp_curve_piece(c0, tn) == tp_curve_piece(c0, t) && fabs(s) <= 1.0e-3 * tn && tn >= 20.0 &&
            tn <= 100000.0) {
            out = tn;
            ok = true;
            live = false;
        }
        return tn;
    };
#pragma nounroll
    for (int it = 0; it < 3; ++it) {
        if (!(liveL || liveR)) break;
        double uL, duL, d2uL, uR, duR, d2uR;
        tp_mix_eval_y<true, NSP>(g, YL, RL, tL, &uL, &duL, &d2uL);
        tp_mix_eval_y<true, NSP>(g, YR, RR, tR, &uR, &duR, &d2uR);
        tL = step(tL, uL, duL, d2uL, eL, TL, liveL, okL);
        tR = step(tR, uR, duR, d2uR, eR, TR, liveR, okR);
    }
}

// A face-side (or cell) state of a mixture in the face frame, in the
// arithmetic of the flux path (double, or float in the mixed form).
template <class R, int NSP>
struct TPCellT {
    R rho, p, T, u, a;
    R v[3];
    R Y[NSP];
};
template <int NSP> using TPCell = TPCellT<double, NSP>;

// Where the values of one stencil position come from: field f (P_* order)
// is base[f * stride] -- the primitive array of this block or of a connected
// block (stride qs), or a fixed inflow state (stride 1); a slip wall negates
// the velocity component `flip` (-1: none).  s points at the position's
// shock indicator (this stage's for a cell of the block, the previous
// detection's for a cell a ghost was copied from; null for an inflow ghost,
// whose indicator is 0).  The same mapping as the ideal-gas stage kernel's
// fetch_ghost.
struct TPRef {
    const double* base;
    long stride;
    int flip;
    const unsigned char* s;
};

__device__ __forceinline__ TPRef tp_ref(const TPArgs& A, int i, int j, int k)
{
    if (i >= 0 && i < A.ni && j >= 0 && j < A.nj && k >= 0 && k < A.nk) {
        const long c = ((long)k * A.nj + j) * A.ni + i;
        return TPRef{A.Pin + c, A.qs, -1, A.S + c};
    }
    int dir, idx, n;
    if (i < 0 || i >= A.ni) { dir = 0; idx = i; n = A.ni; }
    else if (j < 0 || j >= A.nj) { dir = 1; idx = j; n = A.nj; }
    else { dir = 2; idx = k; n = A.nk; }
    const bool plus = idx >= n;
    const int layer = plus ? idx - n + 1 : -idx;
    const int face = (dir == 0) ? (plus ? FACE_EAST : FACE_WEST)
                   : (dir == 1) ? (plus ? FACE_NORTH : FACE_SOUTH)
                                : (plus ? FACE_TOP : FACE_BOTTOM);
    const BCTable* bc = A.bc;
    const int type = bc->type[face];
    if (type == BC_INFLOW) return TPRef{bc->fs_tp[face], 1, -1, nullptr};
    if (type == BC_CONNECTED) {
        // full_face_copy.d:1277-1357: the mapped interior cell of the other block
        const Connection nb = bc->nbr[face];
        const int fa = (face == FACE_NORTH || face == FACE_SOUTH || face >= FACE_TOP) ? i : j;
        const int fb = (face >= FACE_TOP) ? j : k;
        int sc[3];
        connection_source(nb.dims, face, nb.face, nb.orient, nb.n, layer, fa, fb, sc);
        const long c = ((long)sc[2] * nb.n[1] + sc[1]) * nb.n[0] + sc[0];
        return TPRef{A.nbrP[face] + c, nb.qs, -1, A.nbrS[face] + c};
    }
    if (type == BC_EXTRAPOLATE_1 || type == BC_INFLOW_PROFILE) {
        // the face's stored ghost cells: x_order-1 ones, filled at the start of
        // the stage, with shock indicator 0 (the extrapolation does not write
        // it); a static profile's, with the profile's indicators
        const int na = (dir == 0) ? A.nj : A.ni;
        const int fa = (dir == 0) ? j : i, fb = (dir == 2) ? j : k;
        const long nfc = (long)na * ((dir == 2) ? A.nj : A.nk);
        const long g = (layer - 1) * nfc + fa + (long)fb * na;
        return TPRef{A.gx[face] + g, 2 * nfc, -1, (type == BC_INFLOW_PROFILE) ? A.gS[face] + g : nullptr};
    }
    int src;
    if (type == BC_EXTRAPOLATE || type == BC_OUTFLOW_SIMPLE_FLUX) src = plus ? n - 1 : 0;
    else if (type == BC_PERIODIC) src = plus ? layer - 1 : n - layer;
    // slip-wall mirror; a wall without ghost data gets the same values, read
    // only by faces the boundary-face kernel recomputes
    else src = plus ? n - layer : layer - 1;
    int si = i, sj = j, sk = k;
    if (dir == 0) si = src; else if (dir == 1) sj = src; else sk = src;
    const long c = ((long)sk * A.nj + sj) * A.ni + si;
    return TPRef{A.Pin + c, A.qs, (type == BC_SLIP_WALL || type == BC_SLIP_WALL_1) ? dir : -1, A.Sprev + c};
}

__device__ __forceinline__ double tp_val(const TPRef& r, int field) { return r.base[field * r.stride]; }

// velocity component comp (global frame) with the slip-wall reflection
__device__ __forceinline__ double tp_vel(const TPRef& r, int comp)
{
    const double v = r.base[(P_VX + comp) * r.stride];
    return (r.flip == comp) ? -v : v;
}

// The whole state of a stencil position, velocity in the face frame (pn, pt1, pt2).
template <int NSP, class R = double>
__device__ __forceinline__ TPCellT<R, NSP> tp_cell(const TPRef& r, int pn, int pt1, int pt2)
{
    TPCellT<R, NSP> s;
    s.rho = R(tp_val(r, P_RHO)); s.p = R(tp_val(r, P_P)); s.T = R(tp_val(r, P_T)); s.u = R(tp_val(r, P_U));
    s.a = R(tp_val(r, P_A));
    s.v[0] = R(tp_vel(r, pn)); s.v[1] = R(tp_vel(r, pt1)); s.v[2] = R(tp_vel(r, pt2));
#pragma unroll
    for (int sp = 0; sp < NSP; ++sp) s.Y[sp] = R(tp_val(r, P_Y + sp));
    return s;
}

// fluxcalc.d:376-529 with species; velocities in the face frame (v[0] normal).
// The flux functions of the mixture take the arithmetic of M: R = double for
// ModeExact / ModeFast64 and float for ModeMixed's face path.
template <class M, int NSP>
__device__ __forceinline__ void tp_ausmdv(const TPCellT<typename M::R, NSP>& L, const TPCellT<typename M::R, NSP>& Rt,
                                          typename M::R* F)
{
    using R = typename M::R;
    using MD = typename std::conditional<M::kFast, M, ModeExact>::type;
    R rL = L.rho, pL = L.p, pLrL = dv<MD>(pL, rL), uL = L.v[0], vL = L.v[1], wL = L.v[2];
    R eL = L.u, aL = L.a;
    R keL = R(0.5) * (uL * uL + vL * vL + wL * wL);
    R HL = eL + pLrL + keL;
    R rR = Rt.rho, pR = Rt.p, pRrR = dv<MD>(pR, rR), uR = Rt.v[0], vR = Rt.v[1], wR = Rt.v[2];
    R eR = Rt.u, aR = Rt.a;
    R keR = R(0.5) * (uR * uR + vR * vR + wR * wR);
    R HR = eR + pRrR + keR;
    R alphaL = dv<MD>(R(2.0) * pLrL, pLrL + pRrR);
    R alphaR = dv<MD>(R(2.0) * pRrR, pLrL + pRrR);
    R am = mx(aL, aR);
    R ML = dv<MD>(uL, am);
    R MR = dv<MD>(uR, am);
    R pLplus, uLplus;
    R duL = R(0.5) * (uL + ab(uL));
    if (ab(ML) <= R(1.0)) {
        pLplus = pL * (ML + R(1.0)) * (ML + R(1.0)) * (R(2.0) - ML) * R(0.25);
        uLplus = alphaL * (dv<MD>((uL + am) * (uL + am), R(4.0) * am) - duL) + duL;
    } else {
        pLplus = dv<MD>(pL * duL, uL);
        uLplus = duL;
    }
    R pRminus, uRminus;
    R duR = R(0.5) * (uR - ab(uR));
    if (ab(MR) <= R(1.0)) {
        pRminus = pR * (MR - R(1.0)) * (MR - R(1.0)) * (R(2.0) + MR) * R(0.25);
        uRminus = alphaR * (dv<MD>(-(uR - am) * (uR - am), R(4.0) * am) - duR) + duR;
    } else {
        pRminus = dv<MD>(pR * duR, uR);
        uRminus = duR;
    }
    R ru_half = uLplus * rL + uRminus * rR;
    R p_half = pLplus + pRminus;
    R dp = pL - pR;
    dp = dv<MD>(R(10.0) * ab(dp), mn(pL, pR));
    R s = R(0.5) * mn(R(1.0), dp);
    R ru2_AUSMV = uLplus * rL * uL + uRminus * rR * uR;
    R ru2_AUSMD = R(0.5) * (ru_half * (uL + uR) - ab(ru_half) * (uR - uL));
    R ru2_half = (R(0.5) + s) * ru2_AUSMV + (R(0.5) - s) * ru2_AUSMD;
    F[0] = ru_half;
    F[1] = ru2_half + p_half;
    if (ru_half >= R(0.0)) {
        F[2] = ru_half * vL; F[3] = ru_half * wL; F[4] = ru_half * HL;
#pragma unroll
        for (int i = 0; i < NSP; ++i) F[5 + i] = ru_half * L.Y[i];
    } else {
        F[2] = ru_half * vR; F[3] = ru_half * wR; F[4] = ru_half * HR;
#pragma unroll
        for (int i = 0; i < NSP; ++i) F[5 + i] = ru_half * Rt.Y[i];
    }
    bool caseA = ((uL - aL) < R(0.0)) && ((uR - aR) > R(0.0));
    bool caseB = ((uL + aL) < R(0.0)) && ((uR + aR) > R(0.0));
    R d_ua = R(0.0);
    if (caseA && !caseB) d_ua = R(0.125) * ((uR - aR) - (uL - aL));
    if (caseB && !caseA) d_ua = R(0.125) * ((uR + aR) - (uL + aL));
    if (d_ua != R(0.0)) {
        F[0] -= d_ua * (rR - rL);
        F[1] -= d_ua * (rR * uR - rL * uL);
        F[2] -= d_ua * (rR * vR - rL * vL);
        F[3] -= d_ua * (rR * wR - rL * wL);
        F[4] -= d_ua * (rR * HR - rL * HL);
#pragma unroll
        for (int i = 0; i < NSP; ++i) F[5 + i] -= d_ua * (rR * Rt.Y[i] - rL * L.Y[i]);
    }
}

// fluxcalc.d:532-623 with species (Hanel's flux vector splitting).
template <class M, int NSP>
__device__ __forceinline__ void tp_hanel(const TPCellT<typename M::R, NSP>& L, const TPCellT<typename M::R, NSP>& Rt,
                                         typename M::R* F)
{
    using R = typename M::R;
    using MD = typename std::conditional<M::kFast, M, ModeExact>::type;
    const R rL = L.rho, pL = L.p;
    const R pLrL = dv<MD>(pL, rL);
    const R uL = L.v[0], vL = L.v[1], wL = L.v[2];
    const R eL = L.u, aL = L.a;
    const R keL = R(0.5) * (uL * uL + vL * vL + wL * wL);
    const R HL = eL + pLrL + keL;
    const R rR = Rt.rho, pR = Rt.p;
    const R pRrR = dv<MD>(pR, rR);
    const R uR = Rt.v[0], vR = Rt.v[1], wR = Rt.v[2];
    const R eR = Rt.u, aR = Rt.a;
    const R keR = R(0.5) * (uR * uR + vR * vR + wR * wR);
    const R HR = eR + pRrR + keR;
    R pLplus, uLplus;
    if (ab(uL) <= aL) {
        uLplus = dv<MD>(R(1.0), R(4.0) * aL) * (uL + aL) * (uL + aL);
        pLplus = pL * uLplus * (dv<MD>(R(1.0), aL) * (R(2.0) - dv<MD>(uL, aL)));
    } else {
        uLplus = R(0.5) * (uL + ab(uL));
        pLplus = pL * uLplus * dv<MD>(R(1.0), uL);
    }
    R pRminus, uRminus;
    if (ab(uR) <= aR) {
        uRminus = dv<MD>(R(-1.0), R(4.0) * aR) * (uR - aR) * (uR - aR);
        pRminus = pR * uRminus * (dv<MD>(R(1.0), aR) * (R(-2.0) - dv<MD>(uR, aR)));
    } else {
        uRminus = R(0.5) * (uR - ab(uR));
        pRminus = pR * uRminus * dv<MD>(R(1.0), uR);
    }
    const R p_half = pLplus + pRminus;
    F[0] = uLplus * rL + uRminus * rR;
    F[1] = uLplus * rL * uL + uRminus * rR * uR + p_half;
    F[2] = uLplus * rL * vL + uRminus * rR * vR;
    F[3] = uLplus * rL * wL + uRminus * rR * wR;
    F[4] = uLplus * rL * HL + uRminus * rR * HR;
#pragma unroll
    for (int i = 0; i < NSP; ++i) F[5 + i] = uLplus * rL * L.Y[i] + uRminus * rR * Rt.Y[i];
}

// HLLC with species carried by the contact wave (the oracle's hllc).
template <class M, int NSP>
__device__ __forceinline__ void tp_hllc(const TPCellT<typename M::R, NSP>& L, const TPCellT<typename M::R, NSP>& Rt,
                                        typename M::R* F)
{
    using R = typename M::R;
    using MD = typename std::conditional<M::kFast, M, ModeExact>::type;
    R rL = L.rho, pL = L.p, uL = L.v[0], vL = L.v[1], wL = L.v[2], aL = L.a;
    R EL = rL * (L.u + R(0.5) * (uL * uL + vL * vL + wL * wL));
    R rR = Rt.rho, pR = Rt.p, uR = Rt.v[0], vR = Rt.v[1], wR = Rt.v[2], aR = Rt.a;
    R ER = rR * (Rt.u + R(0.5) * (uR * uR + vR * vR + wR * wR));
    R SL = mn(uL - aL, uR - aR);
    R SR = mx(uL + aL, uR + aR);
    R mL = rL * (SL - uL);
    R mR = rR * (SR - uR);
    R Sstar = dv<MD>(pR - pL + mL * uL - mR * uR, mL - mR);
    const bool useL = (R(0.0) <= SL) || (!(SR <= R(0.0)) && (R(0.0) <= Sstar));
    const TPCellT<R, NSP>& K = useL ? L : Rt;
    const R rK = K.rho, uK = K.v[0], vK = K.v[1], wK = K.v[2], pK = K.p, EK = useL ? EL : ER;
    R mf = rK * uK;
    F[0] = mf; F[1] = mf * uK + pK; F[2] = mf * vK; F[3] = mf * wK; F[4] = uK * (EK + pK);
#pragma unroll
    for (int i = 0; i < NSP; ++i) F[5 + i] = mf * K.Y[i];
    if ((R(0.0) <= SL) || (SR <= R(0.0))) return;
    const R SK = useL ? SL : SR;
    R coef = dv<MD>(rK * (SK - uK), SK - Sstar);
    R Us[5];
    Us[0] = coef;
    Us[1] = coef * Sstar;
    Us[2] = coef * vK;
    Us[3] = coef * wK;
    Us[4] = coef * (dv<MD>(EK, rK) + (Sstar - uK) * (Sstar + dv<MD>(pK, rK * (SK - uK))));
    F[0] = F[0] + SK * (Us[0] - rK);
    F[1] = F[1] + SK * (Us[1] - rK * uK);
    F[2] = F[2] + SK * (Us[2] - rK * vK);
    F[3] = F[3] + SK * (Us[3] - rK * wK);
    F[4] = F[4] + SK * (Us[4] - EK);
#pragma unroll
    for (int i = 0; i < NSP; ++i) F[5 + i] = F[5 + i] + SK * (coef * K.Y[i] - rK * K.Y[i]);
}

// gamma = Cp/Cv of a mixture state (gas_model.d:162; therm_perf_gas.d:366-380)
template <class M, int NSP>
__device__ __forceinline__ double tp_gamma(const TPGas& g, const double* Y, double T)
{
    using MD = typename std::conditional<M::kFast, ModeFast64, ModeExact>::type;
    if constexpr (M::kFast) {
        const double R = tp_Rmix<NSP>(g, Y);
        double u, cv;
        tp_mix_u_cv<NSP>(g, Y, R, T, &u, &cv);
        return dv<MD>(cv + R, cv);
    } else {
        double Cp = 0.0, Cv = 0.0;
#pragma unroll
        for (int s = 0; s < NSP; ++s) Cp += Y[s] * cea_Cp<M>(g.curve[s], T);
#pragma unroll
        for (int s = 0; s < NSP; ++s) Cv += Y[s] * (cea_Cp<M>(g.curve[s], T) - g.R[s]);
        return dv<MD>(Cp, Cv);
    }
}

// Cv of a mixture state: sum Y_i (Cp_i(T) - R_i) (therm_perf_gas.d:366-371)
template <class M, int NSP>
__device__ __forceinline__ double tp_cv(const TPGas& g, const double* Y, double T)
{
    if constexpr (M::kFast) {
        const double R = tp_Rmix<NSP>(g, Y);
        double u, cv;
        tp_mix_u_cv<NSP>(g, Y, R, T, &u, &cv);
        return cv;
    } else {
        double Cv = 0.0;
#pragma unroll
        for (int s = 0; s < NSP; ++s) Cv += Y[s] * (cea_Cp<M>(g.curve[s], T) - g.R[s]);
        return Cv;
    }
}

// Cv (what = 0) or gamma = Cp/Cv (what = 1) of a face-side state in FP64
// from the gas model at the side's temperature, whatever the flux path's R.
template <class M, int NSP, class R>
__device__ __forceinline__ double tp_side_thermo(const TPGas& g, const TPCellT<R, NSP>& c, int what)
{
    using MG = typename std::conditional<M::kFast, ModeFast64, ModeExact>::type;
    double Y[NSP];
#pragma unroll
    for (int s = 0; s < NSP; ++s) Y[s] = double(c.Y[s]);
    return what ? tp_gamma<MG, NSP>(g, Y, double(c.T)) : tp_cv<MG, NSP>(g, Y, double(c.T));
}

// efmflx (fluxcalc.d:626-790) with species, which follow the sign of the mass
// flux; Cv of each side from the gas model at the side's temperature.
template <class M, int NSP>
__device__ __forceinline__ void tp_efm(const TPGas& g, const TPCellT<typename M::R, NSP>& L,
                                       const TPCellT<typename M::R, NSP>& Rt, typename M::R* F)
{
    using R = typename M::R;
    using MD = typename std::conditional<M::kFast, M, ModeExact>::type;
    const R dtwspi = R(0.282094792);
    const R rhoL = L.rho, presL = L.p, eL = L.u, tL = L.T;
    const R hL = eL + dv<MD>(presL, rhoL);
    const R vnL = L.v[0], vpL = L.v[1], vqL = L.v[2];
    const R rhoR = Rt.rho, presR = Rt.p, eR = Rt.u, tR = Rt.T;
    const R hR = eR + dv<MD>(presR, rhoR);
    const R vnR = Rt.v[0], vpR = Rt.v[1], vqR = Rt.v[2];
    const R cvL = R(tp_side_thermo<M, NSP>(g, L, 0));
    const R RgasL = dv<MD>(presL, rhoL * tL);
    const R cvR = R(tp_side_thermo<M, NSP>(g, Rt, 0));
    const R RgasR = dv<MD>(presR, rhoR * tR);
    const R rLsqrt = sq_rt<R>(rhoL), rRsqrt = sq_rt<R>(rhoR);
    const R alpha = dv<MD>(rLsqrt, rLsqrt + rRsqrt);
    const R cv = alpha * cvL + (R(1.0) - alpha) * cvR;
    const R Rgas = alpha * RgasL + (R(1.0) - alpha) * RgasR;
    const R cp = cv + Rgas;
    const R gam = dv<MD>(cp, cv);
    const R con = dv<MD>(R(0.5) * (gam + R(1.0)), gam - R(1.0));
    const R rtL = Rgas * tL;
    const R cmpL = sq_rt<R>(R(2.0) * rtL);
    const R hvsqL = R(0.5) * (vnL * vnL + vpL * vpL + vqL * vqL);
    const R snL = dv<MD>(vnL, R(1.0) * cmpL);
    R exL, efL;
    exxef<MD>(snL, exL, efL);
    const R wL = R(0.5) * (R(1.0) + efL);
    const R dL = exL * dtwspi;
    const R rtR = dv<MD>(presR, rhoR);
    const R cmpR = sq_rt<R>(R(2.0) * rtR);
    const R hvsqR = R(0.5) * (vnR * vnR + vpR * vpR + vqR * vqR);
    const R snR = dv<MD>(vnR, R(1.0) * cmpR);
    R exR, efR;
    exxef<MD>(snR, exR, efR);
    const R wR = R(0.5) * (R(1.0) - efR);
    const R dR = -exR * dtwspi;
    const R fmsL = (wL * rhoL * vnL) + (dL * cmpL * rhoL);
    const R fmsR = (wR * rhoR * vnR) + (dR * cmpR * rhoR);
    F[0] = fmsL + fmsR;
    F[1] = fmsL * vnL + fmsR * vnR + wL * presL + wR * presR;
    F[2] = fmsL * vpL + fmsR * vpR;
    F[3] = fmsL * vqL + fmsR * vqR;
    F[4] = (wL * rhoL * vnL) * (hvsqL + hL) + (wR * rhoR * vnR) * (hvsqR + hR) +
           (dL * cmpL * rhoL) * (hvsqL + con * rtL) + (dR * cmpR * rhoR) * (hvsqR + con * rtR);
    const bool left = F[0] > R(0.0);
#pragma unroll
    for (int i = 0; i < NSP; ++i) F[5 + i] = F[0] * (left ? L.Y[i] : Rt.Y[i]);
}

// The face state of a mixture side as the ideal-gas flux functions read it.
template <class R, int NSP>
__device__ __forceinline__ FaceStateT<R> tp_face_state(const TPCellT<R, NSP>& c)
{
    FaceStateT<R> f;
    f.rho = c.rho; f.p = c.p; f.u = c.u; f.a = c.a; f.vn = c.v[0]; f.vt1 = c.v[1]; f.vt2 = c.v[2];
    f.RT = R(0.0); f.T = c.T;
    return f;
}

// ausm_plus_up (fluxcalc.d:860-1037) and roe (fluxcalc.d:1202-1371) with
// species upwinded by the mass flux (ausm_plus_up: M_half > 0 picks the left
// density, ru_half >= 0 the left species; roe: F.mass >= 0); roe's gamma of
// each side from the gas model.  The gas-dynamic part is the ideal-gas
// functions' (they read p, rho, u, a and velocities only; the RT shortcut of
// the fast ideal-gas forms is not used for a mixture).
template <class M, int NSP>
__device__ __forceinline__ void tp_other_flux(const TPGas& g, bool use_roe, double M_inf,
                                              const TPCellT<typename M::R, NSP>& L,
                                              const TPCellT<typename M::R, NSP>& Rt, typename M::R* F)
{
    using R = typename M::R;
    using MD = typename std::conditional<M::kFast, M, ModeExact>::type;
    const FaceStateT<R> Lf = tp_face_state<R, NSP>(L), Rf = tp_face_state<R, NSP>(Rt);
    Flux5T<R> Fl;
    if (use_roe)
        Fl = roe<MD>(Lf, Rf, R(tp_side_thermo<M, NSP>(g, L, 1)), R(tp_side_thermo<M, NSP>(g, Rt, 1)));
    else Fl = ausm_plus_up<MD>(Lf, Rf, R(M_inf));
    F[0] = Fl.mass; F[1] = Fl.mn; F[2] = Fl.mt1; F[3] = Fl.mt2; F[4] = Fl.E;
    const bool left = F[0] >= R(0.0);
#pragma unroll
    for (int i = 0; i < NSP; ++i) F[5 + i] = F[0] * (left ? L.Y[i] : Rt.Y[i]);
}

// The flux calculator of a mixture face in the face frame (fluxcalc.d:61-89):
// FLUX_SHOCK: adaptive_hanel_ausmdv / adaptive_efm_ausmdv (fluxcalc.d:802-836),
// or hanel / efm throughout; FLUX_OTHER: ausm_plus_up or roe.
template <class MA, int NSP, int FLUX>
__device__ __forceinline__ void tp_flux(const TPArgs& A, const TPGas& g, const TPCellT<typename MA::R, NSP>& Lf,
                                        const TPCellT<typename MA::R, NSP>& Rf, bool shock, typename MA::R* F)
{
    using R = typename MA::R;
    using MD = typename std::conditional<MA::kFast, MA, ModeExact>::type;
    if constexpr (FLUX == FLUX_HLLC) {
        tp_hllc<MA, NSP>(Lf, Rf, F);
    } else if constexpr (FLUX == FLUX_SHOCK) {
        const R sound_speed = R(0.5) * (Lf.a + Rf.a);
        const R shear_y = dv<MD>(ab(Lf.v[1] - Rf.v[1]), sound_speed);
        const R shear_z = dv<MD>(ab(Lf.v[2] - Rf.v[2]), sound_speed);
        const bool shear_is_small = mx(shear_y, shear_z) <= R(A.shear_tol);
        if (A.always_hanel || (shock && shear_is_small)) {
            if (A.shock_efm) tp_efm<MA, NSP>(g, Lf, Rf, F);
            else tp_hanel<MA, NSP>(Lf, Rf, F);
        } else {
            tp_ausmdv<MA, NSP>(Lf, Rf, F);
        }
    } else if constexpr (FLUX == FLUX_OTHER) {
        tp_other_flux<MA, NSP>(g, A.other_roe != 0, A.M_inf, Lf, Rf, F);
    } else {
        tp_ausmdv<MA, NSP>(Lf, Rf, F);
    }
}

// The arguments of a batched mixture launch's block entry e: the table of
// TPArgs in device memory read through a uniform, read-only pointer (the
// address-space cast is undone by inlining, so the reads stay scalar loads).
#if defined(__HIP_DEVICE_COMPILE__)
using TPKArgs = const __attribute__((address_space(4))) TPArgs*;
using TPKBatch = const __attribute__((address_space(4))) BatchArgs*;
__device__ __forceinline__ const TPArgs& tp_batch_args(int& entry)
{
    TPKBatch b = (TPKBatch)__builtin_amdgcn_kernarg_segment_ptr();
    entry = __builtin_amdgcn_readfirstlane(b->wg_entry[blockIdx.x]);
    TPKArgs p = (TPKArgs)(b->table + (unsigned long long)entry * sizeof(TPArgs));
    return *(const TPArgs*)p;
}
// Table entry e of a batched launch (a uniform index).
__device__ __forceinline__ const TPArgs& tp_table_args(const BatchArgs& b, int e)
{
    TPKArgs p = (TPKArgs)(b.table + (unsigned long long)__builtin_amdgcn_readfirstlane(e) * sizeof(TPArgs));
    return *(const TPArgs*)p;
}
#else
__device__ __forceinline__ const TPArgs& tp_batch_args(int& entry)
{
    entry = 0;
    return *(const TPArgs*)nullptr;
}
__device__ __forceinline__ const TPArgs& tp_table_args(const BatchArgs&, int)
{
    return *(const TPArgs*)nullptr;
}
#endif

// First face slot of direction DIR in a block's retry list.
template <int DIR>
__device__ __forceinline__ long tp_retry_first(const TPArgs& A)
{
    const long nk = (A.dims == 3) ? A.nk : 1;
    const long nf0 = (long)(A.ni + 1) * A.nj * nk, nf1 = (long)A.ni * (A.nj + 1) * nk;
    return 4 + ((DIR >= 1) ? nf0 : 0) + ((DIR == 2) ? nf1 : 0);
}

// The face kernel's work for face f0 of direction DIR of one block.  LEAN
// (the mixed form's face kernels): the face states' temperatures come from
// tp_face_halley alone, and a face where it does not settle is appended to
// the block's retry list instead of being finished here, which keeps the
// full solve -- and the registers it needs -- out of the kernel that runs
// every face.
template <class M, int NSP, int FLUX, int DIR, bool LEAN = false>
__device__ __forceinline__ void tp_face_body(const TPArgs& A, long f0)
{
    using MA = typename TPArith<M>::type;
    using MF = typename TPFlux<M>::type;
    using R = typename MF::R;
    constexpr bool kPolish = TPArith<M>::polish;
    if (stage_aborted(A.abort_prev)) return;
    const int nfi = A.ni + (DIR == 0), nfj = A.nj + (DIR == 1), nfk = (A.dims == 3) ? A.nk + (DIR == 2) : 1;
    const long nf = (long)nfi * nfj * nfk;
    const bool active = f0 < nf;
    const long f = active ? f0 : nf - 1;   // idle lanes compute the last face and store nothing
    const int i = (int)(f % nfi);
    const int j = (int)((f / nfi) % nfj);
    const int k = (int)(f / ((long)nfi * nfj));
    const int di = (DIR == 0), dj = (DIR == 1), dk = (DIR == 2);
    // face frame permutation: i (x,y,z); j 3D (y,z,x), 2D (y,x,z); k (z,x,y)
    const bool d3 = (A.dims == 3);
    const int pn = DIR, pt1 = (DIR == 0) ? 1 : (DIR == 1) ? (d3 ? 2 : 0) : 0;
    const int pt2 = (DIR == 0) ? 2 : (DIR == 1) ? (d3 ? 0 : 2) : 1;
    // the stencil's four positions
    TPRef r[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) r[m] = tp_ref(A, i + (m - 2) * di, j + (m - 2) * dj, k + (m - 2) * dk);
    TPCellT<R, NSP> Lf, Rf;
    int fe = 0, fb = 0;
    bool okL = false, okR = false, defer = false;
    if (A.high_order) {
        const L2R2WT<R> w = cast_w<R>(A.w[DIR]);
        const bool lim = A.apply_limiter, clip = A.extrema_clipping;
        const int comp[3] = {pn, pt1, pt2};
        // every value the reconstruction and the two solves read, loaded in
        // straight-line code before any is used: one memory round trip for
        // the face instead of one per reconstructed scalar
        constexpr int NV = 5 + NSP;   // velocity (face frame), mass fractions, rho, u
        double v[4][NV];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
#pragma unroll
            for (int d = 0; d < 3; ++d) v[m][d] = tp_vel(r[m], comp[d]);
#pragma unroll
            for (int s = 0; s < NSP; ++s) v[m][3 + s] = tp_val(r[m], P_Y + s);
            v[m][3 + NSP] = tp_val(r[m], P_RHO);
            v[m][4 + NSP] = tp_val(r[m], P_U);
        }
        const double TL = tp_val(r[1], P_T), TR = tp_val(r[2], P_T);
        const double aL = tp_val(r[1], P_A), aR = tp_val(r[2], P_A);
#pragma unroll
        for (int d = 0; d < 3; ++d)
            interp_l2r2_scalar<MF>(w, lim, clip, R(v[0][d]), R(v[1][d]), R(v[2][d]), R(v[3][d]), Lf.v[d], Rf.v[d]);
#pragma unroll
        for (int s = 0; s < NSP; ++s)
            interp_l2r2_scalar<MF>(w, lim, clip, R(v[0][3 + s]), R(v[1][3 + s]), R(v[2][3 + s]), R(v[3][3 + s]),
                                   Lf.Y[s], Rf.Y[s]);
        if (scale_mass_fractions<MF, NSP>(Lf.Y) != 0) {
#pragma unroll
            for (int s = 0; s < NSP; ++s) Lf.Y[s] = R(v[1][3 + s]);
        }
        if (scale_mass_fractions<MF, NSP>(Rf.Y) != 0) {
#pragma unroll
            for (int s = 0; s < NSP; ++s) Rf.Y[s] = R(v[2][3 + s]);
        }
        interp_l2r2_scalar<MF>(w, lim, clip, R(v[0][3 + NSP]), R(v[1][3 + NSP]), R(v[2][3 + NSP]), R(v[3][3 + NSP]),
                               Lf.rho, Rf.rho);
        interp_l2r2_scalar<MF>(w, lim, clip, R(v[0][4 + NSP]), R(v[1][4 + NSP]), R(v[2][4 + NSP]), R(v[3][4 + NSP]),
                               Lf.u, Rf.u);
        // update_thermo_from_rhou on each side (guess: the side's cell T); on
        // failure the side is the cell's state (onedinterp.d:1039-1068); the
        // sound speed is the cell's (sfluidblock.d:2199-2208).  The two sides
        // share one copy of the solver code (a loop that is not unrolled);
        // the lean form solves both sides together.
        if constexpr (LEAN) {
            double YL[NSP], YR[NSP];
#pragma unroll
            for (int s = 0; s < NSP; ++s) { YL[s] = double(Lf.Y[s]); YR[s] = double(Rf.Y[s]); }
            const double RmL = tp_Rmix<NSP>(*A.gas, YL), RmR = tp_Rmix<NSP>(*A.gas, YR);
            double TfL = TL, TfR = TR;
            tp_face_halley_pair<NSP>(*A.gas, YL, YR, RmL, RmR, double(Lf.u), double(Rf.u), TfL, TfR, okL, okR, &fe);
            if (okL) { Lf.T = R(TfL); Lf.p = R(double(Lf.rho) * RmL * TfL); Lf.a = R(aL); }
            if (okR) { Rf.T = R(TfR); Rf.p = R(double(Rf.rho) * RmR * TfR); Rf.a = R(aR); }
            defer = !(okL && okR);
        }
#pragma nounroll
        for (int sd = 0; sd < 2 && !LEAN; ++sd) {
            // the temperature solve in FP64 from the (FP32 in the mixed form)
            // reconstructed density, energy and mass fractions
            TPState<NSP> Q;
            Q.rho = double(sd ? Rf.rho : Lf.rho);
            Q.u = double(sd ? Rf.u : Lf.u);
            Q.T = sd ? TR : TL;
#pragma unroll
            for (int s = 0; s < NSP; ++s) Q.Y[s] = double(sd ? Rf.Y[s] : Lf.Y[s]);
            if (tp_face_eos<MA, NSP, kPolish>(*A.gas, Q, &fe, &fb) == 0) {
                const R a = R(sd ? aR : aL);
                if (sd == 0) { Lf.T = R(Q.T); Lf.u = R(Q.u); Lf.p = R(Q.p); Lf.a = a; okL = true; }
                else { Rf.T = R(Q.T); Rf.u = R(Q.u); Rf.p = R(Q.p); Rf.a = a; okR = true; }
            }
        }
    }
    if constexpr (LEAN) {
        // hand the face on: one atomic per wave for the faces it defers
        const unsigned long long m = __ballot(defer && active);
        if (m) {
            const int lead = __ffsll(m) - 1;
            unsigned int base = 0;
            if (__lane_id() == lead) base = atomicAdd(&A.retry[DIR], (unsigned int)__popcll(m));
            base = __shfl(base, lead);
            if (defer && active) {
                const unsigned int below = (unsigned int)__popcll(m & ((1ull << __lane_id()) - 1ull));
                A.retry[tp_retry_first<DIR>(A) + base + below] = (unsigned int)f;
            }
            fb = (defer && active) ? 1 : 0;
        }
    }
    if (!LEAN || !defer) {
        if (!okL) Lf = tp_cell<NSP, R>(r[1], pn, pt1, pt2);
        if (!okR) Rf = tp_cell<NSP, R>(r[2], pn, pt1, pt2);
        R Fr[5 + NSP];
        const bool shock = (FLUX == FLUX_SHOCK) && ((r[1].s && *r[1].s) || (r[2].s && *r[2].s));
        tp_flux<MF, NSP, FLUX>(A, *A.gas, Lf, Rf, shock, Fr);
        using FS = TPFluxStore<M>;
        FS F[5 + NSP];
    #pragma unroll
        for (int q = 0; q < 5 + NSP; ++q) F[q] = FS(Fr[q]);
        if (active) {
            FS* dst = reinterpret_cast<FS*>(A.F[DIR]);
            const long fq = A.fqs[DIR];
            dst[f] = F[0];
            dst[(1 + pn) * fq + f] = F[1];
            dst[(1 + pt1) * fq + f] = F[2];
            dst[(1 + pt2) * fq + f] = F[3];
            dst[4 * fq + f] = F[4];
    #pragma unroll
            for (int s = 0; s < NSP; ++s) dst[(5 + s) * fq + f] = F[5 + s];
        }
    }
    if (!active) { fe = 0; fb = 0; }
    for (int off = 32; off > 0; off >>= 1) {
        fe += __shfl_xor(fe, off);
        fb += __shfl_xor(fb, off);
    }
    if (__lane_id() == 0) {
        add_stat(A.stats, STAT_FACE_EVALS, (unsigned int)fe);
        add_stat(A.stats, STAT_FACE_FALLBACKS, (unsigned int)fb);
    }
}

// One thread per face of direction DIR; BATCH: a launch over several
// blocks, each workgroup in the block its map entry names.
#ifndef E4_TP_LEAN_WAVES
#define E4_TP_LEAN_WAVES 3
#endif
template <class M> struct TPFaceOccupancy { static constexpr int waves = 3; };
template <> struct TPFaceOccupancy<ModeMixed> { static constexpr int waves = E4_TP_LEAN_WAVES; };

template <class M, int NSP, int FLUX, int DIR, bool BATCH>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(TPFaceOccupancy<M>::waves)))
void tp_face_kernel(typename std::conditional<BATCH, BatchArgs, TPArgs>::type a)
{
    constexpr bool LEAN = std::is_same<M, ModeMixed>::value;
    if constexpr (BATCH) {
        int e;
        const TPArgs& A = tp_batch_args(e);
        tp_face_body<M, NSP, FLUX, DIR, LEAN>(A, (long)((int)blockIdx.x - A.wg0[DIR]) * blockDim.x + threadIdx.x);
    } else {
        tp_face_body<M, NSP, FLUX, DIR, LEAN>(a, (long)blockIdx.x * blockDim.x + threadIdx.x);
    }
}

// The faces of direction DIR that the mixed form's face kernels handed on,
// with the full solve; `part` of `nparts` workgroups of this block take
// every nparts-th group of 128 list entries (the same trip count for every
// lane of a group, idle lanes computing nothing they store).
template <int NSP, int FLUX, int DIR>
__device__ __forceinline__ void tp_retry_dir(const TPArgs& A, int part, int nparts)
{
    const long nk = (A.dims == 3) ? A.nk : 1;
    const long nf = (long)(A.ni + (DIR == 0)) * (A.nj + (DIR == 1)) * (nk + (DIR == 2));
    const long n = (long)A.retry[DIR];
    const unsigned int* list = A.retry + tp_retry_first<DIR>(A);
    for (long g = (long)part * blockDim.x; g < n; g += (long)nparts * blockDim.x) {
        const long idx = g + threadIdx.x;
        const long f = (idx < n) ? (long)list[idx] : nf;
        tp_face_body<ModeMixed, NSP, FLUX, DIR, false>(A, f);
    }
}

// The retry kernel of a mixture block (BATCH: of several, `nparts`
// workgroups per block entry of the table).
template <int NSP, int FLUX, bool BATCH>
__global__ __launch_bounds__(128) void tp_retry_kernel(typename std::conditional<BATCH, BatchArgs, TPArgs>::type a,
                                                       int nparts)
{
    auto run = [&](const TPArgs& A, int part) {
        if (stage_aborted(A.abort_prev) || !A.retry) return;
        tp_retry_dir<NSP, FLUX, 0>(A, part, nparts);
        tp_retry_dir<NSP, FLUX, 1>(A, part, nparts);
        if (A.dims == 3) tp_retry_dir<NSP, FLUX, 2>(A, part, nparts);
    };
    if constexpr (BATCH) {
        run(tp_table_args(a, (int)blockIdx.x / nparts), (int)blockIdx.x % nparts);
    } else {
        run(a, (int)blockIdx.x);
    }
}

// GUESS: A.guess is installed (the learned temperature guess; its LDS staging
// and MLP code exist only in that instantiation).
template <class M, int NSP, bool GUESS>
__device__ __forceinline__ void tp_update_body(const TPArgs& A, long c)
{
    using MA = typename TPArith<M>::type;
    constexpr bool kPolish = TPArith<M>::polish;
    using MD = typename std::conditional<M::kFast, ModeFast64, ModeExact>::type;
    constexpr int NQ = 5 + NSP;
    if (stage_aborted(A.abort_prev)) return;
    // the retry kernel of this stage has run: clear the list for the next stage
    if (A.retry && c == 0) {
        A.retry[0] = 0u;
        A.retry[1] = 0u;
        A.retry[2] = 0u;
    }
    const long ncells = (long)A.ni * A.nj * A.nk;
    const bool active = c < ncells;
    const long qs = A.qs;
    bool invalid = false, decode_fail = false, fell_back = false;
    int fe = 0;
    double Un[NQ];
    TPState<NSP> Q;
    double v[3] = {0.0, 0.0, 0.0};
    double rho = 1.0;
    bool have_state = false;
    if (active) {
        const int i = (int)(c % A.ni), j = (int)((c / A.ni) % A.nj), k = (int)(c / ((long)A.ni * A.nj));
        const long iW = ((long)k * A.nj + j) * (A.ni + 1) + i;
        const long jS = ((long)k * (A.nj + 1) + j) * A.ni + i;
        const long kB = c;
        double integ[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) integ[q] = 0.0;
        using FS = TPFluxStore<M>;
        auto add_face = [&](const double* Fraw, long fq, long fidx, double outsign, double Aface) {
            const FS* Fsrc = reinterpret_cast<const FS*>(Fraw);
            double area = outsign * Aface;
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                if (q == 3 && A.dims != 3) continue;
                integ[q] -= double(Fsrc[q * fq + fidx]) * area;
            }
        };
        double vol_inv = A.vol_inv, qy = 0.0;
        if (A.axi) {
            // 2D axisymmetric: per-radian face areas and volume, and the
            // pressure source of the y-momentum (fvcell.d:2177-2181)
            const AxiGeom g = axi_geom(A.axi, A.ni, A.nj);
            add_face(A.F[1], A.fqs[1], jS + A.ni, 1.0, g.aj[jS + A.ni]);
            add_face(A.F[0], A.fqs[0], iW + 1, 1.0, g.ai[iW + 1]);
            add_face(A.F[1], A.fqs[1], jS, -1.0, g.aj[jS]);
            add_face(A.F[0], A.fqs[0], iW, -1.0, g.ai[iW]);
            vol_inv = 1.0 / g.vol[c];
            qy = 0.0 + A.Pin[c + P_P * qs] * g.areaxy[c] / g.vol[c];
        } else if (A.geo) {
            // a block built from its grid: each face's own area, the cell's volume
            add_face(A.F[1], A.fqs[1], jS + A.ni, 1.0, A.fgeo[1][jS + A.ni]);
            add_face(A.F[0], A.fqs[0], iW + 1, 1.0, A.fgeo[0][iW + 1]);
            add_face(A.F[1], A.fqs[1], jS, -1.0, A.fgeo[1][jS]);
            add_face(A.F[0], A.fqs[0], iW, -1.0, A.fgeo[0][iW]);
            if (A.dims == 3) {
                add_face(A.F[2], A.fqs[2], kB + (long)A.ni * A.nj, 1.0, A.fgeo[2][kB + (long)A.ni * A.nj]);
                add_face(A.F[2], A.fqs[2], kB, -1.0, A.fgeo[2][kB]);
            }
            vol_inv = 1.0 / A.geo[c];
        } else {
            add_face(A.F[1], A.fqs[1], jS + A.ni, 1.0, A.area[1]);
            add_face(A.F[0], A.fqs[0], iW + 1, 1.0, A.area[0]);
            add_face(A.F[1], A.fqs[1], jS, -1.0, A.area[1]);
            add_face(A.F[0], A.fqs[0], iW, -1.0, A.area[0]);
            if (A.dims == 3) {
                add_face(A.F[2], A.fqs[2], kB + (long)A.ni * A.nj, 1.0, A.area[2]);
                add_face(A.F[2], A.fqs[2], kB, -1.0, A.area[2]);
            }
        }
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double Qs = (q == 2) ? qy : 0.0;
            if (A.src) Qs += A.src[c + q * qs];   // user-defined source terms
            double d = vol_inv * integ[q] + Qs;
            double dd0 = (A.stage >= 2) ? A.d0[c + q * qs] : 0.0;
            double dd1 = (A.stage >= 3) ? A.d1[c + q * qs] : 0.0;
            // dU/dt and U are stored now (the low-T fallback below re-stores
            // what it changes), as 64-bit integers: a store of a double might
            // alias the gas model's tables read after it, which would force
            // vector loads of those uniform tables; an integer store cannot
            // (type-based alias analysis), so they stay scalar loads
            if (A.dout) reinterpret_cast<unsigned long long*>(A.dout)[c + q * qs] = __builtin_bit_cast(unsigned long long, d);
            double uold = A.Uold[c + q * qs];
            if (A.stage == 1) Un[q] = uold + A.c1 * d;
            else if (A.stage == 2) Un[q] = uold + A.dt * (A.g1 * dd0 + A.g2 * d);
            else Un[q] = uold + A.dt * (A.g1 * dd0 + A.g2 * dd1 + A.g3 * d);
            reinterpret_cast<unsigned long long*>(A.Unew)[c + q * qs] = __builtin_bit_cast(unsigned long long, Un[q]);
        }
        // decode_conserved (fvcell.d:563-732): density, velocity, u, mass fractions
        rho = Un[0];
        if (!(rho > 0.0)) {
            decode_fail = true;
        } else {
            Q.rho = rho;
            double dinv = dv<MD>(1.0, rho);
            v[0] = Un[1] * dinv; v[1] = Un[2] * dinv; v[2] = Un[3] * dinv;
            double u = Un[4] * dinv;
            double ke = 0.5 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            u -= ke;
            Q.u = u;
#pragma unroll
            for (int s = 0; s < NSP; ++s) Q.Y[s] = Un[5 + s] * dinv;
            if (scale_mass_fractions<MA, NSP>(Q.Y) != 0) decode_fail = true;
            else have_state = true;
        }
    }
    // learned initial guess for the whole wave (all lanes take part)
    double Tguess = 0.0;
    if constexpr (GUESS) {
        __shared__ __attribute__((aligned(16))) __bf16 gm_lds[2 * GM_LDS_PER_WAVE];
        double x[7] = {have_state ? Q.rho : 1.0, have_state ? Q.u : 1.0e6, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < 5 && s < NSP; ++s) x[2 + s] = have_state ? Q.Y[s] : 0.2;
        Tguess = (double)mlp_guess_wave(*A.guess, x, gm_lds + (threadIdx.x / 64) * GM_LDS_PER_WAVE);
    }
    if (have_state) {
        const double e_tgt = Q.u;
        const double Tprev = A.Pin[P_T * qs + c];   // the previous temperature: the bracketed solve's guess
        int rc = -1;
        if constexpr (MA::kFast) {
            // the composition for the polish, the bracketed solve and the sound
            // speed, its segment sums formed at each evaluation from scalar
            // loads of the gas tables: 111-115 VGPRs, four waves per SIMD
            const TPMixY<NSP> X = tp_mix_y<NSP>(*A.gas, Q.Y);
            const double Rmix = tp_Rmix<NSP>(*A.gas, Q.Y);
            if constexpr (GUESS || kPolish) {
                rc = tp_polish_x<MA, NSP>(*A.gas, X, Rmix, Q, e_tgt, GUESS ? Tguess : Tprev, &fe);
                fell_back = rc != 0;
            }
            if (rc != 0) {
                Q.u = e_tgt;
                Q.T = Tprev;
                rc = tp_update_temperature_x<MA, NSP>(*A.gas, X, Rmix, Q, &fe);
            }
            if (rc == 0) {
                Q.p = Q.rho * Rmix * Q.T;
                Q.a = tp_sound_speed_x<MA>(*A.gas, X, Rmix, Q.T);
            }
        } else {
            if constexpr (GUESS) {
                rc = tp_polish<MA, NSP>(*A.gas, Q, e_tgt, Tguess, &fe);
                if (rc == 0) Q.p = Q.rho * tp_Rmix<NSP>(*A.gas, Q.Y) * Q.T;
                else fell_back = true;
            }
            if (rc != 0) {
                Q.u = e_tgt;
                Q.T = Tprev;
                rc = tp_update_thermo_from_rhou<MA, NSP>(*A.gas, Q, &fe);
            }
            if (rc == 0) Q.a = tp_sound_speed<MA, NSP>(*A.gas, Q);
        }
        if (rc != 0) {
            // ignore_low_T_thermo_update_failure: T = 200 K and re-encode U
            Q.T = 200.0;
            tp_update_thermo_from_rhoT<MA, NSP>(*A.gas, Q);
            // (integer stores, as above)
            unsigned long long* Un_out = reinterpret_cast<unsigned long long*>(A.Unew) + c;
            auto bits = [](double x) { return __builtin_bit_cast(unsigned long long, x); };
            Un_out[qs] = bits(rho * v[0]); Un_out[2 * qs] = bits(rho * v[1]); Un_out[3 * qs] = bits(rho * v[2]);
            double ke2 = 0.5 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            Un_out[4 * qs] = bits(rho * (Q.u + ke2));
#pragma unroll
            for (int s = 0; s < NSP; ++s) Un_out[(5 + s) * qs] = bits(rho * Q.Y[s]);
            Q.a = tp_sound_speed<MA, NSP>(*A.gas, Q);
        }
        // check_data (flowstate.d:345-388)
        double fsum = 0.0;
#pragma unroll
        for (int s = 0; s < NSP; ++s) fsum += Q.Y[s];
        if (!isfinite(Q.rho) || Q.rho < 0.0) invalid = true;
        if (!isfinite(Q.T) || Q.T < 1.01) invalid = true;
        if (!isfinite(Q.p) || !isfinite(Q.a)) invalid = true;
        if (fsum < 0.99 || fsum > 1.01 || !isfinite(fsum)) invalid = true;
        if (fabs(v[0]) > 30000.0 || fabs(v[1]) > 30000.0 || fabs(v[2]) > 30000.0) invalid = true;
        if (Q.T > 50000.0) invalid = true;
        double* P = A.Pout;
        P[P_RHO * qs + c] = Q.rho; P[P_VX * qs + c] = v[0]; P[P_VY * qs + c] = v[1]; P[P_VZ * qs + c] = v[2];
        P[P_U * qs + c] = Q.u; P[P_T * qs + c] = Q.T; P[P_P * qs + c] = Q.p; P[P_A * qs + c] = Q.a;
#pragma unroll
        for (int s = 0; s < NSP; ++s) P[(P_Y + s) * qs + c] = Q.Y[s];
    }
    unsigned int nbad = (unsigned int)__popcll(__ballot(invalid));
    unsigned int nfail = (unsigned int)__popcll(__ballot(decode_fail));
    unsigned int nfb = (unsigned int)__popcll(__ballot(fell_back));
    int fsum_w = fe;
    for (int off = 32; off > 0; off >>= 1) fsum_w += __shfl_xor(fsum_w, off);
    if (__lane_id() == 0) {
        count_bad_cells(A.counters, A.abort_this, A.max_invalid, nbad, nfail);
        add_stat(A.stats, STAT_DECODE_EVALS, (unsigned int)fsum_w);
        add_stat(A.stats, STAT_DECODE_FALLBACKS, nfb);
    }
}

#ifndef E4_TP_UPDATE_WAVES
#define E4_TP_UPDATE_WAVES 3
#endif
template <class M, int NSP, bool GUESS, bool BATCH>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(E4_TP_UPDATE_WAVES))) void tp_update_kernel(typename std::conditional<BATCH, BatchArgs, TPArgs>::type a)
{
    if constexpr (BATCH) {
        int e;
        const TPArgs& A = tp_batch_args(e);
        tp_update_body<M, NSP, GUESS>(A, (long)((int)blockIdx.x - A.wg0[3]) * blockDim.x + threadIdx.x);
    } else {
        // rows of cells in order, a contiguous run per XCD: the faces a cell
        // shares with its j and k neighbours are read from one L2 (measured
        // 335 -> 319 us for the mixed form at 128^3; the same mapping made the
        // face kernels and the ideal-gas stage kernel slower, 2-5 %)
        const int w = xcd_contiguous((int)blockIdx.x, (int)gridDim.x);
        tp_update_body<M, NSP, GUESS>(a, (long)w * blockDim.x + threadIdx.x);
    }
}

// Decode of an uploaded U with a temperature guess into the primitive array
// (the restart path: read_solution then decode_conserved).
template <class M, int NSP>
__global__ __launch_bounds__(128) void tp_decode_kernel(TPArgs A, const double* __restrict__ U,
                                                        const double* __restrict__ Tguess)
{
    using MD = typename std::conditional<M::kFast, ModeFast64, ModeExact>::type;
    __shared__ __attribute__((aligned(16))) __bf16 gm_lds[2 * GM_LDS_PER_WAVE];
    const long ncells = (long)A.ni * A.nj * A.nk;
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long qs = A.qs;
    bool bad = false, have_state = false, fell_back = false;
    int fe = 0;
    TPState<NSP> Q;
    double v[3] = {0.0, 0.0, 0.0};
    if (c < ncells) {
        double rho = U[c];
        if (!(rho > 0.0)) {
            bad = true;
        } else {
            Q.rho = rho;
            double dinv = dv<MD>(1.0, rho);
            v[0] = U[qs + c] * dinv; v[1] = U[2 * qs + c] * dinv; v[2] = U[3 * qs + c] * dinv;
            double u = U[4 * qs + c] * dinv;
            u -= 0.5 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            Q.u = u;
#pragma unroll
            for (int s = 0; s < NSP; ++s) Q.Y[s] = U[(5 + s) * qs + c] * dinv;
            if (scale_mass_fractions<M, NSP>(Q.Y) != 0) bad = true;
            else have_state = true;
        }
    }
    double Tg = 0.0;
    if (A.guess) {
        double x[7] = {have_state ? Q.rho : 1.0, have_state ? Q.u : 1.0e6, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < 5 && s < NSP; ++s) x[2 + s] = have_state ? Q.Y[s] : 0.2;
        Tg = (double)mlp_guess_wave(*A.guess, x, gm_lds + (threadIdx.x / 64) * GM_LDS_PER_WAVE);
    }
    if (have_state) {
        const double e_tgt = Q.u;
        int rc = -1;
        if (A.guess) {
            rc = tp_polish<M, NSP>(*A.gas, Q, e_tgt, Tg, &fe);
            if (rc == 0) Q.p = Q.rho * tp_Rmix<NSP>(*A.gas, Q.Y) * Q.T;
            else fell_back = true;
        }
        if (rc != 0) {
            Q.u = e_tgt;
            Q.T = Tguess[c];
            rc = tp_update_thermo_from_rhou<M, NSP>(*A.gas, Q, &fe);
        }
        if (rc != 0) {
            bad = true;
        } else {
            Q.a = tp_sound_speed<M, NSP>(*A.gas, Q);
            double* P = A.Pout;
            P[P_RHO * qs + c] = Q.rho; P[P_VX * qs + c] = v[0]; P[P_VY * qs + c] = v[1]; P[P_VZ * qs + c] = v[2];
            P[P_U * qs + c] = Q.u; P[P_T * qs + c] = Q.T; P[P_P * qs + c] = Q.p; P[P_A * qs + c] = Q.a;
#pragma unroll
            for (int s = 0; s < NSP; ++s) P[(P_Y + s) * qs + c] = Q.Y[s];
        }
    }
    unsigned int nb = (unsigned int)__popcll(__ballot(bad));
    unsigned int nfb = (unsigned int)__popcll(__ballot(fell_back));
    int fsum_w = fe;
    for (int off = 32; off > 0; off >>= 1) fsum_w += __shfl_xor(fsum_w, off);
    if (__lane_id() == 0) {
        if (nb) atomicAdd(&A.counters[0], nb);
        add_stat(A.stats, STAT_DECODE_EVALS, (unsigned int)fsum_w);
        add_stat(A.stats, STAT_DECODE_FALLBACKS, nfb);
    }
}

// fluidblock.d:415-431 with PJ_ShockDetector (shockdetectors.d) for a
// mixture block: the shock indicator of every active cell from the primitive
// state the stage starts from (ghost neighbours as the ghost-cell effects
// give them).
template <class M>
__global__ __launch_bounds__(128) void tp_shock_detect_kernel(TPArgs A)
{
    using MD = typename std::conditional<M::kFast, ModeFast64, ModeExact>::type;
    if (stage_aborted(A.abort_prev)) return;
    const long n = (long)A.ni * A.nj * A.nk;
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int i = (int)(c % A.ni), j = (int)((c / A.ni) % A.nj), k = (int)(c / ((long)A.ni * A.nj));
    const TPRef own = tp_ref(A, i, j, k);
    const double tol = A.compression_tol;
    auto face = [&](const TPRef& L, const TPRef& R, int d) {
        const double uL = tp_vel(L, d), uR = tp_vel(R, d);
        const double aL = tp_val(L, P_A), aR = tp_val(R, P_A);
        const double a_min = (aL < aR) ? aL : aR;
        return dv<MD>(uR - uL, a_min) < tol;
    };
    bool S = false;
    const int ndir = (A.dims == 3) ? 3 : 2;
    const int FM[3] = {FACE_WEST, FACE_SOUTH, FACE_BOTTOM}, FP[3] = {FACE_EAST, FACE_NORTH, FACE_TOP};
    for (int d = 0; d < ndir; ++d) {
        const int di = (d == 0), dj = (d == 1), dk = (d == 2);
        const int pos = (d == 0) ? i : (d == 1 ? j : k), n = (d == 0) ? A.ni : (d == 1 ? A.nj : A.nk);
        // a boundary face without ghost data has the cell on one side only
        // (shockdetectors.d:39-54): its velocity towards the wall
        const double un = tp_vel(own, d), a0 = tp_val(own, P_A);
        if (pos == 0 && (A.noghost & (1 << FM[d]))) S = S || dv<MD>(un, a0) < tol;
        else S = S || face(tp_ref(A, i - di, j - dj, k - dk), own, d);
        if (pos == n - 1 && (A.noghost & (1 << FP[d]))) S = S || dv<MD>(-un, a0) < tol;
        else S = S || face(own, tp_ref(A, i + di, j + dj, k + dk), d);
    }
    A.Snew[c] = S ? 1 : 0;
}

// The faces of a mixture block whose flux is not the two-sided reconstruction
// of tp_face_kernel (sfluidblock.d:2189-2345): next to a wall without ghost
// data (WallBC_WithSlip1) the wall face takes the wall flux of
// fluxcalc.d:130-298 from the wall cell's state (or, unclipped, its linear
// extrapolation, onedinterp.d:820-1037) and the face after it the one-sided
// reconstruction l1r2 / l2r1 (onedinterp.d:165-818); a simple-outflow-flux
// face takes the interior cell's flux (boundary_flux_effect.d:367-546).
// One thread per such face; it overwrites what the face kernel stored.
template <class M, int NSP, int FLUX>
__global__ __launch_bounds__(64) void tp_boundary_face_kernel(TPArgs A)
{
    using MA = typename TPArith<M>::type;
    constexpr bool kPolish = TPArith<M>::polish;
    if (stage_aborted(A.abort_prev)) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.fix_off[6]) return;
    int slab = 0;
    while (slab < 5 && t >= A.fix_off[slab + 1]) ++slab;
    const int d = slab >> 1, side = slab & 1;
    const int n3[3] = {A.ni, A.nj, (A.dims == 3) ? A.nk : 1};
    const int nl = A.fix_layers[slab];
    const int da = (d + 1) % 3, db = (d + 2) % 3;
    const int local = t - A.fix_off[slab];
    const int layer = local % nl, rest = local / nl;
    int c[3];
    const int n = n3[d];
    c[d] = side ? n - layer : layer;            // the face's plus-side cell index along d
    c[da] = rest % n3[da];
    c[db] = rest / n3[da];
    const int i = c[0], j = c[1], k = c[2];
    const int FM[3] = {FACE_WEST, FACE_SOUTH, FACE_BOTTOM}, FP[3] = {FACE_EAST, FACE_NORTH, FACE_TOP};
    const int tm = A.bc->type[FM[d]], tpl = A.bc->type[FP[d]];
    const int pos = c[d];
    const bool d3 = (A.dims == 3);
    const int pn = d, pt1 = (d == 0) ? 1 : (d == 1) ? (d3 ? 2 : 0) : 0;
    const int pt2 = (d == 0) ? 2 : (d == 1) ? (d3 ? 0 : 2) : 1;
    const int di = (d == 0), dj = (d == 1), dk = (d == 2);
    TPRef r[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) r[m] = tp_ref(A, i + (m - 2) * di, j + (m - 2) * dj, k + (m - 2) * dk);
    const TPGas& g = *A.gas;
    double F[5 + NSP];
#pragma unroll
    for (int q = 0; q < 5 + NSP; ++q) F[q] = 0.0;
    int fe = 0, fb = 0;
    const bool ho = A.high_order, lim = A.apply_limiter, clip = A.extrema_clipping;
    if ((pos == 0 && tm == BC_OUTFLOW_SIMPLE_FLUX) || (pos == n && tpl == BC_OUTFLOW_SIMPLE_FLUX)) {
        const bool plus = (pos == n);
        const TPRef& cr = plus ? r[1] : r[2];
        const double rho = tp_val(cr, P_RHO), p = tp_val(cr, P_P), u = tp_val(cr, P_U);
        const double vx = tp_vel(cr, 0), vy = tp_vel(cr, 1), vz = tp_vel(cr, 2);
        const double vn = tp_vel(cr, pn);
        const double mass_flux = rho * vn;
        if ((plus ? 1.0 : -1.0) * mass_flux > 0.0) {
            F[0] = mass_flux;
            F[1] = p + vn * mass_flux;
            F[2] = tp_vel(cr, pt1) * mass_flux;
            F[3] = tp_vel(cr, pt2) * mass_flux;
            const double utot = u + 0.5 * (vx * vx + vy * vy + vz * vz);
            F[4] = mass_flux * utot + p * vn;
#pragma unroll
            for (int s = 0; s < NSP; ++s) F[5 + s] = mass_flux * tp_val(cr, P_Y + s);
        } else {
            F[1] = p;
        }
    } else if ((pos == 0 && tm == BC_SLIP_WALL_1) || (pos == n && tpl == BC_SLIP_WALL_1)) {
        const bool left = (pos == 0);
        const TPRef& c0 = left ? r[2] : r[1];
        const TPRef& c1 = left ? r[3] : r[0];
        TPCell<NSP> S = tp_cell<NSP>(c0, pn, pt1, pt2);
        if (ho && !clip) {
            const double w0 = A.bc->ext_w[d][0], w1 = A.bc->ext_w[d][1];
            TPState<NSP> Q;
            double v[3];
            v[0] = tp_vel(c0, pn) * w0 + tp_vel(c1, pn) * w1;
            v[1] = tp_vel(c0, pt1) * w0 + tp_vel(c1, pt1) * w1;
            v[2] = tp_vel(c0, pt2) * w0 + tp_vel(c1, pt2) * w1;
#pragma unroll
            for (int s = 0; s < NSP; ++s) Q.Y[s] = tp_val(c0, P_Y + s) * w0 + tp_val(c1, P_Y + s) * w1;
            if (scale_mass_fractions<MA, NSP>(Q.Y) != 0) {
#pragma unroll
                for (int s = 0; s < NSP; ++s) Q.Y[s] = tp_val(c0, P_Y + s);
            }
            Q.rho = tp_val(c0, P_RHO) * w0 + tp_val(c1, P_RHO) * w1;
            Q.u = tp_val(c0, P_U) * w0 + tp_val(c1, P_U) * w1;
            Q.T = S.T;
            if (tp_face_eos<MA, NSP, kPolish>(g, Q, &fe, &fb) == 0) {
                S.rho = Q.rho; S.u = Q.u; S.T = Q.T; S.p = Q.p;
                S.v[0] = v[0]; S.v[1] = v[1]; S.v[2] = v[2];
#pragma unroll
                for (int s = 0; s < NSP; ++s) S.Y[s] = Q.Y[s];
            }
        }
        F[1] = wall_pressure(S.rho, S.p, S.a, S.v[0], tp_gamma<MA, NSP>(g, S.Y, S.T), left);
    } else {
        // l1r2 (kind 1: the face after a wall on the minus side) or l2r1 (kind 2)
        const int kind = (pos == 1 && tm == BC_SLIP_WALL_1) ? 1 : 2;
        TPCell<NSP> Lf, Rf;
        bool okL = false, okR = false;
        if (ho) {
            const L2R2W& w = A.w[d];
            auto lin = [&](double q0, double q1) {
                double q = q0 * 0.5 + q1 * 0.5;
                return clip ? clip_limits<MA>(q, q0, q1) : q;
            };
            auto recon = [&](double q0, double q1, double q2, double q3, double& qL, double& qR) {
                double unused;
                if (kind == 1) {
                    interp_l2r2_scalar<MA>(w, lim, clip, q1, q1, q2, q3, unused, qR);
                    const double del = (q2 - q1) * w.two_over_lenR0_plus_lenL0;
                    const double delRplus = (q3 - q2) * w.two_over_lenR1_plus_lenR0;
                    qL = (lim && delRplus * del < 0.0) ? q1 : lin(q1, q2);
                } else {
                    interp_l2r2_scalar<MA>(w, lim, clip, q0, q1, q2, q2, qL, unused);
                    const double delLminus = (q1 - q0) * w.two_over_lenL0_plus_lenL1;
                    const double del = (q2 - q1) * w.two_over_lenR0_plus_lenL0;
                    qR = (lim && delLminus * del < 0.0) ? q2 : lin(q1, q2);
                }
            };
            const int comp[3] = {pn, pt1, pt2};
            for (int dd = 0; dd < 3; ++dd)
                recon(tp_vel(r[0], comp[dd]), tp_vel(r[1], comp[dd]), tp_vel(r[2], comp[dd]), tp_vel(r[3], comp[dd]),
                      Lf.v[dd], Rf.v[dd]);
            for (int s = 0; s < NSP; ++s)
                recon(tp_val(r[0], P_Y + s), tp_val(r[1], P_Y + s), tp_val(r[2], P_Y + s), tp_val(r[3], P_Y + s),
                      Lf.Y[s], Rf.Y[s]);
            if (scale_mass_fractions<MA, NSP>(Lf.Y) != 0)
                for (int s = 0; s < NSP; ++s) Lf.Y[s] = tp_val(r[1], P_Y + s);
            if (scale_mass_fractions<MA, NSP>(Rf.Y) != 0)
                for (int s = 0; s < NSP; ++s) Rf.Y[s] = tp_val(r[2], P_Y + s);
            recon(tp_val(r[0], P_RHO), tp_val(r[1], P_RHO), tp_val(r[2], P_RHO), tp_val(r[3], P_RHO), Lf.rho, Rf.rho);
            recon(tp_val(r[0], P_U), tp_val(r[1], P_U), tp_val(r[2], P_U), tp_val(r[3], P_U), Lf.u, Rf.u);
            for (int sd = 0; sd < 2; ++sd) {
                const TPRef rc = sd ? r[2] : r[1];
                TPState<NSP> Q;
                Q.rho = sd ? Rf.rho : Lf.rho;
                Q.u = sd ? Rf.u : Lf.u;
                Q.T = tp_val(rc, P_T);
                for (int s = 0; s < NSP; ++s) Q.Y[s] = sd ? Rf.Y[s] : Lf.Y[s];
                if (tp_face_eos<MA, NSP, kPolish>(g, Q, &fe, &fb) == 0) {
                    const double a = tp_val(rc, P_A);
                    if (sd == 0) { Lf.T = Q.T; Lf.u = Q.u; Lf.p = Q.p; Lf.a = a; okL = true; }
                    else { Rf.T = Q.T; Rf.u = Q.u; Rf.p = Q.p; Rf.a = a; okR = true; }
                }
            }
        }
        if (!okL) Lf = tp_cell<NSP>(r[1], pn, pt1, pt2);
        if (!okR) Rf = tp_cell<NSP>(r[2], pn, pt1, pt2);
        const bool shock = (FLUX == FLUX_SHOCK) && ((r[1].s && *r[1].s) || (r[2].s && *r[2].s));
        tp_flux<MA, NSP, FLUX>(A, g, Lf, Rf, shock, F);
    }
    const int nfi = A.ni + (d == 0), nfj = A.nj + (d == 1);
    const long f = ((long)k * nfj + j) * nfi + i;
    using FS = TPFluxStore<M>;
    FS* dst = reinterpret_cast<FS*>(A.F[d]);
    const long fq = A.fqs[d];
    dst[f] = FS(F[0]);
    dst[(1 + pn) * fq + f] = FS(F[1]);
    dst[(1 + pt1) * fq + f] = FS(F[2]);
    dst[(1 + pt2) * fq + f] = FS(F[3]);
    dst[4 * fq + f] = FS(F[4]);
#pragma unroll
    for (int s = 0; s < NSP; ++s) dst[(5 + s) * fq + f] = FS(F[5 + s]);
}

// x_order-1 ghost cells of one face of a mixture block (extrapolate_copy.d:
// 102-186): layer 1 linear from the boundary cell c1 and the cell behind it,
// layer 2 from layer 1 and c1, in rho, u, velocity and mass fractions
// (scaled); T by the EOS from the ghost's previous T, then p.  The sound
// speed is not written (the buffer holds the constructed cell's).
template <class M>
__global__ __launch_bounds__(128) void tp_extrap1_kernel(TPArgs A, int face)
{
    using MA = typename TPArith<M>::type;
    constexpr int NSP = 5;
    if (stage_aborted(A.abort_prev)) return;
    const int dir = (face == FACE_EAST || face == FACE_WEST) ? 0 : ((face == FACE_NORTH || face == FACE_SOUTH) ? 1 : 2);
    const bool plus = (face == FACE_EAST || face == FACE_NORTH || face == FACE_TOP);
    const int na = (dir == 0) ? A.nj : A.ni;
    const int nb = (dir == 2) ? A.nj : A.nk;
    const long nfc = (long)na * nb;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nfc) return;
    const int fa = (int)(t % na), fb = (int)(t / na);
    const int n = (dir == 0) ? A.ni : ((dir == 1) ? A.nj : A.nk);
    auto cell = [&](int p) {
        int ci, cj, ck;
        if (dir == 0) { ci = p; cj = fa; ck = fb; }
        else if (dir == 1) { ci = fa; cj = p; ck = fb; }
        else { ci = fa; cj = fb; ck = p; }
        return A.Pin + ((long)ck * A.nj + cj) * A.ni + ci;
    };
    double* G = A.gx[face];
    const long gs = 2 * nfc;
    const int e = plus ? n - 1 : 0, in = plus ? -1 : 1;
    const double* c1 = cell(e);
    for (int layer = 1; layer <= 2; ++layer) {
        const double* p1 = (layer == 1) ? c1 : G + t;
        const long s1 = (layer == 1) ? A.qs : gs;
        const double* p2 = (layer == 1) ? cell(e + in) : c1;
        const long s2 = (layer == 1) ? A.qs : A.qs;
        double* dst = G + (layer - 1) * nfc + t;
        TPState<NSP> Q;
        Q.rho = 2.0 * p1[P_RHO * s1] - p2[P_RHO * s2];
        Q.u = 2.0 * p1[P_U * s1] - p2[P_U * s2];
#pragma unroll
        for (int s = 0; s < NSP; ++s) Q.Y[s] = 2.0 * p1[(P_Y + s) * s1] - p2[(P_Y + s) * s2];
        scale_mass_fractions<MA, NSP>(Q.Y);
        Q.T = dst[P_T * gs];
        Q.p = dst[P_P * gs];
        int fe = 0;
        tp_update_thermo_from_rhou<MA, NSP>(*A.gas, Q, &fe);
        const double vx = 2.0 * p1[P_VX * s1] - p2[P_VX * s2];
        const double vy = 2.0 * p1[P_VY * s1] - p2[P_VY * s2];
        const double vz = 2.0 * p1[P_VZ * s1] - p2[P_VZ * s2];
        dst[P_RHO * gs] = Q.rho; dst[P_U * gs] = Q.u; dst[P_T * gs] = Q.T; dst[P_P * gs] = Q.p;
        dst[P_VX * gs] = vx; dst[P_VY * gs] = vy; dst[P_VZ * gs] = vz;
#pragma unroll
        for (int s = 0; s < NSP; ++s) dst[(P_Y + s) * gs] = Q.Y[s];
    }
}

template <class M>
void launch_tp_extrap1_m(const TPArgs& A, int faces, hipStream_t stream)
{
    for (int f = 0; f < 6; ++f) {
        if (!(faces & (1 << f))) continue;
        const int dir = (f == FACE_EAST || f == FACE_WEST) ? 0 : ((f == FACE_NORTH || f == FACE_SOUTH) ? 1 : 2);
        const long nfc = (long)((dir == 0) ? A.nj : A.ni) * ((dir == 2) ? A.nj : A.nk);
        hipLaunchKernelGGL((tp_extrap1_kernel<M>), dim3((unsigned)((nfc + 127) / 128)), dim3(128), 0, stream, A, f);
    }
}

// Launch of one stage of a mixture block: the face kernels of every
// direction, then the update.
template <class M, int FLUX>
void launch_tp_faces(const TPArgs& A, hipStream_t stream)
{
    const int nfk = (A.dims == 3) ? A.nk : 1;
    const long nf0 = (long)(A.ni + 1) * A.nj * nfk, nf1 = (long)A.ni * (A.nj + 1) * nfk;
    const long nf2 = (long)A.ni * A.nj * (A.nk + 1);
    const dim3 b(128);
    auto g = [](long n) { return dim3((unsigned)((n + 127) / 128)); };
    hipLaunchKernelGGL((tp_face_kernel<M, 5, FLUX, 0, false>), g(nf0), b, 0, stream, A);
    hipLaunchKernelGGL((tp_face_kernel<M, 5, FLUX, 1, false>), g(nf1), b, 0, stream, A);
    if (A.dims == 3) hipLaunchKernelGGL((tp_face_kernel<M, 5, FLUX, 2, false>), g(nf2), b, 0, stream, A);
}

// The retry kernel after the mixed form's face kernels (it must finish the
// handed-on faces before the boundary-face kernel overwrites its faces):
// `parts` workgroups per block.
template <int FLUX, bool BATCH>
void launch_tp_retry_f(const typename std::conditional<BATCH, BatchArgs, TPArgs>::type& a, int nblk, int parts,
                       hipStream_t stream)
{
    hipLaunchKernelGGL((tp_retry_kernel<5, FLUX, BATCH>), dim3((unsigned)(nblk * parts)), dim3(128), 0, stream, a,
                       parts);
}

template <bool BATCH>
void launch_tp_retry(const typename std::conditional<BATCH, BatchArgs, TPArgs>::type& a, int nblk, int parts,
                     hipStream_t stream, int flux)
{
    if (flux == FLUX_HLLC) launch_tp_retry_f<FLUX_HLLC, BATCH>(a, nblk, parts, stream);
    else if (flux == FLUX_SHOCK) launch_tp_retry_f<FLUX_SHOCK, BATCH>(a, nblk, parts, stream);
    else if (flux == FLUX_OTHER) launch_tp_retry_f<FLUX_OTHER, BATCH>(a, nblk, parts, stream);
    else launch_tp_retry_f<FLUX_AUSMDV, BATCH>(a, nblk, parts, stream);
}

// workgroups of the retry kernel per block
constexpr int TP_RETRY_PARTS = 512;
constexpr int TP_RETRY_PARTS_BATCH = 64;

template <class M>
void launch_tp_stage_m(const TPArgs& A, hipStream_t stream, int flux)
{
    if (flux == FLUX_HLLC) launch_tp_faces<M, FLUX_HLLC>(A, stream);
    else if (flux == FLUX_SHOCK) launch_tp_faces<M, FLUX_SHOCK>(A, stream);
    else if (flux == FLUX_OTHER) launch_tp_faces<M, FLUX_OTHER>(A, stream);
    else launch_tp_faces<M, FLUX_AUSMDV>(A, stream);
    if constexpr (std::is_same<M, ModeMixed>::value) launch_tp_retry<false>(A, 1, TP_RETRY_PARTS, stream, flux);
    if (A.fix_off[6] > 0) {
        const dim3 g((unsigned)((A.fix_off[6] + 63) / 64)), b(64);
        if (flux == FLUX_HLLC) hipLaunchKernelGGL((tp_boundary_face_kernel<M, 5, FLUX_HLLC>), g, b, 0, stream, A);
        else if (flux == FLUX_SHOCK) hipLaunchKernelGGL((tp_boundary_face_kernel<M, 5, FLUX_SHOCK>), g, b, 0, stream, A);
        else if (flux == FLUX_OTHER) hipLaunchKernelGGL((tp_boundary_face_kernel<M, 5, FLUX_OTHER>), g, b, 0, stream, A);
        else hipLaunchKernelGGL((tp_boundary_face_kernel<M, 5, FLUX_AUSMDV>), g, b, 0, stream, A);
    }
    const long nc = (long)A.ni * A.nj * ((A.dims == 3) ? A.nk : 1);
    const dim3 grid((unsigned)((nc + 127) / 128)), b(128);
    if (A.guess) hipLaunchKernelGGL((tp_update_kernel<M, 5, true, false>), grid, b, 0, stream, A);
    else hipLaunchKernelGGL((tp_update_kernel<M, 5, false, false>), grid, b, 0, stream, A);
}

// The stage of several mixture blocks of one dimensionality in one launch per
// kernel: the face kernels of each direction over all blocks' faces, then each
// block's boundary-face kernel (when it has faces to recompute; `per_block`
// holds their arguments), then the update over all blocks' cells.  `wg`
// holds each launch's workgroup count and `B[k]` its map (k = direction,
// 3 = update).
template <class M, int FLUX>
void launch_tp_batch_faces(const BatchArgs* B, const int* wg, int dims, hipStream_t stream)
{
    const dim3 b(128);
    hipLaunchKernelGGL((tp_face_kernel<M, 5, FLUX, 0, true>), dim3(wg[0]), b, 0, stream, B[0]);
    hipLaunchKernelGGL((tp_face_kernel<M, 5, FLUX, 1, true>), dim3(wg[1]), b, 0, stream, B[1]);
    if (dims == 3) hipLaunchKernelGGL((tp_face_kernel<M, 5, FLUX, 2, true>), dim3(wg[2]), b, 0, stream, B[2]);
}

template <class M>
void launch_tp_stage_batch_m(const BatchArgs* B, const int* wg, int dims, bool guess, const TPArgs* per_block,
                             int nper, int nblk, hipStream_t stream, int flux)
{
    if (flux == FLUX_HLLC) launch_tp_batch_faces<M, FLUX_HLLC>(B, wg, dims, stream);
    else if (flux == FLUX_SHOCK) launch_tp_batch_faces<M, FLUX_SHOCK>(B, wg, dims, stream);
    else if (flux == FLUX_OTHER) launch_tp_batch_faces<M, FLUX_OTHER>(B, wg, dims, stream);
    else launch_tp_batch_faces<M, FLUX_AUSMDV>(B, wg, dims, stream);
    if constexpr (std::is_same<M, ModeMixed>::value)
        launch_tp_retry<true>(B[0], nblk, TP_RETRY_PARTS_BATCH, stream, flux);
    for (int n = 0; n < nper; ++n) {
        const TPArgs& A = per_block[n];
        const dim3 g((unsigned)((A.fix_off[6] + 63) / 64)), b(64);
        if (flux == FLUX_HLLC) hipLaunchKernelGGL((tp_boundary_face_kernel<M, 5, FLUX_HLLC>), g, b, 0, stream, A);
        else if (flux == FLUX_SHOCK) hipLaunchKernelGGL((tp_boundary_face_kernel<M, 5, FLUX_SHOCK>), g, b, 0, stream, A);
        else if (flux == FLUX_OTHER) hipLaunchKernelGGL((tp_boundary_face_kernel<M, 5, FLUX_OTHER>), g, b, 0, stream, A);
        else hipLaunchKernelGGL((tp_boundary_face_kernel<M, 5, FLUX_AUSMDV>), g, b, 0, stream, A);
    }
    if (guess) hipLaunchKernelGGL((tp_update_kernel<M, 5, true, true>), dim3(wg[3]), dim3(128), 0, stream, B[3]);
    else hipLaunchKernelGGL((tp_update_kernel<M, 5, false, true>), dim3(wg[3]), dim3(128), 0, stream, B[3]);
}

template <class M>
void launch_tp_shock_detect_m(const TPArgs& A, hipStream_t stream)
{
    const long nc = (long)A.ni * A.nj * ((A.dims == 3) ? A.nk : 1);
    hipLaunchKernelGGL((tp_shock_detect_kernel<M>), dim3((unsigned)((nc + 127) / 128)), dim3(128), 0, stream, A);
}

} // namespace e4
