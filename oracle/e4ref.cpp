// e4ref.cpp -- CPU FP64 restatement of Eilmer4's explicit structured-block
// gas-dynamic update.  TEST INFRASTRUCTURE ONLY: this file is the oracle that
// the HIP path is checked against.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it, and only as the checker or as the
// timed CPU baseline.  The product library never links or calls it.
//
// Every function names the reference file:line it restates (paths relative to
// the reference's src/ directory).  Build: oracle/Makefile
// (g++ -O2 -ffp-contract=off, so that no FMA contraction changes roundings).
//
// Scope: uniform Cartesian structured blocks (dx, dy, dz per block), 2D or 3D,
// two ghost-cell layers; ideal gas and thermally perfect mixtures (NASA-9 / CEA
// curves); MUSCL l2r2 reconstruction with van Albada limiting and extrema
// clipping in the face-local frame; AUSMDV and HLLC fluxes; ghost-cell BCs
// (flowstate copy, extrapolate copy, internal copy then reflect, full-face
// copy between blocks in any of the reference's face/orientation pairings, and
// periodic as a block connected to itself); time derivatives, the RK stage
// updates, decode, invalid-cell counting and the CFL signal; several connected
// blocks stepped together.

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>
#include <utility>

extern "C" {

enum { NSP_MAX = 8, NSEG_MAX = 4, NQ_MAX = 5 + NSP_MAX };
enum { GAS_IDEAL = 0, GAS_TP = 1 };
// FluxCalculator values follow the reference's enum order (eilmer/fvcore.d:184-194);
// hllc is the new member appended after roe.
enum { FLUX_AUSMDV = 0, FLUX_HANEL = 1, FLUX_EFM = 2, FLUX_AUSM_PLUS_UP = 3, FLUX_ADAPTIVE_EFM_AUSMDV = 4,
       FLUX_ADAPTIVE_HANEL_AUSMDV = 5, FLUX_ROE = 8, FLUX_HLLC = 9 };
// GasdynamicUpdate (eilmer/fvcore.d:30-41).
enum { UPD_EULER = 0, UPD_PC = 1, UPD_MIDPOINT = 2, UPD_CLASSIC_RK3 = 3, UPD_TVD_RK3 = 4, UPD_DENMAN_RK3 = 5 };
// InterpolateOption (eilmer/fvcore.d:155).
enum { INTERP_PT = 0, INTERP_RHOU = 1, INTERP_RHOP = 2, INTERP_RHOT = 3 };
// Boundary kinds (ghost-cell effects, eilmer/bc/ghost_cell_effect/*).
// 5: OutFlowBC_SimpleFlux (ghosts by x_order-0 extrapolation, boundary-face flux
// by BFE_SimpleOutflowFlux); 6: WallBC_WithSlip1 (no ghost-cell data, one-sided
// reconstruction and the wall flux); 7: OutFlowBC_SimpleExtrapolate x_order 1;
// 8: InFlowBC_StaticProfile (a flow state per ghost cell)
enum { BC_SLIP_WALL = 0, BC_INFLOW = 1, BC_EXTRAPOLATE = 2, BC_PERIODIC = 3, BC_CONNECTED = 4,
       BC_OUTFLOW_SIMPLE_FLUX = 5, BC_SLIP_WALL_1 = 6, BC_EXTRAPOLATE_1 = 7, BC_INFLOW_PROFILE = 8 };
// Face order of a cell (eilmer/fvcore.d Face enum): north, east, south, west, top, bottom.
enum { F_NORTH = 0, F_EAST = 1, F_SOUTH = 2, F_WEST = 3, F_TOP = 4, F_BOTTOM = 5 };

static const double R_universal = 8.31451;     // gas/physical_constants.d:13
static const double T_MIN = 20.0;              // gas/gas_model.d:45
static const double T_MAX = 100000.0;          // gas/gas_model.d:46
static const double epsilon_van_albada = 1.0e-12; // eilmer/onedinterp.d:22
// Blocks (and state arrays) below this many cells run their loops on one
// thread: a thread team per loop costs more than such a loop's work.
static const int OMP_MIN_CELLS = 32768;

struct CEACurve {          // gas/thermo/cea_thermo_curves.d:22-169
    double R;
    int nseg;
    double Tb[NSEG_MAX + 1];
    double Tbl[NSEG_MAX];
    double c[NSEG_MAX][9];
};

struct Gas {
    int type;
    int nsp;
    // ideal gas (gas/ideal_gas.d:31-72)
    double Rgas, gamma, Cv, Cvinv, Cp;
    // thermally perfect mix (gas/therm_perf_gas.d:41-110)
    double R[NSP_MAX];
    CEACurve curve[NSP_MAX];
};

struct GasState { double rho, p, T, u, a; double massf[NSP_MAX]; };
// S: the shock indicator (flowstate.d:44), 0 or 1; copied with the flow state
struct FlowState { GasState gas; double vel[3]; int S; };

// ---------------------------------------------------------------- gas models

Gas* ref_gas_new_ideal(double mMass, double gamma)
{
    Gas* g = (Gas*)calloc(1, sizeof(Gas));
    g->type = GAS_IDEAL; g->nsp = 1;
    g->Rgas = R_universal / mMass;                // ideal_gas.d:65
    g->gamma = gamma;
    g->Cv = g->Rgas / (gamma - 1.0);              // ideal_gas.d:66
    g->Cvinv = 1.0 / g->Cv;                       // ideal_gas.d:67
    g->Cp = g->Rgas * gamma / (gamma - 1.0);      // ideal_gas.d:68
    return g;
}

// coeffs: [nsp][nseg][9]; Tbreaks: [nsp][nseg+1]; Tblends: [nsp][nseg-1]
Gas* ref_gas_new_tp(int nsp, const double* M, int nseg, const double* Tbreaks,
                    const double* Tblends, const double* coeffs)
{
    if (nsp < 1 || nsp > NSP_MAX || nseg < 1 || nseg > NSEG_MAX) return nullptr;
    Gas* g = (Gas*)calloc(1, sizeof(Gas));
    g->type = GAS_TP; g->nsp = nsp;
    for (int s = 0; s < nsp; ++s) {
        g->R[s] = R_universal / M[s];             // therm_perf_gas.d:68
        CEACurve& cv = g->curve[s];
        cv.R = g->R[s]; cv.nseg = nseg;
        for (int b = 0; b <= nseg; ++b) cv.Tb[b] = Tbreaks[s * (nseg + 1) + b];
        for (int b = 0; b < nseg - 1; ++b) cv.Tbl[b] = Tblends[s * (nseg - 1) + b];
        for (int q = 0; q < nseg; ++q)
            for (int j = 0; j < 9; ++j) cv.c[q][j] = coeffs[(s * nseg + q) * 9 + j];
    }
    return g;
}

void ref_gas_free(Gas* g) { free(g); }

// cea_thermo_curves.d:136-169 (returns 0 on success)
static int determineCoefficients(const CEACurve& cv, double T, double a[9])
{
    const int nb = cv.nseg + 1;
    if (T < (cv.Tb[1] - 0.5 * cv.Tbl[0])) { for (int j = 0; j < 9; ++j) a[j] = cv.c[0][j]; return 0; }
    if (T > (cv.Tb[nb - 2] + 0.5 * cv.Tbl[nb - 3])) { for (int j = 0; j < 9; ++j) a[j] = cv.c[cv.nseg - 1][j]; return 0; }
    for (int i = 1; i < nb - 1; ++i) {
        double T_blend_low = cv.Tb[i] - 0.5 * cv.Tbl[i - 1];
        double T_blend_high = cv.Tb[i] + 0.5 * cv.Tbl[i - 1];
        if (T >= T_blend_low && T <= T_blend_high) {
            double wB = (1. / cv.Tbl[i - 1]) * (T - T_blend_low);
            double wA = 1.0 - wB;
            for (int j = 0; j < 9; ++j) a[j] = wA * cv.c[i - 1][j] + wB * cv.c[i][j];
            return 0;
        }
        if (T > T_blend_high && T < (cv.Tb[i + 1] - 0.5 * cv.Tbl[i])) {
            for (int j = 0; j < 9; ++j) a[j] = cv.c[i][j];
            return 0;
        }
    }
    return -1;
}

// cea_thermo_curves.d:48-62
static double eval_Cp(const CEACurve& cv, double T)
{
    if (T < cv.Tb[0]) T = cv.Tb[0];
    if (T > cv.Tb[cv.nseg]) T = cv.Tb[cv.nseg];
    double a[9];
    determineCoefficients(cv, T, a);
    double Cp_on_R = a[0] / (T * T) + a[1] / T + a[2] + a[3] * T;
    Cp_on_R += a[4] * T * T + a[5] * T * T * T + a[6] * T * T * T * T;
    return cv.R * Cp_on_R;
}

// cea_thermo_curves.d:63-91
static double eval_h(const CEACurve& cv, double T)
{
    if (T < cv.Tb[0]) {
        double T_low = cv.Tb[0];
        double h_low = eval_h(cv, T_low);
        double Cp_low = eval_Cp(cv, T_low);
        return h_low - Cp_low * (T_low - T);
    }
    if (T > cv.Tb[cv.nseg]) {
        double T_high = cv.Tb[cv.nseg];
        double h_high = eval_h(cv, T_high);
        double Cp_high = eval_Cp(cv, T_high);
        return h_high + Cp_high * (T - T_high);
    }
    double a[9];
    determineCoefficients(cv, T, a);
    double h_on_RT = -a[0] / (T * T) + a[1] / T * log(T) + a[2] + a[3] * T / 2.0;
    h_on_RT += a[4] * T * T / 3.0 + a[5] * T * T * T / 4.0 + a[6] * T * T * T * T / 5.0 + a[7] / T;
    return cv.R * T * h_on_RT;
}

double ref_cea_eval_h(Gas* g, int isp, double T) { return eval_h(g->curve[isp], T); }
double ref_cea_eval_Cp(Gas* g, int isp, double T) { return eval_Cp(g->curve[isp], T); }

// gas/gas_model.d:330-338
static double mass_average(const Gas* g, const GasState& Q, const double* phi)
{
    double result = 0.0;
    for (int i = 0; i < g->nsp; ++i) result += Q.massf[i] * phi[i];
    return result;
}

// therm_perf_gas_mix_eos.d:63-69
static void tp_update_energy(const Gas* g, GasState& Q)
{
    double vals[NSP_MAX];
    for (int s = 0; s < g->nsp; ++s) vals[s] = eval_h(g->curve[s], Q.T) - g->R[s] * Q.T;
    Q.u = mass_average(g, Q, vals);
}

// nm/newton.d:40-107 specialised to the zeroFn/dzdT closures of
// therm_perf_gas_mix_eos.d:104-123 (including their side effects on Q.T, Q.u).
// Returns 0 and the root in *root, or -1 when the solve throws.
static int tp_newton_solve(Gas* g, GasState& Q, double e_tgt, double x0, double xMin, double xMax,
                           double tol, double* root)
{
    auto f = [&](double T) { Q.T = T; tp_update_energy(g, Q); return e_tgt - Q.u; };
    auto dfdx = [&](double T) {
        double Cv[NSP_MAX];
        for (int s = 0; s < g->nsp; ++s) Cv[s] = eval_Cp(g->curve[s], T) - g->R[s];
        return -1.0 * mass_average(g, Q, Cv);
    };
    const int MAXIT = 30;
    double xL = xMin, xH = xMax;
    double fL = f(xL);
    double fH = f(xH);
    if ((fL > 0.0 && fH > 0.0) || (fL < 0.0 && fH < 0.0)) return -1;
    if (fL == 0.0) { *root = xMin; return 0; }
    if (fH == 0.0) { *root = xMax; return 0; }
    if (fL < 0.0) { xL = xMin; xH = xMax; } else { xH = xMin; xL = xMax; }
    double rts = x0;
    double dxold = (xMax - xMin);
    double dx = dxold;
    double f0 = f(rts);
    double df0 = dfdx(rts);
    for (int j = 0; j < MAXIT; ++j) {
        if ((((rts - xH) * df0 - f0) * ((rts - xL) * df0 - f0) > 0.0) ||
            (fabs(2.0 * f0) > fabs(dxold * df0))) {
            dxold = dx;
            dx = 0.5 * (xH - xL);
            rts = xL + dx;
            if (xL == rts) { *root = rts; return 0; }
        } else {
            dxold = dx;
            dx = f0 / df0;
            double tmp = rts;
            rts -= dx;
            if (tmp == rts) { *root = rts; return 0; }
        }
        if (fabs(dx) < tol) { *root = rts; return 0; }
        f0 = f(rts);
        df0 = dfdx(rts);
        if (f0 < 0.0) xL = rts; else xH = rts;
    }
    return -1;
}

// Generic Newton self-test of nm/newton.d:109-141 (cubic and a*x+sin(x)-exp(x)).
int ref_newton_selftest(int which, double x0, double xMin, double xMax, double* root)
{
    auto f1 = [](double x) { return pow(x, 3) + pow(x, 2) - 3 * x - 3; };
    auto d1 = [](double x) { return 3.0 * pow(x, 2) + 2.0 * x - 3.0; };
    auto f2 = [](double x) { return 3.0 * x + sin(x) - exp(x); };
    auto d2 = [](double x) { return 3.0 + cos(x) - exp(x); };
    double tol = 1.0e-9;
    const int MAXIT = 30;
    double xL = xMin, xH = xMax;
    auto F = [&](double x) { return which == 1 ? f1(x) : f2(x); };
    auto D = [&](double x) { return which == 1 ? d1(x) : d2(x); };
    double fL = F(xL), fH = F(xH);
    if ((fL > 0.0 && fH > 0.0) || (fL < 0.0 && fH < 0.0)) return -1;
    if (fL == 0.0) { *root = xMin; return 0; }
    if (fH == 0.0) { *root = xMax; return 0; }
    if (fL < 0.0) { xL = xMin; xH = xMax; } else { xH = xMin; xL = xMax; }
    double rts = x0, dxold = xMax - xMin, dx = dxold, f0 = F(rts), df0 = D(rts);
    for (int j = 0; j < MAXIT; ++j) {
        if ((((rts - xH) * df0 - f0) * ((rts - xL) * df0 - f0) > 0.0) || (fabs(2.0 * f0) > fabs(dxold * df0))) {
            dxold = dx; dx = 0.5 * (xH - xL); rts = xL + dx;
            if (xL == rts) { *root = rts; return 0; }
        } else {
            dxold = dx; dx = f0 / df0; double tmp = rts; rts -= dx;
            if (tmp == rts) { *root = rts; return 0; }
        }
        if (fabs(dx) < tol) { *root = rts; return 0; }
        f0 = F(rts); df0 = D(rts);
        if (f0 < 0.0) xL = rts; else xH = rts;
    }
    return -1;
}

// therm_perf_gas_mix_eos.d:70-153
static int tp_update_temperature(Gas* g, GasState& Q)
{
    double Tsave = Q.T;
    double TOL = 1.0e-6;
    double e_tgt = Q.u;
    double delT = 1000.0;
    double T1 = fmax(Q.T - 0.5 * delT, T_MIN);
    double T2 = T1 + delT;
    double root;
    if (tp_newton_solve(g, Q, e_tgt, Tsave, T1, T2, TOL, &root) == 0) { Q.T = root; return 0; }
    if (tp_newton_solve(g, Q, e_tgt, Tsave, T_MIN, T_MAX, TOL, &root) == 0) { Q.T = root; return 0; }
    Q.T = Tsave;
    tp_update_energy(g, Q);
    return -1;
}

// thermo/perf_gas_mix_eos.d:43-62, 86-94
static double heavy_R(const Gas* g, const GasState& Q) { return mass_average(g, Q, g->R); }

// ---- GasModel interface (ideal_gas.d:84-144; therm_perf_gas.d:177-196, 345-385)
static int update_thermo_from_rhou(Gas* g, GasState& Q)
{
    if (g->type == GAS_IDEAL) {
        if (Q.u <= 0.0 || Q.rho <= 0.0) return -1;
        Q.T = Q.u * g->Cvinv;
        Q.p = Q.rho * g->Rgas * Q.T;
        return 0;
    }
    if (tp_update_temperature(g, Q) != 0) return -1;
    double Rmix = heavy_R(g, Q);
    Q.p = Q.rho * Rmix * Q.T;
    return 0;
}

static int update_thermo_from_pT(Gas* g, GasState& Q)
{
    if (g->type == GAS_IDEAL) {
        if (Q.T <= 0.0 || Q.p <= 0.0) return -1;
        Q.rho = Q.p / (Q.T * g->Rgas);
        Q.u = g->Cv * Q.T;
        return 0;
    }
    double Rmix = heavy_R(g, Q);
    double denom = Rmix * Q.T;
    Q.rho = Q.p / denom;
    tp_update_energy(g, Q);
    return 0;
}

static int update_thermo_from_rhoT(Gas* g, GasState& Q)
{
    if (g->type == GAS_IDEAL) {
        if (Q.T <= 0.0 || Q.rho <= 0.0) return -1;
        Q.p = Q.rho * g->Rgas * Q.T;
        Q.u = g->Cv * Q.T;
        return 0;
    }
    tp_update_energy(g, Q);
    double Rmix = heavy_R(g, Q);
    Q.p = Q.rho * Rmix * Q.T;
    return 0;
}

static int update_thermo_from_rhop(Gas* g, GasState& Q)
{
    if (g->type == GAS_IDEAL) {
        if (Q.p <= 0.0 || Q.rho <= 0.0) return -1;
        Q.T = Q.p / (Q.rho * g->Rgas);
        Q.u = g->Cv * Q.T;
        return 0;
    }
    double Rmix = heavy_R(g, Q);
    Q.T = Q.p / (Q.rho * Rmix);
    tp_update_energy(g, Q);
    return 0;
}

static int update_sound_speed(Gas* g, GasState& Q)
{
    if (g->type == GAS_IDEAL) {
        if (Q.T <= 0.0) return -1;
        Q.a = sqrt(g->gamma * g->Rgas * Q.T);
        return 0;
    }
    double Cp[NSP_MAX], Cv[NSP_MAX];
    for (int i = 0; i < g->nsp; ++i) Cp[i] = eval_Cp(g->curve[i], Q.T);
    double Cpm = mass_average(g, Q, Cp);
    for (int i = 0; i < g->nsp; ++i) Cv[i] = eval_Cp(g->curve[i], Q.T) - g->R[i];
    double Cvm = mass_average(g, Q, Cv);
    double R = mass_average(g, Q, g->R);
    Q.a = sqrt(Cpm / Cvm * (R * Q.T));
    return 0;
}

// gas/gas_model.d:296-328 (tolerance 0.0, assert_error_tolerance 0.1)
static int scale_mass_fractions(double* massf, int nsp)
{
    if (nsp == 1) {
        if (fabs(massf[0] - 1.0) > 0.1) return -1;
        massf[0] = 1.0;
        return 0;
    }
    double massf_sum = 0.0;
    for (int i = 0; i < nsp; ++i) {
        massf[i] = massf[i] >= 0.0 ? massf[i] : 0.0;
        massf_sum += massf[i];
    }
    if (fabs(massf_sum - 1.0) > 0.1) return -1;
    if (fabs(massf_sum - 1.0) > 0.0) {
        for (int i = 0; i < nsp; ++i) massf[i] /= massf_sum;
    }
    return 0;
}

// Array-of-state entry points for the known-answer tests.  Q layout:
// [rho, p, T, u, a, massf[0..nsp-1]].
static void unpackQ(const Gas* g, const double* q, GasState& Q)
{
    Q.rho = q[0]; Q.p = q[1]; Q.T = q[2]; Q.u = q[3]; Q.a = q[4];
    for (int i = 0; i < g->nsp; ++i) Q.massf[i] = q[5 + i];
}
static void packQ(const Gas* g, const GasState& Q, double* q)
{
    q[0] = Q.rho; q[1] = Q.p; q[2] = Q.T; q[3] = Q.u; q[4] = Q.a;
    for (int i = 0; i < g->nsp; ++i) q[5 + i] = Q.massf[i];
}
int ref_gas_update(Gas* g, int which, double* q)
{
    GasState Q; unpackQ(g, q, Q);
    int rc = -9;
    switch (which) {
    case 0: rc = update_thermo_from_pT(g, Q); break;
    case 1: rc = update_thermo_from_rhou(g, Q); break;
    case 2: rc = update_thermo_from_rhoT(g, Q); break;
    case 3: rc = update_thermo_from_rhop(g, Q); break;
    case 4: rc = update_sound_speed(g, Q); break;
    case 5: tp_update_energy(g, Q); rc = 0; break;
    }
    packQ(g, Q, q);
    return rc;
}
// Batched T(rho,u,Y) decode used to build the temperature tables:
// q is [n][5+nsp]; returns number of failures.
int ref_gas_update_many(Gas* g, int which, int n, double* q)
{
    int bad = 0;
    const int w = 5 + g->nsp;
    #pragma omp parallel for reduction(+:bad) schedule(static) if (n >= OMP_MIN_CELLS)
    for (int i = 0; i < n; ++i) {
        GasState Q; unpackQ(g, q + (size_t)i * w, Q);
        int rc = -9;
        if (which == 1) rc = update_thermo_from_rhou(g, Q);
        else if (which == 0) rc = update_thermo_from_pT(g, Q);
        else if (which == 2) rc = update_thermo_from_rhoT(g, Q);
        else if (which == 4) rc = update_sound_speed(g, Q);
        if (rc != 0) bad++;
        packQ(g, Q, q + (size_t)i * w);
    }
    return bad;
}

// ------------------------------------------------------------ reconstruction

struct L2R2 {   // onedinterp.d:24-42 work variables set by l2r2_prepare
    double aL0, aR0, lenL0_, lenR0_;
    double two_over_lenL0_plus_lenL1, two_over_lenR0_plus_lenL0, two_over_lenR1_plus_lenR0;
    double two_lenL0_plus_lenL1, two_lenR0_plus_lenR1;
};

// onedinterp.d:118-133
static void l2r2_prepare(L2R2& w, double lenL1, double lenL0, double lenR0, double lenR1)
{
    w.lenL0_ = lenL0;
    w.lenR0_ = lenR0;
    w.aL0 = 0.5 * lenL0 / (lenL1 + 2.0 * lenL0 + lenR0);
    w.aR0 = 0.5 * lenR0 / (lenL0 + 2.0 * lenR0 + lenR1);
    w.two_over_lenL0_plus_lenL1 = 2.0 / (lenL0 + lenL1);
    w.two_over_lenR0_plus_lenL0 = 2.0 / (lenR0 + lenL0);
    w.two_over_lenR1_plus_lenR0 = 2.0 / (lenR1 + lenR0);
    w.two_lenL0_plus_lenL1 = (2.0 * lenL0 + lenL1);
    w.two_lenR0_plus_lenR1 = (2.0 * lenR0 + lenR1);
}

// limiters.d:30-38
static inline double clip_to_limits(double q, double A, double B)
{
    double lower_limit = (A <= B) ? A : B;
    double upper_limit = (A > B) ? A : B;
    double qclipped = (q > lower_limit) ? q : lower_limit;
    return (qclipped <= upper_limit) ? qclipped : upper_limit;
}

// onedinterp.d:136-162
static void interp_l2r2_scalar(const L2R2& w, int apply_limiter, int extrema_clipping,
                               double qL1, double qL0, double qR0, double qR1, double& qL, double& qR)
{
    double delLminus = (qL0 - qL1) * w.two_over_lenL0_plus_lenL1;
    double del = (qR0 - qL0) * w.two_over_lenR0_plus_lenL0;
    double delRplus = (qR1 - qR0) * w.two_over_lenR1_plus_lenR0;
    double sL = 1.0;
    double sR = 1.0;
    if (apply_limiter) {
        sL = (delLminus * del + fabs(delLminus * del) + epsilon_van_albada) /
             (delLminus * delLminus + del * del + epsilon_van_albada);
        sR = (del * delRplus + fabs(del * delRplus) + epsilon_van_albada) /
             (del * del + delRplus * delRplus + epsilon_van_albada);
    }
    qL = qL0 + sL * w.aL0 * (del * w.two_lenL0_plus_lenL1 + delLminus * w.lenR0_);
    qR = qR0 - sR * w.aR0 * (delRplus * w.lenL0_ + del * w.two_lenR0_plus_lenR1);
    if (extrema_clipping) {
        qL = clip_to_limits(qL, qL0, qR0);
        qR = clip_to_limits(qR, qL0, qR0);
    }
}

void ref_interp_l2r2_scalar(const double* len4, const double* q4, int apply_limiter, int extrema_clipping,
                            double* qLR)
{
    L2R2 w; l2r2_prepare(w, len4[0], len4[1], len4[2], len4[3]);
    interp_l2r2_scalar(w, apply_limiter, extrema_clipping, q4[0], q4[1], q4[2], q4[3], qLR[0], qLR[1]);
}

// geom/elements/vector3.d:425-450
struct Frame { double n[3], t1[3], t2[3]; };
static inline void to_local(double v[3], const Frame& f)
{
    double vx = v[0] * f.n[0] + v[1] * f.n[1] + v[2] * f.n[2];
    double vy = v[0] * f.t1[0] + v[1] * f.t1[1] + v[2] * f.t1[2];
    double vz = v[0] * f.t2[0] + v[1] * f.t2[1] + v[2] * f.t2[2];
    v[0] = vx; v[1] = vy; v[2] = vz;
}
static inline void to_global(double v[3], const Frame& f)
{
    double vx = v[0] * f.n[0] + v[1] * f.t1[0] + v[2] * f.t2[0];
    double vy = v[0] * f.n[1] + v[1] * f.t1[1] + v[2] * f.t2[1];
    double vz = v[0] * f.n[2] + v[1] * f.t1[2] + v[2] * f.t2[2];
    v[0] = vx; v[1] = vy; v[2] = vz;
}

struct Config {
    int dims;
    int flux_calc;
    int update_scheme;
    int interp_order;
    int apply_limiter;
    int extrema_clipping;
    int thermo_interp;
    int max_invalid_cells;
    double shear_tolerance;         // globalconfig.d:585
    double compression_tolerance;   // globalconfig.d:600
    double M_inf = 0.01;            // ausm_plus_up's reference Mach number squared floor, globalconfig.d:591
};

// onedinterp.d:435-585 together with the EOS mixins of onedinterp.d:1039-1068.
// cL1..cR1 are copies (the reference's in-place rotate/unrotate of the cells'
// velocities is an exact round trip).
static void interp_l2r2(Gas* g, const Config& cfg, const Frame& fr,
                        FlowState cL1, FlowState cL0, FlowState cR0, FlowState cR1,
                        double lenL1, double lenL0, double lenR0, double lenR1,
                        FlowState& Lft, FlowState& Rght)
{
    const int nsp = g->nsp;
    to_local(cL1.vel, fr); to_local(cL0.vel, fr); to_local(cR0.vel, fr); to_local(cR1.vel, fr);
    L2R2 w; l2r2_prepare(w, lenL1, lenL0, lenR0, lenR1);
    for (int d = 0; d < 3; ++d)
        interp_l2r2_scalar(w, cfg.apply_limiter, cfg.extrema_clipping, cL1.vel[d], cL0.vel[d], cR0.vel[d],
                           cR1.vel[d], Lft.vel[d], Rght.vel[d]);
    const GasState &gL1 = cL1.gas, &gL0 = cL0.gas, &gR0 = cR0.gas, &gR1 = cR1.gas;
    if (nsp > 1) {
        for (int s = 0; s < nsp; ++s)
            interp_l2r2_scalar(w, cfg.apply_limiter, cfg.extrema_clipping, gL1.massf[s], gL0.massf[s],
                               gR0.massf[s], gR1.massf[s], Lft.gas.massf[s], Rght.gas.massf[s]);
        if (scale_mass_fractions(Lft.gas.massf, nsp) != 0)
            for (int s = 0; s < nsp; ++s) Lft.gas.massf[s] = gL0.massf[s];
        if (scale_mass_fractions(Rght.gas.massf, nsp) != 0)
            for (int s = 0; s < nsp; ++s) Rght.gas.massf[s] = gR0.massf[s];
    } else {
        Lft.gas.massf[0] = 1.0;
        Rght.gas.massf[0] = 1.0;
    }
    int rcL = 0, rcR = 0;
    switch (cfg.thermo_interp) {
    case INTERP_PT:
        interp_l2r2_scalar(w, cfg.apply_limiter, cfg.extrema_clipping, gL1.p, gL0.p, gR0.p, gR1.p, Lft.gas.p, Rght.gas.p);
        interp_l2r2_scalar(w, cfg.apply_limiter, cfg.extrema_clipping, gL1.T, gL0.T, gR0.T, gR1.T, Lft.gas.T, Rght.gas.T);
        rcL = update_thermo_from_pT(g, Lft.gas);
        if (rcL != 0) { Lft = cL0; }
        rcR = update_thermo_from_pT(g, Rght.gas);
        if (rcR != 0) { Rght = cR0; }
        break;
    case INTERP_RHOU:
        interp_l2r2_scalar(w, cfg.apply_limiter, cfg.extrema_clipping, gL1.rho, gL0.rho, gR0.rho, gR1.rho, Lft.gas.rho, Rght.gas.rho);
        interp_l2r2_scalar(w, cfg.apply_limiter, cfg.extrema_clipping, gL1.u, gL0.u, gR0.u, gR1.u, Lft.gas.u, Rght.gas.u);
        rcL = update_thermo_from_rhou(g, Lft.gas);
        if (rcL != 0) { Lft = cL0; }
        rcR = update_thermo_from_rhou(g, Rght.gas);
        if (rcR != 0) { Rght = cR0; }
        break;
    case INTERP_RHOP:
        interp_l2r2_scalar(w, cfg.apply_limiter, cfg.extrema_clipping, gL1.rho, gL0.rho, gR0.rho, gR1.rho, Lft.gas.rho, Rght.gas.rho);
        interp_l2r2_scalar(w, cfg.apply_limiter, cfg.extrema_clipping, gL1.p, gL0.p, gR0.p, gR1.p, Lft.gas.p, Rght.gas.p);
        rcL = update_thermo_from_rhop(g, Lft.gas);
        if (rcL != 0) { Lft = cL0; }
        rcR = update_thermo_from_rhop(g, Rght.gas);
        if (rcR != 0) { Rght = cR0; }
        break;
    case INTERP_RHOT:
        interp_l2r2_scalar(w, cfg.apply_limiter, cfg.extrema_clipping, gL1.rho, gL0.rho, gR0.rho, gR1.rho, Lft.gas.rho, Rght.gas.rho);
        interp_l2r2_scalar(w, cfg.apply_limiter, cfg.extrema_clipping, gL1.T, gL0.T, gR0.T, gR1.T, Lft.gas.T, Rght.gas.T);
        rcL = update_thermo_from_rhoT(g, Lft.gas);
        if (rcL != 0) { Lft = cL0; }
        rcR = update_thermo_from_rhoT(g, Rght.gas);
        if (rcR != 0) { Rght = cR0; }
        break;
    }
    // The fallback copies above are of the rotated cell states (cL0/cR0 are
    // local-frame copies here), matching the reference, where the fallback
    // copy happens before the transform back to the global frame.
    to_global(Lft.vel, fr);
    to_global(Rght.vel, fr);
}

// ---------------------------------------- one-sided reconstruction (no ghost data)
// onedinterp.d:165-258.  Weights of a linear combination q0*w0 + q1*w1:
// interpolation to the face between two cell centres, or extrapolation past
// q0 to the boundary len0/2 away.
struct Lin { double w0, w1; };
static inline Lin linear_interp_prepare(double len0, double len1)
{
    return Lin{len1 / (len0 + len1), len0 / (len0 + len1)};
}
static inline Lin linear_extrap_prepare(double len0, double len1)
{
    return Lin{(2.0 * len0 + len1) / (len0 + len1), -len0 / (len0 + len1)};
}
static inline double weight_scalar(const Lin& w, int extrema_clipping, double q0, double q1)
{
    double q = q0 * w.w0 + q1 * w.w1;
    if (extrema_clipping) q = clip_to_limits(q, q0, q1);
    return q;
}

// onedinterp.d:178-196: the left state as l2r2 has it; the right state
// interpolated between the two cells, or the cell value where the limiter
// sees an extremum.  (l2r1_prepare sets the same work variables as
// l2r2_prepare for the terms used here.)
static void interp_l2r1_scalar(const L2R2& w, const Lin& lin, int apply_limiter, int extrema_clipping,
                               double qL1, double qL0, double qR0, double& qL, double& qR)
{
    double delLminus = (qL0 - qL1) * w.two_over_lenL0_plus_lenL1;
    double del = (qR0 - qL0) * w.two_over_lenR0_plus_lenL0;
    double sL = 1.0;
    if (apply_limiter)
        sL = (delLminus * del + fabs(delLminus * del) + epsilon_van_albada) /
             (delLminus * delLminus + del * del + epsilon_van_albada);
    qL = qL0 + sL * w.aL0 * (del * w.two_lenL0_plus_lenL1 + delLminus * w.lenR0_);
    if (apply_limiter && (delLminus * del < 0.0)) qR = qR0;
    else qR = weight_scalar(lin, extrema_clipping, qL0, qR0);
    if (extrema_clipping) qL = clip_to_limits(qL, qL0, qR0);
}

// onedinterp.d:209-227: the mirror image of l2r1
static void interp_l1r2_scalar(const L2R2& w, const Lin& lin, int apply_limiter, int extrema_clipping,
                               double qL0, double qR0, double qR1, double& qL, double& qR)
{
    double del = (qR0 - qL0) * w.two_over_lenR0_plus_lenL0;
    double delRplus = (qR1 - qR0) * w.two_over_lenR1_plus_lenR0;
    double sR = 1.0;
    if (apply_limiter)
        sR = (del * delRplus + fabs(del * delRplus) + epsilon_van_albada) /
             (del * del + delRplus * delRplus + epsilon_van_albada);
    qR = qR0 - sR * w.aR0 * (delRplus * w.lenL0_ + del * w.two_lenR0_plus_lenR1);
    if (apply_limiter && (delRplus * del < 0.0)) qL = qL0;
    else qL = weight_scalar(lin, extrema_clipping, qL0, qR0);
    if (extrema_clipping) qR = clip_to_limits(qR, qL0, qR0);
}

static int thermo_update(Gas* g, int thermo_interp, GasState& Q)
{
    switch (thermo_interp) {
    case INTERP_PT: return update_thermo_from_pT(g, Q);
    case INTERP_RHOP: return update_thermo_from_rhop(g, Q);
    case INTERP_RHOT: return update_thermo_from_rhoT(g, Q);
    default: return update_thermo_from_rhou(g, Q);
    }
}

// The two thermodynamic scalars each thermo_interpolator reconstructs.
static double* thermo_pair(GasState& Q, int thermo_interp, int which)
{
    switch (thermo_interp) {
    case INTERP_PT: return which ? &Q.T : &Q.p;
    case INTERP_RHOP: return which ? &Q.p : &Q.rho;
    case INTERP_RHOT: return which ? &Q.T : &Q.rho;
    default: return which ? &Q.u : &Q.rho;
    }
}

// onedinterp.d:587-818: interp_l2r1 (left_full = true; cells cL1, cL0, cR0 =
// cA, cB, cC; the face next to a right boundary without ghost data) and
// interp_l1r2 (left_full = false; cells cL0, cR0, cR1 = cA, cB, cC).
static void interp_l2r1_l1r2(Gas* g, const Config& cfg, const Frame& fr, bool left_full, FlowState cA,
                             FlowState cB, FlowState cC, double len, FlowState& Lft, FlowState& Rght)
{
    const int nsp = g->nsp;
    to_local(cA.vel, fr); to_local(cB.vel, fr); to_local(cC.vel, fr);
    L2R2 w;
    if (left_full) l2r2_prepare(w, len, len, len, len);   // l2r1_prepare(lenL1, lenL0, lenR0)
    else l2r2_prepare(w, len, len, len, len);             // l1r2_prepare(lenL0, lenR0, lenR1)
    const Lin lin = linear_interp_prepare(len, len);
    const int lim = cfg.apply_limiter, clip = cfg.extrema_clipping;
    auto scalar = [&](double a, double b, double c, double& qL, double& qR) {
        if (left_full) interp_l2r1_scalar(w, lin, lim, clip, a, b, c, qL, qR);
        else interp_l1r2_scalar(w, lin, lim, clip, a, b, c, qL, qR);
    };
    const FlowState& cL0 = left_full ? cB : cA;
    const FlowState& cR0 = left_full ? cC : cB;
    for (int d = 0; d < 3; ++d) scalar(cA.vel[d], cB.vel[d], cC.vel[d], Lft.vel[d], Rght.vel[d]);
    if (nsp > 1) {
        for (int s = 0; s < nsp; ++s)
            scalar(cA.gas.massf[s], cB.gas.massf[s], cC.gas.massf[s], Lft.gas.massf[s], Rght.gas.massf[s]);
        if (scale_mass_fractions(Lft.gas.massf, nsp) != 0)
            for (int s = 0; s < nsp; ++s) Lft.gas.massf[s] = cL0.gas.massf[s];
        if (scale_mass_fractions(Rght.gas.massf, nsp) != 0)
            for (int s = 0; s < nsp; ++s) Rght.gas.massf[s] = cR0.gas.massf[s];
    } else {
        Lft.gas.massf[0] = 1.0;
        Rght.gas.massf[0] = 1.0;
    }
    for (int which = 0; which < 2; ++which)
        scalar(*thermo_pair(cA.gas, cfg.thermo_interp, which), *thermo_pair(cB.gas, cfg.thermo_interp, which),
               *thermo_pair(cC.gas, cfg.thermo_interp, which), *thermo_pair(Lft.gas, cfg.thermo_interp, which),
               *thermo_pair(Rght.gas, cfg.thermo_interp, which));
    if (thermo_update(g, cfg.thermo_interp, Lft.gas) != 0) Lft = cL0;
    if (thermo_update(g, cfg.thermo_interp, Rght.gas) != 0) Rght = cR0;
    to_global(Lft.vel, fr);
    to_global(Rght.vel, fr);
}

// onedinterp.d:820-1037: interp_l2r0 / interp_l0r2, the state at a wall face
// without ghost data, extrapolated linearly from the two cells next to it
// (c0 at the wall, c1 behind it).  With extrema clipping the reference
// returns at once and the face keeps the copy of the wall cell's state.
static void interp_wall_side(Gas* g, const Config& cfg, const Frame& fr, FlowState c0, FlowState c1, double len,
                             FlowState& side)
{
    if (cfg.extrema_clipping) return;
    const int nsp = g->nsp;
    to_local(c0.vel, fr); to_local(c1.vel, fr);
    const Lin w = linear_extrap_prepare(len, len);
    for (int d = 0; d < 3; ++d) side.vel[d] = weight_scalar(w, 0, c0.vel[d], c1.vel[d]);
    if (nsp > 1) {
        for (int s = 0; s < nsp; ++s) side.gas.massf[s] = weight_scalar(w, 0, c0.gas.massf[s], c1.gas.massf[s]);
        if (scale_mass_fractions(side.gas.massf, nsp) != 0)
            for (int s = 0; s < nsp; ++s) side.gas.massf[s] = c0.gas.massf[s];
    } else {
        side.gas.massf[0] = 1.0;
    }
    for (int which = 0; which < 2; ++which)
        *thermo_pair(side.gas, cfg.thermo_interp, which) =
            weight_scalar(w, 0, *thermo_pair(c0.gas, cfg.thermo_interp, which),
                          *thermo_pair(c1.gas, cfg.thermo_interp, which));
    if (thermo_update(g, cfg.thermo_interp, side.gas) != 0) side = c0;
    to_global(side.vel, fr);
}

// ------------------------------------------------------------------- fluxes
// Conserved layout: [mass, mom_x, mom_y, mom_z, total_energy, rhoY_0..rhoY_{nsp-1}]
// (species only when nsp > 1).

// fluxcalc.d:376-529 (velocities already in the face frame)
static void ausmdv(const Gas* g, const FlowState& Lft, const FlowState& Rght, double* F)
{
    double rL = Lft.gas.rho;
    double pL = Lft.gas.p;
    double pLrL = pL / rL;
    double uL = Lft.vel[0];
    double vL = Lft.vel[1];
    double wL = Lft.vel[2];
    double eL = Lft.gas.u;
    double aL = Lft.gas.a;
    double keL = 0.5 * (uL * uL + vL * vL + wL * wL);
    double HL = eL + pLrL + keL;
    double rR = Rght.gas.rho;
    double pR = Rght.gas.p;
    double pRrR = pR / rR;
    double uR = Rght.vel[0];
    double vR = Rght.vel[1];
    double wR = Rght.vel[2];
    double eR = Rght.gas.u;
    double aR = Rght.gas.a;
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
    const int nsp = g->nsp;
    F[0] = ru_half;
    if (ru_half >= 0.0) {
        F[1] = ru2_half + p_half; F[2] = ru_half * vL; F[3] = ru_half * wL;
        F[4] = ru_half * HL;
        if (nsp > 1) for (int i = 0; i < nsp; ++i) F[5 + i] = ru_half * Lft.gas.massf[i];
    } else {
        F[1] = ru2_half + p_half; F[2] = ru_half * vR; F[3] = ru_half * wR;
        F[4] = ru_half * HR;
        if (nsp > 1) for (int i = 0; i < nsp; ++i) F[5 + i] = ru_half * Rght.gas.massf[i];
    }
    const double C_EFIX = 0.125;
    bool caseA = ((uL - aL) < 0.0) && ((uR - aR) > 0.0);
    bool caseB = ((uL + aL) < 0.0) && ((uR + aR) > 0.0);
    double d_ua = 0.0;
    if (caseA && !caseB) d_ua = C_EFIX * ((uR - aR) - (uL - aL));
    if (caseB && !caseA) d_ua = C_EFIX * ((uR + aR) - (uL + aL));
    if (d_ua != 0.0) {
        F[0] -= d_ua * (rR - rL);
        F[1] -= d_ua * (rR * uR - rL * uL);
        F[2] -= d_ua * (rR * vR - rL * vL);
        F[3] -= d_ua * (rR * wR - rL * wL);
        F[4] -= d_ua * (rR * HR - rL * HL);
        if (nsp > 1)
            for (int i = 0; i < nsp; ++i)
                F[5 + i] -= d_ua * (rR * Rght.gas.massf[i] - rL * Lft.gas.massf[i]);
    }
}

// HLLC (no counterpart in the reference; Toro 2009, "Riemann Solvers and
// Numerical Methods for Fluid Dynamics", 3rd ed., section 10.4, eqs. 10.37-10.39
// with the pressure-free contact speed S* of eq. 10.37 and the Davis wave-speed
// bounds SL = min(uL-aL, uR-aR), SR = max(uL+aL, uR+aR) of eq. 10.48).  The
// total energy uses the face internal energy (the gas model's u), so the
// scheme works for any caloric EOS; species are carried by the contact wave.
static void hllc(const Gas* g, const FlowState& Lft, const FlowState& Rght, double* F)
{
    const int nsp = g->nsp;
    double rL = Lft.gas.rho, pL = Lft.gas.p, uL = Lft.vel[0], vL = Lft.vel[1], wL = Lft.vel[2];
    double aL = Lft.gas.a;
    double EL = rL * (Lft.gas.u + 0.5 * (uL * uL + vL * vL + wL * wL));
    double rR = Rght.gas.rho, pR = Rght.gas.p, uR = Rght.vel[0], vR = Rght.vel[1], wR = Rght.vel[2];
    double aR = Rght.gas.a;
    double ER = rR * (Rght.gas.u + 0.5 * (uR * uR + vR * vR + wR * wR));
    double SL = fmin(uL - aL, uR - aR);
    double SR = fmax(uL + aL, uR + aR);
    double mL = rL * (SL - uL);
    double mR = rR * (SR - uR);
    double Sstar = (pR - pL + mL * uL - mR * uR) / (mL - mR);
    double Fl[NQ_MAX], Fr[NQ_MAX];
    // physical fluxes
    double mfL = rL * uL, mfR = rR * uR;
    Fl[0] = mfL; Fl[1] = mfL * uL + pL; Fl[2] = mfL * vL; Fl[3] = mfL * wL; Fl[4] = uL * (EL + pL);
    Fr[0] = mfR; Fr[1] = mfR * uR + pR; Fr[2] = mfR * vR; Fr[3] = mfR * wR; Fr[4] = uR * (ER + pR);
    if (nsp > 1) for (int i = 0; i < nsp; ++i) { Fl[5 + i] = mfL * Lft.gas.massf[i]; Fr[5 + i] = mfR * Rght.gas.massf[i]; }
    const int nq = 5 + (nsp > 1 ? nsp : 0);
    if (0.0 <= SL) { for (int q = 0; q < nq; ++q) F[q] = Fl[q]; return; }
    if (SR <= 0.0) { for (int q = 0; q < nq; ++q) F[q] = Fr[q]; return; }
    const bool left = (0.0 <= Sstar);
    const double rK = left ? rL : rR, uK = left ? uL : uR, vK = left ? vL : vR, wK = left ? wL : wR;
    const double pK = left ? pL : pR, EK = left ? EL : ER, SK = left ? SL : SR;
    const double* FK = left ? Fl : Fr;
    const FlowState& K = left ? Lft : Rght;
    double coef = rK * (SK - uK) / (SK - Sstar);
    double Us[NQ_MAX];
    Us[0] = coef;
    Us[1] = coef * Sstar;
    Us[2] = coef * vK;
    Us[3] = coef * wK;
    Us[4] = coef * (EK / rK + (Sstar - uK) * (Sstar + pK / (rK * (SK - uK))));
    if (nsp > 1) for (int i = 0; i < nsp; ++i) Us[5 + i] = coef * K.gas.massf[i];
    double UK[NQ_MAX];
    UK[0] = rK; UK[1] = rK * uK; UK[2] = rK * vK; UK[3] = rK * wK; UK[4] = EK;
    if (nsp > 1) for (int i = 0; i < nsp; ++i) UK[5 + i] = rK * K.gas.massf[i];
    for (int q = 0; q < nq; ++q) F[q] = FK[q] + SK * (Us[q] - UK[q]);
}

// fluxcalc.d:532-623 (Hanel's flux vector splitting; velocities in the face frame)
static void hanel(const Gas* g, const FlowState& Lft, const FlowState& Rght, double* F)
{
    const double rL = Lft.gas.rho, pL = Lft.gas.p;
    const double pLrL = pL / rL;
    const double uL = Lft.vel[0], vL = Lft.vel[1], wL = Lft.vel[2];
    const double eL = Lft.gas.u, aL = Lft.gas.a;
    const double keL = 0.5 * (uL * uL + vL * vL + wL * wL);
    const double HL = eL + pLrL + keL;
    const double rR = Rght.gas.rho, pR = Rght.gas.p;
    const double pRrR = pR / rR;
    const double uR = Rght.vel[0], vR = Rght.vel[1], wR = Rght.vel[2];
    const double eR = Rght.gas.u, aR = Rght.gas.a;
    const double keR = 0.5 * (uR * uR + vR * vR + wR * wR);
    const double HR = eR + pRrR + keR;
    // left: pressure and velocity splitting (Wada and Liou 1997, eqs. 7, 9)
    double pLplus, uLplus;
    if (fabs(uL) <= aL) {
        uLplus = 1.0 / (4.0 * aL) * (uL + aL) * (uL + aL);
        pLplus = pL * uLplus * (1.0 / aL * (2.0 - uL / aL));
    } else {
        uLplus = 0.5 * (uL + fabs(uL));
        pLplus = pL * uLplus * (1.0 / uL);
    }
    double pRminus, uRminus;
    if (fabs(uR) <= aR) {
        uRminus = -1.0 / (4.0 * aR) * (uR - aR) * (uR - aR);
        pRminus = pR * uRminus * (1.0 / aR * (-2.0 - uR / aR));
    } else {
        uRminus = 0.5 * (uR - fabs(uR));
        pRminus = pR * uRminus * (1.0 / uR);
    }
    const double p_half = pLplus + pRminus;
    F[0] = uLplus * rL + uRminus * rR;
    F[1] = uLplus * rL * uL + uRminus * rR * uR + p_half;
    F[2] = uLplus * rL * vL + uRminus * rR * vR;
    F[3] = uLplus * rL * wL + uRminus * rR * wR;
    F[4] = uLplus * rL * HL + uRminus * rR * HR;
    const int nsp = g->nsp;
    if (nsp > 1)
        for (int i = 0; i < nsp; ++i)
            F[5 + i] = uLplus * rL * Lft.gas.massf[i] + uRminus * rR * Rght.gas.massf[i];
}

// fluxcalc.d:822-836: Hanel across a shock (either side flagged) where the
// tangential velocity jump is small, AUSMDV elsewhere.
static void adaptive_hanel_ausmdv(const Gas* g, const Config& cfg, const FlowState& Lft, const FlowState& Rght,
                                  double* F)
{
    const double sound_speed = 0.5 * (Lft.gas.a + Rght.gas.a);
    const double shear_y = fabs(Lft.vel[1] - Rght.vel[1]) / sound_speed;
    const double shear_z = fabs(Lft.vel[2] - Rght.vel[2]) / sound_speed;
    const bool shear_is_small = fmax(shear_y, shear_z) <= cfg.shear_tolerance;
    if ((Lft.S == 1 || Rght.S == 1) && shear_is_small) hanel(g, Lft, Rght, F);
    else ausmdv(g, Lft, Rght, F);
}

static double gas_gamma(const Gas* g, const GasState& Q);
static double gas_Cv(const Gas* g, const GasState& Q);

// fluxcalc.d:792-820: exp(-x^2) and erf(x) by the reference's polynomial
// approximation (Abramowitz and Stegun 7.1.26), clamped beyond |x| = 5.
static void exxef(double sn, double& exx, double& ef)
{
    const double P = 0.327591100;
    const double A1 = 0.254829592;
    const double A2 = -0.284496736;
    const double A3 = 1.421413741;
    const double A4 = -1.453152027;
    const double A5 = 1.061405429;
    const double LIMIT = 5.0;
    const double EXLIM = 0.138879e-10;
    const double EFLIM = 1.0;
    double ef1;
    if (fabs(sn) > LIMIT) {
        exx = EXLIM;
        ef1 = EFLIM;
    } else {
        const double snsq = sn * sn;
        exx = exp(-snsq);
        const double y = 1.0 / (1.0 + P * fabs(sn));
        ef1 = 1.0 - y * (A1 + y * (A2 + y * (A3 + y * (A4 + A5 * y)))) * exx;
    }
    ef = copysign(ef1, sn);
}

// fluxcalc.d:626-790: Macrossan and Pullin's equilibrium flux method.  The
// gas "constants" come from the two states (Cv of the gas model, R = p/(rho T)),
// weighted by sqrt(rho); the right state's RT is p/rho where the left one's is
// R T, as the reference has it.  Species follow the sign of the mass flux.
static void efm(const Gas* g, const FlowState& Lft, const FlowState& Rght, double* F)
{
    const double PHI = 1.0;
    const double dtwspi = 0.282094792;
    const double rhoL = Lft.gas.rho, presL = Lft.gas.p, eL = Lft.gas.u;
    const double hL = eL + presL / rhoL;
    const double tL = Lft.gas.T;
    const double vnL = Lft.vel[0], vpL = Lft.vel[1], vqL = Lft.vel[2];
    const double rhoR = Rght.gas.rho, presR = Rght.gas.p, eR = Rght.gas.u;
    const double hR = eR + presR / rhoR;
    const double tR = Rght.gas.T;
    const double vnR = Rght.vel[0], vpR = Rght.vel[1], vqR = Rght.vel[2];
    const double cvL = gas_Cv(g, Lft.gas);
    const double RgasL = presL / (rhoL * tL);
    const double cvR = gas_Cv(g, Rght.gas);
    const double RgasR = presR / (rhoR * tR);
    const double rLsqrt = sqrt(rhoL);
    const double rRsqrt = sqrt(rhoR);
    const double alpha = rLsqrt / (rLsqrt + rRsqrt);
    const double cv = alpha * cvL + (1.0 - alpha) * cvR;
    const double Rgas = alpha * RgasL + (1.0 - alpha) * RgasR;
    const double cp = cv + Rgas;
    const double gam = cp / cv;
    const double con = 0.5 * (gam + 1.0) / (gam - 1.0);
    const double rtL = Rgas * tL;
    const double cmpL = sqrt(2.0 * rtL);
    const double hvsqL = 0.5 * (vnL * vnL + vpL * vpL + vqL * vqL);
    const double snL = vnL / (PHI * cmpL);
    double exL, efL;
    exxef(snL, exL, efL);
    const double wL = 0.5 * (1.0 + efL);
    const double dL = exL * dtwspi;
    const double rtR = presR / rhoR;
    const double cmpR = sqrt(2.0 * rtR);
    const double hvsqR = 0.5 * (vnR * vnR + vpR * vpR + vqR * vqR);
    const double snR = vnR / (PHI * cmpR);
    double exR, efR;
    exxef(snR, exR, efR);
    const double wR = 0.5 * (1.0 - efR);
    const double dR = -exR * dtwspi;
    const double fmsL = (wL * rhoL * vnL) + (dL * cmpL * rhoL);
    const double fmsR = (wR * rhoR * vnR) + (dR * cmpR * rhoR);
    F[0] = fmsL + fmsR;
    F[1] = fmsL * vnL + fmsR * vnR + wL * presL + wR * presR;
    F[2] = fmsL * vpL + fmsR * vpR;
    F[3] = fmsL * vqL + fmsR * vqR;
    F[4] = (wL * rhoL * vnL) * (hvsqL + hL) + (wR * rhoR * vnR) * (hvsqR + hR) +
           (dL * cmpL * rhoL) * (hvsqL + con * rtL) + (dR * cmpR * rhoR) * (hvsqR + con * rtR);
    const int nsp = g->nsp;
    if (nsp > 1) {
        const FlowState& up = (F[0] > 0.0) ? Lft : Rght;
        for (int i = 0; i < nsp; ++i) F[5 + i] = F[0] * up.gas.massf[i];
    }
}

// fluxcalc.d:802-820: EFM across a shock (either side flagged) where the
// tangential velocity jump is small, AUSMDV elsewhere.
static void adaptive_efm_ausmdv(const Gas* g, const Config& cfg, const FlowState& Lft, const FlowState& Rght,
                                double* F)
{
    const double sound_speed = 0.5 * (Lft.gas.a + Rght.gas.a);
    const double shear_y = fabs(Lft.vel[1] - Rght.vel[1]) / sound_speed;
    const double shear_z = fabs(Lft.vel[2] - Rght.vel[2]) / sound_speed;
    const bool shear_is_small = fmax(shear_y, shear_z) <= cfg.shear_tolerance;
    if ((Lft.S == 1 || Rght.S == 1) && shear_is_small) efm(g, Lft, Rght, F);
    else ausmdv(g, Lft, Rght, F);
}

// fluxcalc.d:860-1037: Liou's AUSM+-up (2006) with the interface sound speed
// the mean of the two sides' and the reference's M_inf floor on M0^2.
static double aup_M1plus(double M) { return 0.5 * (M + fabs(M)); }
static double aup_M1minus(double M) { return 0.5 * (M - fabs(M)); }
static double aup_M2plus(double M) { return 0.25 * (M + 1.0) * (M + 1.0); }
static double aup_M2minus(double M) { return -0.25 * (M - 1.0) * (M - 1.0); }
static double aup_M4plus(double M, double beta)
{
    if (fabs(M) >= 1.0) return aup_M1plus(M);
    const double M2p = aup_M2plus(M), M2m = aup_M2minus(M);
    return M2p * (1.0 - 16.0 * beta * M2m);
}
static double aup_M4minus(double M, double beta)
{
    if (fabs(M) >= 1.0) return aup_M1minus(M);
    const double M2p = aup_M2plus(M), M2m = aup_M2minus(M);
    return M2m * (1.0 + 16.0 * beta * M2p);
}
static double aup_P5plus(double M, double alpha)
{
    if (fabs(M) >= 1.0) return (1.0 / M) * aup_M1plus(M);
    const double M2p = aup_M2plus(M), M2m = aup_M2minus(M);
    return M2p * ((2.0 - M) - 16.0 * alpha * M * M2m);
}
static double aup_P5minus(double M, double alpha)
{
    if (fabs(M) >= 1.0) return (1.0 / M) * aup_M1minus(M);
    const double M2p = aup_M2plus(M), M2m = aup_M2minus(M);
    return M2m * ((-2.0 - M) + 16.0 * alpha * M * M2p);
}

static void ausm_plus_up(const Gas* g, const FlowState& Lft, const FlowState& Rght, double M_inf, double* F)
{
    const double rL = Lft.gas.rho, pL = Lft.gas.p;
    const double uL = Lft.vel[0], vL = Lft.vel[1], wL = Lft.vel[2];
    const double eL = Lft.gas.u, aL = Lft.gas.a;
    const double keL = 0.5 * (uL * uL + vL * vL + wL * wL);
    const double HL = eL + pL / rL + keL;
    const double rR = Rght.gas.rho, pR = Rght.gas.p;
    const double uR = Rght.vel[0], vR = Rght.vel[1], wR = Rght.vel[2];
    const double eR = Rght.gas.u, aR = Rght.gas.a;
    const double keR = 0.5 * (uR * uR + vR * vR + wR * wR);
    const double HR = eR + pR / rR + keR;
    const double a_half = 0.5 * (aR + aL);
    const double ML = uL / a_half;
    const double MR = uR / a_half;
    const double MbarSq = (uL * uL + uR * uR) / (2.0 * a_half * a_half);
    const double M0Sq = fmin(1.0, fmax(MbarSq, M_inf));
    const double fa = sqrt(M0Sq) * (2.0 - sqrt(M0Sq));
    const double alpha = 0.1875 * (-4.0 + 5 * fa * fa);
    const double beta = 0.125;
    const double M4plus_ML = aup_M4plus(ML, beta);
    const double P5plus_ML = aup_P5plus(ML, alpha);
    const double M4minus_MR = aup_M4minus(MR, beta);
    const double P5minus_MR = aup_P5minus(MR, alpha);
    const double KP = 0.25;
    const double KU = 0.75;
    const double SIGMA = 1.0;
    const double r_half = 0.5 * (rL + rR);
    const double Mp = -KP / fa * fmax((1.0 - SIGMA * MbarSq), 0.0) * (pR - pL) / (r_half * a_half * a_half);
    const double Pu = -KU * P5plus_ML * P5minus_MR * (rL + rR) * fa * a_half * (uR - uL);
    const double M_half = M4plus_ML + M4minus_MR + Mp;
    double ru_half = a_half * M_half;
    if (M_half > 0.0) ru_half *= rL;
    else ru_half *= rR;
    const double p_half = P5plus_ML * pL + P5minus_MR * pR + Pu;
    const double ru2_half = (ru_half >= 0.0) ? ru_half * uL : ru_half * uR;
    const int nsp = g->nsp;
    F[0] = ru_half;
    if (ru_half >= 0.0) {
        F[1] = ru2_half + p_half; F[2] = ru_half * vL; F[3] = ru_half * wL;
        F[4] = ru_half * HL;
        if (nsp > 1) for (int i = 0; i < nsp; ++i) F[5 + i] = ru_half * Lft.gas.massf[i];
    } else {
        F[1] = ru2_half + p_half; F[2] = ru_half * vR; F[3] = ru_half * wR;
        F[4] = ru_half * HR;
        if (nsp > 1) for (int i = 0; i < nsp; ++i) F[5 + i] = ru_half * Rght.gas.massf[i];
    }
}

// fluxcalc.d:1202-1371: Roe-Pike (Toro 2009, section 11.3) with the
// Kermani-Plett entropy fix on the acoustic eigenvalues; gamma of each side
// from the gas model (Cp/Cv), averaged.  The sums run over all five waves
// in the reference's order, zero eigenvector entries included.
static void roe(const Gas* g, const FlowState& Lft, const FlowState& Rght, double* F)
{
    const double rL = Lft.gas.rho, pL = Lft.gas.p;
    const double pLrL = pL / rL;
    const double uL = Lft.vel[0], vL = Lft.vel[1], wL = Lft.vel[2];
    const double eL = Lft.gas.u, aL = Lft.gas.a;
    const double keL = 0.5 * (uL * uL + vL * vL + wL * wL);
    const double HL = eL + pLrL + keL;
    const double rR = Rght.gas.rho, pR = Rght.gas.p;
    const double pRrR = pR / rR;
    const double uR = Rght.vel[0], vR = Rght.vel[1], wR = Rght.vel[2];
    const double eR = Rght.gas.u, aR = Rght.gas.a;
    const double keR = 0.5 * (uR * uR + vR * vR + wR * wR);
    const double HR = eR + pRrR + keR;
    const double gL = gas_gamma(g, Lft.gas);
    const double gR = gas_gamma(g, Rght.gas);
    const double ghat = (gR + gL) / 2.0;
    const double rhat = sqrt(rL * rR);
    const double uhat = (sqrt(rL) * uL + sqrt(rR) * uR) / (sqrt(rL) + sqrt(rR));
    const double vhat = (sqrt(rL) * vL + sqrt(rR) * vR) / (sqrt(rL) + sqrt(rR));
    const double what = (sqrt(rL) * wL + sqrt(rR) * wR) / (sqrt(rL) + sqrt(rR));
    const double Hhat = (sqrt(rL) * HL + sqrt(rR) * HR) / (sqrt(rL) + sqrt(rR));
    const double ahat2 = (ghat - 1.0) * (Hhat - 0.5 * (uhat * uhat + vhat * vhat + what * what));
    const double ahat = sqrt(ahat2);
    double lambda[5] = {uhat - ahat, uhat, uhat, uhat, uhat + ahat};
    const double lambdaL = uL - aL;
    const double lambdaR = uR - aR;
    const double eps = 1.0e-12;
    const double EPS = 4.0 * fmax(0.0, (lambdaR - lambdaL));
    if (fabs(lambda[0]) < EPS) lambda[0] = (lambda[0] * lambda[0] + EPS * EPS) / (2.0 * EPS + eps);
    if (fabs(lambda[4]) < EPS) lambda[4] = (lambda[4] * lambda[4] + EPS * EPS) / (2.0 * EPS + eps);
    const double K[5][5] = {
        {1.0, uhat - ahat, vhat, what, Hhat - uhat * ahat},
        {1.0, uhat, vhat, what, 0.5 * (uhat * uhat + vhat * vhat + what * what)},
        {0.0, 0.0, 1.0, 0.0, vhat},
        {0.0, 0.0, 0.0, 1.0, what},
        {1.0, uhat + ahat, vhat, what, Hhat + uhat * ahat}};
    double alpha[5];
    alpha[0] = 1.0 / (2.0 * ahat2) * ((pR - pL) - rhat * ahat * (uR - uL));
    alpha[1] = (rR - rL) - (pR - pL) / ahat2;
    alpha[2] = rhat * (vR - vL);
    alpha[3] = rhat * (wR - wL);
    alpha[4] = 1.0 / (2.0 * ahat2) * ((pR - pL) + rhat * ahat * (uR - uL));
    double FL, FR;
    FL = rL * uL;
    FR = rR * uR;
    F[0] = 0.5 * (FL + FR);
    for (int i = 0; i < 5; ++i) F[0] -= 0.5 * alpha[i] * fabs(lambda[i]) * K[i][0];
    FL = pL + rL * uL * uL;
    FR = pR + rR * uR * uR;
    F[1] = 0.5 * (FL + FR);
    for (int i = 0; i < 5; ++i) F[1] -= 0.5 * alpha[i] * fabs(lambda[i]) * K[i][1];
    FL = rL * uL * vL;
    FR = rR * uR * vR;
    F[2] = 0.5 * (FL + FR);
    for (int i = 0; i < 5; ++i) F[2] -= 0.5 * alpha[i] * fabs(lambda[i]) * K[i][2];
    FL = rL * uL * wL;
    FR = rR * uR * wR;
    F[3] = 0.5 * (FL + FR);
    for (int i = 0; i < 5; ++i) F[3] -= 0.5 * alpha[i] * fabs(lambda[i]) * K[i][3];
    FL = (rL * eL + rL * (uL * uL + vL * vL + wL * wL) / 2.0 + pL) * uL;
    FR = (rR * eR + rR * (uR * uR + vR * vR + wR * wR) / 2.0 + pR) * uR;
    F[4] = 0.5 * (FL + FR);
    for (int i = 0; i < 5; ++i) F[4] -= 0.5 * alpha[i] * fabs(lambda[i]) * K[i][4];
    const int nsp = g->nsp;
    if (nsp > 1) {
        const FlowState& up = (F[0] >= 0.0) ? Lft : Rght;
        for (int i = 0; i < nsp; ++i) F[5 + i] = F[0] * up.gas.massf[i];
    }
}

// The flux calculator of cfg in the face frame (fluxcalc.d:61-89).
static void flux_in_face_frame(const Gas* g, const Config& cfg, const FlowState& Lft, const FlowState& Rght,
                               double* F)
{
    switch (cfg.flux_calc) {
    case FLUX_HLLC: hllc(g, Lft, Rght, F); break;
    case FLUX_HANEL: hanel(g, Lft, Rght, F); break;
    case FLUX_EFM: efm(g, Lft, Rght, F); break;
    case FLUX_AUSM_PLUS_UP: ausm_plus_up(g, Lft, Rght, cfg.M_inf, F); break;
    case FLUX_ADAPTIVE_EFM_AUSMDV: adaptive_efm_ausmdv(g, cfg, Lft, Rght, F); break;
    case FLUX_ADAPTIVE_HANEL_AUSMDV: adaptive_hanel_ausmdv(g, cfg, Lft, Rght, F); break;
    case FLUX_ROE: roe(g, Lft, Rght, F); break;
    default: ausmdv(g, Lft, Rght, F); break;
    }
}

// fluxcalc.d:26-128 on a fixed grid (gvel = 0, omegaz = 0).
static void compute_interface_flux(const Gas* g, const Config& cfg, FlowState Lft, FlowState Rght,
                                   const Frame& fr, double* F)
{
    to_local(Lft.vel, fr);
    to_local(Rght.vel, fr);
    flux_in_face_frame(g, cfg, Lft, Rght, F);
    // The grid-velocity terms of fluxcalc.d:112-121 add exact zeros on a fixed grid.
    double m[3] = {F[1], F[2], F[3]};
    to_global(m, fr);
    F[1] = m[0]; F[2] = m[1]; F[3] = m[2];
}

// gas_model.d:159 with ideal_gas.d:160-163, therm_perf_gas.d:366-371: Cv of the
// state (ideal gas: the constant _Cv; mixture: sum Y_i (Cp_i(T) - R_i))
static double gas_Cv(const Gas* g, const GasState& Q)
{
    if (g->type == GAS_IDEAL) return g->Cv;
    double Cv[NSP_MAX];
    for (int i = 0; i < g->nsp; ++i) Cv[i] = eval_Cp(g->curve[i], Q.T) - g->R[i];
    return mass_average(g, Q, Cv);
}

// gas_model.d:162: gamma = Cp/Cv of the state (ideal gas: _Cp/_Cv)
static double gas_gamma(const Gas* g, const GasState& Q)
{
    if (g->type == GAS_IDEAL) return g->Cp / g->Cv;
    double Cp[NSP_MAX], Cv[NSP_MAX];
    for (int i = 0; i < g->nsp; ++i) Cp[i] = eval_Cp(g->curve[i], Q.T);
    double Cpm = mass_average(g, Q, Cp);
    for (int i = 0; i < g->nsp; ++i) Cv[i] = eval_Cp(g->curve[i], Q.T) - g->R[i];
    double Cvm = mass_average(g, Q, Cv);
    return Cpm / Cvm;
}

// fluxcalc.d:130-298, compute_flux_at_left_wall (left = true, S the state on
// the right of the face) and compute_flux_at_right_wall, on a fixed grid
// (wall velocity vstar = 0): the wall pressure from the Riemann invariant of
// the wave that brings the gas to rest, an isentropic estimate refined by a
// secant iteration when it is a shock, capped at ten times the gas pressure.
static void wall_flux(const Gas* g, FlowState S, const Frame& fr, bool left, int nq, double* F)
{
    to_local(S.vel, fr);
    const double vstar = 0.0;
    const double a = S.gas.a, v = S.vel[0];
    const double gam = gas_gamma(g, S.gas);
    const double rho = S.gas.rho, p = S.gas.p;
    double tmp;
    if (left) {
        double Ubar = v - 2.0 * a / (gam - 1.0);
        tmp = (vstar - Ubar) * (gam - 1.0) / (2.0 * sqrt(gam)) * sqrt(rho / pow(p, 1.0 / gam));
    } else {
        double Ubar = v + 2.0 * a / (gam - 1.0);
        tmp = (Ubar - vstar) * (gam - 1.0) / (2.0 * sqrt(gam)) * sqrt(rho / pow(p, 1.0 / gam));
    }
    const double ptiny = 0.1;   // FlowStateLimits.min_pressure, fvcore.d:26
    double pstar = (tmp > 0.0) ? pow(tmp, 2.0 * gam / (gam - 1.0)) : ptiny;
    if (pstar > 1.1 * p) {
        auto f = [&](double ps) {
            double xi = ps / p;
            double M1sq = 1.0 + (gam + 1.0) / 2.0 / gam * (xi - 1.0);
            double u1 = sqrt(M1sq) * a;
            double u2 = u1 * ((gam - 1.0) * M1sq + 2.0) / ((gam + 1.0) * M1sq);
            return left ? vstar - u1 + u2 - v : vstar + u1 - u2 - v;
        };
        int count = 0;
        double incr;
        do {
            double f0 = f(pstar);
            double dp = 0.001 * pstar;
            double f1 = f(pstar + dp);
            incr = -f0 * dp / (f1 - f0);
            pstar += incr;
            count += 1;
        } while (fabs(incr) / pstar > 0.01 && count < 10);
    }
    pstar = fmin(pstar, p * 10.0);
    for (int q = 0; q < nq; ++q) F[q] = 0.0;
    double m[3] = {pstar, 0.0, 0.0};
    to_global(m, fr);
    F[1] = m[0]; F[2] = m[1]; F[3] = m[2];
    F[4] = pstar * vstar;
}

// boundary_flux_effect.d:481-546 (BFE_SimpleOutflowFlux.compute_outflow_flux) on
// a fixed grid: the flux of the interior cell's own state when it leaves the
// domain (outsign: +1 on a plus face, -1 on a minus face), else a closed face
// carrying only the pressure.
static void outflow_flux(const Gas* g, const FlowState& fs, int outsign, const Frame& fr, int nq, double* F)
{
    const double* n = fr.n;
    const double vn = fs.vel[0] * n[0] + fs.vel[1] * n[1] + fs.vel[2] * n[2];
    const double mass_flux = fs.gas.rho * vn;
    for (int q = 0; q < nq; ++q) F[q] = 0.0;
    if ((outsign * mass_flux) > 0.0) {
        F[0] = mass_flux;
        for (int d = 0; d < 3; ++d) F[1 + d] = fs.gas.p * n[d] + fs.vel[d] * mass_flux;
        double utot = fs.gas.u + 0.5 * (fs.vel[0] * fs.vel[0] + fs.vel[1] * fs.vel[1] + fs.vel[2] * fs.vel[2]);
        F[4] = mass_flux * utot + fs.gas.p * vn;
        if (g->nsp > 1)
            for (int s = 0; s < g->nsp; ++s) F[5 + s] = mass_flux * fs.gas.massf[s];
    } else {
        for (int d = 0; d < 3; ++d) F[1 + d] = n[d] * fs.gas.p;
        F[4] = fs.gas.p * 0.0;
    }
}

// Stand-alone flux entry for tests.  fs layout: [rho,p,T,u,a,vx,vy,vz,massf..]
// in the face frame already (n, t1, t2) = (x, y, z).
int ref_flux(Gas* g, int calc, const double* L, const double* R, double* F)
{
    FlowState l, r;
    auto unpack = [&](const double* s, FlowState& f) {
        f.gas.rho = s[0]; f.gas.p = s[1]; f.gas.T = s[2]; f.gas.u = s[3]; f.gas.a = s[4];
        f.vel[0] = s[5]; f.vel[1] = s[6]; f.vel[2] = s[7];
        for (int i = 0; i < g->nsp; ++i) f.gas.massf[i] = s[8 + i];
    };
    unpack(L, l); unpack(R, r);
    l.S = r.S = 0;
    Config cfg{};
    cfg.flux_calc = calc; cfg.shear_tolerance = 0.20; cfg.M_inf = 0.01;
    flux_in_face_frame(g, cfg, l, r, F);
    return 0;
}

// exxef (fluxcalc.d:792-820) for tests.
void ref_exxef(double sn, double* exx, double* ef) { exxef(sn, *exx, *ef); }

// The same with the sides' shock indicators and the calculators' options.
int ref_flux_opt(Gas* g, int calc, const double* L, const double* R, int SL, int SR, double shear_tolerance,
                 double M_inf, double* F)
{
    FlowState l, r;
    auto unpack = [&](const double* s, FlowState& f) {
        f.gas.rho = s[0]; f.gas.p = s[1]; f.gas.T = s[2]; f.gas.u = s[3]; f.gas.a = s[4];
        f.vel[0] = s[5]; f.vel[1] = s[6]; f.vel[2] = s[7];
        for (int i = 0; i < g->nsp; ++i) f.gas.massf[i] = s[8 + i];
    };
    unpack(L, l); unpack(R, r);
    l.S = SL; r.S = SR;
    Config cfg{};
    cfg.flux_calc = calc; cfg.shear_tolerance = shear_tolerance; cfg.M_inf = M_inf;
    flux_in_face_frame(g, cfg, l, r, F);
    return 0;
}

// Stand-alone boundary fluxes for tests, in the frame (n, t1, t2) = (x, y, z):
// the wall flux (left != 0: the wall is on the left of the gas state) and the
// simple outflow flux through a face of outward sign outsign.
int ref_wall_flux(Gas* g, int left, const double* st, double* F)
{
    FlowState f{};
    f.gas.rho = st[0]; f.gas.p = st[1]; f.gas.T = st[2]; f.gas.u = st[3]; f.gas.a = st[4];
    f.vel[0] = st[5]; f.vel[1] = st[6]; f.vel[2] = st[7];
    for (int i = 0; i < g->nsp; ++i) f.gas.massf[i] = st[8 + i];
    const Frame fr{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    wall_flux(g, f, fr, left != 0, 5 + (g->nsp > 1 ? g->nsp : 0), F);
    return 0;
}

int ref_outflow_flux(Gas* g, int outsign, const double* st, double* F)
{
    FlowState f{};
    f.gas.rho = st[0]; f.gas.p = st[1]; f.gas.T = st[2]; f.gas.u = st[3]; f.gas.a = st[4];
    f.vel[0] = st[5]; f.vel[1] = st[6]; f.vel[2] = st[7];
    for (int i = 0; i < g->nsp; ++i) f.gas.massf[i] = st[8 + i];
    const Frame fr{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    outflow_flux(g, f, outsign, fr, 5 + (g->nsp > 1 ? g->nsp : 0), F);
    return 0;
}

// One-sided scalar reconstructions on a uniform grid of cell length len:
// kind 0 l2r1 (q = qL1, qL0, qR0), 1 l1r2 (q = qL0, qR0, qR1), 2 the wall-side
// linear extrapolation (q = wall cell, the cell behind it).  out = (qL, qR),
// or the extrapolated value in out[0].
void ref_interp_onesided_scalar(int kind, double len, const double* q, int apply_limiter, int extrema_clipping,
                                double* out)
{
    L2R2 w; l2r2_prepare(w, len, len, len, len);
    if (kind == 0) interp_l2r1_scalar(w, linear_interp_prepare(len, len), apply_limiter, extrema_clipping, q[0], q[1], q[2], out[0], out[1]);
    else if (kind == 1) interp_l1r2_scalar(w, linear_interp_prepare(len, len), apply_limiter, extrema_clipping, q[0], q[1], q[2], out[0], out[1]);
    else out[0] = weight_scalar(linear_extrap_prepare(len, len), 0, q[0], q[1]);
}

// ------------------------------------------------------------------- blocks

// Large per-slot arrays are sized without value-initialisation and then
// filled by all threads, so that their pages are first touched in parallel
// (a 256^3 block holds about 10 GB of them).
extern "C++" {
template <class T>
struct NoInitAlloc : std::allocator<T> {
    template <class U> struct rebind { using other = NoInitAlloc<U>; };
    NoInitAlloc() = default;
    template <class U> NoInitAlloc(const NoInitAlloc<U>&) {}
    template <class U, class... A> void construct(U* p, A&&... a)
    {
        if constexpr (sizeof...(A) == 0) ::new ((void*)p) U;
        else ::new ((void*)p) U(std::forward<A>(a)...);
    }
};
template <class T> using BigVec = std::vector<T, NoInitAlloc<T>>;

template <class T>
static void par_fill(BigVec<T>& v, size_t n, const T& value)
{
    v.resize(n);
    T* d = v.data();
    #pragma omp parallel for schedule(static) if (n >= 1024 * OMP_MIN_CELLS)
    for (long long i = 0; i < (long long)n; ++i) d[i] = value;
}
}   // extern "C++"

struct Block;
// profile: the ghost states of a static profile, [layer][a + b * na] over the
// face's in-face axes (north/south (i, k), east/west (j, k), top/bottom (i, j))
struct BC { int type; FlowState fs; Block* other; int other_face, orientation; std::vector<FlowState> profile; };

struct Block {
    Gas* gas;
    Config cfg;
    int ni, nj, nk, nidim, njdim, nkdim;
    int imin, imax, jmin, jmax, kmin, kmax;
    int nq;
    double dx, dy, dz, vol, area[3];
    Frame frame[3];
    BigVec<FlowState> fs;
    BigVec<double> U[4], dUdt[4];
    BigVec<double> F[3];          // face fluxes, indexed by the cell on the + side
    std::vector<char> bad;
    BC bc[6];
    // 2D axisymmetric geometry (per radian): cell volumes and (x,y)-plane
    // areas [ni*nj], i-face areas [(ni+1)*nj], j-face areas [ni*(nj+1)]
    bool axi = false;
    std::vector<double> axi_vol, axi_areaxy, axi_ai, axi_aj;
    // user-defined source terms per cell slot [slot * nq + q] (empty: none)
    std::vector<double> udf_src;
    bool stringent_cfl = false;     // GlobalConfig.stringent_cfl (globalconfig.d:729)
    // geometry computed from the grid vertices (ref_block_new_grid): per slot
    // the cell volume and i/j/k lengths (ghost slots: the lengths the
    // reference gives its ghost cells), per face of direction d the area and
    // the n, t1, t2 frame, indexed by the slot of the cell on the + side
    bool general = false;
    std::vector<double> gvol, glen[3], garea[3];
    std::vector<Frame> gfr[3];
};

static inline size_t cidx(const Block* b, int i, int j, int k) { return ((size_t)k * b->njdim + j) * b->nidim + i; }
static void apply_bcs(Block* b);

// Geometry of a face of direction dir (indexed by its + side slot) and of a
// cell slot: the block's uniform values, or the per-face / per-cell arrays of
// a block built from grid vertices.
static inline const Frame& face_frame(const Block* b, int dir, size_t slot)
{
    return b->general ? b->gfr[dir][slot] : b->frame[dir];
}
static inline double face_area(const Block* b, int dir, size_t slot)
{
    return b->general ? b->garea[dir][slot] : b->area[dir];
}
static inline double cell_len(const Block* b, int dir, size_t slot)
{
    if (b->general) return b->glen[dir][slot];
    return (dir == 0) ? b->dx : (dir == 1) ? b->dy : b->dz;
}
static inline double cell_vol(const Block* b, size_t slot) { return b->general ? b->gvol[slot] : b->vol; }

// ------------------------------------------------- cell and face geometry
// geom/elements/vector3.d, properties.d, fvinterface.d in the reference's
// operation order.
struct V3 { double x, y, z; };

// vector3.d:386-407: unit vector, small components flushed to zero
static void v3_normalize(double v[3])
{
    double magnitude = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (magnitude > 0.0) {
        v[0] /= magnitude; v[1] /= magnitude; v[2] /= magnitude;
    } else {
        v[0] = v[1] = v[2] = 0.0;
    }
    const double small = 1.0e-30;
    for (int q = 0; q < 3; ++q) if (fabs(v[q]) < small) v[q] = 0.0;
}

// vector3.d:604-609: v3 = v1 x v2
static void v3_cross(double v3[3], const double v1[3], const double v2[3])
{
    v3[0] = v1[1] * v2[2] - v2[1] * v1[2];
    v3[1] = v2[0] * v1[2] - v1[0] * v2[2];
    v3[2] = v1[0] * v2[1] - v2[0] * v1[1];
}

// properties.d:97-139: area and n, t1, t2 of the quadrilateral p0 p1 p2 p3
// (normal by the right-hand rule over the cycle)
static void quad_properties(const V3& p0, const V3& p1, const V3& p2, const V3& p3, Frame& fr, double& area)
{
    double p01x = p1.x - p0.x + p2.x - p3.x;
    double p01y = p1.y - p0.y + p2.y - p3.y;
    double p01z = p1.z - p0.z + p2.z - p3.z;
    double p03x = p3.x - p0.x + p2.x - p1.x;
    double p03y = p3.y - p0.y + p2.y - p1.y;
    double p03z = p3.z - p0.z + p2.z - p1.z;
    // vector3.d:634-641 cross_product
    double vax = p01y * p03z - p01z * p03y;
    double vay = p01z * p03x - p01x * p03z;
    double vaz = p01x * p03y - p01y * p03x;
    vax *= 0.25; vay *= 0.25; vaz *= 0.25;
    area = sqrt(vax * vax + vay * vay + vaz * vaz);
    if (area > 1.0e-20) {
        fr.n[0] = vax / area; fr.n[1] = vay / area; fr.n[2] = vaz / area;
        fr.t1[0] = p01x; fr.t1[1] = p01y; fr.t1[2] = p01z;
        v3_normalize(fr.t1);
        v3_cross(fr.t2, fr.n, fr.t1);
        v3_normalize(fr.t2);
    } else {
        fr = Frame{{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
    }
}

// properties.d:293-316: volume and centroid of the tetrahedron p0 p1 p2 p3
static void tetrahedron_properties(const V3& p0, const V3& p1, const V3& p2, const V3& p3, V3& centroid,
                                   double& volume)
{
    centroid = V3{0.25 * (p0.x + p1.x + p2.x + p3.x), 0.25 * (p0.y + p1.y + p2.y + p3.y),
                  0.25 * (p0.z + p1.z + p2.z + p3.z)};
    double dx01 = p1.x - p0.x, dy01 = p1.y - p0.y, dz01 = p1.z - p0.z;
    double dx02 = p2.x - p0.x, dy02 = p2.y - p0.y, dz02 = p2.z - p0.z;
    double cx = dy01 * dz02 - dz01 * dy02;
    double cy = dz01 * dx02 - dx01 * dz02;
    double cz = dx01 * dy02 - dy01 * dx02;
    double dx03 = p3.x - p0.x, dy03 = p3.y - p0.y, dz03 = p3.z - p0.z;
    volume = (dx03 * cx + dy03 * cy + dz03 * cz) / 6.0;
}

// properties.d:390-419: the quadrilateral-based pyramid p0 p1 p2 p3 / p4 as
// four tetrahedra about the base mid-point
static void pyramid_properties(const V3& p0, const V3& p1, const V3& p2, const V3& p3, const V3& p4, V3& centroid,
                               double& volume)
{
    V3 pmB{0.25 * (p0.x + p1.x + p2.x + p3.x), 0.25 * (p0.y + p1.y + p2.y + p3.y), 0.25 * (p0.z + p1.z + p2.z + p3.z)};
    volume = 0.0;
    V3 moment{0.0, 0.0, 0.0}, tc;
    double tv;
    const V3* base[5] = {&p0, &p1, &p2, &p3, &p0};
    for (int t = 0; t < 4; ++t) {
        tetrahedron_properties(*base[t], *base[t + 1], pmB, p4, tc, tv);
        volume += tv;
        tc.x *= tv; tc.y *= tv; tc.z *= tv;
        moment.x += tc.x; moment.y += tc.y; moment.z += tc.z;
    }
    if (fabs(volume) > 0.0) { moment.x /= volume; moment.y /= volume; moment.z /= volume; }
    centroid = moment;
}

// properties.d:568-650: volume (six pyramids on the faces about the vertex
// mean) and the i, j, k widths between opposite face mid-points
static void hex_cell_properties(const V3 p[8], double& volume, double& iLen, double& jLen, double& kLen)
{
    V3 c{0.125 * (p[0].x + p[1].x + p[2].x + p[3].x + p[4].x + p[5].x + p[6].x + p[7].x),
         0.125 * (p[0].y + p[1].y + p[2].y + p[3].y + p[4].y + p[5].y + p[6].y + p[7].y),
         0.125 * (p[0].z + p[1].z + p[2].z + p[3].z + p[4].z + p[5].z + p[6].z + p[7].z)};
    auto mid = [&](int a, int b2, int c2, int d) {
        return V3{0.25 * (p[a].x + p[b2].x + p[c2].x + p[d].x), 0.25 * (p[a].y + p[b2].y + p[c2].y + p[d].y),
                  0.25 * (p[a].z + p[b2].z + p[c2].z + p[d].z)};
    };
    V3 pmN = mid(3, 2, 6, 7), pmE = mid(1, 2, 6, 5), pmS = mid(0, 1, 5, 4);
    V3 pmW = mid(0, 3, 7, 4), pmT = mid(4, 5, 6, 7), pmB = mid(0, 1, 2, 3);
    double dx = pmE.x - pmW.x, dy = pmE.y - pmW.y, dz = pmE.z - pmW.z;
    iLen = sqrt(dx * dx + dy * dy + dz * dz);
    dx = pmN.x - pmS.x; dy = pmN.y - pmS.y; dz = pmN.z - pmS.z;
    jLen = sqrt(dx * dx + dy * dy + dz * dz);
    dx = pmT.x - pmB.x; dy = pmT.y - pmB.y; dz = pmT.z - pmB.z;
    kLen = sqrt(dx * dx + dy * dy + dz * dz);
    static const int pyr[6][4] = {{6, 7, 3, 2}, {5, 6, 2, 1}, {4, 5, 1, 0}, {7, 4, 0, 3}, {7, 6, 5, 4}, {0, 1, 2, 3}};
    volume = 0.0;
    V3 sc;
    double sv;
    for (int q = 0; q < 6; ++q) {
        pyramid_properties(p[pyr[q][0]], p[pyr[q][1]], p[pyr[q][2]], p[pyr[q][3]], c, sc, sv);
        volume += sv;
    }
    if ((volume < 0.0 && fabs(volume) < 1.0e-12) || (volume >= 0.0 && volume < 1.0e-20)) volume = 0.0;
}

// properties.d:177-237: (x,y)-plane area and i, j widths of the quadrilateral
// cell p0 p1 p2 p3
static void xyplane_quad_cell_properties(const V3& p0, const V3& p1, const V3& p2, const V3& p3, double& area,
                                         double& iLen, double& jLen)
{
    double xA = p1.x, yA = p1.y, xB = p2.x, yB = p2.y, xC = p3.x, yC = p3.y, xD = p0.x, yD = p0.y;
    area = 0.5 * ((xB + xA) * (yB - yA) + (xC + xB) * (yC - yB) + (xD + xC) * (yD - yC) + (xA + xD) * (yA - yD));
    double xN = 0.5 * (xC + xB), yN = 0.5 * (yC + yB);
    double xS = 0.5 * (xD + xA), yS = 0.5 * (yD + yA);
    double xE = 0.5 * (xA + xB), yE = 0.5 * (yA + yB);
    double xW = 0.5 * (xD + xC), yW = 0.5 * (yD + yC);
    double dx = xN - xS, dy = yN - yS;
    jLen = sqrt(dx * dx + dy * dy);
    dx = xE - xW; dy = yE - yW;
    iLen = sqrt(dx * dx + dy * dy);
}

// fvinterface.d:265-297 (planar): the face from A to B, normal to the right
static void face_2d_properties(const V3& A, const V3& B, Frame& fr, double& area)
{
    double LAB = sqrt((B.x - A.x) * (B.x - A.x) + (B.y - A.y) * (B.y - A.y));
    if (LAB > 1.0e-12) {
        fr.n[0] = (B.y - A.y) / LAB; fr.n[1] = -(B.x - A.x) / LAB; fr.n[2] = 0.0;
        fr.t2[0] = 0.0; fr.t2[1] = 0.0; fr.t2[2] = 1.0;
        v3_cross(fr.t1, fr.n, fr.t2);
        area = LAB;
    } else {
        fr = Frame{{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
        area = 0.0;
    }
}

// Ghost-cell lengths of compute_primary_cell_geometric_data
// (sfluidblock.d:673-713): the two ghost layers beside each face copy the
// lengths of the first and second interior cells.
static void mirror_ghost_lengths(Block* b)
{
    const int ndir = (b->cfg.dims == 3) ? 3 : 2;
    const int lo[3] = {b->imin, b->jmin, b->kmin}, hi[3] = {b->imax, b->jmax, b->kmax};
    for (int dir = 0; dir < ndir; ++dir) {
        const int a1 = (dir + 1) % 3, a2 = (dir + 2) % 3;
        for (int q2 = lo[a2]; q2 <= hi[a2]; ++q2)
            for (int q1 = lo[a1]; q1 <= hi[a1]; ++q1)
                for (int side = 0; side < 2; ++side)
                    for (int layer = 1; layer <= 2; ++layer) {
                        int g[3], s[3];
                        g[a1] = s[a1] = q1; g[a2] = s[a2] = q2;
                        g[dir] = side ? hi[dir] + layer : lo[dir] - layer;
                        s[dir] = side ? hi[dir] - layer + 1 : lo[dir] + layer - 1;
                        const size_t gc = cidx(b, g[0], g[1], g[2]), sc = cidx(b, s[0], s[1], s[2]);
                        for (int d = 0; d < 3; ++d) b->glen[d][gc] = b->glen[d][sc];
                    }
    }
}

// sfluidblock.d:94-126, 205-209 (ghost layers = 2); geometry for a uniform
// Cartesian grid (the values properties.d:97-139, 568-650 produce for it).
Block* ref_block_new(Gas* g, int dims, int ni, int nj, int nk, double dx, double dy, double dz)
{
    Block* b = new Block();
    b->gas = g;
    // the reference's defaults (globalconfig.d:440-600): adaptive_hanel_ausmdv, pc, ...
    b->cfg = Config{dims, FLUX_ADAPTIVE_HANEL_AUSMDV, UPD_PC, 2, 1, 1, INTERP_RHOU, 0, 0.20, -0.30};
    b->ni = ni; b->nj = nj; b->nk = (dims == 3) ? nk : 1;
    b->nidim = ni + 4; b->njdim = nj + 4; b->nkdim = (dims == 3) ? nk + 4 : 1;
    b->imin = 2; b->imax = ni + 1; b->jmin = 2; b->jmax = nj + 1;
    b->kmin = (dims == 3) ? 2 : 0; b->kmax = (dims == 3) ? nk + 1 : 0;
    b->nq = 5 + (g->nsp > 1 ? g->nsp : 0);
    b->dx = dx; b->dy = dy; b->dz = dz;
    if (dims == 3) {
        b->vol = dx * dy * dz;
        b->area[0] = dy * dz; b->area[1] = dz * dx; b->area[2] = dx * dy;
        b->frame[0] = Frame{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        b->frame[1] = Frame{{0, 1, 0}, {0, 0, 1}, {1, 0, 0}};
        b->frame[2] = Frame{{0, 0, 1}, {1, 0, 0}, {0, 1, 0}};
    } else {
        b->vol = dx * dy;
        b->area[0] = dy; b->area[1] = dx; b->area[2] = 0.0;
        b->frame[0] = Frame{{1, -0.0, 0}, {-0.0, -1, 0}, {0, 0, 1}};   // fvinterface.d:265-296
        b->frame[1] = Frame{{0, 1, 0}, {1, 0, 0}, {0, 0, 1}};
        b->frame[2] = Frame{{0, 0, 1}, {1, 0, 0}, {0, 1, 0}};
    }
    size_t n = (size_t)b->nidim * b->njdim * b->nkdim;
    // every cell, ghost cells included, starts as the flow state FVCell's
    // constructor makes (fvcell.d:170, gas_state.d:64-88): p = 100 kPa,
    // T = 300 K, at rest, all of the first species, with its sound speed; a
    // ghost cell that no ghost-cell effect writes keeps it
    FlowState init{};
    init.gas.p = 100.0e3; init.gas.T = 300.0; init.gas.massf[0] = 1.0;
    update_thermo_from_pT(g, init.gas);
    update_sound_speed(g, init.gas);
    par_fill(b->fs, n, init);
    for (int t = 0; t < 4; ++t) { par_fill(b->U[t], n * b->nq, 0.0); par_fill(b->dUdt[t], n * b->nq, 0.0); }
    for (int d = 0; d < 3; ++d) par_fill(b->F[d], n * b->nq, 0.0);
    b->bad.assign(n, 0);
    for (int f = 0; f < 6; ++f) b->bc[f].type = BC_SLIP_WALL;   // prep.lua default WallBC_WithSlip
    return b;
}

void ref_block_free(Block* b) { delete b; }

// A block whose geometry is computed from its grid vertices
// (compute_primary_cell_geometric_data, sfluidblock.d:650-842, with the cell
// and face vertex orders of sfluidblock.d:486-645): V holds (ni+1)(nj+1)(nk+1)
// vertices (x, y, z), i fastest (3D; in 2D (ni+1)(nj+1) with z ignored).
// Returns null when a cell has no positive volume.
Block* ref_block_new_grid(Gas* g, int dims, int ni, int nj, int nk, const double* V)
{
    Block* b = ref_block_new(g, dims, ni, nj, nk, 1.0, 1.0, 1.0);
    b->general = true;
    const size_t n = (size_t)b->nidim * b->njdim * b->nkdim;
    b->gvol.assign(n, 0.0);
    for (int d = 0; d < 3; ++d) {
        b->glen[d].assign(n, 0.0);
        b->garea[d].assign(n, 0.0);
        b->gfr[d].assign(n, Frame{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}});
    }
    const int nkv = (dims == 3) ? nk + 1 : 1;
    (void)nkv;
    auto vtx = [&](int i, int j, int k) {
        const double* q = V + 3 * (((size_t)k * (nj + 1) + j) * (ni + 1) + i);
        return V3{q[0], q[1], (dims == 3) ? q[2] : 0.0};
    };
    const int nkc = (dims == 3) ? nk : 1;
    for (int k = 0; k < nkc; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                const size_t c = cidx(b, i + b->imin, j + b->jmin, k + b->kmin);
                double vol, iL, jL, kL = 0.0;
                if (dims == 3) {
                    const V3 p[8] = {vtx(i, j, k), vtx(i + 1, j, k), vtx(i + 1, j + 1, k), vtx(i, j + 1, k),
                                     vtx(i, j, k + 1), vtx(i + 1, j, k + 1), vtx(i + 1, j + 1, k + 1),
                                     vtx(i, j + 1, k + 1)};
                    hex_cell_properties(p, vol, iL, jL, kL);
                } else {
                    xyplane_quad_cell_properties(vtx(i, j, 0), vtx(i + 1, j, 0), vtx(i + 1, j + 1, 0), vtx(i, j + 1, 0),
                                                 vol, iL, jL);
                }
                if (!(vol > 0.0)) { delete b; return nullptr; }
                b->gvol[c] = vol;
                b->glen[0][c] = iL; b->glen[1][c] = jL; b->glen[2][c] = kL;
            }
    // i-faces: west faces of cells i (normal towards +i); j-faces south, k-faces bottom
    for (int k = 0; k < nkc; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i <= ni; ++i) {
                const size_t f = cidx(b, i + b->imin, j + b->jmin, k + b->kmin);
                if (dims == 3) quad_properties(vtx(i, j, k), vtx(i, j + 1, k), vtx(i, j + 1, k + 1), vtx(i, j, k + 1),
                                               b->gfr[0][f], b->garea[0][f]);
                else face_2d_properties(vtx(i, j, 0), vtx(i, j + 1, 0), b->gfr[0][f], b->garea[0][f]);
            }
    for (int k = 0; k < nkc; ++k)
        for (int j = 0; j <= nj; ++j)
            for (int i = 0; i < ni; ++i) {
                const size_t f = cidx(b, i + b->imin, j + b->jmin, k + b->kmin);
                if (dims == 3) quad_properties(vtx(i, j, k), vtx(i, j, k + 1), vtx(i + 1, j, k + 1), vtx(i + 1, j, k),
                                               b->gfr[1][f], b->garea[1][f]);
                else face_2d_properties(vtx(i + 1, j, 0), vtx(i, j, 0), b->gfr[1][f], b->garea[1][f]);
            }
    if (dims == 3)
        for (int k = 0; k <= nk; ++k)
            for (int j = 0; j < nj; ++j)
                for (int i = 0; i < ni; ++i) {
                    const size_t f = cidx(b, i + b->imin, j + b->jmin, k + b->kmin);
                    quad_properties(vtx(i, j, k), vtx(i + 1, j, k), vtx(i + 1, j + 1, k), vtx(i, j + 1, k),
                                    b->gfr[2][f], b->garea[2][f]);
                }
    mirror_ghost_lengths(b);
    return b;
}

// Geometry of a block built from its grid: cell volumes and i/j/k lengths
// [4][ncells] and, per direction, face areas and frames [10][nfaces(d)]
// (area, n, t1, t2), faces of direction d indexed i fastest over
// (ni + [d==0]) x (nj + [d==1]) x (nk + [d==2]).
int ref_block_get_geometry(const Block* b, double* cells, double* faces0, double* faces1, double* faces2)
{
    if (!b->general) return -1;
    const int nkc = (b->cfg.dims == 3) ? b->nk : 1;
    const size_t nc = (size_t)b->ni * b->nj * nkc;
    size_t c = 0;
    for (int k = 0; k < nkc; ++k)
        for (int j = 0; j < b->nj; ++j)
            for (int i = 0; i < b->ni; ++i, ++c) {
                const size_t s = cidx(b, i + b->imin, j + b->jmin, k + b->kmin);
                cells[c] = b->gvol[s];
                for (int d = 0; d < 3; ++d) cells[(1 + d) * nc + c] = b->glen[d][s];
            }
    double* out[3] = {faces0, faces1, faces2};
    const int ndir = (b->cfg.dims == 3) ? 3 : 2;
    for (int d = 0; d < ndir; ++d) {
        if (!out[d]) continue;
        const int fi = b->ni + (d == 0), fj = b->nj + (d == 1), fk = nkc + (d == 2);
        const size_t nf = (size_t)fi * fj * fk;
        size_t f = 0;
        for (int k = 0; k < fk; ++k)
            for (int j = 0; j < fj; ++j)
                for (int i = 0; i < fi; ++i, ++f) {
                    const size_t s = cidx(b, i + b->imin, j + b->jmin, k + b->kmin);
                    const Frame& fr = b->gfr[d][s];
                    out[d][f] = b->garea[d][s];
                    for (int q = 0; q < 3; ++q) {
                        out[d][(1 + q) * nf + f] = fr.n[q];
                        out[d][(4 + q) * nf + f] = fr.t1[q];
                        out[d][(7 + q) * nf + f] = fr.t2[q];
                    }
                }
    }
    return 0;
}

void ref_block_set_shock_tolerances(Block* b, double shear_tolerance, double compression_tolerance)
{
    b->cfg.shear_tolerance = shear_tolerance;
    b->cfg.compression_tolerance = compression_tolerance;
}

void ref_block_set_M_inf(Block* b, double M_inf) { b->cfg.M_inf = M_inf; }

void ref_block_set_config(Block* b, int flux_calc, int update_scheme, int interp_order, int apply_limiter,
                          int extrema_clipping, int thermo_interp, int max_invalid_cells)
{
    b->cfg.flux_calc = flux_calc; b->cfg.update_scheme = update_scheme; b->cfg.interp_order = interp_order;
    b->cfg.apply_limiter = apply_limiter; b->cfg.extrema_clipping = extrema_clipping;
    b->cfg.thermo_interp = thermo_interp; b->cfg.max_invalid_cells = max_invalid_cells;
}

// A static profile: states [2][face cells][9 + nsp] = rho, p, T, u, a, vx, vy,
// vz, S, massf...
void ref_block_set_bc_profile(Block* b, int face, const double* states)
{
    const int dir = (face == F_EAST || face == F_WEST) ? 0 : (face == F_NORTH || face == F_SOUTH) ? 1 : 2;
    const size_t nfc = (size_t)((dir == 0) ? b->nj : b->ni) * ((dir == 2) ? b->nj : b->nk);
    const int w = 9 + b->gas->nsp;
    BC& bc = b->bc[face];
    bc.type = BC_INFLOW_PROFILE;
    bc.profile.assign(2 * nfc, FlowState{});
    for (size_t g = 0; g < 2 * nfc; ++g) {
        const double* st = states + g * w;
        FlowState& f = bc.profile[g];
        f.gas.rho = st[0]; f.gas.p = st[1]; f.gas.T = st[2]; f.gas.u = st[3]; f.gas.a = st[4];
        f.vel[0] = st[5]; f.vel[1] = st[6]; f.vel[2] = st[7];
        f.S = st[8] != 0.0;
        for (int s = 0; s < b->gas->nsp; ++s) f.gas.massf[s] = st[9 + s];
    }
}

static void face_ghost_lengths(Block* b, int face);

// fs vector layout: [rho,p,T,u,a,vx,vy,vz,massf...].  Returns -1 for a
// boundary the oracle does not provide on a block built from its grid (those
// without ghost data, whose one-sided reconstruction takes one cell length).
int ref_block_set_bc(Block* b, int face, int type, const double* fs)
{
    if (b->general && type == BC_SLIP_WALL_1) return -1;
    b->bc[face].type = type;
    if (b->general) face_ghost_lengths(b, face);
    if (fs) {
        FlowState& f = b->bc[face].fs;
        f.gas.rho = fs[0]; f.gas.p = fs[1]; f.gas.T = fs[2]; f.gas.u = fs[3]; f.gas.a = fs[4];
        f.vel[0] = fs[5]; f.vel[1] = fs[6]; f.vel[2] = fs[7];
        for (int i = 0; i < b->gas->nsp; ++i) f.gas.massf[i] = fs[8 + i];
    }
    return 0;
}

// fvcell.d:494-560 (ftl = 0 unless given)
static void encode_conserved(const Block* b, const FlowState& fs, double* U)
{
    double myrho = fs.gas.rho;
    U[0] = myrho;
    U[1] = fs.gas.rho * fs.vel[0]; U[2] = fs.gas.rho * fs.vel[1]; U[3] = fs.gas.rho * fs.vel[2];
    double u = fs.gas.u;
    double ke = 0.5 * (fs.vel[0] * fs.vel[0] + fs.vel[1] * fs.vel[1] + fs.vel[2] * fs.vel[2]);
    U[4] = fs.gas.rho * (u + ke);
    if (b->gas->nsp > 1)
        for (int i = 0; i < b->gas->nsp; ++i) U[5 + i] = fs.gas.rho * fs.gas.massf[i];
}

// fvcell.d:563-732.  Returns 0, -1 (non-positive mass), -2 (bad massf or thermo).
static int decode_conserved(Block* b, FlowState& fs, double* U)
{
    Gas* g = b->gas;
    if (!(U[0] > 0.0)) return -1;
    double rho = U[0];
    fs.gas.rho = rho;
    double dinv = 1.0 / rho;
    fs.vel[0] = U[1] * dinv; fs.vel[1] = U[2] * dinv; fs.vel[2] = U[3] * dinv;
    double rE = U[4];
    double u = rE * dinv;
    double ke = 0.5 * (fs.vel[0] * fs.vel[0] + fs.vel[1] * fs.vel[1] + fs.vel[2] * fs.vel[2]);
    u -= ke;
    fs.gas.u = u;
    if (g->nsp > 1) {
        for (int i = 0; i < g->nsp; ++i) fs.gas.massf[i] = U[5 + i] * dinv;
        if (scale_mass_fractions(fs.gas.massf, g->nsp) != 0) return -2;
    } else {
        fs.gas.massf[0] = U[0] * dinv;   // U.massf[0] tracks U.mass exactly for one species
    }
    if (update_thermo_from_rhou(g, fs.gas) != 0) {
        // ignore_low_T_thermo_update_failure (globalconfig.d:510, default true),
        // suggested_low_T_value 200 K (globalconfig.d:511)
        if (rho > 0.0) {
            fs.gas.T = 200.0;
            if (update_thermo_from_rhoT(g, fs.gas) != 0) return -2;
            encode_conserved(b, fs, U);
        } else {
            return -2;
        }
    }
    if (update_sound_speed(g, fs.gas) != 0) return -2;
    return 0;
}

// flowstate.d:345-388 with gas_state.d:199-241 and FlowStateLimits (fvcore.d:18-27)
static bool check_data(const Block* b, const FlowState& fs)
{
    bool ok = true;
    const GasState& q = fs.gas;
    if (!std::isfinite(q.rho) || q.rho < 1.01 * 0.0) ok = false;
    if (!std::isfinite(q.T) || q.T < 1.01 * 1.0) ok = false;
    if (!std::isfinite(q.p)) ok = false;
    if (!std::isfinite(q.a)) ok = false;
    double f_sum = 0.0;
    for (int i = 0; i < b->gas->nsp; ++i) f_sum += q.massf[i];
    if (f_sum < 0.99 || f_sum > 1.01 || !std::isfinite(f_sum)) ok = false;
    if (fabs(fs.vel[0]) > 30000.0 || fabs(fs.vel[1]) > 30000.0 || fabs(fs.vel[2]) > 30000.0) ok = false;
    if (q.T < 0.0) ok = false;
    if (q.T > 50000.0) ok = false;
    return ok;
}

// Initial state from (p, T, vel, massf) per active cell, as the initial flow
// writer does (flowstate.d:177-216 constructor from p,T: update_thermo_from_pT,
// update_sound_speed), then encode.  Arrays are [ncells] in i-fastest order;
// massf is [nsp][ncells] (may be null for one species).
int ref_block_set_from_pT(Block* b, const double* p, const double* T, const double* vx, const double* vy,
                          const double* vz, const double* massf)
{
    const int nsp = b->gas->nsp;
    const size_t nc = (size_t)b->ni * b->nj * b->nk;
    size_t c = 0;
    for (int k = b->kmin; k <= b->kmax; ++k)
        for (int j = b->jmin; j <= b->jmax; ++j)
            for (int i = b->imin; i <= b->imax; ++i, ++c) {
                FlowState& f = b->fs[cidx(b, i, j, k)];
                f.gas.p = p[c]; f.gas.T = T[c];
                f.vel[0] = vx[c]; f.vel[1] = vy[c]; f.vel[2] = vz[c];
                for (int s = 0; s < nsp; ++s) f.gas.massf[s] = massf ? massf[s * nc + c] : 1.0;
                if (update_thermo_from_pT(b->gas, f.gas) != 0) return -1;
                if (update_sound_speed(b->gas, f.gas) != 0) return -1;
                encode_conserved(b, f, &b->U[0][cidx(b, i, j, k) * b->nq]);
            }
    return 0;
}

// Restart path: U0 given ([nq][ncells]); decode into fs.  Tguess (may be null)
// seeds the thermally perfect Newton as a restart's stored T would.
int ref_block_set_U(Block* b, const double* U, const double* Tguess)
{
    const size_t nc = (size_t)b->ni * b->nj * b->nk;
    int bad = 0;
    #pragma omp parallel for collapse(2) schedule(static) reduction(+:bad) if (b->ni * b->nj * b->nk >= OMP_MIN_CELLS)
    for (int k = b->kmin; k <= b->kmax; ++k)
        for (int j = b->jmin; j <= b->jmax; ++j) {
            size_t c = ((size_t)(k - b->kmin) * b->nj + (j - b->jmin)) * b->ni;
            for (int i = b->imin; i <= b->imax; ++i, ++c) {
                size_t ci = cidx(b, i, j, k);
                double* u = &b->U[0][ci * b->nq];
                for (int q = 0; q < b->nq; ++q) u[q] = U[q * nc + c];
                if (Tguess) b->fs[ci].gas.T = Tguess[c];
                if (decode_conserved(b, b->fs[ci], u) != 0) bad++;
            }
        }
    return bad ? -1 : 0;
}

void ref_block_get_U(const Block* b, int ftl, double* U)
{
    const size_t nc = (size_t)b->ni * b->nj * b->nk;
    #pragma omp parallel for collapse(2) schedule(static) if (b->ni * b->nj * b->nk >= OMP_MIN_CELLS)
    for (int k = b->kmin; k <= b->kmax; ++k)
        for (int j = b->jmin; j <= b->jmax; ++j) {
            size_t c = ((size_t)(k - b->kmin) * b->nj + (j - b->jmin)) * b->ni;
            for (int i = b->imin; i <= b->imax; ++i, ++c) {
                const double* u = &b->U[ftl][cidx(b, i, j, k) * b->nq];
                for (int q = 0; q < b->nq; ++q) U[q * nc + c] = u[q];
            }
        }
}

// shock indicators of the active cells, i fastest
void ref_block_get_S(const Block* b, int* out)
{
    size_t c = 0;
    for (int k = b->kmin; k <= b->kmax; ++k)
        for (int j = b->jmin; j <= b->jmax; ++j)
            for (int i = b->imin; i <= b->imax; ++i, ++c) out[c] = b->fs[cidx(b, i, j, k)].S;
}

// fs fields: rho,p,T,u,a,vx,vy,vz,massf... as [nfield][ncells]
void ref_block_get_fs(const Block* b, double* out)
{
    const size_t nc = (size_t)b->ni * b->nj * b->nk;
    #pragma omp parallel for collapse(2) schedule(static) if (b->ni * b->nj * b->nk >= OMP_MIN_CELLS)
    for (int k = b->kmin; k <= b->kmax; ++k)
        for (int j = b->jmin; j <= b->jmax; ++j) {
            size_t c = ((size_t)(k - b->kmin) * b->nj + (j - b->jmin)) * b->ni;
            for (int i = b->imin; i <= b->imax; ++i, ++c) {
                const FlowState& f = b->fs[cidx(b, i, j, k)];
                out[0 * nc + c] = f.gas.rho; out[1 * nc + c] = f.gas.p; out[2 * nc + c] = f.gas.T;
                out[3 * nc + c] = f.gas.u; out[4 * nc + c] = f.gas.a;
                out[5 * nc + c] = f.vel[0]; out[6 * nc + c] = f.vel[1]; out[7 * nc + c] = f.vel[2];
                for (int s = 0; s < b->gas->nsp; ++s) out[(8 + s) * nc + c] = f.gas.massf[s];
            }
        }
}

// Every cell's flow state, ghost cells included ([8 + nsp][nidim*njdim*nkdim],
// the reference's own slot order), after the ghost cells are filled, and the
// shock indicators of all slots.
void ref_block_get_fs_full(Block* b, int fill, double* out, int* S)
{
    if (fill) apply_bcs(b);
    const size_t n = b->fs.size();
    for (size_t c = 0; c < n; ++c) {
        const FlowState& f = b->fs[c];
        out[0 * n + c] = f.gas.rho; out[1 * n + c] = f.gas.p; out[2 * n + c] = f.gas.T;
        out[3 * n + c] = f.gas.u; out[4 * n + c] = f.gas.a;
        out[5 * n + c] = f.vel[0]; out[6 * n + c] = f.vel[1]; out[7 * n + c] = f.vel[2];
        for (int s = 0; s < b->gas->nsp; ++s) out[(8 + s) * n + c] = f.gas.massf[s];
        if (S) S[c] = f.S;
    }
}

// ------------------------------------------------------- block connections
//
// full_face_copy.d:142-962 (set_up_cell_mapping_phase0): the ghost cell at
// layer L (1, 2) outside face f, at in-face position (a, b), copies the
// interior cell of the other block at depth L-1 inside face g.  The in-face
// coordinates of a face are the index axes the reference loops over:
// north/south (i, k), east/west (j, k), top/bottom (i, j); 2D faces use the
// first one only.  The reference lists the mapping case by case; it is
// restated here by the rule the cases follow:
//   2D: traversing both faces counter-clockwise around their own blocks, the
//       positions run in opposite senses, so the coordinate is reversed when
//       f and g are both in {south, east} or both in {north, west};
//   3D: with x~ = n - 1 - x, orientation o gives (a', b') = (a, b), (b~, a),
//       (a~, b~), (b, a~) for o = 0..3; the first output is reversed when the
//       two faces have the same handedness (sign of (a x b) . outward normal:
//       +1 for east, south, top, -1 for north, west, bottom); and o = 1, 3
//       trade places when f is right-handed.
static const int FACE_AX[6][2] = {{0, 2}, {1, 2}, {0, 2}, {1, 2}, {0, 1}, {0, 1}};
static const int FACE_NORMAL[6] = {1, 0, 1, 0, 2, 2};
static const int FACE_PLUS[6] = {1, 1, 0, 0, 1, 0};
static const int FACE_HAND[6] = {-1, 1, 1, -1, 1, -1};

// Source cell (0-based active indices src[3]) in a block of sizes n[3] for the
// ghost at (layer, a, b) of face f connected to face g with orientation o.
void ref_connection_source(int dims, int f, int g, int o, const int* n, int layer, int a, int b, int* src)
{
    int out0, out1 = 0;
    if (dims == 2) {
        const bool grp_f = (f == F_SOUTH || f == F_EAST), grp_g = (g == F_SOUTH || g == F_EAST);
        const int ax = (g == F_NORTH || g == F_SOUTH) ? 0 : 1;
        out0 = (grp_f == grp_g) ? n[ax] - 1 - a : a;
        src[ax] = out0;
        src[1 - ax] = FACE_PLUS[g] ? n[1 - ax] - layer : layer - 1;
        src[2] = 0;
        return;
    }
    const int A0 = FACE_AX[g][0], A1 = FACE_AX[g][1];
    if (FACE_HAND[f] == 1 && (o & 1)) o = (o + 2) % 4;
    switch (o) {
    case 0: out0 = a; out1 = b; break;
    case 1: out0 = n[A0] - 1 - b; out1 = a; break;
    case 2: out0 = n[A0] - 1 - a; out1 = n[A1] - 1 - b; break;
    default: out0 = b; out1 = n[A1] - 1 - a; break;
    }
    if (FACE_HAND[f] == FACE_HAND[g]) out0 = n[A0] - 1 - out0;
    src[A0] = out0;
    src[A1] = out1;
    const int dn = FACE_NORMAL[g];
    src[dn] = FACE_PLUS[g] ? n[dn] - layer : layer - 1;
}

// Connect face `face` of b to face `other_face` of other (one direction; the
// caller connects both sides, as the reference's connectBlocks does).
void ref_block_connect(Block* b, int face, Block* other, int other_face, int orientation)
{
    b->bc[face].type = BC_CONNECTED;
    b->bc[face].other = other;
    b->bc[face].other_face = other_face;
    b->bc[face].orientation = orientation;
    if (b->general) face_ghost_lengths(b, face);
}

// The ghost-cell lengths of one face of a block built from its grid: mirrored
// from the interior (sfluidblock.d:673-713), or, across a full-face
// connection (a connected or periodic face), the i/j/k lengths of the mapped
// cell itself, as exchange_geometry_phase2 copies them
// (full_face_copy.d:1042-1085, CopyDataOption.grid, fvcell.d:220-230).
static void face_ghost_lengths(Block* b, int face)
{
    const int dims = b->cfg.dims;
    const int dir = (face == F_EAST || face == F_WEST) ? 0 : (face == F_NORTH || face == F_SOUTH) ? 1 : 2;
    const bool plus = (face == F_EAST || face == F_NORTH || face == F_TOP);
    const int lo[3] = {b->imin, b->jmin, b->kmin}, hi[3] = {b->imax, b->jmax, b->kmax};
    const int a1 = (dir + 1) % 3, a2 = (dir + 2) % 3;
    const BC& bc = b->bc[face];
    for (int q2 = lo[a2]; q2 <= hi[a2]; ++q2)
        for (int q1 = lo[a1]; q1 <= hi[a1]; ++q1)
            for (int layer = 1; layer <= 2; ++layer) {
                int g[3], s[3];
                g[a1] = s[a1] = q1; g[a2] = s[a2] = q2;
                g[dir] = plus ? hi[dir] + layer : lo[dir] - layer;
                const size_t gc = cidx(b, g[0], g[1], g[2]);
                const Block* src_blk = b;
                size_t sc;
                if (bc.type == BC_CONNECTED && bc.other) {
                    const Block* o = bc.other;
                    const int on[3] = {o->ni, o->nj, o->nk};
                    const int gidx[3] = {g[0] - b->imin, g[1] - b->jmin, g[2] - b->kmin};
                    const int fa = gidx[FACE_AX[face][0]];
                    const int fb = (dims == 3) ? gidx[FACE_AX[face][1]] : 0;
                    int src[3];
                    ref_connection_source(dims, face, bc.other_face, bc.orientation, on, layer, fa, fb, src);
                    src_blk = o;
                    sc = cidx(o, src[0] + o->imin, src[1] + o->jmin, src[2] + o->kmin);
                } else if (bc.type == BC_PERIODIC) {
                    s[dir] = plus ? lo[dir] + layer - 1 : hi[dir] - layer + 1;
                    sc = cidx(b, s[0], s[1], s[2]);
                } else {
                    s[dir] = plus ? hi[dir] - layer + 1 : lo[dir] + layer - 1;
                    sc = cidx(b, s[0], s[1], s[2]);
                }
                for (int d = 0; d < 3; ++d) b->glen[d][gc] = cell_len(src_blk, d, sc);
            }
}

// Ghost-cell effects (bc/ghost_cell_effect): flow_state_copy.d:66-169,
// extrapolate_copy.d:82-… (x_order 0), internal_copy_then_reflect.d:94-310 with
// ghost_cell.d:32-43, full_face_copy.d:1277-1357 for a block connected to
// itself across opposite faces (periodic).
static void apply_bcs(Block* b)
{
    const int dims = b->cfg.dims;
    const int nfaces = (dims == 3) ? 6 : 4;
    for (int face = 0; face < nfaces; ++face) {
        const BC& bc = b->bc[face];
        int dir = (face == F_EAST || face == F_WEST) ? 0 : (face == F_NORTH || face == F_SOUTH) ? 1 : 2;
        bool plus = (face == F_EAST || face == F_NORTH || face == F_TOP);
        int lo[3] = {b->imin, b->jmin, b->kmin}, hi[3] = {b->imax, b->jmax, b->kmax};
        int n[3] = {b->ni, b->nj, b->nk};
        int a1 = (dir + 1) % 3, a2 = (dir + 2) % 3;
        for (int q2 = lo[a2]; q2 <= hi[a2]; ++q2)
            for (int q1 = lo[a1]; q1 <= hi[a1]; ++q1)
                for (int layer = 1; layer <= 2; ++layer) {
                    int g[3], s[3];
                    g[a1] = s[a1] = q1; g[a2] = s[a2] = q2;
                    g[dir] = plus ? hi[dir] + layer : lo[dir] - layer;
                    FlowState& dest = b->fs[cidx(b, g[0], g[1], g[2])];
                    switch (bc.type) {
                    case BC_INFLOW:
                        dest = bc.fs;
                        break;
                    case BC_INFLOW_PROFILE: {
                        // flow_state_copy_from_profile.d: the nearest profile
                        // point's state, matched by the caller
                        const int gi[3] = {g[0] - b->imin, g[1] - b->jmin, g[2] - b->kmin};
                        const int na = (dir == 0) ? b->nj : b->ni;
                        const int fa = (dir == 0) ? gi[1] : gi[0], fb = (dir == 2) ? gi[1] : ((dims == 3) ? gi[2] : 0);
                        const size_t nfc = (size_t)na * ((dir == 2) ? b->nj : b->nk);
                        dest = bc.profile[(layer - 1) * nfc + fa + (size_t)fb * na];
                        break;
                    }
                    case BC_EXTRAPOLATE:
                    case BC_OUTFLOW_SIMPLE_FLUX:
                        s[dir] = plus ? hi[dir] : lo[dir];
                        dest = b->fs[cidx(b, s[0], s[1], s[2])];
                        break;
                    case BC_EXTRAPOLATE_1: {
                        // extrapolate_copy.d:102-186: linear in the primitive
                        // variables from the two cells inside (for layer 2: the
                        // first ghost and the boundary cell), then T and p from
                        // the EOS; the sound speed and shock indicator are not
                        // written
                        int s1[3] = {s[0], s[1], s[2]}, s2[3] = {s[0], s[1], s[2]};
                        const int edge = plus ? hi[dir] : lo[dir], in = plus ? -1 : 1;
                        s1[dir] = (layer == 1) ? edge : edge - in;
                        s2[dir] = (layer == 1) ? edge + in : edge;
                        const FlowState c1 = b->fs[cidx(b, s1[0], s1[1], s1[2])];
                        const FlowState c2 = b->fs[cidx(b, s2[0], s2[1], s2[2])];
                        dest.gas.rho = 2.0 * c1.gas.rho - c2.gas.rho;
                        dest.gas.u = 2.0 * c1.gas.u - c2.gas.u;
                        const int nsp = b->gas->nsp;
                        if (nsp > 1) {
                            for (int q = 0; q < nsp; ++q) dest.gas.massf[q] = 2.0 * c1.gas.massf[q] - c2.gas.massf[q];
                            scale_mass_fractions(dest.gas.massf, nsp);
                        } else {
                            dest.gas.massf[0] = 1.0;
                        }
                        update_thermo_from_rhou(b->gas, dest.gas);
                        for (int d = 0; d < 3; ++d) dest.vel[d] = 2.0 * c1.vel[d] - c2.vel[d];
                        break;
                    }
                    case BC_SLIP_WALL_1:
                        break;   // no ghost-cell data
                    case BC_PERIODIC:
                        s[dir] = plus ? lo[dir] + layer - 1 : hi[dir] - layer + 1;
                        dest = b->fs[cidx(b, s[0], s[1], s[2])];
                        break;
                    case BC_CONNECTED: {
                        // full_face_copy.d:1277-1357, 1633-1646: copy the mapped
                        // interior flow state of the other block
                        const Block* o = bc.other;
                        const int on[3] = {o->ni, o->nj, o->nk};
                        const int gidx[3] = {g[0] - b->imin, g[1] - b->jmin, g[2] - b->kmin};
                        const int fa = gidx[FACE_AX[face][0]];
                        const int fb = (dims == 3) ? gidx[FACE_AX[face][1]] : 0;
                        int src[3];
                        ref_connection_source(dims, face, bc.other_face, bc.orientation, on, layer, fa, fb, src);
                        dest = o->fs[cidx(o, src[0] + o->imin, src[1] + o->jmin, src[2] + o->kmin)];
                        break;
                    }
                    case BC_SLIP_WALL:
                    default: {
                        s[dir] = plus ? hi[dir] - layer + 1 : lo[dir] + layer - 1;
                        dest = b->fs[cidx(b, s[0], s[1], s[2])];
                        // reflected about the boundary face of the boundary cell
                        int fpos[3] = {g[0], g[1], g[2]};
                        fpos[dir] = plus ? hi[dir] + 1 : lo[dir];
                        const Frame& fr = face_frame(b, dir, cidx(b, fpos[0], fpos[1], fpos[2]));
                        to_local(dest.vel, fr);
                        dest.vel[0] = -(dest.vel[0]);
                        to_global(dest.vel, fr);
                        break;
                    }
                    }
                    (void)n;
                }
    }
}

// fluidblock.d:415-431 with shockdetectors.d (PJ_ShockDetector): a face is a
// shock point when the normal velocity jump across it, scaled by the smaller
// sound speed, is below the compression tolerance; a cell is a shock point
// when any of its faces is.  The ghost cells keep the indicator they were
// filled with (none for x_order-1 extrapolation, which does not write it).
static const int MINUS_FACE[3] = {F_WEST, F_SOUTH, F_BOTTOM};
static const int PLUS_FACE[3] = {F_EAST, F_NORTH, F_TOP};
// bc.lua: WallBC_WithSlip1 is the one boundary here without ghost-cell data
static inline bool ghost_data_available(const Block* b, int face) { return b->bc[face].type != BC_SLIP_WALL_1; }

static inline int pj_shock(const FlowState& cL, const FlowState& cR, const Frame& fr, double tol)
{
    const double uL = cL.vel[0] * fr.n[0] + cL.vel[1] * fr.n[1] + cL.vel[2] * fr.n[2];
    const double uR = cR.vel[0] * fr.n[0] + cR.vel[1] * fr.n[1] + cR.vel[2] * fr.n[2];
    const double aL = cL.gas.a, aR = cR.gas.a;
    const double a_min = (aL < aR) ? aL : aR;
    return ((uR - uL) / a_min) < tol;
}

static void detect_shock_points(Block* b)
{
    const int ndir = (b->cfg.dims == 3) ? 3 : 2;
    const double tol = b->cfg.compression_tolerance;
    const size_t st[3] = {1, (size_t)b->nidim, (size_t)b->nidim * b->njdim};
    std::vector<int> S((size_t)b->nidim * b->njdim * b->nkdim, 0);
    for (int k = b->kmin; k <= b->kmax; ++k)
        for (int j = b->jmin; j <= b->jmax; ++j)
            for (int i = b->imin; i <= b->imax; ++i) {
                const size_t c = cidx(b, i, j, k);
                const int pos[3] = {i, j, k}, lo[3] = {b->imin, b->jmin, b->kmin}, hi[3] = {b->imax, b->jmax, b->kmax};
                int s = 0;
                for (int d = 0; d < ndir; ++d) {
                    const Frame& fm = face_frame(b, d, c);
                    const Frame& fp = face_frame(b, d, c + st[d]);
                    const FlowState& cs = b->fs[c];
                    // a boundary face without ghost data has one cell:
                    // shockdetectors.d:39-54 compares its velocity towards the wall
                    if (pos[d] == lo[d] && !ghost_data_available(b, MINUS_FACE[d])) {
                        const double un = cs.vel[0] * fm.n[0] + cs.vel[1] * fm.n[1] + cs.vel[2] * fm.n[2];
                        s |= (un / cs.gas.a) < tol;
                    } else {
                        s |= pj_shock(b->fs[c - st[d]], cs, fm, tol);   // minus face
                    }
                    if (pos[d] == hi[d] && !ghost_data_available(b, PLUS_FACE[d])) {
                        const double un = cs.vel[0] * fp.n[0] + cs.vel[1] * fp.n[1] + cs.vel[2] * fp.n[2];
                        s |= ((-un) / cs.gas.a) < tol;
                    } else {
                        s |= pj_shock(cs, b->fs[c + st[d]], fp, tol);   // plus face
                    }
                }
                S[c] = s;
            }
    for (int k = b->kmin; k <= b->kmax; ++k)
        for (int j = b->jmin; j <= b->jmax; ++j)
            for (int i = b->imin; i <= b->imax; ++i) b->fs[cidx(b, i, j, k)].S = S[cidx(b, i, j, k)];
}

// sfluidblock.d:2016-2346, the general branch: l2r2 reconstruction, the
// one-sided variants and wall fluxes next to boundaries without ghost data,
// faces whose flux the boundary computes skipped; then the boundary fluxes
// (postConvFluxAction SimpleOutflowFlux, boundary_flux_effect.d:367-546).
static void convective_flux(Block* b)
{
    Gas* g = b->gas;
    const Config& cfg = b->cfg;
    const bool do_reconstruction = (cfg.interp_order > 1);
    const int nq = b->nq;
    const int ndir = (cfg.dims == 3) ? 3 : 2;
    const double len[3] = {b->dx, b->dy, b->dz};
    for (int dir = 0; dir < ndir; ++dir) {
        int st[3] = {0, 0, 0}; st[dir] = 1;
        const size_t stride = (size_t)st[0] + (size_t)st[1] * b->nidim + (size_t)st[2] * b->nidim * b->njdim;
        const int ihi = b->imax + st[0], jhi = b->jmax + st[1], khi = b->kmax + st[2];
        const bool ng_m = !ghost_data_available(b, MINUS_FACE[dir]), ng_p = !ghost_data_available(b, PLUS_FACE[dir]);
        const bool bf_m = b->bc[MINUS_FACE[dir]].type == BC_OUTFLOW_SIMPLE_FLUX;
        const bool bf_p = b->bc[PLUS_FACE[dir]].type == BC_OUTFLOW_SIMPLE_FLUX;
        const int lo = (dir == 0) ? b->imin : (dir == 1) ? b->jmin : b->kmin;
        const int hi = (dir == 0) ? b->imax : (dir == 1) ? b->jmax : b->kmax;
        #pragma omp parallel for collapse(2) schedule(static) if (b->ni * b->nj * b->nk >= OMP_MIN_CELLS)
        for (int k = b->kmin; k <= khi; ++k)
            for (int j = b->jmin; j <= jhi; ++j)
                for (int i = b->imin; i <= ihi; ++i) {
                    const int pos = (dir == 0) ? i : (dir == 1) ? j : k;
                    size_t cR0 = cidx(b, i, j, k);
                    size_t cL0 = cR0 - stride, cL1 = cL0 - stride, cR1 = cR0 + stride;
                    const Frame& fr = face_frame(b, dir, cR0);
                    const bool wall_m = (pos == lo) && ng_m, wall_p = (pos == hi + 1) && ng_p;
                    FlowState Lft = wall_m ? b->fs[cR0] : b->fs[cL0];
                    FlowState Rght = wall_p ? b->fs[cL0] : b->fs[cR0];
                    if (do_reconstruction) {
                        const double l = len[dir];
                        if (wall_m) interp_wall_side(g, cfg, fr, b->fs[cR0], b->fs[cR1], l, Rght);
                        else if (pos == lo + 1 && ng_m)
                            interp_l2r1_l1r2(g, cfg, fr, false, b->fs[cL0], b->fs[cR0], b->fs[cR1], l, Lft, Rght);
                        else if (pos == hi && ng_p)
                            interp_l2r1_l1r2(g, cfg, fr, true, b->fs[cL1], b->fs[cL0], b->fs[cR0], l, Lft, Rght);
                        else if (wall_p) interp_wall_side(g, cfg, fr, b->fs[cL0], b->fs[cL1], l, Lft);
                        else if (b->general)
                            interp_l2r2(g, cfg, fr, b->fs[cL1], b->fs[cL0], b->fs[cR0], b->fs[cR1], cell_len(b, dir, cL1),
                                        cell_len(b, dir, cL0), cell_len(b, dir, cR0), cell_len(b, dir, cR1), Lft, Rght);
                        else interp_l2r2(g, cfg, fr, b->fs[cL1], b->fs[cL0], b->fs[cR0], b->fs[cR1], l, l, l, l,
                                         Lft, Rght);
                    }
                    if ((pos == lo && bf_m) || (pos == hi + 1 && bf_p)) continue;
                    double Fq[NQ_MAX];
                    if (wall_m) wall_flux(g, Rght, fr, true, nq, Fq);
                    else if (wall_p) wall_flux(g, Lft, fr, false, nq, Fq);
                    else compute_interface_flux(g, cfg, Lft, Rght, fr, Fq);
                    double* dst = &b->F[dir][cR0 * nq];
                    for (int q = 0; q < nq; ++q) dst[q] = Fq[q];
                }
        // boundary faces whose flux the boundary condition computes, from the
        // interior cell next to them
        for (int side = 0; side < 2; ++side) {
            if (!(side ? bf_p : bf_m)) continue;
            for (int k = b->kmin; k <= b->kmax; ++k)
                for (int j = b->jmin; j <= b->jmax; ++j)
                    for (int i = b->imin; i <= b->imax; ++i) {
                        const int pos = (dir == 0) ? i : (dir == 1) ? j : k;
                        if (pos != (side ? hi : lo)) continue;
                        const size_t c = cidx(b, i, j, k);
                        const size_t face = side ? c + stride : c;
                        outflow_flux(g, b->fs[c], side ? 1 : -1, face_frame(b, dir, face), nq, &b->F[dir][face * nq]);
                    }
        }
    }
}

// fvcell.d:735-875 with the source vector Q built as simcore.d:2442-2465
// builds it: cleared, then add_inviscid_source_vector (fvcell.d:2156-2199;
// the axisymmetric pressure term of the y-momentum equation, p * areaxy /
// volume), then the user-defined source terms of the cell
// (user_defined_source_terms.d:23-56 adds them to Q); dUdt = integral / V + Q.
static void time_derivatives(Block* b, size_t c, double* dUdt)
{
    const int nq = b->nq;
    const bool three_d = (b->cfg.dims == 3);
    const size_t sj = b->nidim, sk = (size_t)b->nidim * b->njdim;
    // face order north, east, south, west, top, bottom (sfluidblock.d:507-518)
    const double* Fs[6] = {&b->F[1][(c + sj) * nq], &b->F[0][(c + 1) * nq], &b->F[1][c * nq],
                           &b->F[0][c * nq], three_d ? &b->F[2][(c + sk) * nq] : nullptr,
                           three_d ? &b->F[2][c * nq] : nullptr};
    const double outsign[6] = {1.0, 1.0, -1.0, -1.0, 1.0, -1.0};
    double A[6] = {face_area(b, 1, c + sj), face_area(b, 0, c + 1), face_area(b, 1, c), face_area(b, 0, c),
                   three_d ? face_area(b, 2, c + sk) : 0.0, three_d ? face_area(b, 2, c) : 0.0};
    double volume = cell_vol(b, c), Qy = 0.0;
    if (b->axi) {
        const int i = (int)(c % b->nidim) - b->imin, j = (int)(c / b->nidim) - b->jmin;
        const size_t cc = (size_t)j * b->ni + i;
        A[0] = b->axi_aj[(size_t)(j + 1) * b->ni + i];
        A[1] = b->axi_ai[(size_t)j * (b->ni + 1) + i + 1];
        A[2] = b->axi_aj[(size_t)j * b->ni + i];
        A[3] = b->axi_ai[(size_t)j * (b->ni + 1) + i];
        volume = b->axi_vol[cc];
        Qy = 0.0 + b->fs[c].gas.p * b->axi_areaxy[cc] / volume;
    }
    double integral[NQ_MAX];
    for (int q = 0; q < nq; ++q) integral[q] = 0.0;
    const int nf = three_d ? 6 : 4;
    for (int f = 0; f < nf; ++f) {
        const double* myF = Fs[f];
        double area = outsign[f] * A[f];
        integral[0] -= myF[0] * area;
        integral[1] -= myF[1] * area;
        integral[2] -= myF[2] * area;
        if (three_d) integral[3] -= myF[3] * area;
        integral[4] -= myF[4] * area;
        for (int q = 5; q < nq; ++q) integral[q] -= myF[q] * area;
    }
    double vol_inv = 1.0 / volume;
    const double* udf = b->udf_src.empty() ? nullptr : &b->udf_src[c * nq];
    for (int q = 0; q < nq; ++q) {
        double Q = (q == 2) ? Qy : 0.0;
        if (udf) Q += udf[q];
        dUdt[q] = vol_inv * integral[q] + Q;
    }
}

// User-defined source terms, constant in time: Q [nq][ncells] in the block's
// conserved order (mass, x/y/z momentum, total energy, then species
// densities), or NULL for none.
void ref_block_set_source_terms(Block* b, const double* Q)
{
    if (!Q) { b->udf_src.clear(); return; }
    const size_t nc = (size_t)b->ni * b->nj * b->nk;
    b->udf_src.assign(b->U[0].size(), 0.0);
    size_t c = 0;
    for (int k = b->kmin; k <= b->kmax; ++k)
        for (int j = b->jmin; j <= b->jmax; ++j)
            for (int i = b->imin; i <= b->imax; ++i, ++c) {
                const size_t ci = cidx(b, i, j, k);
                for (int q = 0; q < b->nq; ++q) b->udf_src[ci * b->nq + q] = Q[q * nc + c];
            }
}

// Axisymmetric geometry of a 2D block whose lower-left vertex is (x0, y0),
// vertex (i, j) at (x0 + i dx, y0 + j dy): xyplane_quad_cell_properties
// (properties.d:177-237, volume = area * centroid y, fvcell.d:314-351) and
// FVInterface.update_2D_geometric_data (fvinterface.d:265-297, area = length *
// mid-point y); face vertex orders sfluidblock.d:538-645.
void ref_block_set_axisymmetric(Block* b, double x0, double y0)
{
    const int ni = b->ni, nj = b->nj;
    auto X = [&](int i) { return x0 + i * b->dx; };
    auto Y = [&](int j) { return y0 + j * b->dy; };
    b->axi = true;
    b->axi_vol.assign((size_t)ni * nj, 0.0);
    b->axi_areaxy.assign((size_t)ni * nj, 0.0);
    b->axi_ai.assign((size_t)(ni + 1) * nj, 0.0);
    b->axi_aj.assign((size_t)ni * (nj + 1), 0.0);
    for (int j = 0; j < nj; ++j)
        for (int i = 0; i < ni; ++i) {
            // p0 = D (i, j), p1 = A (i+1, j), p2 = B (i+1, j+1), p3 = C (i, j+1)
            const double xA = X(i + 1), yA = Y(j), xB = X(i + 1), yB = Y(j + 1);
            const double xC = X(i), yC = Y(j + 1), xD = X(i), yD = Y(j);
            const double area = 0.5 * ((xB + xA) * (yB - yA) + (xC + xB) * (yC - yB) +
                                       (xD + xC) * (yD - yC) + (xA + xD) * (yA - yD));
            const double yc = -1.0 / (area * 6.0) *
                              ((xB - xA) * (yA * yA + yA * yB + yB * yB) + (xC - xB) * (yB * yB + yB * yC + yC * yC) +
                               (xD - xC) * (yC * yC + yC * yD + yD * yD) + (xA - xD) * (yD * yD + yD * yA + yA * yA));
            b->axi_vol[(size_t)j * ni + i] = area * yc;
            b->axi_areaxy[(size_t)j * ni + i] = area;
        }
    auto face_area = [](double xA, double yA, double xB, double yB) {
        const double LAB = sqrt((xB - xA) * (xB - xA) + (yB - yA) * (yB - yA));
        const double Ybar = 0.5 * (yA + yB);
        return LAB * Ybar;
    };
    for (int j = 0; j < nj; ++j)
        for (int i = 0; i <= ni; ++i) b->axi_ai[(size_t)j * (ni + 1) + i] = face_area(X(i), Y(j), X(i), Y(j + 1));
    for (int j = 0; j <= nj; ++j)
        for (int i = 0; i < ni; ++i) b->axi_aj[(size_t)j * ni + i] = face_area(X(i + 1), Y(j), X(i), Y(j));
}

static int nstages(int scheme)
{
    switch (scheme) {
    case UPD_EULER: return 1;
    case UPD_PC: case UPD_MIDPOINT: return 2;
    default: return 3;
    }
}

// fvcell.d:1153-1393
static void stage_update(Block* b, int stage, size_t c, double dt)
{
    const int nq = b->nq;
    const int sch = b->cfg.update_scheme;
    double* d0 = &b->dUdt[0][c * nq];
    double* d1 = &b->dUdt[1][c * nq];
    double* d2 = &b->dUdt[2][c * nq];
    if (stage == 1) {
        double g1 = 1.0;
        if (sch == UPD_MIDPOINT || sch == UPD_CLASSIC_RK3) g1 = 0.5;
        if (sch == UPD_DENMAN_RK3) g1 = 8.0 / 15.0;
        const double* U0 = &b->U[0][c * nq];
        double* U1 = &b->U[1][c * nq];
        for (int q = 0; q < nq; ++q) U1[q] = U0[q] + dt * g1 * d0[q];
    } else if (stage == 2) {
        double g1 = 0.5, g2 = 0.5;
        switch (sch) {
        case UPD_MIDPOINT: g1 = 0.0; g2 = 1.0; break;
        case UPD_CLASSIC_RK3: g1 = -1.0; g2 = 2.0; break;
        case UPD_TVD_RK3: g1 = 0.25; g2 = 0.25; break;
        case UPD_DENMAN_RK3: g1 = -17.0 / 60.0; g2 = 5.0 / 12.0; break;
        default: break;
        }
        const double* Uo = &b->U[(sch == UPD_DENMAN_RK3) ? 1 : 0][c * nq];
        double* U2 = &b->U[2][c * nq];
        for (int q = 0; q < nq; ++q) U2[q] = Uo[q] + dt * (g1 * d0[q] + g2 * d1[q]);
    } else {
        double g1 = 1.0 / 6.0, g2 = 1.0 / 6.0, g3 = 4.0 / 6.0;
        if (sch == UPD_CLASSIC_RK3) { g1 = 1.0 / 6.0; g2 = 4.0 / 6.0; g3 = 1.0 / 6.0; }
        if (sch == UPD_DENMAN_RK3) { g1 = 0.0; g2 = -5.0 / 12.0; g3 = 3.0 / 4.0; }
        const double* Uo = &b->U[(sch == UPD_DENMAN_RK3) ? 2 : 0][c * nq];
        double* U3 = &b->U[3][c * nq];
        for (int q = 0; q < nq; ++q) U3[q] = Uo[q] + dt * (g1 * d0[q] + g2 * d1[q] + g3 * d2[q]);
    }
}

// One gasdynamic_explicit_increment_with_fixed_grid (simcore.d:2303-2952) for
// one block.  Returns 0; -1 for a decode that throws (non-positive density or
// failed thermo); -2 when the invalid-cell count exceeds max_invalid_cells.
// nbad[stage] receives the per-stage invalid-cell count.
// Flux, derivatives, update and decode of one stage of one block whose ghost
// cells are filled.
static void block_stage(Block* b, int stage, double dt, int& decode_fail_out, int& invalid_out)
{
    const int nq = b->nq;
    const int ftl = stage - 1;
    convective_flux(b);
    int decode_fail = 0, invalid = 0;
    {
        #pragma omp parallel for collapse(2) schedule(static) reduction(+:decode_fail, invalid) if (b->ni * b->nj * b->nk >= OMP_MIN_CELLS)
        for (int k = b->kmin; k <= b->kmax; ++k)
            for (int j = b->jmin; j <= b->jmax; ++j)
                for (int i = b->imin; i <= b->imax; ++i) {
                    size_t c = cidx(b, i, j, k);
                    time_derivatives(b, c, &b->dUdt[ftl][c * nq]);
                    stage_update(b, stage, c, dt);
                    if (decode_conserved(b, b->fs[c], &b->U[ftl + 1][c * nq]) != 0) { decode_fail++; continue; }
                    if (!check_data(b, b->fs[c])) invalid++;
                }
    }
    decode_fail_out = decode_fail;
    invalid_out = invalid;
}

int ref_multi_step(Block** blks, int nblk, double dt, int* nbad);

int ref_block_step(Block* b, double dt, int* nbad)
{
    return ref_multi_step(&b, 1, dt, nbad);
}

// gasdynamic_explicit_increment_with_fixed_grid over several blocks: per stage,
// every block's ghost cells are filled first (exchange_ghost_cell_boundary_data,
// simcore.d:1599-1668, then the pre-reconstruction BCs) and then every block is
// updated.  nbad[stage] receives the invalid-cell count summed over blocks;
// the invalid-cell limit applies per block.
int ref_multi_step(Block** blks, int nblk, double dt, int* nbad)
{
    const int ns = nstages(blks[0]->cfg.update_scheme);
    // the adaptive flux calculators switch the shock detector on (globalconfig.d:1942-1954)
    const bool detect = blks[0]->cfg.flux_calc == FLUX_ADAPTIVE_HANEL_AUSMDV ||
                        blks[0]->cfg.flux_calc == FLUX_ADAPTIVE_EFM_AUSMDV;
    for (int stage = 1; stage <= ns; ++stage) {
        for (int n = 0; n < nblk; ++n) apply_bcs(blks[n]);
        // simcore.d:2352-2358: after the ghost-cell fill, before the fluxes
        if (detect)
            for (int n = 0; n < nblk; ++n) detect_shock_points(blks[n]);
        int total = 0, any_fail = 0, too_many = 0;
        for (int n = 0; n < nblk; ++n) {
            int decode_fail = 0, invalid = 0;
            block_stage(blks[n], stage, dt, decode_fail, invalid);
            total += invalid + decode_fail;
            if (decode_fail) any_fail = 1;
            if (invalid > blks[n]->cfg.max_invalid_cells) too_many = 1;
        }
        if (nbad) nbad[stage - 1] = total;
        if (any_fail) return -1;
        if (too_many) return -2;
    }
    for (int n = 0; n < nblk; ++n) std::swap(blks[n]->U[0], blks[n]->U[ns]);   // simcore.d:2936-2941
    return 0;
}

// fvcell.d:1624-1705 (non-stringent CFL) and fluidblock.d:776-862 without
// super-time-stepping or local time-stepping.  out = {dt_allow, cfl_max, signal_max}.
void ref_block_dt(Block* b, double dt_current, double cfl_value, int check_cfl, double* out)
{
    const bool three_d = (b->cfg.dims == 3);
    double cfl_allow = 0.9;
    switch (nstages(b->cfg.update_scheme)) { case 1: cfl_allow = 0.9; break; case 2: cfl_allow = 1.2; break; default: cfl_allow = 1.6; }
    bool first = true;
    double dt_allow = 0, cfl_min = 0, cfl_max = 0, signal = 0, smax = 0;
    const size_t sj = b->nidim, sk = (size_t)b->nidim * b->njdim;
    for (int k = b->kmin; k <= b->kmax; ++k)
        for (int j = b->jmin; j <= b->jmax; ++j)
            for (int i = b->imin; i <= b->imax; ++i) {
                const size_t c = cidx(b, i, j, k);
                const FlowState& fs = b->fs[c];
                const double* v = fs.vel;
                const double iLength = cell_len(b, 0, c), jLength = cell_len(b, 1, c), kLength = cell_len(b, 2, c);
                const Frame& fN = face_frame(b, 1, c + sj);
                const Frame& fE = face_frame(b, 0, c + 1);
                const Frame& fT = three_d ? face_frame(b, 2, c + sk) : b->frame[2];
                double un_N = fabs(v[0] * fN.n[0] + v[1] * fN.n[1] + v[2] * fN.n[2]);
                double un_E = fabs(v[0] * fE.n[0] + v[1] * fE.n[1] + v[2] * fE.n[2]);
                double un_T = three_d ? fabs(v[0] * fT.n[0] + v[1] * fT.n[1] + v[2] * fT.n[2]) : 0.0;
                signal = 0;
                if (b->stringent_cfl) {
                    // fvcell.d:1642-1650: the fastest normal speed over the shortest length
                    double un_max = fmax(un_N, un_E);
                    double minLength = fmin(iLength, jLength);
                    if (three_d) {
                        un_max = fmax(un_max, un_T);
                        minLength = fmin(minLength, kLength);
                    }
                    signal = (un_max + fs.gas.a) / minLength;
                } else {
                    double signalN = (un_N + fs.gas.a) / jLength;
                    signal = fmax(signal, signalN);
                    double signalE = (un_E + fs.gas.a) / iLength;
                    signal = fmax(signal, signalE);
                    if (three_d) {
                        double signalT = (un_T + fs.gas.a) / kLength;
                        signal = fmax(signal, signalT);
                    }
                }
                smax = fmax(smax, signal);
                double cfl_local = dt_current * signal;
                double dt_local = cfl_value / signal;
                if (first) { cfl_min = cfl_local; cfl_max = cfl_local; dt_allow = dt_local; first = false; }
                else { cfl_min = fmin(cfl_min, cfl_local); cfl_max = fmax(cfl_max, cfl_local); dt_allow = fmin(dt_allow, dt_local); }
            }
    if (check_cfl && (cfl_max < 0.0 || cfl_max > cfl_allow)) {
        cfl_max = 0.5 * cfl_allow;
        dt_allow = cfl_max / signal;   // the last cell's signal, as in the reference
    }
    (void)cfl_min;
    out[0] = dt_allow; out[1] = cfl_max; out[2] = smax;
}

int ref_block_nq(const Block* b) { return b->nq; }

void ref_block_set_stringent_cfl(Block* b, int on) { b->stringent_cfl = (on != 0); }

} // extern "C"
