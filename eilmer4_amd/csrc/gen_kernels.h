// gen_kernels.h -- the stage of a block built from its grid (general
// structured geometry): per-face areas and n, t1, t2 frames, per-cell volumes
// and i/j/k lengths, as compute_primary_cell_geometric_data gives them
// (eilmer/sfluidblock.d:650-842).  One thread per face evaluates the flux the
// reference's convective_flux_phase0 general branch evaluates
// (sfluidblock.d:2189-2345):
//   the four stencil cells' velocities into the face frame (vector3.d:425-450),
//   slip-wall ghosts reflected about their boundary face first
//   (internal_copy_then_reflect.d:94-310, ghost_cell.d:32-43);
//   l2r2 with the weights of the four cells' lengths (onedinterp.d:118-162,
//   435-585; ghost lengths mirrored, or across a connection the mapped cell's);
//   the face EOS and the flux calculator in the face frame, the momentum flux
//   turned back to the global frame (fluxcalc.d:26-128);
//   a simple-outflow-flux boundary face from its interior cell
//   (boundary_flux_effect.d:481-546).
// Face fluxes go to HBM ([5+nsp][faces] per direction); the update kernel then
// sums them with the face areas and divides by the cell volume (fvcell.d:735-875).
//
// NSP = 0 is an ideal gas, whose flow state array holds the P_* fields
// without mass fractions; NSP > 0 a thermally perfect mixture (the
// reconstruction of mass fractions and the temperature solve of the mixture
// kernels).  In fp64_exact every expression keeps the reference's operation
// order, including the face states' round trip through the global frame
// between the reconstruction and the flux (onedinterp.d:584 rotates them back,
// fluxcalc.d:34-40 forward again), so the result is the CPU oracle's bit for
// bit.  The fast modes skip that round trip (a rounding-level difference).
#pragma once
#include "tp_kernels.h"

namespace e4 {

struct GFrame { double n[3], t1[3], t2[3]; };

// area (row 0) and frame (rows 1-9) of face f of direction d
__device__ __forceinline__ GFrame gen_frame(const TPArgs& A, int d, long f)
{
    const double* g = A.fgeo[d] + f;
    const long s = A.fqs[d];
    GFrame fr;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        fr.n[q] = g[(1 + q) * s];
        fr.t1[q] = g[(4 + q) * s];
        fr.t2[q] = g[(7 + q) * s];
    }
    return fr;
}

// vector3.d:425-450
__device__ __forceinline__ void gen_to_local(double v[3], const GFrame& f)
{
    double vx = v[0] * f.n[0] + v[1] * f.n[1] + v[2] * f.n[2];
    double vy = v[0] * f.t1[0] + v[1] * f.t1[1] + v[2] * f.t1[2];
    double vz = v[0] * f.t2[0] + v[1] * f.t2[1] + v[2] * f.t2[2];
    v[0] = vx; v[1] = vy; v[2] = vz;
}
__device__ __forceinline__ void gen_to_global(double v[3], const GFrame& f)
{
    double vx = v[0] * f.n[0] + v[1] * f.t1[0] + v[2] * f.t2[0];
    double vy = v[0] * f.n[1] + v[1] * f.t1[1] + v[2] * f.t2[1];
    double vz = v[0] * f.n[2] + v[1] * f.t1[2] + v[2] * f.t2[2];
    v[0] = vx; v[1] = vy; v[2] = vz;
}

// Index of face (i, j, k) of direction d (faces i fastest over
// (ni + [d==0]) x (nj + [d==1]) x nk').
__device__ __forceinline__ long gen_face_index(const TPArgs& A, int d, int i, int j, int k)
{
    const int nfi = A.ni + (d == 0), nfj = A.nj + (d == 1);
    return ((long)k * nfj + j) * nfi + i;
}

// Length in direction d of the stencil position (i, j, k), which may be a
// ghost cell one or two layers outside the block: the ghost cells of a face
// copy the lengths of the interior cells they mirror (sfluidblock.d:673-713),
// those of a periodic face the wrapped cells', those of a connected face the
// mapped cell's own i/j/k lengths (exchange_geometry_phase2,
// full_face_copy.d:1042-1085).
__device__ __forceinline__ double gen_len(const TPArgs& A, int d, int i, int j, int k)
{
    if (i >= 0 && i < A.ni && j >= 0 && j < A.nj && k >= 0 && k < A.nk)
        return A.geo[(1 + d) * A.qs + ((long)k * A.nj + j) * A.ni + i];
    int dir, idx, n;
    if (i < 0 || i >= A.ni) { dir = 0; idx = i; n = A.ni; }
    else if (j < 0 || j >= A.nj) { dir = 1; idx = j; n = A.nj; }
    else { dir = 2; idx = k; n = A.nk; }
    const bool plus = idx >= n;
    const int layer = plus ? idx - n + 1 : -idx;
    const int face = (dir == 0) ? (plus ? FACE_EAST : FACE_WEST)
                   : (dir == 1) ? (plus ? FACE_NORTH : FACE_SOUTH)
                                : (plus ? FACE_TOP : FACE_BOTTOM);
    const int type = A.bc->type[face];
    if (type == BC_CONNECTED) {
        const Connection nb = A.bc->nbr[face];
        const int fa = (face == FACE_NORTH || face == FACE_SOUTH || face >= FACE_TOP) ? i : j;
        const int fb = (face >= FACE_TOP) ? j : k;
        int sc[3];
        connection_source(nb.dims, face, nb.face, nb.orient, nb.n, layer, fa, fb, sc);
        const long c = ((long)sc[2] * nb.n[1] + sc[1]) * nb.n[0] + sc[0];
        return A.nbrGeo[face][(1 + d) * nb.qs + c];
    }
    const int src = (type == BC_PERIODIC) ? (plus ? layer - 1 : n - layer) : (plus ? n - layer : layer - 1);
    int si = i, sj = j, sk = k;
    if (dir == 0) si = src; else if (dir == 1) sj = src; else sk = src;
    return A.geo[(1 + d) * A.qs + ((long)sk * A.nj + sj) * A.ni + si];
}

// Global velocity of a stencil position reached along direction d from the
// block; a slip-wall ghost (r.flip >= 0) is its mirrored cell's velocity
// reflected about the boundary face at the ghost's in-face position.
__device__ __forceinline__ void gen_vel(const TPArgs& A, const TPRef& r, int d, int i, int j, int k, double v[3])
{
    v[0] = r.base[P_VX * r.stride];
    v[1] = r.base[P_VY * r.stride];
    v[2] = r.base[P_VZ * r.stride];
    if (r.flip >= 0) {
        const int idx = (d == 0) ? i : (d == 1) ? j : k;
        const int n = (d == 0) ? A.ni : (d == 1) ? A.nj : A.nk;
        const int fpos = (idx < 0) ? 0 : n;
        const long f = gen_face_index(A, d, (d == 0) ? fpos : i, (d == 1) ? fpos : j, (d == 2) ? fpos : k);
        const GFrame fw = gen_frame(A, d, f);
        gen_to_local(v, fw);
        v[0] = -(v[0]);
        gen_to_global(v, fw);
    }
}

// l2r2_prepare (onedinterp.d:118-133) for one face
template <class R>
__device__ __forceinline__ L2R2WT<R> gen_l2r2_weights(double lenL1, double lenL0, double lenR0, double lenR1)
{
    L2R2WT<R> w;
    w.lenL0 = R(lenL0);
    w.lenR0 = R(lenR0);
    w.aL0 = R(0.5 * lenL0 / (lenL1 + 2.0 * lenL0 + lenR0));
    w.aR0 = R(0.5 * lenR0 / (lenL0 + 2.0 * lenR0 + lenR1));
    w.two_over_lenL0_plus_lenL1 = R(2.0 / (lenL0 + lenL1));
    w.two_over_lenR0_plus_lenL0 = R(2.0 / (lenR0 + lenL0));
    w.two_over_lenR1_plus_lenR0 = R(2.0 / (lenR1 + lenR0));
    w.two_lenL0_plus_lenL1 = R(2.0 * lenL0 + lenL1);
    w.two_lenR0_plus_lenR1 = R(2.0 * lenR0 + lenR1);
    return w;
}

// BFE_SimpleOutflowFlux.compute_outflow_flux (boundary_flux_effect.d:481-546) on
// a fixed grid, in the global frame: the interior cell's own flux when the
// flow leaves (outsign * mass flux > 0), else the face closed with its pressure.
template <int NSP>
__device__ __forceinline__ void gen_outflow_flux(const TPRef& r, int outsign, const GFrame& fr, double* F)
{
    const double rho = r.base[P_RHO * r.stride], p = r.base[P_P * r.stride], u = r.base[P_U * r.stride];
    const double v[3] = {r.base[P_VX * r.stride], r.base[P_VY * r.stride], r.base[P_VZ * r.stride]};
    const double vn = v[0] * fr.n[0] + v[1] * fr.n[1] + v[2] * fr.n[2];
    const double mass_flux = rho * vn;
#pragma unroll
    for (int q = 0; q < 5 + NSP; ++q) F[q] = 0.0;
    if ((outsign * mass_flux) > 0.0) {
        F[0] = mass_flux;
#pragma unroll
        for (int d = 0; d < 3; ++d) F[1 + d] = p * fr.n[d] + v[d] * mass_flux;
        const double utot = u + 0.5 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        F[4] = mass_flux * utot + p * vn;
#pragma unroll
        for (int s = 0; s < NSP; ++s) F[5 + s] = mass_flux * r.base[(P_Y + s) * r.stride];
    } else {
#pragma unroll
        for (int d = 0; d < 3; ++d) F[1 + d] = fr.n[d] * p;
        F[4] = p * 0.0;
    }
}

// One face state: density, velocity in the face frame, internal energy,
// pressure, temperature, sound speed and mass fractions.
template <int NSP>
struct GSide {
    double rho, p, T, u, a;
    double v[3];
    double Y[NSP > 0 ? NSP : 1];
};

template <int NSP>
__device__ __forceinline__ void gen_cell_side(const TPRef& r, const double vloc[3], GSide<NSP>& s)
{
    s.rho = r.base[P_RHO * r.stride]; s.p = r.base[P_P * r.stride]; s.T = r.base[P_T * r.stride];
    s.u = r.base[P_U * r.stride]; s.a = r.base[P_A * r.stride];
    s.v[0] = vloc[0]; s.v[1] = vloc[1]; s.v[2] = vloc[2];
#pragma unroll
    for (int q = 0; q < NSP; ++q) s.Y[q] = r.base[(P_Y + q) * r.stride];
}

// Flux of the face calculator on two face-frame states, F[5+NSP] in the face
// frame (mass, normal and two tangential momenta, energy, species).
template <class M, int NSP, int FLUX>
__device__ __forceinline__ void gen_flux(const TPArgs& A, const GSide<NSP>& L, const GSide<NSP>& Rt, bool shock,
                                         double* F)
{
    if constexpr (NSP == 0) {
        using R = typename M::R;
        FaceStateT<R> Lf, Rf;
        Lf.rho = R(L.rho); Lf.p = R(L.p); Lf.u = R(L.u); Lf.a = R(L.a);
        Lf.vn = R(L.v[0]); Lf.vt1 = R(L.v[1]); Lf.vt2 = R(L.v[2]);
        Rf.rho = R(Rt.rho); Rf.p = R(Rt.p); Rf.u = R(Rt.u); Rf.a = R(Rt.a);
        Rf.vn = R(Rt.v[0]); Rf.vt1 = R(Rt.v[1]); Rf.vt2 = R(Rt.v[2]);
        Lf.T = R(L.T); Rf.T = R(Rt.T);
        if constexpr (M::kFast) {
            Lf.RT = R(A.ig.Rgas) * R(L.T);
            Rf.RT = R(A.ig.Rgas) * R(Rt.T);
        } else {
            Lf.RT = Rf.RT = R(0);
        }
        const FluxOpt<R> opt{R(A.shear_tol), A.always_hanel != 0, A.shock_efm != 0, A.other_roe != 0, R(A.M_inf)};
        Flux5T<R> Fl = flux_calc<M, FLUX>(Lf, Rf, cast_gas<R>(A.ig), shock, opt);
        F[0] = Fl.mass; F[1] = Fl.mn; F[2] = Fl.mt1; F[3] = Fl.mt2; F[4] = Fl.E;
    } else {
        using MA = typename TPArith<M>::type;
        TPCell<NSP> Lc, Rc;
        Lc.rho = L.rho; Lc.p = L.p; Lc.T = L.T; Lc.u = L.u; Lc.a = L.a;
        Rc.rho = Rt.rho; Rc.p = Rt.p; Rc.T = Rt.T; Rc.u = Rt.u; Rc.a = Rt.a;
#pragma unroll
        for (int q = 0; q < 3; ++q) { Lc.v[q] = L.v[q]; Rc.v[q] = Rt.v[q]; }
#pragma unroll
        for (int q = 0; q < NSP; ++q) { Lc.Y[q] = L.Y[q]; Rc.Y[q] = Rt.Y[q]; }
        tp_flux<MA, NSP, FLUX>(A, *A.gas, Lc, Rc, shock, F);
    }
}

// One face of direction DIR per thread.
template <class M, int NSP, int FLUX, int DIR>
__global__ __launch_bounds__(128) void gen_face_kernel(TPArgs A)
{
    using MR = typename std::conditional<NSP == 0, M, typename TPArith<M>::type>::type;
    using R = typename MR::R;
    constexpr bool kPolish = (NSP > 0) && TPArith<M>::polish;
    constexpr bool kExact = !M::kFast;
    constexpr int NQ = 5 + NSP;
    if (stage_aborted(A.abort_prev)) return;
    const int nfi = A.ni + (DIR == 0), nfj = A.nj + (DIR == 1), nfk = (A.dims == 3) ? A.nk + (DIR == 2) : 1;
    const long nf = (long)nfi * nfj * nfk;
    const long f0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = f0 < nf;
    const long f = active ? f0 : nf - 1;   // idle lanes compute the last face and store nothing
    const int i = (int)(f % nfi);
    const int j = (int)((f / nfi) % nfj);
    const int k = (int)(f / ((long)nfi * nfj));
    const int di = (DIR == 0), dj = (DIR == 1), dk = (DIR == 2);
    const int pos = (DIR == 0) ? i : (DIR == 1) ? j : k;
    const int nD = (DIR == 0) ? A.ni : (DIR == 1) ? A.nj : A.nk;
    const GFrame fr = gen_frame(A, DIR, f);
    double F[NQ];
    int fe = 0, fb = 0;
    const int face_m = (DIR == 0) ? FACE_WEST : (DIR == 1) ? FACE_SOUTH : FACE_BOTTOM;
    const int face_p = (DIR == 0) ? FACE_EAST : (DIR == 1) ? FACE_NORTH : FACE_TOP;
    const bool out_m = pos == 0 && A.bc->type[face_m] == BC_OUTFLOW_SIMPLE_FLUX;
    const bool out_p = pos == nD && A.bc->type[face_p] == BC_OUTFLOW_SIMPLE_FLUX;
    if (out_m || out_p) {
        // the boundary computes this face's flux from the cell inside it
        const TPRef rc = out_m ? tp_ref(A, i, j, k) : tp_ref(A, i - di, j - dj, k - dk);
        gen_outflow_flux<NSP>(rc, out_m ? -1 : 1, fr, F);
    } else {
        TPRef r[4];
        double v[4][3];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int si = i + (m - 2) * di, sj = j + (m - 2) * dj, sk = k + (m - 2) * dk;
            r[m] = tp_ref(A, si, sj, sk);
            gen_vel(A, r[m], DIR, si, sj, sk, v[m]);
            gen_to_local(v[m], fr);
        }
        GSide<NSP> Lf, Rf;
        bool okL = false, okR = false;
        if (A.high_order) {
            double len[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) len[m] = gen_len(A, DIR, i + (m - 2) * di, j + (m - 2) * dj, k + (m - 2) * dk);
            const L2R2WT<R> w = gen_l2r2_weights<R>(len[0], len[1], len[2], len[3]);
            const bool lim = A.apply_limiter, clip = A.extrema_clipping;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                R qL, qR;
                interp_l2r2_scalar<MR>(w, lim, clip, R(v[0][d]), R(v[1][d]), R(v[2][d]), R(v[3][d]), qL, qR);
                Lf.v[d] = qL; Rf.v[d] = qR;
            }
            if constexpr (NSP > 0) {
                // (R is double for mixtures)
#pragma unroll
                for (int s = 0; s < NSP; ++s)
                    interp_l2r2_scalar<MR>(w, lim, clip, tp_val(r[0], P_Y + s), tp_val(r[1], P_Y + s),
                                           tp_val(r[2], P_Y + s), tp_val(r[3], P_Y + s), Lf.Y[s], Rf.Y[s]);
                if (scale_mass_fractions<MR, NSP>(Lf.Y) != 0) {
#pragma unroll
                    for (int s = 0; s < NSP; ++s) Lf.Y[s] = tp_val(r[1], P_Y + s);
                }
                if (scale_mass_fractions<MR, NSP>(Rf.Y) != 0) {
#pragma unroll
                    for (int s = 0; s < NSP; ++s) Rf.Y[s] = tp_val(r[2], P_Y + s);
                }
            }
            {
                R qL, qR;
                interp_l2r2_scalar<MR>(w, lim, clip, R(tp_val(r[0], P_RHO)), R(tp_val(r[1], P_RHO)),
                                       R(tp_val(r[2], P_RHO)), R(tp_val(r[3], P_RHO)), qL, qR);
                Lf.rho = qL; Rf.rho = qR;
                interp_l2r2_scalar<MR>(w, lim, clip, R(tp_val(r[0], P_U)), R(tp_val(r[1], P_U)), R(tp_val(r[2], P_U)),
                                       R(tp_val(r[3], P_U)), qL, qR);
                Lf.u = qL; Rf.u = qR;
            }
            // update_thermo_from_rhou on each side; on failure the side is its
            // cell's state (onedinterp.d:1039-1068); the sound speed is the
            // cell's (sfluidblock.d:2199-2208)
            if constexpr (NSP == 0) {
                if (Lf.u > 0.0 && Lf.rho > 0.0) {
                    Lf.T = Lf.u * A.ig.Cvinv; Lf.p = Lf.rho * A.ig.Rgas * Lf.T; Lf.a = tp_val(r[1], P_A); okL = true;
                }
                if (Rf.u > 0.0 && Rf.rho > 0.0) {
                    Rf.T = Rf.u * A.ig.Cvinv; Rf.p = Rf.rho * A.ig.Rgas * Rf.T; Rf.a = tp_val(r[2], P_A); okR = true;
                }
            } else {
                using MA = typename TPArith<M>::type;
#pragma nounroll
                for (int sd = 0; sd < 2; ++sd) {
                    const TPRef rc = sd ? r[2] : r[1];
                    TPState<NSP> Q;
                    Q.rho = sd ? Rf.rho : Lf.rho;
                    Q.u = sd ? Rf.u : Lf.u;
                    Q.T = tp_val(rc, P_T);
#pragma unroll
                    for (int s = 0; s < NSP; ++s) Q.Y[s] = sd ? Rf.Y[s] : Lf.Y[s];
                    if (tp_face_eos<MA, NSP, kPolish>(*A.gas, Q, &fe, &fb) == 0) {
                        const double a = tp_val(rc, P_A);
                        if (sd == 0) { Lf.T = Q.T; Lf.u = Q.u; Lf.p = Q.p; Lf.a = a; okL = true; }
                        else { Rf.T = Q.T; Rf.u = Q.u; Rf.p = Q.p; Rf.a = a; okR = true; }
                    }
                }
            }
            if (!okL) gen_cell_side<NSP>(r[1], v[1], Lf);
            if (!okR) gen_cell_side<NSP>(r[2], v[2], Rf);
            if constexpr (kExact) {
                // the reconstruction hands its states over in the global frame
                // and the flux calculator turns them back into the face frame
                gen_to_global(Lf.v, fr); gen_to_global(Rf.v, fr);
                gen_to_local(Lf.v, fr); gen_to_local(Rf.v, fr);
            }
        } else {
            gen_cell_side<NSP>(r[1], v[1], Lf);
            gen_cell_side<NSP>(r[2], v[2], Rf);
        }
        const bool shock = (FLUX == FLUX_SHOCK) && ((r[1].s && *r[1].s) || (r[2].s && *r[2].s));
        gen_flux<M, NSP, FLUX>(A, Lf, Rf, shock, F);
        double m[3] = {F[1], F[2], F[3]};
        gen_to_global(m, fr);
        F[1] = m[0]; F[2] = m[1]; F[3] = m[2];
    }
    if (active) {
        // mixture blocks keep the flux in the element type the mixture update
        // reads (FP32 in the mixed form)
        using FS = typename std::conditional<(NSP > 0), TPFluxStore<M>, double>::type;
        FS* dst = reinterpret_cast<FS*>(A.F[DIR]);
        const long fq = A.fqs[DIR];
#pragma unroll
        for (int q = 0; q < NQ; ++q) dst[q * fq + f] = FS(F[q]);
    }
    if constexpr (NSP > 0) {
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
}

// Ideal-gas decode_conserved into the flow-state array (fvcell.d:563-732,
// ideal_gas.d:98-144), with ignore_low_T_thermo_update_failure: u <= 0 gives
// T = 200 K, u = Cv T and the re-encoded U (fvcell.d:685-697).  Returns false
// when rho <= 0 (the decode would throw).  FP64 in every mode.
template <class M>
__device__ __forceinline__ bool ig_decode_cell(const IdealGas& g, double Un[5], double& rho, double v[3], double& u,
                                               double& T, double& p, double& a)
{
    using MD = typename std::conditional<M::kFast, ModeFast64, ModeExact>::type;
    rho = Un[0];
    if (!(rho > 0.0)) return false;
    const double dinv = dv<MD>(1.0, rho);
    v[0] = Un[1] * dinv; v[1] = Un[2] * dinv; v[2] = Un[3] * dinv;
    u = Un[4] * dinv;
    const double ke = 0.5 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    u -= ke;
    if (u <= 0.0) {
        T = 200.0;
        u = g.Cv * T;
        p = rho * g.Rgas * T;
        Un[1] = rho * v[0]; Un[2] = rho * v[1]; Un[3] = rho * v[2];
        Un[4] = rho * (u + 0.5 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
    } else {
        T = u * g.Cvinv;
        p = rho * g.Rgas * T;
    }
    a = sqrt(g.gammaR * T);
    return true;
}

// check_data (flowstate.d:345-388, gas_state.d:199-241) of an ideal-gas state
__device__ __forceinline__ bool ig_invalid(double rho, const double v[3], double T, double p, double a)
{
    bool invalid = false;
    if (!isfinite(rho) || rho < 0.0) invalid = true;
    if (!isfinite(T) || T < 1.01) invalid = true;
    if (!isfinite(p) || !isfinite(a)) invalid = true;
    if (fabs(v[0]) > 30000.0 || fabs(v[1]) > 30000.0 || fabs(v[2]) > 30000.0) invalid = true;
    if (T > 50000.0) invalid = true;
    return invalid;
}

__device__ __forceinline__ void ig_store(const TPArgs& A, long c, double rho, const double v[3], double u, double T,
                                         double p, double a)
{
    double* P = A.Pout;
    const long qs = A.qs;
    P[P_RHO * qs + c] = rho; P[P_VX * qs + c] = v[0]; P[P_VY * qs + c] = v[1]; P[P_VZ * qs + c] = v[2];
    P[P_U * qs + c] = u; P[P_T * qs + c] = T; P[P_P * qs + c] = p; P[P_A * qs + c] = a;
}

// The flux integral of cell (i, j, k) over its faces in the reference's order
// N, E, S, W, T, B with outsign * area (fvcell.d:789-825), and its volume.
template <int NQ>
__device__ __forceinline__ void gen_integral(const TPArgs& A, long c, int i, int j, int k, double integ[NQ])
{
    const long iW = ((long)k * A.nj + j) * (A.ni + 1) + i;
    const long jS = ((long)k * (A.nj + 1) + j) * A.ni + i;
    const long kB = c;
#pragma unroll
    for (int q = 0; q < NQ; ++q) integ[q] = 0.0;
    auto add_face = [&](int d, long fidx, double outsign) {
        const double area = outsign * A.fgeo[d][fidx];
        const double* Fs = A.F[d];
        const long fq = A.fqs[d];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            if (q == 3 && A.dims != 3) continue;
            integ[q] -= Fs[q * fq + fidx] * area;
        }
    };
    add_face(1, jS + A.ni, 1.0);
    add_face(0, iW + 1, 1.0);
    add_face(1, jS, -1.0);
    add_face(0, iW, -1.0);
    if (A.dims == 3) {
        add_face(2, kB + (long)A.ni * A.nj, 1.0);
        add_face(2, kB, -1.0);
    }
}

// Ideal-gas update of a block built from its grid: time derivative, stage
// update, decode into the flow-state array, check_data, write of U.
template <class M>
__global__ __launch_bounds__(128) void ig_update_kernel(TPArgs A)
{
    if (stage_aborted(A.abort_prev)) return;
    const long ncells = (long)A.ni * A.nj * A.nk;
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = c < ncells;
    bool invalid = false, decode_fail = false;
    if (active) {
        const long qs = A.qs;
        const int i = (int)(c % A.ni), j = (int)((c / A.ni) % A.nj), k = (int)(c / ((long)A.ni * A.nj));
        double integ[5];
        gen_integral<5>(A, c, i, j, k, integ);
        const double vol_inv = 1.0 / A.geo[c];
        double Un[5];
        // every value the update reads, loaded before the first store
        double uold[5], dd0[5], dd1[5], srcq[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            uold[q] = A.Uold[c + q * qs];
            dd0[q] = (A.stage >= 2) ? A.d0[c + q * qs] : 0.0;
            dd1[q] = (A.stage >= 3) ? A.d1[c + q * qs] : 0.0;
            srcq[q] = A.src ? A.src[c + q * qs] : 0.0;   // user-defined source terms
        }
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            double Qs = 0.0;
            if (A.src) Qs += srcq[q];
            const double d = vol_inv * integ[q] + Qs;
            if (A.dout) A.dout[c + q * qs] = d;
            if (A.stage == 1) Un[q] = uold[q] + A.c1 * d;
            else if (A.stage == 2) Un[q] = uold[q] + A.dt * (A.g1 * dd0[q] + A.g2 * d);
            else Un[q] = uold[q] + A.dt * (A.g1 * dd0[q] + A.g2 * dd1[q] + A.g3 * d);
        }
        double rho, v[3], u, T, p, a;
        if (!ig_decode_cell<M>(A.ig, Un, rho, v, u, T, p, a)) {
            decode_fail = true;
        } else {
            invalid = ig_invalid(rho, v, T, p, a);
            ig_store(A, c, rho, v, u, T, p, a);
        }
#pragma unroll
        for (int q = 0; q < 5; ++q) A.Unew[c + q * qs] = Un[q];
    }
    const unsigned int nbad = (unsigned int)__popcll(__ballot(invalid));
    const unsigned int nfail = (unsigned int)__popcll(__ballot(decode_fail));
    if (__lane_id() == 0) count_bad_cells(A.counters, A.abort_this, A.max_invalid, nbad, nfail);
}

// Decode of an uploaded U into the flow-state array of an ideal-gas block
// built from its grid (counters[0]: cells that do not decode to a valid state).
template <class M>
__global__ __launch_bounds__(128) void ig_decode_kernel(TPArgs A, double* __restrict__ U)
{
    const long ncells = (long)A.ni * A.nj * A.nk;
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    if (c < ncells) {
        double Un[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) Un[q] = U[c + q * A.qs];
        double rho, v[3], u, T, p, a;
        if (!ig_decode_cell<M>(A.ig, Un, rho, v, u, T, p, a)) {
            bad = true;
        } else {
            bad = ig_invalid(rho, v, T, p, a);
            ig_store(A, c, rho, v, u, T, p, a);
#pragma unroll
            for (int q = 1; q < 5; ++q) U[c + q * A.qs] = Un[q];
        }
    }
    const unsigned int nbad = (unsigned int)__popcll(__ballot(bad));
    if (__lane_id() == 0 && nbad) atomicAdd(&A.counters[0], nbad);
}

// Shock detector of a block built from its grid (fluidblock.d:415-431,
// shockdetectors.d): the normal velocities along each face's own normal.
template <class M>
__global__ __launch_bounds__(128) void gen_shock_detect_kernel(TPArgs A)
{
    using MD = typename std::conditional<M::kFast, ModeFast64, ModeExact>::type;
    if (stage_aborted(A.abort_prev)) return;
    const long n = (long)A.ni * A.nj * A.nk;
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int i = (int)(c % A.ni), j = (int)((c / A.ni) % A.nj), k = (int)(c / ((long)A.ni * A.nj));
    const TPRef own = tp_ref(A, i, j, k);
    const double tol = A.compression_tol;
    bool S = false;
    const int ndir = (A.dims == 3) ? 3 : 2;
    for (int d = 0; d < ndir; ++d) {
        const int di = (d == 0), dj = (d == 1), dk = (d == 2);
        double vo[3];
        gen_vel(A, own, d, i, j, k, vo);
        const double ao = tp_val(own, P_A);
        for (int side = 0; side < 2; ++side) {
            const int s = side ? 1 : -1;
            const int ni_ = i + s * di, nj_ = j + s * dj, nk_ = k + s * dk;
            const TPRef nb = tp_ref(A, ni_, nj_, nk_);
            double vn[3];
            gen_vel(A, nb, d, ni_, nj_, nk_, vn);
            const long fidx = gen_face_index(A, d, i + (side ? di : 0), j + (side ? dj : 0), k + (side ? dk : 0));
            const GFrame fr = gen_frame(A, d, fidx);
            const double* vL = side ? vo : vn;
            const double* vR = side ? vn : vo;
            const double uL = vL[0] * fr.n[0] + vL[1] * fr.n[1] + vL[2] * fr.n[2];
            const double uR = vR[0] * fr.n[0] + vR[1] * fr.n[1] + vR[2] * fr.n[2];
            const double aL = side ? ao : tp_val(nb, P_A), aR = side ? tp_val(nb, P_A) : ao;
            const double a_min = (aL < aR) ? aL : aR;
            S = S || dv<MD>(uR - uL, a_min) < tol;
        }
    }
    A.Snew[c] = S ? 1 : 0;
}

// determine_time_step_size's signal of a block built from its grid
// (fvcell.d:1624-1705): normal velocities along the north, east and top faces'
// normals over the cell's j, i and k lengths; block maximum by atomic max on
// the bit pattern, the last cell's signal kept.
static __global__ __launch_bounds__(256) void gen_cfl_kernel(TPArgs A, int stringent, unsigned long long* smax_bits,
                                                      double* last_signal)
{
    __shared__ double wave_max[4];
    const long ncells = (long)A.ni * A.nj * A.nk;
    const long qs = A.qs;
    double sig = 0.0;
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += (long)gridDim.x * blockDim.x) {
        const int i = (int)(c % A.ni), j = (int)((c / A.ni) % A.nj), k = (int)(c / ((long)A.ni * A.nj));
        const double* P = A.Pin;
        const double v[3] = {P[P_VX * qs + c], P[P_VY * qs + c], P[P_VZ * qs + c]};
        const double a = P[P_A * qs + c];
        const GFrame fN = gen_frame(A, 1, gen_face_index(A, 1, i, j + 1, k));
        const GFrame fE = gen_frame(A, 0, gen_face_index(A, 0, i + 1, j, k));
        const double iLength = A.geo[qs + c], jLength = A.geo[2 * qs + c], kLength = A.geo[3 * qs + c];
        const double un_N = fabs(v[0] * fN.n[0] + v[1] * fN.n[1] + v[2] * fN.n[2]);
        const double un_E = fabs(v[0] * fE.n[0] + v[1] * fE.n[1] + v[2] * fE.n[2]);
        double un_T = 0.0;
        if (A.dims == 3) {
            const GFrame fT = gen_frame(A, 2, gen_face_index(A, 2, i, j, k + 1));
            un_T = fabs(v[0] * fT.n[0] + v[1] * fT.n[1] + v[2] * fT.n[2]);
        }
        double s = 0.0;
        if (stringent) {
            double un_max = fmax(un_N, un_E);
            double minLength = fmin(iLength, jLength);
            if (A.dims == 3) {
                un_max = fmax(un_max, un_T);
                minLength = fmin(minLength, kLength);
            }
            s = (un_max + a) / minLength;
        } else {
            s = fmax(s, (un_N + a) / jLength);
            s = fmax(s, (un_E + a) / iLength);
            if (A.dims == 3) s = fmax(s, (un_T + a) / kLength);
        }
        if (c == ncells - 1) *last_signal = s;
        sig = fmax(sig, s);
    }
    for (int off = 32; off > 0; off >>= 1) sig = fmax(sig, __shfl_xor(sig, off));
    if (__lane_id() == 0) wave_max[threadIdx.x / 64] = sig;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double m = fmax(fmax(wave_max[0], wave_max[1]), fmax(wave_max[2], wave_max[3]));
        atomicMax(smax_bits, (unsigned long long)__double_as_longlong(m));
    }
}

// One stage of a block built from its grid: the face kernels of every
// direction, then the update (ideal gas: ig_update_kernel; mixture: the
// mixture update kernel, which reads the face areas and cell volumes).
template <class M, int NSP, int FLUX>
void launch_gen_faces(const TPArgs& A, hipStream_t stream)
{
    const int nfk = (A.dims == 3) ? A.nk : 1;
    const long nf0 = (long)(A.ni + 1) * A.nj * nfk, nf1 = (long)A.ni * (A.nj + 1) * nfk;
    const long nf2 = (long)A.ni * A.nj * (A.nk + 1);
    const dim3 b(128);
    auto g = [](long n) { return dim3((unsigned)((n + 127) / 128)); };
    hipLaunchKernelGGL((gen_face_kernel<M, NSP, FLUX, 0>), g(nf0), b, 0, stream, A);
    hipLaunchKernelGGL((gen_face_kernel<M, NSP, FLUX, 1>), g(nf1), b, 0, stream, A);
    if (A.dims == 3) hipLaunchKernelGGL((gen_face_kernel<M, NSP, FLUX, 2>), g(nf2), b, 0, stream, A);
}

template <class M, int NSP>
void launch_gen_stage_m(const TPArgs& A, hipStream_t stream, int flux)
{
    if (flux == FLUX_HLLC) launch_gen_faces<M, NSP, FLUX_HLLC>(A, stream);
    else if (flux == FLUX_SHOCK) launch_gen_faces<M, NSP, FLUX_SHOCK>(A, stream);
    else if (flux == FLUX_OTHER) launch_gen_faces<M, NSP, FLUX_OTHER>(A, stream);
    else launch_gen_faces<M, NSP, FLUX_AUSMDV>(A, stream);
    const long nc = (long)A.ni * A.nj * ((A.dims == 3) ? A.nk : 1);
    const dim3 grid((unsigned)((nc + 127) / 128)), b(128);
    if constexpr (NSP == 0) {
        hipLaunchKernelGGL((ig_update_kernel<M>), grid, b, 0, stream, A);
    } else {
        if (A.guess) hipLaunchKernelGGL((tp_update_kernel<M, NSP, true, false>), grid, b, 0, stream, A);
        else hipLaunchKernelGGL((tp_update_kernel<M, NSP, false, false>), grid, b, 0, stream, A);
    }
}

template <class M>
void launch_gen_aux_m(const TPArgs& A, hipStream_t stream, int what, double* U)
{
    const long nc = (long)A.ni * A.nj * ((A.dims == 3) ? A.nk : 1);
    const dim3 grid((unsigned)((nc + 127) / 128)), b(128);
    if (what == 0) hipLaunchKernelGGL((gen_shock_detect_kernel<M>), grid, b, 0, stream, A);
    else hipLaunchKernelGGL((ig_decode_kernel<M>), grid, b, 0, stream, A, U);
}

} // namespace e4
