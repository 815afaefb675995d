// stage_kernels.h -- the fused stage kernel of the explicit update for
// ideal-gas structured blocks on gfx950, templated on the arithmetic mode.
//
// One launch performs, for every active cell of a block, what
// gasdynamic_explicit_increment_with_fixed_grid does per stage
// (eilmer/simcore.d:2338-2478): ghost-cell fill (done on the fly for the
// stencil cells that fall outside the block), convective_flux_phase0
// (sfluidblock.d:2189-2345), time_derivatives (fvcell.d:735-875), the stage
// update (fvcell.d:1153-1393), decode_conserved (fvcell.d:563-732) with the
// low-T fallback, and count_invalid_cells (fluidblock.d:433-499).
//
// Layout: conserved state is SoA [5][ncells] in HBM, i fastest.  A workgroup
// owns a TX x TY column of cells and marches up a chunk of k-planes.  Per
// plane:
//   1. the plane's primitive state (rho, vel, u, a), with two halo cells in i
//      and j, is decoded once into LDS;
//   2. every cell reconstructs its two face states per direction (the left
//      state of its plus face and the right state of its minus face) from one
//      limiter evaluation per scalar: on the uniform grid the van Albada
//      factor of both is the same expression of the same two differences, so
//      the results are those of the per-face evaluation of
//      onedinterp.d:136-162, bit for bit.  Right states go to LDS; the
//      k-direction stencil lives in a three-plane register queue;
//   3. every cell evaluates the flux of its plus faces (i, j from LDS, k from
//      registers); fluxes of the tile's west and south faces are extra items;
//   4. flux divergence, stage update, decode check and the write of the new
//      state happen in registers.
#pragma once
#include "stage_args.h"

namespace e4 {

// The stride between conserved quantities, opaque to the optimiser: without
// it the address bases A.X + q*qs of every array and quantity are hoisted out
// of the k-loop into scalar registers, more than the register file holds.
__device__ __forceinline__ long opaque_qs(long qs)
{
    asm volatile("" : "+s"(qs));
    return qs;
}

__device__ __forceinline__ int opaque_i(int x)
{
    asm volatile("" : "+v"(x));
    return x;
}

// The kernel's arguments read through a pointer to the kernarg segment that
// the optimiser cannot see through: each phase of the plane loop re-reads what
// it uses (scalar loads) instead of every argument being held in scalar
// registers across the loop, which overflows the scalar register file.
// (The host pass of the compiler only parses the kernel; it sees a plain
// pointer.)
// A batched launch's kernel argument is a BatchArgs; each workgroup reads its
// block's StageArgs from the table through the same kind of pointer (a
// read-only, uniform address, so scalar loads).
#if defined(__HIP_DEVICE_COMPILE__)
using KArgs = const __attribute__((address_space(4))) StageArgs*;
using KBatch = const __attribute__((address_space(4))) BatchArgs*;

__device__ __forceinline__ KArgs kargs()
{
    KArgs p = (KArgs)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(p));
    return p;
}

__device__ __forceinline__ KArgs kargs_entry(int e)
{
    KBatch b = (KBatch)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(b));
    KArgs p = (KArgs)(b->table + (unsigned long long)e * sizeof(StageArgs));
    asm volatile("" : "+s"(p));
    return p;
}

__device__ __forceinline__ int batch_entry()
{
    KBatch b = (KBatch)__builtin_amdgcn_kernarg_segment_ptr();
    return __builtin_amdgcn_readfirstlane(b->wg_entry[blockIdx.x]);
}

// word idx of a batched launch's workgroup map (uniform)
__device__ __forceinline__ int batch_word(int idx)
{
    KBatch b = (KBatch)__builtin_amdgcn_kernarg_segment_ptr();
    return __builtin_amdgcn_readfirstlane(b->wg_entry[idx]);
}

// a batched launch's work queue (null: one tile per workgroup) and tile count
__device__ __forceinline__ unsigned int* batch_queue()
{
    KBatch b = (KBatch)__builtin_amdgcn_kernarg_segment_ptr();
    return b->queue;
}

__device__ __forceinline__ int batch_ntiles()
{
    KBatch b = (KBatch)__builtin_amdgcn_kernarg_segment_ptr();
    return b->ntiles;
}
#else
using KArgs = const StageArgs*;
__device__ __forceinline__ KArgs kargs() { return nullptr; }
__device__ __forceinline__ KArgs kargs_entry(int) { return nullptr; }
__device__ __forceinline__ int batch_entry() { return 0; }
__device__ __forceinline__ int batch_word(int) { return 0; }
__device__ __forceinline__ unsigned int* batch_queue() { return nullptr; }
__device__ __forceinline__ int batch_ntiles() { return 0; }
#endif

// Decode of the cell (i,j,k); positions one or two cells outside the block in
// one index direction are ghost cells and take the value the face's ghost-cell
// effect gives them.  Positions further out return a harmless dummy state.
template <class M>
__device__ __forceinline__ Prim load_decode(KArgs A, long c)
{
    const long qs = opaque_qs(A->qs);
    const double* p = A->Us + c;
    const IdealGas g = A->gas;
    return decode_ideal<M>(p[0], p[qs], p[2 * qs], p[3 * qs], p[4 * qs], g);
}

__device__ __forceinline__ long cell_index(KArgs A, int i, int j, int k)
{
    return ((long)k * A->nj + j) * A->ni + i;
}

// With SH the shock indicator travels in the sign of the sound speed (a < 0:
// shock point): interior cells carry this stage's indicator, ghost cells the
// one of the cell they were copied from at the previous detection (the
// ghost-cell fill precedes the detector, simcore.d:2338-2358), inflow ghosts 0.
// STORED: ghost states stored per face (static profiles, x_order-1
// extrapolation) are read; without it those faces give the slip-wall mirror,
// which only faces of the cells the boundary kernel finishes read (the stage
// kernel leaves two layers next to such faces to it).
template <class M, bool SH = false, bool STORED = false>
__device__ Prim fetch_ghost(KArgs A, int i, int j, int k)
{
    Prim dummy{1.0, 0.0, 0.0, 0.0, 1.0, 1.0};
    int dir, idx, n;
    if (i < 0 || i >= A->ni) {
        if (j < 0 || j >= A->nj || k < 0 || k >= A->nk) return dummy;
        dir = 0; idx = i; n = A->ni;
    } else if (j < 0 || j >= A->nj) {
        if (k < 0 || k >= A->nk) return dummy;
        dir = 1; idx = j; n = A->nj;
    } else {
        dir = 2; idx = k; n = A->nk;
    }
    const bool plus = idx >= n;
    const int layer = plus ? idx - n + 1 : -idx;
    if (layer > 2) return dummy;
    const int face = (dir == 0) ? (plus ? FACE_EAST : FACE_WEST)
                   : (dir == 1) ? (plus ? FACE_NORTH : FACE_SOUTH)
                                : (plus ? FACE_TOP : FACE_BOTTOM);
    const int type = A->bc->type[face];
    if (type == BC_INFLOW) return A->bc->fs[face];
    if (type == BC_CONNECTED) {
        // full_face_copy.d:1277-1357: the mapped interior cell of the other block
        const Connection nb = A->bc->nbr[face];
        const int fa = (face == FACE_NORTH || face == FACE_SOUTH || face >= FACE_TOP) ? i : j;
        const int fb = (face >= FACE_TOP) ? j : k;
        int src[3];
        connection_source(nb.dims, face, nb.face, nb.orient, nb.n, layer, fa, fb, src);
        const long c = ((long)src[2] * nb.n[1] + src[1]) * nb.n[0] + src[0];
        const double* p = A->nbrU[face] + c;
        const long qs = nb.qs;
        const IdealGas g = A->gas;
        Prim P = decode_ideal<M>(p[0], p[qs], p[2 * qs], p[3 * qs], p[4 * qs], g);
        if (SH && A->nbrS[face][c]) P.a = -(P.a);
        return P;
    }
    int si = i, sj = j, sk = k;
    if (STORED && (type == BC_INFLOW_PROFILE || type == BC_EXTRAPOLATE_1)) {
        // stored ghost states: a static profile's (flow_state_copy_from_profile.d),
        // with its points' shock indicators, or the x_order-1 extrapolation
        // filled at the start of the stage (extrapolate_ghosts_kernel), whose
        // indicator is 0
        const int na = (dir == 0) ? A->nj : A->ni;
        const int fa = (dir == 0) ? j : i, fb = (dir == 2) ? j : k;
        const long nfc = (long)na * ((dir == 2) ? A->nj : A->nk);
        const long g = (layer - 1) * nfc + fa + (long)fb * na;
        const double* G = A->gx[face] + g;
        const long gs = 2 * nfc;
        Prim P{G[P_RHO * gs], G[P_VX * gs], G[P_VY * gs], G[P_VZ * gs], G[P_U * gs], G[P_A * gs]};
        if (SH && type == BC_INFLOW_PROFILE && A->gS[face][g]) P.a = -(P.a);
        return P;
    }
    int src;
    if (type == BC_EXTRAPOLATE || type == BC_OUTFLOW_SIMPLE_FLUX) src = plus ? n - 1 : 0;
    else if (type == BC_PERIODIC) src = plus ? layer - 1 : n - layer;
    // slip-wall mirror; a wall without ghost data (BC_SLIP_WALL_1) gets the
    // same values, which only faces the boundary kernel recomputes read
    else src = plus ? n - layer : layer - 1;
    if (dir == 0) si = src; else if (dir == 1) sj = src; else sk = src;
    const long sc = cell_index(A, si, sj, sk);
    Prim P = load_decode<M>(A, sc);
    if (SH && A->Sprev[sc]) P.a = -(P.a);
    if (type == BC_SLIP_WALL || type == BC_SLIP_WALL_1) {
        // reflect_normal_velocity (ghost_cell.d:32-43): negate the face-normal component
        if (dir == 0) P.v0 = -(P.v0); else if (dir == 1) P.v1 = -(P.v1); else P.v2 = -(P.v2);
    }
    return P;
}

template <class M, bool SH = false, bool STORED = false>
__device__ __forceinline__ PrimT<typename M::R> fetch(KArgs A, int i, int j, int k)
{
    using R = typename M::R;
    Prim P;
    if (i >= 0 && i < A->ni && j >= 0 && j < A->nj && k >= 0 && k < A->nk) {
        const long c = cell_index(A, i, j, k);
        P = load_decode<M>(A, c);
        if (SH && A->S[c]) P.a = -(P.a);
    } else {
        P = fetch_ghost<M, SH, STORED>(A, i, j, k);
    }
    return PrimT<R>{R(P.rho), R(P.v0), R(P.v1), R(P.v2), R(P.u), R(P.a)};
}

// The update formulas of fvcell.d:1153-1393 for one conserved quantity.
__device__ __forceinline__ double stage_update(KArgs A, double uold, double dd0, double dd1, double d)
{
    if (A->stage == 1) return uold + A->c1 * d;
    if (A->stage == 2) return uold + A->dt * (A->g1 * dd0 + A->g2 * d);
    return uold + A->dt * (A->g1 * dd0 + A->g2 * dd1 + A->g3 * d);
}

__device__ __forceinline__ unsigned int wave_count(bool pred)
{
    return (unsigned int)__popcll(__ballot(pred));
}

// A reconstructed face state of a cell: density, velocity in the global frame,
// internal energy and the cell-centre sound speed (sfluidblock.d:2199-2208).
template <class R>
struct SideT {
    R rho, v0, v1, v2, u, a;
};

// onedinterp.d:136-162 for the two faces of one cell on a uniform grid.  For
// the plus face (qL1, qL0, qR0) = (qm, q0, qp) and for the minus face
// (qL0, qR0, qR1) = (qm, q0, qp); with equal cell lengths the three scalings
// are the same number, so delLminus of the plus face and del of the minus face
// are both dm, del of the plus face and delRplus of the minus face are both
// dp, and sL of the plus face equals sR of the minus face.
template <class M>
__device__ __forceinline__ void interp_l2r2_cell(const L2R2UT<typename M::R>& w, bool apply_limiter,
                                                 bool extrema_clipping, typename M::R qm, typename M::R q0,
                                                 typename M::R qp, typename M::R& qLp, typename M::R& qRm)
{
    using R = typename M::R;
    const R eps = R(1.0e-12);   // onedinterp.d:22
    R dm = (q0 - qm) * w.c;
    R dp = (qp - q0) * w.c;
    R s = R(1.0);
    if (apply_limiter) s = dv<M>(dm * dp + ab(dm * dp) + eps, dm * dm + dp * dp + eps);
    qLp = q0 + s * w.a * (dp * w.len3 + dm * w.len);
    qRm = q0 - s * w.a * (dp * w.len + dm * w.len3);
    if (extrema_clipping) {
        qLp = clip_limits<M>(qLp, q0, qp);
        qRm = clip_limits<M>(qRm, qm, q0);
    }
}

// Both face states of the cell Q0 in one direction (neighbours Qm, Qp), with
// the EOS fall-back of onedinterp.d:1039-1068: a face state with u <= 0 or
// rho <= 0 is replaced by the cell state.  With order 1 both are the cell state.
template <class M>
__device__ __forceinline__ void recon_cell(const L2R2UT<typename M::R>& w, bool high_order, bool apply_limiter,
                                           bool extrema_clipping, const PrimT<typename M::R>& Qm,
                                           const PrimT<typename M::R>& Q0, const PrimT<typename M::R>& Qp,
                                           SideT<typename M::R>& Lp, SideT<typename M::R>& Rm)
{
    using R = typename M::R;
    Lp = SideT<R>{Q0.rho, Q0.v0, Q0.v1, Q0.v2, Q0.u, Q0.a};
    Rm = Lp;
    if (high_order) {
        SideT<R> l, r;
        interp_l2r2_cell<M>(w, apply_limiter, extrema_clipping, Qm.v0, Q0.v0, Qp.v0, l.v0, r.v0);
        interp_l2r2_cell<M>(w, apply_limiter, extrema_clipping, Qm.v1, Q0.v1, Qp.v1, l.v1, r.v1);
        interp_l2r2_cell<M>(w, apply_limiter, extrema_clipping, Qm.v2, Q0.v2, Qp.v2, l.v2, r.v2);
        interp_l2r2_cell<M>(w, apply_limiter, extrema_clipping, Qm.rho, Q0.rho, Qp.rho, l.rho, r.rho);
        interp_l2r2_cell<M>(w, apply_limiter, extrema_clipping, Qm.u, Q0.u, Qp.u, l.u, r.u);
        // the fall-back is rare: a branch around it, not selects
        const SideT<R> cell = Lp;
        Lp = SideT<R>{l.rho, l.v0, l.v1, l.v2, l.u, Q0.a};
        Rm = SideT<R>{r.rho, r.v0, r.v1, r.v2, r.u, Q0.a};
        const bool badL = l.u <= R(0) || l.rho <= R(0), badR = r.u <= R(0) || r.rho <= R(0);
        if (__builtin_expect(badL || badR, 0)) {
            if (badL) Lp = cell;
            if (badR) Rm = cell;
        }
    }
}

template <int C, class R>
__device__ __forceinline__ R vcomp(const SideT<R>& S)
{
    if constexpr (C == 0) return S.v0;
    else if constexpr (C == 1) return S.v1;
    else return S.v2;
}

// Flux through a face from its left and right states; the face frame is the
// permutation (PN, PT1, PT2) of the global axes: 3D i-face (x,y,z), j-face
// (y,z,x), k-face (z,x,y); 2D i-face (x,y,z), j-face (y,x,z)
// (sfluidblock.d:527-588, fvinterface.d:265-296).  The 2D i-face's t1 = -y is
// an exact sign flip that every operation of the face path commutes with, so
// it is not applied.  The face EOS is p = rho Rgas (u / Cv) (ideal_gas.d:98-106).
// Fg receives the flux in the global frame (mass, x-, y-, z-momentum, energy).
// With FLUX_SHOCK the sides' sound speeds carry their cells' shock indicators
// in the sign (fetch).
template <class M, int FLUX, int PN, int PT1, int PT2>
__device__ __forceinline__ void side_flux(const SideT<typename M::R>& L, const SideT<typename M::R>& Rs,
                                          const IdealGasT<typename M::R>& g, typename M::R Fg[5],
                                          FluxOpt<typename M::R> opt = {})
{
    using R = typename M::R;
    FaceStateT<R> Lf, Rf;
    bool shock = false;
    Lf.rho = L.rho; Lf.u = L.u; Lf.a = L.a;
    Rf.rho = Rs.rho; Rf.u = Rs.u; Rf.a = Rs.a;
    if constexpr (FLUX == FLUX_SHOCK) {
        shock = L.a < R(0) || Rs.a < R(0);
        Lf.a = ab(L.a);
        Rf.a = ab(Rs.a);
    }
    Lf.vn = vcomp<PN>(L); Lf.vt1 = vcomp<PT1>(L); Lf.vt2 = vcomp<PT2>(L);
    Rf.vn = vcomp<PN>(Rs); Rf.vt1 = vcomp<PT1>(Rs); Rf.vt2 = vcomp<PT2>(Rs);
    R TL = Lf.u * g.Cvinv;
    R TR = Rf.u * g.Cvinv;
    Lf.T = TL;
    Rf.T = TR;
    if constexpr (M::kFast) {
        Lf.RT = g.Rgas * TL;
        Rf.RT = g.Rgas * TR;
        Lf.p = Lf.rho * Lf.RT;
        Rf.p = Rf.rho * Rf.RT;
    } else {
        Lf.p = Lf.rho * g.Rgas * TL;
        Rf.p = Rf.rho * g.Rgas * TR;
        Lf.RT = Rf.RT = R(0);
    }
    Flux5T<R> F = flux_calc<M, FLUX>(Lf, Rf, g, shock, opt);
    Fg[0] = F.mass;
    Fg[1 + PN] = F.mn;
    Fg[1 + PT1] = F.mt1;
    Fg[1 + PT2] = F.mt2;
    Fg[4] = F.E;
}

template <class R>
__device__ __forceinline__ FluxOpt<R> shock_opt(KArgs A)
{
    return FluxOpt<R>{R(A->shear_tol), A->always_hanel != 0, A->shock_efm != 0, A->other_roe != 0, R(A->M_inf)};
}

template <class R>
__device__ __forceinline__ PrimT<R> lds_prim(const R* sP, int plane, int s)
{
    return PrimT<R>{sP[s], sP[plane + s], sP[2 * plane + s], sP[3 * plane + s], sP[4 * plane + s],
                    sP[5 * plane + s]};
}

template <class R>
__device__ __forceinline__ void lds_put_side(R* dst, int nslot, int s, const SideT<R>& S)
{
    dst[s] = S.rho; dst[nslot + s] = S.v0; dst[2 * nslot + s] = S.v1;
    dst[3 * nslot + s] = S.v2; dst[4 * nslot + s] = S.u; dst[5 * nslot + s] = S.a;
}

template <class R>
__device__ __forceinline__ SideT<R> lds_side(const R* src, int nslot, int s)
{
    return SideT<R>{src[s], src[nslot + s], src[2 * nslot + s], src[3 * nslot + s], src[4 * nslot + s],
                    src[5 * nslot + s]};
}

template <class R>
__device__ __forceinline__ void lds_put_flux(R* dst, int nslot, int s, const R F[5])
{
#pragma unroll
    for (int q = 0; q < 5; ++q) dst[q * nslot + s] = F[q];
}

// The rest of a stage for one cell from its flux integral (the sum over its
// faces of -F * outsign * area), its inverse volume and the y-momentum source
// (the axisymmetric pressure term, 0 for planar flow): time derivative
// (fvcell.d:826-832) with the block's user-defined source terms added to the
// source vector after that term (simcore.d:2442-2465), stage update, decode checks
// with the low-T fallback (fvcell.d:570-589, 680-697), check_data
// (flowstate.d:345-388, gas_state.d:199-241), write, and the wave's counts of
// invalid cells and decode failures.  All in FP64.
template <class M>
__device__ __forceinline__ void finish_cell(KArgs ap, long c, const double integ[5], double vol_inv, double qy)
{
    const long qs = opaque_qs(ap->qs);
    const double* __restrict__ src = ap->src;
    double Un[5];
    // every value the update reads, loaded before the first store (the
    // compiler cannot move a load of one array across a store to another, so
    // interleaved they would cost a memory round trip per quantity)
    double uold[5], dd0[5], dd1[5], srcq[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        uold[q] = ap->Uold[c + q * qs];
        dd0[q] = (ap->stage >= 2) ? ap->d0[c + q * qs] : 0.0;
        dd1[q] = (ap->stage >= 3) ? ap->d1[c + q * qs] : 0.0;
        srcq[q] = src ? src[c + q * qs] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        double Q = (q == 2) ? qy : 0.0;
        if (src) Q += srcq[q];
        double d = vol_inv * integ[q] + Q;
        if (ap->dout) ap->dout[c + q * qs] = d;
        Un[q] = stage_update(ap, uold[q], dd0[q], dd1[q], d);
    }
    bool decode_fail = false, invalid = false;
    double rho = Un[0];
    if (!(rho > 0.0)) {
        decode_fail = true;
    } else {
        using MD = typename std::conditional<M::kFast, ModeFast64, ModeExact>::type;
        double dinv = dv<MD>(1.0, rho);
        double vx = Un[1] * dinv, vy = Un[2] * dinv, vz = Un[3] * dinv;
        double u = Un[4] * dinv;
        double ke = 0.5 * (vx * vx + vy * vy + vz * vz);
        u -= ke;
        double T;
        if (u <= 0.0) {
            T = 200.0;
            u = ap->gas.Cv * T;
            Un[1] = rho * vx; Un[2] = rho * vy; Un[3] = rho * vz;
            Un[4] = rho * (u + ke);
        } else {
            T = u * ap->gas.Cvinv;
        }
        double p = rho * ap->gas.Rgas * T;
        if (!isfinite(rho)) invalid = true;
        if (!isfinite(T) || T < 1.01 * 1.0) invalid = true;
        if (!isfinite(p)) invalid = true;
        if (fabs(vx) > 30000.0 || fabs(vy) > 30000.0 || fabs(vz) > 30000.0) invalid = true;
        if (T > 50000.0) invalid = true;
    }
#pragma unroll
    for (int q = 0; q < 5; ++q) ap->Unew[c + q * qs] = Un[q];
    unsigned int nbad = wave_count(invalid), nfail = wave_count(decode_fail);
    if ((nbad | nfail) && (__lane_id() == __ffsll((unsigned long long)__ballot(true)) - 1))
        count_bad_cells(ap->counters, ap->abort_this, ap->max_invalid, nbad, nfail);
}

// Whether the cell is in the layers next to a one-sided boundary that the
// boundary kernel finishes.
template <int DIMS>
__device__ __forceinline__ bool left_to_boundary(KArgs A, int i, int j, int k)
{
    bool f = i < A->fix_lo[0] || i >= A->ni - A->fix_hi[0] || j < A->fix_lo[1] || j >= A->nj - A->fix_hi[1];
    if constexpr (DIMS == 3) f = f || k < A->fix_lo[2] || k >= A->nk - A->fix_hi[2];
    return f;
}

// Waves per SIMD the register allocation of the stage kernel aims for
// (measured at 256^3: three for the FP64 forms, whose few spills cost less
// than the occupancy gains, and which the 48 KiB of LDS per workgroup allows;
// four for the mixed form, which then fits in 128 registers).
#ifndef E4_STAGE_WAVES_EXACT
#define E4_STAGE_WAVES_EXACT 3
#endif
#ifndef E4_STAGE_WAVES_FAST64
#define E4_STAGE_WAVES_FAST64 3
#endif
#ifndef E4_STAGE_WAVES_MIXED
#define E4_STAGE_WAVES_MIXED 4
#endif
template <class M> struct StageOccupancy { static constexpr int waves = E4_STAGE_WAVES_FAST64; };
template <> struct StageOccupancy<ModeExact> { static constexpr int waves = E4_STAGE_WAVES_EXACT; };
template <> struct StageOccupancy<ModeMixed> { static constexpr int waves = E4_STAGE_WAVES_MIXED; };

// DEF: the default reconstruction (order 2, limiter, extrema clipping) fixed
// at compile time; otherwise the three options are read from the arguments.
// One tile of the stage: with BATCH, tile wg of the launch's tile list (the
// workgroup's own index, or the one it took from the work queue); without,
// the workgroup's grid position.
template <class M, int TX, int TY, int DIMS, int FLUX, bool DEF, bool BATCH>
__device__ __forceinline__ void stage_tile(int wg)
{
    // the arguments of this workgroup's block, re-read per phase
    const int entry = BATCH ? batch_word(wg) : 0;
    auto args = [&]() { return BATCH ? kargs_entry(entry) : kargs(); };
    using R = typename M::R;
    constexpr bool SH = (FLUX == FLUX_SHOCK);
    constexpr int NT = TX * TY;
    constexpr int PX = TX + 4, PY = TY + 4, PLANE = PX * PY;
    constexpr int NSI = (TX + 1) * TY;       // i-face slots: face fc (0..TX) of row r is r*(TX+1)+fc
    constexpr int NSJ = TX * (TY + 1);       // j-face slots: face fr (0..TY) of column c is fr*TX+c
    constexpr int NHALO = 2 * (TX + TY);     // face states of halo cells
    constexpr int NBND = TX + TY;            // west and south faces of the tile
    static_assert(NHALO <= NT && NBND <= NT, "tile too narrow for its halo work");
    // j-face frame
    constexpr int JN = 1, JT1 = (DIMS == 3) ? 2 : 0, JT2 = (DIMS == 3) ? 0 : 2;
    __shared__ R sP[6 * PLANE];              // rho, vx, vy, vz, u, a of plane k with halos
    __shared__ R sI[6 * NSI];                // right states of i-faces, then i-face fluxes (global frame)
    __shared__ R sJ[6 * NSJ];                // right states of j-faces, then j-face fluxes
    __shared__ R sLw[6 * TY];                // left states of the tile's west faces
    __shared__ R sLs[6 * TX];                // left states of the tile's south faces

    KArgs ap = args();
    if (stage_aborted(ap->abort_prev)) return;
    int bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
    if constexpr (BATCH) {
        const int w = wg - ap->wg0;
        bx = w % ap->tiles_x;
        by = (w / ap->tiles_x) % ap->tiles_y;
        bz = w / (ap->tiles_x * ap->tiles_y);
    }
    const int tid = threadIdx.x;
    const int tx = tid % TX, ty = tid / TX;
    const int i0 = bx * TX, j0 = by * TY;
    const int i = i0 + tx, j = j0 + ty;
    const bool own_ij = (i < ap->ni) && (j < ap->nj);
    const int k0 = (DIMS == 3) ? bz * ap->kchunk : 0;
    const int k1 = (DIMS == 3) ? min(k0 + ap->kchunk, ap->nk) : 1;
    const bool ho = DEF || ap->high_order != 0, lim = DEF || ap->apply_limiter != 0;
    const bool clip = DEF || ap->extrema_clipping != 0;

    // k-direction register queue at the top of iteration k: own = plane k,
    // qa = k+1, qb = k+2; LpC is the left state of the top face of cell k+1,
    // Ft and Fb the fluxes of the own cell's top and bottom faces.  The
    // k-direction work of the next plane is done at the end of an iteration,
    // where the i/j face states are no longer live.
    PrimT<R> own, qa, qb;
    SideT<R> LpC;
    R Ft[5] = {R(0), R(0), R(0), R(0), R(0)}, Fb[5] = {R(0), R(0), R(0), R(0), R(0)};
    if constexpr (DIMS == 3) {
        const L2R2UT<R> wk = cast_wu<R>(L2R2U(ap->wu[2]));
        const IdealGasT<R> gR = cast_gas<R>(IdealGas(ap->gas));
        PrimT<R> qm2 = fetch<M, SH>(ap, i, j, k0 - 2);
        PrimT<R> qm1 = fetch<M, SH>(ap, i, j, k0 - 1);
        own = fetch<M, SH>(ap, i, j, k0);
        qa = fetch<M, SH>(ap, i, j, k0 + 1);
        qb = fetch<M, SH>(ap, i, j, k0 + 2);
        SideT<R> Lbelow, Rdummy, Lown, Rown, Rabove;
        recon_cell<M>(wk, ho, lim, clip, qm2, qm1, own, Lbelow, Rdummy);
        recon_cell<M>(wk, ho, lim, clip, qm1, own, qa, Lown, Rown);
        if (own_ij) side_flux<M, FLUX, 2, 0, 1>(Lbelow, Rown, gR, Fb, shock_opt<R>(ap));
        recon_cell<M>(wk, ho, lim, clip, own, qa, qb, LpC, Rabove);
        if (own_ij) side_flux<M, FLUX, 2, 0, 1>(Lown, Rabove, gR, Ft, shock_opt<R>(ap));
    }

    for (int k = k0; k < k1; ++k) {
        // per-thread offsets are recomputed in every iteration rather than
        // held in registers across the loop
        ap = args();
        const int tidk = opaque_i(tid);
        const int txk = tidk % TX, tyk = tidk / TX;
        const int ik = i0 + txk, jk = j0 + tyk;
        const bool own_k = (ik < ap->ni) && (jk < ap->nj);
        if constexpr (DIMS == 2) own = fetch<M, SH>(ap, ik, jk, 0);
        // conserved state of plane k+3, loaded now and decoded at the end
        double raw[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        const bool raw_ok = (DIMS == 3) && own_k && (k + 1 < k1) && (k + 3 < ap->nk);
        if (raw_ok) {
            const long qs = opaque_qs(ap->qs);
            const double* p = ap->Us + cell_index(ap, ik, jk, k + 3);
#pragma unroll
            for (int q = 0; q < 5; ++q) raw[q] = p[q * qs];
        }
        // 1. plane k with halos into LDS
        const int s0 = (tyk + 2) * PX + (txk + 2);
        sP[0 * PLANE + s0] = own.rho; sP[1 * PLANE + s0] = own.v0; sP[2 * PLANE + s0] = own.v1;
        sP[3 * PLANE + s0] = own.v2; sP[4 * PLANE + s0] = own.u; sP[5 * PLANE + s0] = own.a;
        for (int t = tidk; t < 4 * (TX + TY); t += NT) {
            int ii, jj;
            if (t < 4 * TY) {
                int side = t / (2 * TY), rem = t % (2 * TY);
                int col = rem / TY, row = rem % TY;
                ii = side ? TX + 2 + col : col;
                jj = row + 2;
            } else {
                int t2 = t - 4 * TY;
                int side = t2 / (2 * TX), rem = t2 % (2 * TX);
                int row = rem / TX, col = rem % TX;
                jj = side ? TY + 2 + row : row;
                ii = col + 2;
            }
            PrimT<R> h = fetch<M, SH>(ap, i0 + ii - 2, j0 + jj - 2, k);
            const int s = jj * PX + ii;
            sP[0 * PLANE + s] = h.rho; sP[1 * PLANE + s] = h.v0; sP[2 * PLANE + s] = h.v1;
            sP[3 * PLANE + s] = h.v2; sP[4 * PLANE + s] = h.u; sP[5 * PLANE + s] = h.a;
        }
        __syncthreads();
        // 2. face states: own cell in i and j (right states to LDS, left states
        // kept) and the halo cells' states the tile's boundary faces need.
        ap = args();
        const L2R2UT<R> wi = cast_wu<R>(L2R2U(ap->wu[0])), wj = cast_wu<R>(L2R2U(ap->wu[1]));
        {
            const int e = tidk - (NT - NHALO);
            if (e >= 0) {
                // e < TY: west neighbour column (left state), e < 2TY: east neighbour
                // column (right state), then south row (left), north row (right)
                int s, step, dst_kind, slot;
                if (e < TY) { s = (e + 2) * PX + 1; step = 1; dst_kind = 0; slot = e; }
                else if (e < 2 * TY) { s = (e - TY + 2) * PX + TX + 2; step = 1; dst_kind = 1; slot = (e - TY) * (TX + 1) + TX; }
                else if (e < 2 * TY + TX) { s = PX + (e - 2 * TY) + 2; step = PX; dst_kind = 2; slot = e - 2 * TY; }
                else { s = (TY + 2) * PX + (e - 2 * TY - TX) + 2; step = PX; dst_kind = 3; slot = TY * TX + (e - 2 * TY - TX); }
                SideT<R> L, Rr;
                if (step == 1)
                    recon_cell<M>(wi, ho, lim, clip, lds_prim(sP, PLANE, s - 1), lds_prim(sP, PLANE, s),
                                  lds_prim(sP, PLANE, s + 1), L, Rr);
                else
                    recon_cell<M>(wj, ho, lim, clip, lds_prim(sP, PLANE, s - PX), lds_prim(sP, PLANE, s),
                                  lds_prim(sP, PLANE, s + PX), L, Rr);
                if (dst_kind == 0) lds_put_side(sLw, TY, slot, L);
                else if (dst_kind == 1) lds_put_side(sI, NSI, slot, Rr);
                else if (dst_kind == 2) lds_put_side(sLs, TX, slot, L);
                else lds_put_side(sJ, NSJ, slot, Rr);
            }
        }
        SideT<R> LpI, LpJ;
        {
            SideT<R> Rm;
            recon_cell<M>(wi, ho, lim, clip, lds_prim(sP, PLANE, s0 - 1), own, lds_prim(sP, PLANE, s0 + 1), LpI, Rm);
            lds_put_side(sI, NSI, tyk * (TX + 1) + txk, Rm);
            recon_cell<M>(wj, ho, lim, clip, lds_prim(sP, PLANE, s0 - PX), own, lds_prim(sP, PLANE, s0 + PX), LpJ, Rm);
            lds_put_side(sJ, NSJ, tyk * TX + txk, Rm);
        }
        __syncthreads();
        // 3. fluxes of the own cell's plus faces and of the tile's west and
        // south faces; each face's flux overwrites its right state in LDS
        // (the same thread read it).
        ap = args();
        const IdealGasT<R> gR = cast_gas<R>(IdealGas(ap->gas));
        if (own_k) {
            R Fx[5];
            const int si = tyk * (TX + 1) + txk + 1, sj = (tyk + 1) * TX + txk;
            side_flux<M, FLUX, 0, 1, 2>(LpI, lds_side(sI, NSI, si), gR, Fx, shock_opt<R>(ap));
            lds_put_flux(sI, NSI, si, Fx);
            side_flux<M, FLUX, JN, JT1, JT2>(LpJ, lds_side(sJ, NSJ, sj), gR, Fx, shock_opt<R>(ap));
            lds_put_flux(sJ, NSJ, sj, Fx);
        }
        if (tidk < NBND) {
            R Fx[5];
            if (tidk < TY) {
                if (j0 + tidk < ap->nj) {
                    const int si = tidk * (TX + 1);
                    side_flux<M, FLUX, 0, 1, 2>(lds_side(sLw, TY, tidk), lds_side(sI, NSI, si), gR, Fx,
                                                shock_opt<R>(ap));
                    lds_put_flux(sI, NSI, si, Fx);
                }
            } else {
                const int c = tidk - TY;
                if (i0 + c < ap->ni) {
                    side_flux<M, FLUX, JN, JT1, JT2>(lds_side(sLs, TX, c), lds_side(sJ, NSJ, c), gR, Fx,
                                                     shock_opt<R>(ap));
                    lds_put_flux(sJ, NSJ, c, Fx);
                }
            }
        }
        __syncthreads();
        // 4. flux divergence in face order N, E, S, W, T, B (fvcell.d:789-825),
        // stage update, decode check, write; all in FP64.  Cells next to a
        // one-sided boundary are the boundary kernel's.
        ap = args();
        if (own_k && !(ap->fix_off[6] != 0 && left_to_boundary<DIMS>(ap, ik, jk, k))) {
            const int iW = tyk * (TX + 1) + txk, jS = tyk * TX + txk;
            const int iE = iW + 1, jN = jS + TX;
            double integ[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
            auto add_face = [&](const R* Fsrc, int nface, int fidx, double outsign, double Aface) {
                double area = outsign * Aface;
                integ[0] -= double(Fsrc[0 * nface + fidx]) * area;
                integ[1] -= double(Fsrc[1 * nface + fidx]) * area;
                integ[2] -= double(Fsrc[2 * nface + fidx]) * area;
                if (DIMS == 3) integ[3] -= double(Fsrc[3 * nface + fidx]) * area;
                integ[4] -= double(Fsrc[4 * nface + fidx]) * area;
            };
            bool axi = false;
            if constexpr (DIMS == 2) {
                if (ap->axi) {
                    // per-radian face areas, volume and the pressure source
                    axi = true;
                    const AxiGeom g = axi_geom(ap->axi, ap->ni, ap->nj);
                    const long c = cell_index(ap, ik, jk, 0);
                    add_face(sJ, NSJ, jN, 1.0, g.aj[(long)(jk + 1) * ap->ni + ik]);
                    add_face(sI, NSI, iE, 1.0, g.ai[(long)jk * (ap->ni + 1) + ik + 1]);
                    add_face(sJ, NSJ, jS, -1.0, g.aj[(long)jk * ap->ni + ik]);
                    add_face(sI, NSI, iW, -1.0, g.ai[(long)jk * (ap->ni + 1) + ik]);
                    const Prim P = load_decode<M>(ap, c);
                    const double p = P.rho * ap->gas.Rgas * (P.u * ap->gas.Cvinv);
                    const double vol = g.vol[c];
                    finish_cell<M>(ap, c, integ, 1.0 / vol, 0.0 + p * g.areaxy[c] / vol);
                }
            }
            if (!axi) {
                add_face(sJ, NSJ, jN, 1.0, ap->area[1]);
                add_face(sI, NSI, iE, 1.0, ap->area[0]);
                add_face(sJ, NSJ, jS, -1.0, ap->area[1]);
                add_face(sI, NSI, iW, -1.0, ap->area[0]);
                if constexpr (DIMS == 3) {
                    add_face(Ft, 1, 0, 1.0, ap->area[2]);
                    add_face(Fb, 1, 0, -1.0, ap->area[2]);
                }
                finish_cell<M>(ap, cell_index(ap, ik, jk, k), integ, ap->vol_inv, 0.0);
            }
        }
        // 5. k-direction work for plane k+1: face states of cell k+2 and the
        // flux of the top face of cell k+1
        if constexpr (DIMS == 3) {
            if (k + 1 < k1) {
                ap = args();
                const IdealGasT<R> gR = cast_gas<R>(IdealGas(ap->gas));
                PrimT<R> q3;
                if (raw_ok) {
                    Prim P = decode_ideal<M>(raw[0], raw[1], raw[2], raw[3], raw[4], IdealGas(ap->gas));
                    if (SH && ap->S[cell_index(ap, ik, jk, k + 3)]) P.a = -(P.a);
                    q3 = PrimT<R>{R(P.rho), R(P.v0), R(P.v1), R(P.v2), R(P.u), R(P.a)};
                } else {
                    q3 = fetch<M, SH>(ap, ik, jk, k + 3);
                }
                SideT<R> Lnext, Rnext;
                recon_cell<M>(cast_wu<R>(L2R2U(ap->wu[2])), ho, lim, clip, qa, qb, q3, Lnext, Rnext);
#pragma unroll
                for (int q = 0; q < 5; ++q) Fb[q] = Ft[q];
                if (own_k) side_flux<M, FLUX, 2, 0, 1>(LpC, Rnext, gR, Ft, shock_opt<R>(ap));
                LpC = Lnext;
                own = qa; qa = qb; qb = q3;
            }
        }
    }
}

// The stage kernel.  A batched launch with a work queue (BatchArgs::queue) is
// persistent: each workgroup takes tiles from the queue (largest blocks'
// tiles first, as the host orders them) until none is left, and the last
// workgroup to finish clears the queue for the next launch on the stream;
// without a queue each workgroup is one tile.
template <class M, int TX, int TY, int DIMS, int FLUX, bool DEF, bool BATCH>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(StageOccupancy<M>::waves)))
void stage_kernel(typename std::conditional<BATCH, BatchArgs, StageArgs>::type)
{
    if constexpr (BATCH) {
        unsigned int* q = batch_queue();
        if (q) {
            __shared__ int s_tile;
            const int ntiles = batch_ntiles();
            for (;;) {
                if (threadIdx.x == 0) s_tile = (int)atomicAdd(q, 1u);
                __syncthreads();
                const int t = s_tile;
                __syncthreads();
                if (t >= ntiles) break;
                stage_tile<M, TX, TY, DIMS, FLUX, DEF, BATCH>(t);
                __syncthreads();
            }
            if (threadIdx.x == 0) {
                __threadfence();
                if (atomicAdd(q + 1, 1u) == gridDim.x - 1) {
                    atomicExch(q, 0u);
                    atomicExch(q + 1, 0u);
                }
            }
            return;
        }
        stage_tile<M, TX, TY, DIMS, FLUX, DEF, BATCH>((int)blockIdx.x);
    } else {
        stage_tile<M, TX, TY, DIMS, FLUX, DEF, BATCH>(0);
    }
}

// ------------------------------------------------------------ boundaries
// Faces whose flux the reference computes otherwise than from two
// reconstructed sides (sfluidblock.d:2189-2345): next to a boundary without
// ghost-cell data (WallBC_WithSlip1) the boundary face takes the wall flux of
// fluxcalc.d:130-298 and the face after it the one-sided reconstruction
// l1r2/l2r1 of onedinterp.d:165-818; on an OutFlowBC_SimpleFlux face the
// flux is the interior cell's (boundary_flux_effect.d:367-546).  The stage
// kernel leaves the cells beside such faces (two layers, one layer) to
// boundary_cell_kernel, which evaluates all faces of each such cell, interior
// ones with the stage kernel's functions, and finishes the cell.

__device__ __forceinline__ double prim_comp(const Prim& P, int d) { return d == 0 ? P.v0 : (d == 1 ? P.v1 : P.v2); }

// The cell (i, j, k) with its index along direction D replaced by p.
template <int D>
__device__ __forceinline__ void along(int& i, int& j, int& k, int p)
{
    if constexpr (D == 0) i = p;
    else if constexpr (D == 1) j = p;
    else k = p;
}

// The side of an l1r2 (or, mirrored, l2r1) face next to a boundary without
// ghost data, per scalar (onedinterp.d:187-191, 218-222): the mean of the two
// cells beside the face (linear_interp_prepare on equal lengths gives weights
// 1/2), or the boundary-side cell's value where the limiter sees an extremum
// over (boundary-side, far, beyond).  The reference's product of the two
// scaled differences is formed from the negated pair on the l2r1 side, the
// same number.
template <class M>
__device__ __forceinline__ typename M::R lin_side(const L2R2UT<typename M::R>& w, bool lim, bool clip,
                                                 typename M::R qn, typename M::R qf, typename M::R qb)
{
    using R = typename M::R;
    const R del = (qf - qn) * w.c;
    const R delp = (qb - qf) * w.c;
    if (lim && delp * del < R(0)) return qn;
    R q = qn * R(0.5) + qf * R(0.5);
    if (clip) q = clip_limits<M>(q, qn, qf);
    return q;
}

template <class M>
__device__ __forceinline__ SideT<typename M::R> lin_side_state(const L2R2UT<typename M::R>& w, bool lim, bool clip,
                                                               const PrimT<typename M::R>& n,
                                                               const PrimT<typename M::R>& f,
                                                               const PrimT<typename M::R>& b)
{
    using R = typename M::R;
    SideT<R> s{lin_side<M>(w, lim, clip, n.rho, f.rho, b.rho), lin_side<M>(w, lim, clip, n.v0, f.v0, b.v0),
               lin_side<M>(w, lim, clip, n.v1, f.v1, b.v1), lin_side<M>(w, lim, clip, n.v2, f.v2, b.v2),
               lin_side<M>(w, lim, clip, n.u, f.u, b.u), n.a};
    // the EOS fall-back to the cell's state (onedinterp.d:1039-1068)
    if (s.u <= R(0) || s.rho <= R(0)) s = SideT<R>{n.rho, n.v0, n.v1, n.v2, n.u, n.a};
    return s;
}

// Flux (global frame, FP64) of the direction-D face whose plus-side cell is
// (i, j, k); pos is that cell's index along D (0 .. n).
template <class M, int FLUX, int D, int DIMS>
__device__ __forceinline__ void boundary_face_flux(KArgs A, int i, int j, int k, double F[5])
{
    using R = typename M::R;
    constexpr bool SH = (FLUX == FLUX_SHOCK);
    constexpr int PN = D, PT1 = (D == 0) ? 1 : (D == 1 ? ((DIMS == 3) ? 2 : 0) : 0);
    constexpr int PT2 = (D == 0) ? 2 : (D == 1 ? ((DIMS == 3) ? 0 : 2) : 1);
    const int n = (D == 0) ? A->ni : (D == 1 ? A->nj : A->nk);
    const int pos = (D == 0) ? i : (D == 1 ? j : k);
    const int fm = (D == 0) ? FACE_WEST : (D == 1 ? FACE_SOUTH : FACE_BOTTOM);
    const int fp = (D == 0) ? FACE_EAST : (D == 1 ? FACE_NORTH : FACE_TOP);
    const int tm = A->bc->type[fm], tp = A->bc->type[fp];
    auto cell64 = [&](int p) {
        int ci = i, cj = j, ck = k;
        along<D>(ci, cj, ck, p);
        return load_decode<M>(A, cell_index(A, ci, cj, ck));
    };
    auto cellR = [&](int p) {
        int ci = i, cj = j, ck = k;
        along<D>(ci, cj, ck, p);
        return fetch<M, SH, true>(A, ci, cj, ck);
    };
#pragma unroll
    for (int q = 0; q < 5; ++q) F[q] = 0.0;
    const bool ho = A->high_order != 0, lim = A->apply_limiter != 0, clip = A->extrema_clipping != 0;
    const IdealGas g = A->gas;
    // outflow flux of the interior cell's state (its p = rho Rgas T, T = u/Cv,
    // as the decode has it)
    if ((pos == 0 && tm == BC_OUTFLOW_SIMPLE_FLUX) || (pos == n && tp == BC_OUTFLOW_SIMPLE_FLUX)) {
        const bool plus = (pos == n);
        const Prim c = cell64(plus ? n - 1 : 0);
        const double p = c.rho * g.Rgas * (c.u * g.Cvinv);
        const double vn = prim_comp(c, D);
        const double mass_flux = c.rho * vn;
        if ((plus ? 1.0 : -1.0) * mass_flux > 0.0) {
            F[0] = mass_flux;
            F[1] = c.v0 * mass_flux; F[2] = c.v1 * mass_flux; F[3] = c.v2 * mass_flux;
            F[1 + D] = p + vn * mass_flux;
            const double utot = c.u + 0.5 * (c.v0 * c.v0 + c.v1 * c.v1 + c.v2 * c.v2);
            F[4] = mass_flux * utot + p * vn;
        } else {
            F[1 + D] = p;
        }
        return;
    }
    // wall face: the wall cell's state, or with reconstruction and no
    // clipping its linear extrapolation (onedinterp.d:820-1037)
    if ((pos == 0 && tm == BC_SLIP_WALL_1) || (pos == n && tp == BC_SLIP_WALL_1)) {
        const bool left = (pos == 0);
        const Prim c0 = cell64(left ? 0 : n - 1);
        double rho = c0.rho, u = c0.u, vn = prim_comp(c0, D);
        if (ho && !clip) {
            const Prim c1 = cell64(left ? 1 : n - 2);
            const double w0 = A->bc->ext_w[D][0], w1 = A->bc->ext_w[D][1];
            const double r = c0.rho * w0 + c1.rho * w1, e = c0.u * w0 + c1.u * w1;
            if (e > 0.0 && r > 0.0) {
                rho = r; u = e;
                vn = prim_comp(c0, D) * w0 + prim_comp(c1, D) * w1;
            }
        }
        const double p = rho * g.Rgas * (u * g.Cvinv);
        F[1 + D] = wall_pressure(rho, p, c0.a, vn, A->bc->wall_gamma, left);
        return;
    }
    const L2R2UT<R> w = cast_wu<R>(L2R2U(A->wu[D]));
    SideT<R> L, Rr, unused;
    if (pos == 1 && tm == BC_SLIP_WALL_1) {
        const PrimT<R> q0 = cellR(0), q1 = cellR(1), q2 = cellR(2);
        recon_cell<M>(w, ho, lim, clip, q0, q1, q2, unused, Rr);
        L = ho ? lin_side_state<M>(w, lim, clip, q0, q1, q2) : SideT<R>{q0.rho, q0.v0, q0.v1, q0.v2, q0.u, q0.a};
    } else if (pos == n - 1 && tp == BC_SLIP_WALL_1) {
        const PrimT<R> qa = cellR(n - 3), qb = cellR(n - 2), qc = cellR(n - 1);
        recon_cell<M>(w, ho, lim, clip, qa, qb, qc, L, unused);
        Rr = ho ? lin_side_state<M>(w, lim, clip, qc, qb, qa) : SideT<R>{qc.rho, qc.v0, qc.v1, qc.v2, qc.u, qc.a};
    } else {
        const PrimT<R> qm2 = cellR(pos - 2), qm1 = cellR(pos - 1), q0 = cellR(pos), qp1 = cellR(pos + 1);
        recon_cell<M>(w, ho, lim, clip, qm2, qm1, q0, L, unused);
        recon_cell<M>(w, ho, lim, clip, qm1, q0, qp1, unused, Rr);
    }
    R Fg[5];
    side_flux<M, FLUX, PN, PT1, PT2>(L, Rr, cast_gas<R>(g), Fg, shock_opt<R>(A));
#pragma unroll
    for (int q = 0; q < 5; ++q) F[q] = double(Fg[q]);
}

// The cells of the slabs beside one-sided boundaries (each cell once: a cell
// in several slabs belongs to the first), 64 per workgroup: wave f evaluates
// face f (north, east, south, west, top, bottom) of all 64 cells, so every
// lane computes one face flux rather than all of its cell's; the fluxes meet
// in LDS and the first wave sums them in face order and finishes the cells.
// A batched launch covers the blocks of a batched stage: each workgroup reads
// its block's arguments from the table (bwg0: the block's first workgroup).
template <class M, int DIMS, int FLUX, bool BATCH>
__global__ __launch_bounds__(128 * DIMS) void boundary_cell_kernel(
    typename std::conditional<BATCH, BatchArgs, StageArgs>::type)
{
    constexpr int NF = 2 * DIMS;
    __shared__ double sF[NF][5][64];
    KArgs A;
    int wg = (int)blockIdx.x;
    if constexpr (BATCH) {
        A = kargs_entry(batch_entry());
        wg -= A->bwg0;
    } else {
        A = kargs();
    }
    if (stage_aborted(A->abort_prev)) return;
    const int lane = threadIdx.x & 63;
    const int f = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int t = wg * 64 + lane;
    bool valid = t < A->fix_off[6];
    int i = 0, j = 0, k = 0;
    if (valid) {
        int slab = 0;
        while (slab < 5 && t >= A->fix_off[slab + 1]) ++slab;
        const int d = slab >> 1, side = slab & 1;
        const int n[3] = {A->ni, A->nj, (DIMS == 3) ? A->nk : 1};
        const int L = side ? ((d == 0) ? A->fix_hi[0] : (d == 1 ? A->fix_hi[1] : A->fix_hi[2]))
                           : ((d == 0) ? A->fix_lo[0] : (d == 1 ? A->fix_lo[1] : A->fix_lo[2]));
        const int da = (d + 1) % 3, db = (d + 2) % 3;
        const int local = t - A->fix_off[slab];
        const int layer = local % L, rest = local / L;
        int c[3];
        c[d] = side ? n[d] - 1 - layer : layer;
        c[da] = rest % n[da];
        c[db] = rest / n[da];
        for (int s2 = 0; s2 < slab; ++s2) {
            const int d2 = s2 >> 1;
            const int L2 = (s2 & 1) ? A->fix_hi[d2] : A->fix_lo[d2];
            if ((s2 & 1) ? (c[d2] >= n[d2] - L2) : (c[d2] < L2)) valid = false;
        }
        i = c[0]; j = c[1]; k = c[2];
    }
    if (valid) {
        double F[5];
        switch (f) {
        case 0: boundary_face_flux<M, FLUX, 1, DIMS>(A, i, j + 1, k, F); break;
        case 1: boundary_face_flux<M, FLUX, 0, DIMS>(A, i + 1, j, k, F); break;
        case 2: boundary_face_flux<M, FLUX, 1, DIMS>(A, i, j, k, F); break;
        case 3: boundary_face_flux<M, FLUX, 0, DIMS>(A, i, j, k, F); break;
        case 4: if constexpr (DIMS == 3) boundary_face_flux<M, FLUX, 2, DIMS>(A, i, j, k + 1, F); break;
        default: if constexpr (DIMS == 3) boundary_face_flux<M, FLUX, 2, DIMS>(A, i, j, k, F); break;
        }
#pragma unroll
        for (int q = 0; q < 5; ++q) sF[f][q][lane] = F[q];
    }
    __syncthreads();
    if (f != 0 || !valid) return;
    // flux divergence in face order N, E, S, W, T, B (fvcell.d:789-825)
    double integ[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    auto add_face = [&](int face, double outsign, double Aface) {
        const double area = outsign * Aface;
        integ[0] -= sF[face][0][lane] * area;
        integ[1] -= sF[face][1][lane] * area;
        integ[2] -= sF[face][2][lane] * area;
        if (DIMS == 3) integ[3] -= sF[face][3][lane] * area;
        integ[4] -= sF[face][4][lane] * area;
    };
    const long cc = cell_index(A, i, j, k);
    if (DIMS == 2 && A->axi) {
        const AxiGeom g = axi_geom(A->axi, A->ni, A->nj);
        add_face(0, 1.0, g.aj[(long)(j + 1) * A->ni + i]);
        add_face(1, 1.0, g.ai[(long)j * (A->ni + 1) + i + 1]);
        add_face(2, -1.0, g.aj[(long)j * A->ni + i]);
        add_face(3, -1.0, g.ai[(long)j * (A->ni + 1) + i]);
        const Prim P = load_decode<M>(A, cc);
        const double p = P.rho * A->gas.Rgas * (P.u * A->gas.Cvinv);
        const double vol = g.vol[cc];
        finish_cell<M>(A, cc, integ, 1.0 / vol, 0.0 + p * g.areaxy[cc] / vol);
        return;
    }
    add_face(0, 1.0, A->area[1]);
    add_face(1, 1.0, A->area[0]);
    add_face(2, -1.0, A->area[1]);
    add_face(3, -1.0, A->area[0]);
    if constexpr (DIMS == 3) {
        add_face(4, 1.0, A->area[2]);
        add_face(5, -1.0, A->area[2]);
    }
    finish_cell<M>(A, cc, integ, A->vol_inv, 0.0);
}

// x_order-1 ghost cells of one face of an ideal-gas block
// (extrapolate_copy.d:102-186), stored per face: layer 1 linear in rho, u and
// velocity from the boundary cell c1 and the cell behind it, layer 2 from
// layer 1 and c1; the sound speed stays the constructed cell's (the buffer's
// initial value), which the extrapolation does not write.
// A batched launch covers the x_order-1 faces of several blocks: word 2w of
// the workgroup map holds (table entry << 3 | face) and word 2w+1 the
// workgroup's index within that face.
template <class M, bool BATCH>
__global__ __launch_bounds__(128) void extrapolate_ghosts_kernel(
    typename std::conditional<BATCH, BatchArgs, StageArgs>::type, int face, double* G)
{
    KArgs A;
    int wg = (int)blockIdx.x;
    if constexpr (BATCH) {
        const int code = batch_word(2 * (int)blockIdx.x);
        wg = batch_word(2 * (int)blockIdx.x + 1);
        A = kargs_entry(code >> 3);
        face = code & 7;
        G = const_cast<double*>(A->gx[face]);
    } else {
        A = kargs();
    }
    if (stage_aborted(A->abort_prev)) return;
    const int dir = (face == FACE_EAST || face == FACE_WEST) ? 0 : ((face == FACE_NORTH || face == FACE_SOUTH) ? 1 : 2);
    const bool plus = (face == FACE_EAST || face == FACE_NORTH || face == FACE_TOP);
    const int na = (dir == 0) ? A->nj : A->ni;
    const long nfc = (long)na * ((dir == 2) ? A->nj : A->nk);
    const long t = (long)wg * blockDim.x + threadIdx.x;
    if (t >= nfc) return;
    const int fa = (int)(t % na), fb = (int)(t / na);
    const int n = (dir == 0) ? A->ni : ((dir == 1) ? A->nj : A->nk);
    auto cell = [&](int p) {
        int ci, cj, ck;
        if (dir == 0) { ci = p; cj = fa; ck = fb; }
        else if (dir == 1) { ci = fa; cj = p; ck = fb; }
        else { ci = fa; cj = fb; ck = p; }
        return load_decode<M>(A, cell_index(A, ci, cj, ck));
    };
    const int e = plus ? n - 1 : 0, in = plus ? -1 : 1;
    const Prim c1 = cell(e), c2 = cell(e + in);
    const long gs = 2 * nfc;
    auto put = [&](int layer, const Prim& a, const Prim& b) {
        double* d = G + (layer - 1) * nfc + t;
        d[P_RHO * gs] = 2.0 * a.rho - b.rho;
        d[P_VX * gs] = 2.0 * a.v0 - b.v0;
        d[P_VY * gs] = 2.0 * a.v1 - b.v1;
        d[P_VZ * gs] = 2.0 * a.v2 - b.v2;
        d[P_U * gs] = 2.0 * a.u - b.u;
    };
    put(1, c1, c2);
    const Prim g1{2.0 * c1.rho - c2.rho, 2.0 * c1.v0 - c2.v0, 2.0 * c1.v1 - c2.v1, 2.0 * c1.v2 - c2.v2,
                  2.0 * c1.u - c2.u, 0.0};
    put(2, g1, c1);
}

template <class M>
void launch_extrapolate_ghosts_mode(const StageArgs& A, int faces, hipStream_t stream)
{
    for (int f = 0; f < 6; ++f) {
        if (!(faces & (1 << f))) continue;
        const int dir = (f == FACE_EAST || f == FACE_WEST) ? 0 : ((f == FACE_NORTH || f == FACE_SOUTH) ? 1 : 2);
        const long nfc = (long)((dir == 0) ? A.nj : A.ni) * ((dir == 2) ? A.nj : A.nk);
        hipLaunchKernelGGL((extrapolate_ghosts_kernel<M, false>), dim3((unsigned)((nfc + 127) / 128)), dim3(128), 0,
                           stream, A, f, const_cast<double*>(A.gx[f]));
    }
}

template <class M>
void launch_extrapolate_ghosts_batch_mode(const BatchArgs& B, int nwg, hipStream_t stream)
{
    if (nwg <= 0) return;
    hipLaunchKernelGGL((extrapolate_ghosts_kernel<M, true>), dim3((unsigned)nwg), dim3(128), 0, stream, B, 0,
                       (double*)nullptr);
}

template <class M>
void launch_boundary_cells_mode(const StageArgs& A, hipStream_t stream, int dims, int flux)
{
    const int n = A.fix_off[6];
    if (n <= 0) return;
    const dim3 block(128 * dims), grid((unsigned)((n + 63) / 64));
    if (dims == 3) {
        if (flux == FLUX_HLLC) hipLaunchKernelGGL((boundary_cell_kernel<M, 3, FLUX_HLLC, false>), grid, block, 0, stream, A);
        else if (flux == FLUX_SHOCK) hipLaunchKernelGGL((boundary_cell_kernel<M, 3, FLUX_SHOCK, false>), grid, block, 0, stream, A);
        else if (flux == FLUX_OTHER) hipLaunchKernelGGL((boundary_cell_kernel<M, 3, FLUX_OTHER, false>), grid, block, 0, stream, A);
        else hipLaunchKernelGGL((boundary_cell_kernel<M, 3, FLUX_AUSMDV, false>), grid, block, 0, stream, A);
    } else {
        if (flux == FLUX_HLLC) hipLaunchKernelGGL((boundary_cell_kernel<M, 2, FLUX_HLLC, false>), grid, block, 0, stream, A);
        else if (flux == FLUX_SHOCK) hipLaunchKernelGGL((boundary_cell_kernel<M, 2, FLUX_SHOCK, false>), grid, block, 0, stream, A);
        else if (flux == FLUX_OTHER) hipLaunchKernelGGL((boundary_cell_kernel<M, 2, FLUX_OTHER, false>), grid, block, 0, stream, A);
        else hipLaunchKernelGGL((boundary_cell_kernel<M, 2, FLUX_AUSMDV, false>), grid, block, 0, stream, A);
    }
}

// nwg workgroups: the sum over the batch's blocks of ceil(boundary cells / 64),
// each mapped to its block by B.wg_entry.
template <class M>
void launch_boundary_cells_batch_mode(const BatchArgs& B, int nwg, hipStream_t stream, int dims, int flux)
{
    if (nwg <= 0) return;
    const dim3 block(128 * dims), grid((unsigned)nwg);
    if (dims == 3) {
        if (flux == FLUX_HLLC) hipLaunchKernelGGL((boundary_cell_kernel<M, 3, FLUX_HLLC, true>), grid, block, 0, stream, B);
        else if (flux == FLUX_SHOCK) hipLaunchKernelGGL((boundary_cell_kernel<M, 3, FLUX_SHOCK, true>), grid, block, 0, stream, B);
        else if (flux == FLUX_OTHER) hipLaunchKernelGGL((boundary_cell_kernel<M, 3, FLUX_OTHER, true>), grid, block, 0, stream, B);
        else hipLaunchKernelGGL((boundary_cell_kernel<M, 3, FLUX_AUSMDV, true>), grid, block, 0, stream, B);
    } else {
        if (flux == FLUX_HLLC) hipLaunchKernelGGL((boundary_cell_kernel<M, 2, FLUX_HLLC, true>), grid, block, 0, stream, B);
        else if (flux == FLUX_SHOCK) hipLaunchKernelGGL((boundary_cell_kernel<M, 2, FLUX_SHOCK, true>), grid, block, 0, stream, B);
        else if (flux == FLUX_OTHER) hipLaunchKernelGGL((boundary_cell_kernel<M, 2, FLUX_OTHER, true>), grid, block, 0, stream, B);
        else hipLaunchKernelGGL((boundary_cell_kernel<M, 2, FLUX_AUSMDV, true>), grid, block, 0, stream, B);
    }
}

template <class M, int DIMS, int FLUX>
void launch_stage_df(const StageArgs& A, dim3 grid, hipStream_t stream)
{
    const dim3 block(TILE_X * TILE_Y);
    // the less common calculators (FLUX_OTHER) read the reconstruction options
    // from the arguments in every case
    if constexpr (FLUX != FLUX_OTHER) {
        if (A.high_order && A.apply_limiter && A.extrema_clipping) {
            hipLaunchKernelGGL((stage_kernel<M, TILE_X, TILE_Y, DIMS, FLUX, true, false>), grid, block, 0, stream, A);
            return;
        }
    }
    hipLaunchKernelGGL((stage_kernel<M, TILE_X, TILE_Y, DIMS, FLUX, false, false>), grid, block, 0, stream, A);
}

template <class M, int DIMS, int FLUX>
void launch_stage_batch_df(const BatchArgs& B, int nwg, bool def, hipStream_t stream)
{
    const dim3 block(TILE_X * TILE_Y), grid(nwg);
    if constexpr (FLUX != FLUX_OTHER) {
        if (def) {
            hipLaunchKernelGGL((stage_kernel<M, TILE_X, TILE_Y, DIMS, FLUX, true, true>), grid, block, 0, stream, B);
            return;
        }
    }
    hipLaunchKernelGGL((stage_kernel<M, TILE_X, TILE_Y, DIMS, FLUX, false, true>), grid, block, 0, stream, B);
}

template <class M>
void launch_stage_batch_mode(const BatchArgs& B, int nwg, bool def, hipStream_t stream, int dims, int flux)
{
    if (dims == 3) {
        if (flux == FLUX_HLLC) launch_stage_batch_df<M, 3, FLUX_HLLC>(B, nwg, def, stream);
        else if (flux == FLUX_SHOCK) launch_stage_batch_df<M, 3, FLUX_SHOCK>(B, nwg, def, stream);
        else if (flux == FLUX_OTHER) launch_stage_batch_df<M, 3, FLUX_OTHER>(B, nwg, def, stream);
        else launch_stage_batch_df<M, 3, FLUX_AUSMDV>(B, nwg, def, stream);
    } else {
        if (flux == FLUX_HLLC) launch_stage_batch_df<M, 2, FLUX_HLLC>(B, nwg, def, stream);
        else if (flux == FLUX_SHOCK) launch_stage_batch_df<M, 2, FLUX_SHOCK>(B, nwg, def, stream);
        else if (flux == FLUX_OTHER) launch_stage_batch_df<M, 2, FLUX_OTHER>(B, nwg, def, stream);
        else launch_stage_batch_df<M, 2, FLUX_AUSMDV>(B, nwg, def, stream);
    }
}

// Launch of the stage kernel of mode M for a block of dims dimensions.
template <class M>
void launch_stage_mode(const StageArgs& A, dim3 grid, hipStream_t stream, int dims, int flux)
{
    if (dims == 3) {
        if (flux == FLUX_HLLC) launch_stage_df<M, 3, FLUX_HLLC>(A, grid, stream);
        else if (flux == FLUX_SHOCK) launch_stage_df<M, 3, FLUX_SHOCK>(A, grid, stream);
        else if (flux == FLUX_OTHER) launch_stage_df<M, 3, FLUX_OTHER>(A, grid, stream);
        else launch_stage_df<M, 3, FLUX_AUSMDV>(A, grid, stream);
    } else {
        if (flux == FLUX_HLLC) launch_stage_df<M, 2, FLUX_HLLC>(A, grid, stream);
        else if (flux == FLUX_SHOCK) launch_stage_df<M, 2, FLUX_SHOCK>(A, grid, stream);
        else if (flux == FLUX_OTHER) launch_stage_df<M, 2, FLUX_OTHER>(A, grid, stream);
        else launch_stage_df<M, 2, FLUX_AUSMDV>(A, grid, stream);
    }
}

// fluidblock.d:415-431 with PJ_ShockDetector (shockdetectors.d) for every
// active cell from the stage's stencil state: a face is a shock point when
// (u_n,right - u_n,left) / min(a_left, a_right) < compression tolerance (the
// face normal of this axis-aligned grid is a unit axis); a cell is one when
// any of its faces is.  Ghost neighbours are the ghost-cell effects' states.
template <class M>
__device__ __forceinline__ Prim prim_at(KArgs A, int i, int j, int k)
{
    if (i >= 0 && i < A->ni && j >= 0 && j < A->nj && k >= 0 && k < A->nk)
        return load_decode<M>(A, cell_index(A, i, j, k));
    return fetch_ghost<M, false, true>(A, i, j, k);
}

__device__ __forceinline__ double prim_vel(const Prim& P, int d) { return d == 0 ? P.v0 : (d == 1 ? P.v1 : P.v2); }

// A workgroup owns a 32 x 8 column of cells and marches a chunk of k-planes:
// each plane, with a one-cell ring in i and j, is decoded once into LDS
// (velocity and sound speed; double-buffered so that plane k+1 loads while
// plane k is tested); the k-direction neighbours come from the next plane
// and from the thread's previous cell.
constexpr int SD_TX = 32, SD_TY = 8, SD_KC = 16;

template <class M, int DIMS>
__global__ __launch_bounds__(256) void shock_detect_kernel(StageArgs)
{
    using MD = typename std::conditional<M::kFast, ModeFast64, ModeExact>::type;
    constexpr int TX = SD_TX, TY = SD_TY, PX = TX + 2, PY = TY + 2, PL = PX * PY;
    __shared__ double sv[2][4][PL];          // vx, vy, vz, a of a plane and its ring
    KArgs A = kargs();
    if (stage_aborted(A->abort_prev)) return;
    const int i0 = blockIdx.x * TX, j0 = blockIdx.y * TY;
    const int k0 = (DIMS == 3) ? blockIdx.z * SD_KC : 0;
    const int k1 = (DIMS == 3) ? min(k0 + SD_KC, A->nk) : 1;
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int i = i0 + tx, j = j0 + ty;
    const bool own = i < A->ni && j < A->nj;
    const int h = (ty + 1) * PX + (tx + 1);
    auto load_plane = [&](int buf, int k) {
        for (int t = threadIdx.x; t < PL; t += blockDim.x) {
            const int hx = t % PX, hy = t / PX;
            // ring corners are never read; positions outside in two
            // directions give a dummy state
            const Prim P = prim_at<M>(A, i0 + hx - 1, j0 + hy - 1, k);
            sv[buf][0][t] = P.v0; sv[buf][1][t] = P.v1; sv[buf][2][t] = P.v2; sv[buf][3][t] = P.a;
        }
    };
    const double tol = A->compression_tol;
    auto face = [&](double uL, double aL, double uR, double aR) {
        const double a_min = (aL < aR) ? aL : aR;
        return dv<MD>(uR - uL, a_min) < tol;
    };
    double vz_below = 0.0, a_below = 1.0;
    if constexpr (DIMS == 3) {
        if (own) {
            const Prim P = prim_at<M>(A, i, j, k0 - 1);
            vz_below = P.v2;
            a_below = P.a;
        }
    }
    load_plane(0, k0);
    int buf = 0;
    for (int k = k0; k < k1; ++k) {
        if constexpr (DIMS == 3) load_plane(buf ^ 1, k + 1);
        __syncthreads();
        if (own) {
            const double (*c)[PL] = sv[buf];
            // a boundary face without ghost data has the cell on one side only
            // (shockdetectors.d:39-54): its velocity towards the wall
            const int ng = A->noghost;
            const double a0 = c[3][h];
            auto minus_face = [&](bool wall, double uL, double aL, double u0) {
                return wall ? dv<MD>(u0, a0) < tol : face(uL, aL, u0, a0);
            };
            auto plus_face = [&](bool wall, double u0, double uR, double aR) {
                return wall ? dv<MD>(-u0, a0) < tol : face(u0, a0, uR, aR);
            };
            bool S = minus_face(i == 0 && (ng & (1 << FACE_WEST)), c[0][h - 1], c[3][h - 1], c[0][h]) ||
                     plus_face(i == A->ni - 1 && (ng & (1 << FACE_EAST)), c[0][h], c[0][h + 1], c[3][h + 1]) ||
                     minus_face(j == 0 && (ng & (1 << FACE_SOUTH)), c[1][h - PX], c[3][h - PX], c[1][h]) ||
                     plus_face(j == A->nj - 1 && (ng & (1 << FACE_NORTH)), c[1][h], c[1][h + PX], c[3][h + PX]);
            if constexpr (DIMS == 3) {
                const double (*n)[PL] = sv[buf ^ 1];
                S = S || minus_face(k == 0 && (ng & (1 << FACE_BOTTOM)), vz_below, a_below, c[2][h]) ||
                    plus_face(k == A->nk - 1 && (ng & (1 << FACE_TOP)), c[2][h], n[2][h], n[3][h]);
                vz_below = c[2][h];
                a_below = c[3][h];
            }
            A->Snew[cell_index(A, i, j, k)] = S ? 1 : 0;
        }
        __syncthreads();
        buf ^= 1;
    }
}

template <class M>
void launch_shock_detect_mode(const StageArgs& A, int dims, hipStream_t stream)
{
    const dim3 block(SD_TX * SD_TY);
    const dim3 grid((A.ni + SD_TX - 1) / SD_TX, (A.nj + SD_TY - 1) / SD_TY,
                    (dims == 3) ? (A.nk + SD_KC - 1) / SD_KC : 1);
    if (dims == 3) hipLaunchKernelGGL((shock_detect_kernel<M, 3>), grid, block, 0, stream, A);
    else hipLaunchKernelGGL((shock_detect_kernel<M, 2>), grid, block, 0, stream, A);
}

} // namespace e4
