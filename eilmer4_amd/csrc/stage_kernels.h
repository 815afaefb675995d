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
// owns a TX x TY column of cells and marches up a chunk of k-planes.  Each
// plane's primitive state (rho, vel, u, a), with two halo cells in i and j, is
// decoded once into LDS; i- and j-faces are reconstructed from LDS, each face
// once per plane; the k-direction stencil of a column lives in a four-plane
// register queue, and the k-face flux of the top of one plane is carried as
// the bottom of the next.  The flux divergence, stage update, decode check and
// the write of the new state all happen in registers.
#pragma once
#include "stage_args.h"

namespace e4 {

// Decode of the cell (i,j,k); positions one or two cells outside the block in
// one index direction are ghost cells and take the value the face's ghost-cell
// effect gives them.  Positions further out return a harmless dummy state.
template <class M>
__device__ __forceinline__ Prim load_decode(const StageArgs& A, long c)
{
    return decode_ideal<M>(A.Us[c], A.Us[c + A.qs], A.Us[c + 2 * A.qs], A.Us[c + 3 * A.qs], A.Us[c + 4 * A.qs],
                           A.gas);
}

__device__ __forceinline__ long cell_index(const StageArgs& A, int i, int j, int k)
{
    return ((long)k * A.nj + j) * A.ni + i;
}

template <class M>
__device__ Prim fetch_ghost(const StageArgs& A, int i, int j, int k)
{
    Prim dummy{1.0, 0.0, 0.0, 0.0, 1.0, 1.0};
    int dir, idx, n;
    if (i < 0 || i >= A.ni) {
        if (j < 0 || j >= A.nj || k < 0 || k >= A.nk) return dummy;
        dir = 0; idx = i; n = A.ni;
    } else if (j < 0 || j >= A.nj) {
        if (k < 0 || k >= A.nk) return dummy;
        dir = 1; idx = j; n = A.nj;
    } else {
        dir = 2; idx = k; n = A.nk;
    }
    const bool plus = idx >= n;
    const int layer = plus ? idx - n + 1 : -idx;
    if (layer > 2) return dummy;
    const int face = (dir == 0) ? (plus ? FACE_EAST : FACE_WEST)
                   : (dir == 1) ? (plus ? FACE_NORTH : FACE_SOUTH)
                                : (plus ? FACE_TOP : FACE_BOTTOM);
    const int type = A.bc_type[face];
    if (type == BC_INFLOW) return A.bc_fs[face];
    int src;
    if (type == BC_EXTRAPOLATE) src = plus ? n - 1 : 0;
    else if (type == BC_PERIODIC) src = plus ? layer - 1 : n - layer;
    else src = plus ? n - layer : layer - 1;     // slip wall mirror
    int si = i, sj = j, sk = k;
    if (dir == 0) si = src; else if (dir == 1) sj = src; else sk = src;
    Prim P = load_decode<M>(A, cell_index(A, si, sj, sk));
    if (type == BC_SLIP_WALL) {
        // reflect_normal_velocity (ghost_cell.d:32-43): negate the face-normal component
        if (dir == 0) P.v0 = -(P.v0); else if (dir == 1) P.v1 = -(P.v1); else P.v2 = -(P.v2);
    }
    return P;
}

template <class M>
__device__ __forceinline__ PrimT<typename M::R> fetch(const StageArgs& A, int i, int j, int k)
{
    using R = typename M::R;
    Prim P;
    if (i >= 0 && i < A.ni && j >= 0 && j < A.nj && k >= 0 && k < A.nk) P = load_decode<M>(A, cell_index(A, i, j, k));
    else P = fetch_ghost<M>(A, i, j, k);
    return PrimT<R>{R(P.rho), R(P.v0), R(P.v1), R(P.v2), R(P.u), R(P.a)};
}

// Face frames as permutations of (x,y,z): 3D i-face (x,y,z), j-face (y,z,x),
// k-face (z,x,y); 2D i-face (x,y,z), j-face (y,x,z) (sfluidblock.d:527-588,
// fvinterface.d:265-296).  The 2D i-face's t1 = -y is an exact sign flip that
// every operation of the face path commutes with, so it is not applied.
template <int DIMS>
__device__ __forceinline__ void face_perm(int dir, int& pn, int& pt1, int& pt2)
{
    if (dir == 0) { pn = 0; pt1 = 1; pt2 = 2; }
    else if (dir == 1) { if (DIMS == 3) { pn = 1; pt1 = 2; pt2 = 0; } else { pn = 1; pt1 = 0; pt2 = 2; } }
    else { pn = 2; pt1 = 0; pt2 = 1; }
}

__device__ __forceinline__ unsigned int wave_count(bool pred)
{
    return (unsigned int)__popcll(__ballot(pred));
}

// The update formulas of fvcell.d:1153-1393 for one conserved quantity.
__device__ __forceinline__ double stage_update(const StageArgs& A, double uold, double dd0, double dd1, double d)
{
    if (A.stage == 1) return uold + A.c1 * d;
    if (A.stage == 2) return uold + A.dt * (A.g1 * dd0 + A.g2 * d);
    return uold + A.dt * (A.g1 * dd0 + A.g2 * dd1 + A.g3 * d);
}

template <class M, int TX, int TY, int DIMS, int FLUX>
__global__ __launch_bounds__(TX * TY) void stage_kernel(StageArgs A)
{
    using R = typename M::R;
    constexpr int NT = TX * TY;
    constexpr int PX = TX + 4, PY = TY + 4, PLANE = PX * PY;
    constexpr int NFI = TY * (TX + 1), NFJ = (TY + 1) * TX;
    __shared__ R sP[6 * PLANE];              // rho, vx, vy, vz, u, a of plane k with halos
    __shared__ R sFi[5 * NFI];               // i-face fluxes (global frame)
    __shared__ R sFj[5 * NFJ];               // j-face fluxes (global frame)

    const int tid = threadIdx.x;
    const int tx = tid % TX, ty = tid / TX;
    const int i0 = blockIdx.x * TX, j0 = blockIdx.y * TY;
    const int i = i0 + tx, j = j0 + ty;
    const bool own_ij = (i < A.ni) && (j < A.nj);
    const int k0 = (DIMS == 3) ? blockIdx.z * A.kchunk : 0;
    const int k1 = (DIMS == 3) ? min(k0 + A.kchunk, A.nk) : 1;
    const bool ho = A.high_order != 0, lim = A.apply_limiter != 0, clip = A.extrema_clipping != 0;
    const IdealGasT<R> gR = cast_gas<R>(A.gas);

    // k-face frame (z, x, y)
    auto kface = [](const PrimT<R>& P) { return PrimT<R>{P.rho, P.v2, P.v0, P.v1, P.u, P.a}; };

    // k-direction register queue (k-1, k, k+1, k+2) and the carried bottom-face flux
    PrimT<R> q0, q1, q2, q3;
    R Fb[5] = {R(0), R(0), R(0), R(0), R(0)};
    if constexpr (DIMS == 3) {
        PrimT<R> qa = fetch<M>(A, i, j, k0 - 2);
        q0 = fetch<M>(A, i, j, k0 - 1);
        q1 = fetch<M>(A, i, j, k0);
        q2 = fetch<M>(A, i, j, k0 + 1);
        if (own_ij) {
            Flux5T<R> F = face_flux_ideal<M, FLUX>(kface(qa), kface(q0), kface(q1), kface(q2), cast_w<R>(A.w[2]), gR,
                                                   ho, lim, clip);
            Fb[0] = F.mass; Fb[1] = F.mt1; Fb[2] = F.mt2; Fb[3] = F.mn; Fb[4] = F.E;
        }
        q3 = fetch<M>(A, i, j, k0 + 2);
    }

    for (int k = k0; k < k1; ++k) {
        PrimT<R> own = (DIMS == 3) ? q1 : fetch<M>(A, i, j, 0);
        // 1. plane k with halos into LDS
        {
            const int s = (ty + 2) * PX + (tx + 2);
            sP[0 * PLANE + s] = own.rho; sP[1 * PLANE + s] = own.v0; sP[2 * PLANE + s] = own.v1;
            sP[3 * PLANE + s] = own.v2; sP[4 * PLANE + s] = own.u; sP[5 * PLANE + s] = own.a;
        }
        for (int t = tid; t < 4 * (TX + TY); t += NT) {
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
            PrimT<R> h = fetch<M>(A, i0 + ii - 2, j0 + jj - 2, k);
            const int s = jj * PX + ii;
            sP[0 * PLANE + s] = h.rho; sP[1 * PLANE + s] = h.v0; sP[2 * PLANE + s] = h.v1;
            sP[3 * PLANE + s] = h.v2; sP[4 * PLANE + s] = h.u; sP[5 * PLANE + s] = h.a;
        }
        __syncthreads();
        // 2. i- and j-faces of plane k: work item w < NT is the west face of cell
        // (tx,ty), NT <= w < 2NT the south face, and the rest the east faces of
        // the last column and the north faces of the top row.
        for (int wI = tid; wI < 2 * NT + TX + TY; wI += NT) {
            int dir, fr, fc;   // face direction, face row, face column (tile-local)
            if (wI < NT) { dir = 0; fr = ty; fc = tx; }
            else if (wI < 2 * NT) { dir = 1; fr = ty; fc = tx; }
            else {
                int e = wI - 2 * NT;
                if (e < TY) { dir = 0; fr = e; fc = TX; } else { dir = 1; fr = TY; fc = e - TY; }
            }
            const int gi = i0 + fc, gj = j0 + fr;
            const bool valid = (dir == 0) ? (gi <= A.ni && gj < A.nj) : (gj <= A.nj && gi < A.ni);
            if (!valid) continue;
            int pn, pt1, pt2;
            face_perm<DIMS>(dir, pn, pt1, pt2);
            // stencil L1 position (LDS): i-face fc sits between columns fc+1, fc+2
            const int step = (dir == 0) ? 1 : PX;
            const int base = (dir == 0) ? (fr + 2) * PX + fc : fr * PX + (fc + 2);
            PrimT<R> S[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int s = base + m * step;
                S[m].rho = sP[0 * PLANE + s];
                S[m].v0 = sP[(1 + pn) * PLANE + s];
                S[m].v1 = sP[(1 + pt1) * PLANE + s];
                S[m].v2 = sP[(1 + pt2) * PLANE + s];
                S[m].u = sP[4 * PLANE + s];
                S[m].a = sP[5 * PLANE + s];
            }
            Flux5T<R> F = face_flux_ideal<M, FLUX>(S[0], S[1], S[2], S[3], cast_w<R>(A.w[dir]), gR, ho, lim, clip);
            R* dst = (dir == 0) ? sFi : sFj;
            const int nface = (dir == 0) ? NFI : NFJ;
            const int fidx = (dir == 0) ? fr * (TX + 1) + fc : fr * TX + fc;
            dst[0 * nface + fidx] = F.mass;
            dst[(1 + pn) * nface + fidx] = F.mn;
            dst[(1 + pt1) * nface + fidx] = F.mt1;
            dst[(1 + pt2) * nface + fidx] = F.mt2;
            dst[4 * nface + fidx] = F.E;
        }
        // 3. k-face at the top of cell k from the register queue
        R Ft[5] = {R(0), R(0), R(0), R(0), R(0)};
        if constexpr (DIMS == 3) {
            if (own_ij) {
                Flux5T<R> F = face_flux_ideal<M, FLUX>(kface(q0), kface(q1), kface(q2), kface(q3), cast_w<R>(A.w[2]),
                                                       gR, ho, lim, clip);
                Ft[0] = F.mass; Ft[1] = F.mt1; Ft[2] = F.mt2; Ft[3] = F.mn; Ft[4] = F.E;
            }
        }
        __syncthreads();
        // 4. flux divergence in face order N, E, S, W, T, B (fvcell.d:789-825),
        // stage update, decode check, write; all in FP64.
        if (own_ij) {
            const int iW = ty * (TX + 1) + tx, iE = iW + 1;
            const int jS = ty * TX + tx, jN = jS + TX;
            double integ[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
            auto add_face = [&](const R* Fsrc, int nface, int fidx, double outsign, double Aface) {
                double area = outsign * Aface;
                integ[0] -= double(Fsrc[0 * nface + fidx]) * area;
                integ[1] -= double(Fsrc[1 * nface + fidx]) * area;
                integ[2] -= double(Fsrc[2 * nface + fidx]) * area;
                if (DIMS == 3) integ[3] -= double(Fsrc[3 * nface + fidx]) * area;
                integ[4] -= double(Fsrc[4 * nface + fidx]) * area;
            };
            add_face(sFj, NFJ, jN, 1.0, A.area[1]);
            add_face(sFi, NFI, iE, 1.0, A.area[0]);
            add_face(sFj, NFJ, jS, -1.0, A.area[1]);
            add_face(sFi, NFI, iW, -1.0, A.area[0]);
            if constexpr (DIMS == 3) {
                add_face(Ft, 1, 0, 1.0, A.area[2]);
                add_face(Fb, 1, 0, -1.0, A.area[2]);
            }
            const long c = cell_index(A, i, j, k);
            const long qs = A.qs;
            double Un[5];
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                double d = A.vol_inv * integ[q] + 0.0;
                double dd0 = (A.stage >= 2) ? A.d0[c + q * qs] : 0.0;
                double dd1 = (A.stage >= 3) ? A.d1[c + q * qs] : 0.0;
                if (A.dout) A.dout[c + q * qs] = d;
                Un[q] = stage_update(A, A.Uold[c + q * qs], dd0, dd1, d);
            }
            // decode_conserved checks and the low-T fallback (fvcell.d:570-589, 680-697)
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
                    u = A.gas.Cv * T;
                    Un[1] = rho * vx; Un[2] = rho * vy; Un[3] = rho * vz;
                    Un[4] = rho * (u + ke);
                } else {
                    T = u * A.gas.Cvinv;
                }
                double p = rho * A.gas.Rgas * T;
                // check_data (flowstate.d:345-388, gas_state.d:199-241)
                if (!isfinite(rho)) invalid = true;
                if (!isfinite(T) || T < 1.01 * 1.0) invalid = true;
                if (!isfinite(p)) invalid = true;
                if (fabs(vx) > 30000.0 || fabs(vy) > 30000.0 || fabs(vz) > 30000.0) invalid = true;
                if (T > 50000.0) invalid = true;
            }
#pragma unroll
            for (int q = 0; q < 5; ++q) A.Unew[c + q * qs] = Un[q];
            unsigned int nbad = wave_count(invalid), nfail = wave_count(decode_fail);
            if ((nbad | nfail) && (__lane_id() == __ffsll((unsigned long long)__ballot(true)) - 1)) {
                if (nbad) atomicAdd(&A.counters[0], nbad);
                if (nfail) atomicAdd(&A.counters[1], nfail);
            }
        }
        if constexpr (DIMS == 3) {
#pragma unroll
            for (int q = 0; q < 5; ++q) Fb[q] = Ft[q];
            q0 = q1; q1 = q2; q2 = q3;
            if (k + 1 < k1) q3 = fetch<M>(A, i, j, k + 3);
        }
    }
}

} // namespace e4
