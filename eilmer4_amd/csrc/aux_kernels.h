// aux_kernels.h -- CFL signal reduction, decode/validation and counter reset
// kernels (host API translation unit only).
#pragma once
#include "stage_args.h"

namespace e4 {

__device__ __forceinline__ unsigned int wave_count_aux(bool pred)
{
    return (unsigned int)__popcll(__ballot(pred));
}

// CFL signal (fvcell.d:1624-1661; stringent: the fastest normal speed over
// the shortest cell length) and its maximum over the block; the maximum of a
// non-negative double is taken on its bit pattern.
struct CflArgs {
    const double* __restrict__ U;
    long qs;
    int ni, nj, nk, dims, stringent;
    double iLength, jLength, kLength;
    IdealGas gas;
    unsigned long long* smax_bits;
    double* last_signal;
};

template <class M>
__global__ __launch_bounds__(256) void cfl_kernel(CflArgs A)
{
    __shared__ double wave_max[4];
    const long ncells = (long)A.ni * A.nj * A.nk;
    double signal = 0.0;
    // grid-stride loop: a few thousand workgroups, one atomic each
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += (long)gridDim.x * blockDim.x) {
        Prim P = decode_ideal<M>(A.U[c], A.U[c + A.qs], A.U[c + 2 * A.qs], A.U[c + 3 * A.qs], A.U[c + 4 * A.qs], A.gas);
        double s = 0.0;
        double un_N = fabs(P.v1);
        double un_E = fabs(P.v0);
        double un_T = (A.dims == 3) ? fabs(P.v2) : 0.0;
        if (A.stringent) {
            double un_max = fmax(un_N, un_E);
            double minLength = fmin(A.iLength, A.jLength);
            if (A.dims == 3) {
                un_max = fmax(un_max, un_T);
                minLength = fmin(minLength, A.kLength);
            }
            s = (un_max + P.a) / minLength;
        } else {
            double signalN = (un_N + P.a) / A.jLength;
            s = fmax(s, signalN);
            double signalE = (un_E + P.a) / A.iLength;
            s = fmax(s, signalE);
            if (A.dims == 3) {
                double signalT = (un_T + P.a) / A.kLength;
                s = fmax(s, signalT);
            }
        }
        if (c == ncells - 1) *A.last_signal = s;
        signal = fmax(signal, s);
    }
    for (int off = 32; off > 0; off >>= 1) signal = fmax(signal, __shfl_xor(signal, off));
    if (__lane_id() == 0) wave_max[threadIdx.x / 64] = signal;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = fmax(fmax(wave_max[0], wave_max[1]), fmax(wave_max[2], wave_max[3]));
        atomicMax(A.smax_bits, (unsigned long long)__double_as_longlong(m));
    }
}

// Decode every active cell into flow-state arrays (rho,p,T,u,a,vx,vy,vz) and
// count cells that would fail decode or check_data.
struct DecodeArgs {
    const double* __restrict__ U;
    double* fs;       // may be null
    long qs, ncells;
    IdealGas gas;
    unsigned int* counters;
};

__global__ __launch_bounds__(256) void decode_kernel(DecodeArgs A)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    if (c < A.ncells) {
        double rho = A.U[c];
        if (!(rho > 0.0)) {
            bad = true;
        } else {
            Prim P = decode_ideal<ModeExact>(rho, A.U[c + A.qs], A.U[c + 2 * A.qs], A.U[c + 3 * A.qs], A.U[c + 4 * A.qs], A.gas);
            double T = P.u * A.gas.Cvinv;
            double p = P.rho * A.gas.Rgas * T;
            if (!(P.u > 0.0) || !isfinite(T) || T < 1.01 || T > 50000.0 || !isfinite(p) || !isfinite(P.a) ||
                fabs(P.v0) > 30000.0 || fabs(P.v1) > 30000.0 || fabs(P.v2) > 30000.0)
                bad = true;
            if (A.fs) {
                const long n = A.ncells;
                A.fs[c] = P.rho; A.fs[n + c] = p; A.fs[2 * n + c] = T; A.fs[3 * n + c] = P.u;
                A.fs[4 * n + c] = P.a; A.fs[5 * n + c] = P.v0; A.fs[6 * n + c] = P.v1; A.fs[7 * n + c] = P.v2;
            }
        }
    }
    unsigned int nb = wave_count_aux(bad);
    if (nb && __lane_id() == 0) atomicAdd(&A.counters[0], nb);
}

__global__ void reset_counters_kernel(unsigned int* counters, int n, unsigned long long* smax)
{
    for (int t = threadIdx.x; t < n; t += blockDim.x) counters[t] = 0u;
    if (threadIdx.x == 0 && smax) *smax = 0ull;
}


// CFL signal from the primitive array (fvcell.d:1624-1661).
__global__ __launch_bounds__(256) void tp_cfl_kernel(const double* __restrict__ P, long qs, int ni, int nj, int nk,
                                                     int dims, int stringent, double iL, double jL, double kL,
                                                     unsigned long long* smax_bits, double* last_signal)
{
    __shared__ double wave_max[4];
    const long ncells = (long)ni * nj * nk;
    double sig = 0.0;
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += (long)gridDim.x * blockDim.x) {
        double a = P[P_A * qs + c];
        double s = 0.0;
        if (stringent) {
            double un_max = fmax(fabs(P[P_VY * qs + c]), fabs(P[P_VX * qs + c]));
            double minLength = fmin(iL, jL);
            if (dims == 3) {
                un_max = fmax(un_max, fabs(P[P_VZ * qs + c]));
                minLength = fmin(minLength, kL);
            }
            s = (un_max + a) / minLength;
        } else {
            s = fmax(s, (fabs(P[P_VY * qs + c]) + a) / jL);
            s = fmax(s, (fabs(P[P_VX * qs + c]) + a) / iL);
            if (dims == 3) s = fmax(s, (fabs(P[P_VZ * qs + c]) + a) / kL);
        }
        if (c == ncells - 1) *last_signal = s;
        sig = fmax(sig, s);
    }
    for (int off = 32; off > 0; off >>= 1) sig = fmax(sig, __shfl_xor(sig, off));
    if (__lane_id() == 0) wave_max[threadIdx.x / 64] = sig;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = fmax(fmax(wave_max[0], wave_max[1]), fmax(wave_max[2], wave_max[3]));
        atomicMax(smax_bits, (unsigned long long)__double_as_longlong(m));
    }
}

} // namespace e4
