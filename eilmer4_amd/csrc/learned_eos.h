// learned_eos.h -- the learned initial guess of the thermally perfect
// temperature solve: a 7 -> 64 -> 64 -> 64 -> 1 SiLU network with BF16 weights,
// evaluated for the 64 cells of a wave with v_mfma_f32_16x16x32_bf16 (FP32
// accumulation), followed by an FP64 Newton polish accepted by its energy
// residual.
//
// Operand layout of v_mfma_f32_16x16x32_bf16 (wave64): lane l supplies
//   A (16 x 32, rows x k): row l%16, k = 8*(l/16) .. 8*(l/16)+7
//   B (32 x 16, k x cols): col l%16, k = 8*(l/16) .. 8*(l/16)+7
//   C (16 x 16):           col l%16, rows 4*(l/16) .. 4*(l/16)+3
// A wave's 64 cells are four 16-row tiles; hidden layers of 64 units are four
// 16-column tiles and two 32-wide k-chunks.  Activations pass between layers
// through LDS as BF16 rows (one row per cell), in one buffer per wave that each
// layer overwrites in place: a layer loads all its A operands into registers
// before it stores any output (a wave's LDS operations run in program order),
// which halves the LDS of the kernels that evaluate the network.
#pragma once
#include <hip/hip_runtime.h>
#include "tp_gas_device.h"

namespace e4 {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int GM_NH = 64;
constexpr int GM_K1 = 32;            // input features padded to one k-chunk
constexpr int GM_RS = 72;            // LDS row stride (bf16) of the activation buffers
constexpr int GM_LDS_PER_WAVE = 64 * GM_RS;       // one activation buffer, bf16 elements

struct GuessModel {
    float in_mean[8], in_std[8];     // feature normalisation (rho, u, Y0..Y4)
    float T_mean, T_std;             // output scaling
    float b1[GM_NH], b2[GM_NH], b3[GM_NH], w4[GM_NH], b4;
    const __bf16* W1;                // [64][32] (features 0..6, zero padded)
    const __bf16* W2;                // [64][64]
    const __bf16* W3;                // [64][64]
};

__device__ __forceinline__ float gm_silu(float x) { return x * __builtin_amdgcn_rcpf(1.0f + __expf(-x)); }

// One dense layer of the wave, in place: buf = silu(buf . W^T + b) for 64 rows
// (K input columns, 64 output columns).  buf is not restrict-qualified: the
// stores must stay after the loads.
template <int K>
__device__ __forceinline__ void gm_layer(__bf16* buf, const __bf16* __restrict__ W, const float* __restrict__ b,
                                         int lane)
{
    const int r16 = lane & 15, kq = lane >> 4;
    bf16x8 a[4][K / 32];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int kc = 0; kc < K / 32; ++kc)
            a[rt][kc] = *reinterpret_cast<const bf16x8*>(buf + (rt * 16 + r16) * GM_RS + kc * 32 + 8 * kq);
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kc = 0; kc < K / 32; ++kc) {
                bf16x8 bw = *reinterpret_cast<const bf16x8*>(W + (nt * 16 + r16) * K + kc * 32 + 8 * kq);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[rt][kc], bw, acc, 0, 0, 0);
            }
            const int col = nt * 16 + r16;
            const float bias = b[col];
#pragma unroll
            for (int q = 0; q < 4; ++q) buf[(rt * 16 + 4 * kq + q) * GM_RS + col] = (__bf16)gm_silu(acc[q] + bias);
        }
    }
}

// Temperature guesses for the 64 cells of the calling wave.  Every lane of the
// workgroup must call it (it synchronises the workgroup); lds points at this
// wave's GM_LDS_PER_WAVE bf16 elements.  x = (rho, u, Y0..Y4) of the lane's cell.
__device__ __forceinline__ float mlp_guess_wave(const GuessModel& G, const double* x, __bf16* lds)
{
    const int lane = __lane_id();
    __bf16* buf = lds;
#pragma unroll
    for (int f = 0; f < GM_K1; ++f) {
        float z = (f < 7) ? ((float)x[f] - G.in_mean[f]) * G.in_std[f] : 0.0f;
        buf[lane * GM_RS + f] = (__bf16)z;
    }
    __syncthreads();
    gm_layer<GM_K1>(buf, G.W1, G.b1, lane);
    __syncthreads();
    gm_layer<GM_NH>(buf, G.W2, G.b2, lane);
    __syncthreads();
    gm_layer<GM_NH>(buf, G.W3, G.b3, lane);
    __syncthreads();
    float y = G.b4;
#pragma unroll 8
    for (int n = 0; n < GM_NH; ++n) y += G.w4[n] * (float)buf[lane * GM_RS + n];
    __syncthreads();
    return G.T_mean + G.T_std * y;
}

// FP64 polish from a temperature guess (the learned guess, or the stored or
// the face's cell temperature): two iterations without bracketing, then the
// acceptance test: the energy residual at the result, divided by the
// derivative there, estimates the remaining temperature error, and the polish
// is accepted when that estimate is at most 1e-11 T.  Q.u is left at the
// energy of the accepted temperature (at the target energy after an early
// acceptance, below).  Returns 0 when accepted, -1 to fall back to the reference's
// bracketed solve.
//
// Fast modes (which require curves with shared breakpoints): Halley steps
// (cubic convergence) on the exact u, du/dT and d2u/dT2 of the blended
// curves, the species sums of every segment formed once per solve.  From a face's cell
// temperature (typically tens to a few hundred kelvin from the face root) two
// Halley steps reach the acceptance bound; Newton would need three.  A step s
// with |s| <= 1e-4 T that stays on one smooth piece of the curve is accepted
// without evaluating its result: the cubic convergence leaves the new
// temperature within about |s|^3 / T^2 <= 1e-12 T of the root, inside the
// bound the residual test accepts.  ModeExact:
// Newton steps with the reference's dzdT, sum_s Y_s (Cp_s - R_s), which inside a blend is not the
// derivative of the blended energy (convergence there is only linear, hence a
// test on the residual rather than on the size of the last step).
template <class M, int NSP, class MX>
__device__ __forceinline__ int tp_polish_x(const TPGas& g, const MX& X, double Rmix, TPState<NSP>& Q, double e_tgt,
                                           double T0, int* fevals)
{
    using MD = typename std::conditional<M::kFast, ModeFast64, ModeExact>::type;
    if (!(T0 >= 20.0 && T0 <= 100000.0)) return -1;
    if constexpr (M::kFast) {
        double T = T0;
#pragma nounroll
        for (int it = 0;; ++it) {
            if (it == 2 && !(T >= 20.0 && T <= 100000.0)) return -1;
            double u, du, d2u;
            tp_mix_eval<true>(X, g.curve[0], Rmix, T, &u, &du, &d2u);
            (*fevals)++;
            const double f = u - e_tgt;
            if (it == 2) {
                if (!(fabs(f) <= 1.0e-11 * T * fabs(du))) return -1;
                Q.T = T;
                Q.u = u;
                return 0;
            }
            const double s = dv<MD>(2.0 * f * du, 2.0 * du * du - f * d2u);
            const double Tn = T - s;
            if (fabs(s) <= 1.0e-4 * Tn && Tn >= 20.0 && Tn <= 100000.0 &&
                tp_curve_piece(g.curve[0], Tn) == tp_curve_piece(g.curve[0], T)) {
                Q.T = Tn;
                Q.u = e_tgt;
                return 0;
            }
            T = Tn;
        }
    } else {
        // two Newton steps, then the energy at the result (one evaluation site)
        double T = T0, df = 1.0;
#pragma nounroll
        for (int it = 0;; ++it) {
            if (it == 2 && !(T >= 20.0 && T <= 100000.0)) return -1;
            double u, dz;
            tp_u_dzdT<M, NSP>(g, Q.Y, T, it < 2, &u, &dz);
            (*fevals)++;
            if (it == 2) {
                if (!(fabs(e_tgt - u) <= 1.0e-11 * T * fabs(df))) return -1;
                Q.T = T;
                Q.u = u;
                return 0;
            }
            df = dz;
            T -= dv<MD>(e_tgt - u, df);
        }
    }
}

template <class M, int NSP>
__device__ __forceinline__ int tp_polish(const TPGas& g, TPState<NSP>& Q, double e_tgt, double T0, int* fevals)
{
    if constexpr (M::kFast) {
        const TPMixY<NSP> X = tp_mix_y<NSP>(g, Q.Y);
        return tp_polish_x<M, NSP>(g, X, tp_Rmix<NSP>(g, Q.Y), Q, e_tgt, T0, fevals);
    } else {
        TPMix X;   // unused by the reference form
        return tp_polish_x<M, NSP>(g, X, 0.0, Q, e_tgt, T0, fevals);
    }
}

} // namespace e4
