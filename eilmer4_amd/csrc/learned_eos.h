// learned_eos.h -- the learned initial guess of the thermally perfect
// temperature solve: a 7 -> 64 -> 64 -> 64 -> 1 SiLU network with BF16 weights,
// evaluated for the 64 cells of a wave with v_mfma_f32_16x16x32_bf16 (FP32
// accumulation), followed by an FP64 Newton polish judged by its energy
// residual.
//
// Operand layout of v_mfma_f32_16x16x32_bf16 (wave64): lane l supplies
//   A (16 x 32, rows x k): row l%16, k = 8*(l/16) .. 8*(l/16)+7
//   B (32 x 16, k x cols): col l%16, k = 8*(l/16) .. 8*(l/16)+7
//   C (16 x 16):           col l%16, rows 4*(l/16) .. 4*(l/16)+3
// A wave's 64 cells are four 16-row tiles; hidden layers of 64 units are four
// 16-column tiles and two 32-wide k-chunks.  Activations pass between layers
// through LDS as BF16 rows (one row per cell).
#pragma once
#include <hip/hip_runtime.h>
#include "tp_gas_device.h"

namespace e4 {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int GM_NH = 64;
constexpr int GM_K1 = 32;            // input features padded to one k-chunk
constexpr int GM_RS = 72;            // LDS row stride (bf16) of the activation buffers
constexpr int GM_LDS_PER_WAVE = 2 * 64 * GM_RS;   // two activation buffers, bf16 elements

struct GuessModel {
    float in_mean[8], in_std[8];     // feature normalisation (rho, u, Y0..Y4)
    float T_mean, T_std;             // output scaling
    float b1[GM_NH], b2[GM_NH], b3[GM_NH], w4[GM_NH], b4;
    const __bf16* W1;                // [64][32] (features 0..6, zero padded)
    const __bf16* W2;                // [64][64]
    const __bf16* W3;                // [64][64]
};

__device__ __forceinline__ float gm_silu(float x) { return x * __builtin_amdgcn_rcpf(1.0f + __expf(-x)); }

// One dense layer of the wave: out = silu(in . W^T + b) for 64 rows.
template <int K>
__device__ __forceinline__ void gm_layer(const __bf16* __restrict__ in, __bf16* __restrict__ out,
                                         const __bf16* __restrict__ W, const float* __restrict__ b, int lane)
{
    const int r16 = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kc = 0; kc < K / 32; ++kc) {
                bf16x8 a = *reinterpret_cast<const bf16x8*>(in + (rt * 16 + r16) * GM_RS + kc * 32 + 8 * kq);
                bf16x8 bw = *reinterpret_cast<const bf16x8*>(W + (nt * 16 + r16) * K + kc * 32 + 8 * kq);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw, acc, 0, 0, 0);
            }
            const int col = nt * 16 + r16;
            const float bias = b[col];
#pragma unroll
            for (int q = 0; q < 4; ++q) out[(rt * 16 + 4 * kq + q) * GM_RS + col] = (__bf16)gm_silu(acc[q] + bias);
        }
    }
}

// Temperature guesses for the 64 cells of the calling wave.  Every lane of the
// workgroup must call it (it synchronises the workgroup); lds points at this
// wave's GM_LDS_PER_WAVE bf16 elements.  x = (rho, u, Y0..Y4) of the lane's cell.
__device__ __forceinline__ float mlp_guess_wave(const GuessModel& G, const double* x, __bf16* lds)
{
    const int lane = __lane_id();
    __bf16* bufA = lds;
    __bf16* bufB = lds + 64 * GM_RS;
#pragma unroll
    for (int f = 0; f < GM_K1; ++f) {
        float z = (f < 7) ? ((float)x[f] - G.in_mean[f]) * G.in_std[f] : 0.0f;
        bufA[lane * GM_RS + f] = (__bf16)z;
    }
    __syncthreads();
    gm_layer<GM_K1>(bufA, bufB, G.W1, G.b1, lane);
    __syncthreads();
    gm_layer<GM_NH>(bufB, bufA, G.W2, G.b2, lane);
    __syncthreads();
    gm_layer<GM_NH>(bufA, bufB, G.W3, G.b3, lane);
    __syncthreads();
    float y = G.b4;
#pragma unroll 8
    for (int n = 0; n < GM_NH; ++n) y += G.w4[n] * (float)bufB[lane * GM_RS + n];
    __syncthreads();
    return G.T_mean + G.T_std * y;
}

// FP64 Newton polish from the learned guess: two Newton steps on
// f(T) = e_tgt - u(T) without bracketing, then the judge: the energy residual
// at the result, divided by the last derivative, estimates the remaining
// temperature error; the polish is accepted when that estimate is at most
// 1e-11 T.  (Inside the curves' blend ranges the derivative sum_s Y_s (Cp_s -
// R_s) is not the exact derivative of the blended energy, so convergence there
// is only linear and the size of the last step alone would not bound the
// error.)  Q.u is left at the energy of the last evaluated temperature, which
// here is the result.  Returns 0 when accepted, -1 to fall back to the
// reference's bracketed solve.
template <class M, int NSP>
__device__ int tp_polish(const TPGas& g, TPState<NSP>& Q, double e_tgt, double T0, int* fevals)
{
    using MD = typename std::conditional<M::kFast, ModeFast64, ModeExact>::type;
    if (!(T0 >= 20.0 && T0 <= 100000.0)) return -1;
    // energy and its T derivative at one temperature (one mixture evaluation
    // in the fast modes)
    auto u_dzdT = [&](double T, double& u, double& dz) {
        if constexpr (M::kFast) {
            if (g.shared) {
                double cv;
                tp_mix_u_cv<NSP>(g, Q.Y, tp_Rmix<NSP>(g, Q.Y), T, &u, &cv);
                dz = -1.0 * cv;
                return;
            }
        }
        u = tp_energy<M, NSP>(g, Q.Y, T);
        dz = tp_dzdT<M, NSP>(g, Q.Y, T);
    };
    double T = T0, df = 1.0;
    for (int it = 0; it < 2; ++it) {
        double u;
        u_dzdT(T, u, df);
        (*fevals)++;
        double f = e_tgt - u;
        T -= dv<MD>(f, df);
    }
    if (!(T >= 20.0 && T <= 100000.0)) return -1;
    double u = tp_energy<M, NSP>(g, Q.Y, T);
    (*fevals)++;
    double err = fabs(dv<MD>(e_tgt - u, df));
    if (!(err <= 1.0e-11 * T)) return -1;
    Q.T = T;
    Q.u = u;
    return 0;
}

} // namespace e4
