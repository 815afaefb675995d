// tp_fast.hip -- thermally perfect kernel instantiations for ModeFast64 and ModeMixed, and the
// learned-guess inference kernel.
// Built with -ffp-contract=fast.
#include "tp_kernels.h"

namespace e4 {

void launch_tp_stage_fast64(const TPArgs& A, hipStream_t stream, int flux)
{
    launch_tp_stage_m<ModeFast64>(A, stream, flux);
}

void launch_tp_stage_mixed(const TPArgs& A, hipStream_t stream, int flux)
{
    launch_tp_stage_m<ModeMixed>(A, stream, flux);
}

void launch_tp_stage_batch_fast64(const BatchArgs* B, const int* wg, int dims, bool guess, const TPArgs* per_block,
                                  int nper, int nblk, hipStream_t stream, int flux)
{
    launch_tp_stage_batch_m<ModeFast64>(B, wg, dims, guess, per_block, nper, nblk, stream, flux);
}

void launch_tp_stage_batch_mixed(const BatchArgs* B, const int* wg, int dims, bool guess, const TPArgs* per_block,
                                 int nper, int nblk, hipStream_t stream, int flux)
{
    launch_tp_stage_batch_m<ModeMixed>(B, wg, dims, guess, per_block, nper, nblk, stream, flux);
}

void launch_tp_shock_detect_fast64(const TPArgs& A, hipStream_t stream)
{
    launch_tp_shock_detect_m<ModeFast64>(A, stream);
}

void launch_tp_extrap1_fast64(const TPArgs& A, int faces, hipStream_t stream)
{
    launch_tp_extrap1_m<ModeFast64>(A, faces, stream);
}

void launch_tp_decode_fast64(const TPArgs& A, const double* U, const double* Tguess, hipStream_t stream)
{
    const long nc = (long)A.ni * A.nj * ((A.dims == 3) ? A.nk : 1);
    hipLaunchKernelGGL((tp_decode_kernel<ModeFast64, 5>), dim3((unsigned)((nc + 127) / 128)), dim3(128), 0, stream, A, U,
                       Tguess);
}

// Learned temperature guesses for n states x[n][7] = (rho, u, Y0..Y4): the
// inference the decode runs, exposed for evaluation.
__global__ __launch_bounds__(128) void mlp_guess_kernel(const GuessModel* G, const double* __restrict__ x, int n,
                                                        double* __restrict__ T)
{
    __shared__ __attribute__((aligned(16))) __bf16 lds[2 * GM_LDS_PER_WAVE];
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double xi[7] = {1.0, 1.0e6, 0.2, 0.2, 0.2, 0.2, 0.2};
    if (c < n)
        for (int f = 0; f < 7; ++f) xi[f] = x[c * 7 + f];
    float t = mlp_guess_wave(*G, xi, lds + (threadIdx.x / 64) * GM_LDS_PER_WAVE);
    if (c < n) T[c] = (double)t;
}

void launch_mlp_guess(const GuessModel* G, const double* x, int n, double* T, hipStream_t stream)
{
    hipLaunchKernelGGL(mlp_guess_kernel, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, stream, G, x, n, T);
}

} // namespace e4
