// stage_fast.hip -- stage-kernel instantiations for ModeFast64, ModeMixed.
// Built with -ffp-contract=fast.
#include "stage_kernels.h"
#include "tp_kernels.h"

namespace e4 {

void launch_stage_fast64(const StageArgs& A, dim3 grid, hipStream_t stream, int dims, int flux)
{
    launch_stage_mode<ModeFast64>(A, grid, stream, dims, flux);
}

void launch_stage_mixed(const StageArgs& A, dim3 grid, hipStream_t stream, int dims, int flux)
{
    launch_stage_mode<ModeMixed>(A, grid, stream, dims, flux);
}

void launch_stage_batch_fast64(const BatchArgs& B, int nwg, bool def, hipStream_t stream, int dims, int flux)
{
    launch_stage_batch_mode<ModeFast64>(B, nwg, def, stream, dims, flux);
}

void launch_stage_batch_mixed(const BatchArgs& B, int nwg, bool def, hipStream_t stream, int dims, int flux)
{
    launch_stage_batch_mode<ModeMixed>(B, nwg, def, stream, dims, flux);
}

template <class M>
static void launch_tp_stage_mode(const TPArgs& A, hipStream_t stream, int flux)
{
    const int nfk = (A.dims == 3) ? A.nk : 1;
    const long nf0 = (long)(A.ni + 1) * A.nj * nfk, nf1 = (long)A.ni * (A.nj + 1) * nfk;
    const long nf2 = (long)A.ni * A.nj * (A.nk + 1), nc = (long)A.ni * A.nj * nfk;
    const dim3 b(128);
    auto g = [](long n) { return dim3((unsigned)((n + 127) / 128)); };
    if (flux == FLUX_HLLC) {
        hipLaunchKernelGGL((tp_face_kernel<M, 5, FLUX_HLLC, 0>), g(nf0), b, 0, stream, A);
        hipLaunchKernelGGL((tp_face_kernel<M, 5, FLUX_HLLC, 1>), g(nf1), b, 0, stream, A);
        if (A.dims == 3) hipLaunchKernelGGL((tp_face_kernel<M, 5, FLUX_HLLC, 2>), g(nf2), b, 0, stream, A);
    } else {
        hipLaunchKernelGGL((tp_face_kernel<M, 5, FLUX_AUSMDV, 0>), g(nf0), b, 0, stream, A);
        hipLaunchKernelGGL((tp_face_kernel<M, 5, FLUX_AUSMDV, 1>), g(nf1), b, 0, stream, A);
        if (A.dims == 3) hipLaunchKernelGGL((tp_face_kernel<M, 5, FLUX_AUSMDV, 2>), g(nf2), b, 0, stream, A);
    }
    if (A.guess) hipLaunchKernelGGL((tp_update_kernel<M, 5, true>), g(nc), b, 0, stream, A);
    else hipLaunchKernelGGL((tp_update_kernel<M, 5, false>), g(nc), b, 0, stream, A);
}

void launch_tp_stage_fast64(const TPArgs& A, hipStream_t stream, int flux)
{
    launch_tp_stage_mode<ModeFast64>(A, stream, flux);
}

void launch_tp_stage_mixed(const TPArgs& A, hipStream_t stream, int flux)
{
    launch_tp_stage_mode<ModeMixed>(A, stream, flux);
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
    __shared__ __bf16 lds[2 * GM_LDS_PER_WAVE];
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
