// stage_exact.hip -- stage-kernel instantiations for ModeExact.
// Built with -ffp-contract=off (reference roundings).
#include "stage_kernels.h"
#include "tp_kernels.h"

namespace e4 {

void launch_stage_exact(const StageArgs& A, dim3 grid, hipStream_t stream, int dims, int flux)
{
    launch_stage_mode<ModeExact>(A, grid, stream, dims, flux);
}

void launch_stage_batch_exact(const BatchArgs& B, int nwg, bool def, hipStream_t stream, int dims, int flux)
{
    launch_stage_batch_mode<ModeExact>(B, nwg, def, stream, dims, flux);
}

void launch_tp_stage_exact(const TPArgs& A, hipStream_t stream, int flux)
{
    const int nfk = (A.dims == 3) ? A.nk : 1;
    const long nf0 = (long)(A.ni + 1) * A.nj * nfk, nf1 = (long)A.ni * (A.nj + 1) * nfk;
    const long nf2 = (long)A.ni * A.nj * (A.nk + 1), nc = (long)A.ni * A.nj * nfk;
    const dim3 b(128);
    auto g = [](long n) { return dim3((unsigned)((n + 127) / 128)); };
    if (flux == FLUX_HLLC) {
        hipLaunchKernelGGL((tp_face_kernel<ModeExact, 5, FLUX_HLLC, 0>), g(nf0), b, 0, stream, A);
        hipLaunchKernelGGL((tp_face_kernel<ModeExact, 5, FLUX_HLLC, 1>), g(nf1), b, 0, stream, A);
        if (A.dims == 3) hipLaunchKernelGGL((tp_face_kernel<ModeExact, 5, FLUX_HLLC, 2>), g(nf2), b, 0, stream, A);
    } else {
        hipLaunchKernelGGL((tp_face_kernel<ModeExact, 5, FLUX_AUSMDV, 0>), g(nf0), b, 0, stream, A);
        hipLaunchKernelGGL((tp_face_kernel<ModeExact, 5, FLUX_AUSMDV, 1>), g(nf1), b, 0, stream, A);
        if (A.dims == 3) hipLaunchKernelGGL((tp_face_kernel<ModeExact, 5, FLUX_AUSMDV, 2>), g(nf2), b, 0, stream, A);
    }
    if (A.guess) hipLaunchKernelGGL((tp_update_kernel<ModeExact, 5, true>), g(nc), b, 0, stream, A);
    else hipLaunchKernelGGL((tp_update_kernel<ModeExact, 5, false>), g(nc), b, 0, stream, A);
}

void launch_tp_decode_exact(const TPArgs& A, const double* U, const double* Tguess, hipStream_t stream)
{
    const long nc = (long)A.ni * A.nj * ((A.dims == 3) ? A.nk : 1);
    hipLaunchKernelGGL((tp_decode_kernel<ModeExact, 5>), dim3((unsigned)((nc + 127) / 128)), dim3(128), 0, stream, A, U,
                       Tguess);
}

} // namespace e4
