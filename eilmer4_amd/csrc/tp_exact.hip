// tp_exact.hip -- thermally perfect kernel instantiations for ModeExact.
// Built with -ffp-contract=off (reference roundings).
#include "tp_kernels.h"

namespace e4 {

void launch_tp_stage_exact(const TPArgs& A, hipStream_t stream, int flux)
{
    launch_tp_stage_m<ModeExact>(A, stream, flux);
}

void launch_tp_stage_batch_exact(const BatchArgs* B, const int* wg, int dims, bool guess, const TPArgs* per_block,
                                 int nper, int nblk, hipStream_t stream, int flux)
{
    launch_tp_stage_batch_m<ModeExact>(B, wg, dims, guess, per_block, nper, nblk, stream, flux);
}

void launch_tp_shock_detect_exact(const TPArgs& A, hipStream_t stream)
{
    launch_tp_shock_detect_m<ModeExact>(A, stream);
}

void launch_tp_extrap1_exact(const TPArgs& A, int faces, hipStream_t stream)
{
    launch_tp_extrap1_m<ModeExact>(A, faces, stream);
}

void launch_tp_decode_exact(const TPArgs& A, const double* U, const double* Tguess, hipStream_t stream)
{
    const long nc = (long)A.ni * A.nj * ((A.dims == 3) ? A.nk : 1);
    hipLaunchKernelGGL((tp_decode_kernel<ModeExact, 5>), dim3((unsigned)((nc + 127) / 128)), dim3(128), 0, stream, A, U,
                       Tguess);
}

} // namespace e4
