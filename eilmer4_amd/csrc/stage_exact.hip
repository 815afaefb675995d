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

void launch_extrapolate_ghosts_batch_exact(const BatchArgs& B, int nwg, hipStream_t stream)
{
    launch_extrapolate_ghosts_batch_mode<ModeExact>(B, nwg, stream);
}

void launch_extrapolate_ghosts_exact(const StageArgs& A, int faces, hipStream_t stream)
{
    launch_extrapolate_ghosts_mode<ModeExact>(A, faces, stream);
}

void launch_boundary_cells_batch_exact(const BatchArgs& B, int nwg, hipStream_t stream, int dims, int flux)
{
    launch_boundary_cells_batch_mode<ModeExact>(B, nwg, stream, dims, flux);
}

void launch_boundary_cells_exact(const StageArgs& A, hipStream_t stream, int dims, int flux)
{
    launch_boundary_cells_mode<ModeExact>(A, stream, dims, flux);
}

void launch_tp_stage_exact(const TPArgs& A, hipStream_t stream, int flux)
{
    launch_tp_stage_m<ModeExact>(A, stream, flux);
}

void launch_shock_detect_exact(const StageArgs& A, int dims, hipStream_t stream)
{
    launch_shock_detect_mode<ModeExact>(A, dims, stream);
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
