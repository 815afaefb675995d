// stage_exact.hip -- ideal-gas kernel instantiations for ModeExact.
// Built with -ffp-contract=off (reference roundings).
#include "stage_kernels.h"

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

void launch_shock_detect_exact(const StageArgs& A, int dims, hipStream_t stream)
{
    launch_shock_detect_mode<ModeExact>(A, dims, stream);
}

} // namespace e4
