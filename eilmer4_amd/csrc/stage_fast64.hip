// stage_fast64.hip -- stage-kernel instantiations for ModeFast64 (the plain FP64 form).
// Built with -ffp-contract=fast.
#include "stage_kernels.h"

namespace e4 {

void launch_stage_fast64(const StageArgs& A, dim3 grid, hipStream_t stream, int dims, int flux)
{
    launch_stage_mode<ModeFast64>(A, grid, stream, dims, flux);
}

void launch_stage_batch_fast64(const BatchArgs& B, int nwg, bool def, hipStream_t stream, int dims, int flux)
{
    launch_stage_batch_mode<ModeFast64>(B, nwg, def, stream, dims, flux);
}

void launch_extrapolate_ghosts_batch_fast64(const BatchArgs& B, int nwg, hipStream_t stream)
{
    launch_extrapolate_ghosts_batch_mode<ModeFast64>(B, nwg, stream);
}

void launch_extrapolate_ghosts_fast64(const StageArgs& A, int faces, hipStream_t stream)
{
    launch_extrapolate_ghosts_mode<ModeFast64>(A, faces, stream);
}

void launch_boundary_cells_batch_fast64(const BatchArgs& B, int nwg, hipStream_t stream, int dims, int flux)
{
    launch_boundary_cells_batch_mode<ModeFast64>(B, nwg, stream, dims, flux);
}

void launch_boundary_cells_fast64(const StageArgs& A, hipStream_t stream, int dims, int flux)
{
    launch_boundary_cells_mode<ModeFast64>(A, stream, dims, flux);
}

void launch_shock_detect_fast64(const StageArgs& A, int dims, hipStream_t stream)
{
    launch_shock_detect_mode<ModeFast64>(A, dims, stream);
}

} // namespace e4
