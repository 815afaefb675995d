// stage_mixed.hip -- stage-kernel instantiations for ModeMixed (FP32 reconstruction and flux).
// Built with -ffp-contract=fast.
#include "stage_kernels.h"

namespace e4 {

void launch_stage_mixed(const StageArgs& A, dim3 grid, hipStream_t stream, int dims, int flux)
{
    launch_stage_mode<ModeMixed>(A, grid, stream, dims, flux);
}

void launch_stage_batch_mixed(const BatchArgs& B, int nwg, bool def, hipStream_t stream, int dims, int flux)
{
    launch_stage_batch_mode<ModeMixed>(B, nwg, def, stream, dims, flux);
}

void launch_boundary_cells_batch_mixed(const BatchArgs& B, int nwg, hipStream_t stream, int dims, int flux)
{
    launch_boundary_cells_batch_mode<ModeMixed>(B, nwg, stream, dims, flux);
}

void launch_boundary_cells_mixed(const StageArgs& A, hipStream_t stream, int dims, int flux)
{
    launch_boundary_cells_mode<ModeMixed>(A, stream, dims, flux);
}

} // namespace e4
