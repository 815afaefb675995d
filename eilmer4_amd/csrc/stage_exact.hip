// stage_exact.hip -- stage-kernel instantiations for ModeExact.
// Built with -ffp-contract=off (reference roundings).
#include "stage_kernels.h"

namespace e4 {

void launch_stage_exact(const StageArgs& A, dim3 grid, hipStream_t stream, int dims, int flux)
{
    const dim3 block(TILE_X * TILE_Y);
    if (dims == 3) {
        if (flux == FLUX_HLLC) hipLaunchKernelGGL((stage_kernel<ModeExact, TILE_X, TILE_Y, 3, FLUX_HLLC>), grid, block, 0, stream, A);
        else hipLaunchKernelGGL((stage_kernel<ModeExact, TILE_X, TILE_Y, 3, FLUX_AUSMDV>), grid, block, 0, stream, A);
    } else {
        if (flux == FLUX_HLLC) hipLaunchKernelGGL((stage_kernel<ModeExact, TILE_X, TILE_Y, 2, FLUX_HLLC>), grid, block, 0, stream, A);
        else hipLaunchKernelGGL((stage_kernel<ModeExact, TILE_X, TILE_Y, 2, FLUX_AUSMDV>), grid, block, 0, stream, A);
    }
}

} // namespace e4
