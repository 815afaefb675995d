// stage_fast.hip -- stage-kernel instantiations for ModeFast64, ModeMixed.
// Built with -ffp-contract=fast.
#include "stage_kernels.h"

namespace e4 {

void launch_stage_fast64(const StageArgs& A, dim3 grid, hipStream_t stream, int dims, int flux)
{
    const dim3 block(TILE_X * TILE_Y);
    if (dims == 3) {
        if (flux == FLUX_HLLC) hipLaunchKernelGGL((stage_kernel<ModeFast64, TILE_X, TILE_Y, 3, FLUX_HLLC>), grid, block, 0, stream, A);
        else hipLaunchKernelGGL((stage_kernel<ModeFast64, TILE_X, TILE_Y, 3, FLUX_AUSMDV>), grid, block, 0, stream, A);
    } else {
        if (flux == FLUX_HLLC) hipLaunchKernelGGL((stage_kernel<ModeFast64, TILE_X, TILE_Y, 2, FLUX_HLLC>), grid, block, 0, stream, A);
        else hipLaunchKernelGGL((stage_kernel<ModeFast64, TILE_X, TILE_Y, 2, FLUX_AUSMDV>), grid, block, 0, stream, A);
    }
}

void launch_stage_mixed(const StageArgs& A, dim3 grid, hipStream_t stream, int dims, int flux)
{
    const dim3 block(TILE_X * TILE_Y);
    if (dims == 3) {
        if (flux == FLUX_HLLC) hipLaunchKernelGGL((stage_kernel<ModeMixed, TILE_X, TILE_Y, 3, FLUX_HLLC>), grid, block, 0, stream, A);
        else hipLaunchKernelGGL((stage_kernel<ModeMixed, TILE_X, TILE_Y, 3, FLUX_AUSMDV>), grid, block, 0, stream, A);
    } else {
        if (flux == FLUX_HLLC) hipLaunchKernelGGL((stage_kernel<ModeMixed, TILE_X, TILE_Y, 2, FLUX_HLLC>), grid, block, 0, stream, A);
        else hipLaunchKernelGGL((stage_kernel<ModeMixed, TILE_X, TILE_Y, 2, FLUX_AUSMDV>), grid, block, 0, stream, A);
    }
}

} // namespace e4
