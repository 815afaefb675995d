// gen_fast.hip -- kernels of blocks built from their grid, ModeFast64 and
// ModeMixed instantiations.  Built with -ffp-contract=fast.
#include "gen_kernels.h"

namespace e4 {

void launch_gen_stage_fast64(const TPArgs& A, hipStream_t stream, int flux, int nsp)
{
    if (nsp == 0) launch_gen_stage_m<ModeFast64, 0>(A, stream, flux);
    else launch_gen_stage_m<ModeFast64, 5>(A, stream, flux);
}

void launch_gen_stage_mixed(const TPArgs& A, hipStream_t stream, int flux, int nsp)
{
    if (nsp == 0) launch_gen_stage_m<ModeMixed, 0>(A, stream, flux);
    else launch_gen_stage_m<ModeMixed, 5>(A, stream, flux);
}

void launch_gen_aux_fast64(const TPArgs& A, hipStream_t stream, int what, double* U)
{
    launch_gen_aux_m<ModeFast64>(A, stream, what, U);
}

} // namespace e4
