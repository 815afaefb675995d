// gen_exact.hip -- kernels of blocks built from their grid, ModeExact
// instantiations and the CFL signal kernel.
// Built with -ffp-contract=off (reference roundings).
#include "gen_kernels.h"

namespace e4 {

void launch_gen_stage_exact(const TPArgs& A, hipStream_t stream, int flux, int nsp)
{
    if (nsp == 0) launch_gen_stage_m<ModeExact, 0>(A, stream, flux);
    else launch_gen_stage_m<ModeExact, 5>(A, stream, flux);
}

void launch_gen_aux_exact(const TPArgs& A, hipStream_t stream, int what, double* U)
{
    launch_gen_aux_m<ModeExact>(A, stream, what, U);
}

void launch_gen_cfl(const TPArgs& A, int stringent, unsigned long long* smax_bits, double* last_signal, int nwg,
                    hipStream_t stream)
{
    hipLaunchKernelGGL(gen_cfl_kernel, dim3((unsigned)nwg), dim3(256), 0, stream, A, stringent, smax_bits, last_signal);
}

} // namespace e4
