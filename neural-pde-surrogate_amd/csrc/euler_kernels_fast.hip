// Fast-arithmetic build of the stage kernels: the same kernels as the exact build (stage_kernel.hpp)
// with euler_numerics.hpp's `fast` policy (reciprocal-based divisions without the final rounding
// correction, a shortened square root) and compiled with -ffp-contract=fast.  Results differ from the
// reference solver's by rounding-level amounts per step; selected per handle with
// npde_set_arithmetic(NPDE_ARITH_FAST).  The exact build stays the default.
#define NPDE_FAST_ARITH 1
#include "stage_kernel.hpp"

namespace npde {

hipError_t launch_stage_fast(int scheme, int lim, int stage, bool corr, int waves_per_block, const InstDesc* desc,
                             Ctl* ctl, const Task* tasks, int ntasks, Partial* part, int* ndone, hipStream_t s) {
    const int blocks = (ntasks + waves_per_block - 1) / waves_per_block;
    hipLaunchKernelGGL(pick_stage(scheme, lim, stage, corr), dim3(blocks), dim3(waves_per_block * kWave), 0, s,
                       desc, ctl, tasks, ntasks, part, ndone);
    return hipGetLastError();
}

}  // namespace npde
