// Fast-arithmetic build of the stage kernels: the same kernels as the exact build (stage_kernel.hpp)
// with euler_numerics.hpp's `fast` policy (reciprocal-based divisions without the final rounding
// correction, a shortened square root) and compiled with -ffp-contract=fast.  Results differ from the
// reference solver's by rounding-level amounts per step; selected per handle with
// npde_set_arithmetic(NPDE_ARITH_FAST).  The exact build stays the default.
#define NPDE_FAST_ARITH 1
#include "stage_kernel.hpp"

namespace npde {

// seam_bytes: the seam rows' dynamic LDS; lds_reserve: LDS bytes per workgroup under an occupancy cap, or 0
hipError_t launch_stage_fast(int scheme, int lim, int stage, bool corr, int waves_per_block, int mode, int row_waves,
                             int compact, int seam_rows, unsigned seam_bytes, unsigned lds_reserve,
                             const InstDesc* desc, Ctl* ctl, const Task* tasks, int ntasks, Partial* part, int* ndone,
                             hipStream_t s) {
    const dim3 grid = row_waves > 0 ? dim3(unsigned(ntasks))
                                    : dim3(unsigned((ntasks + waves_per_block - 1) / waves_per_block));
    const dim3 block = dim3(unsigned((row_waves > 0 ? row_waves : waves_per_block) * kWave));
    const StageFn fn = pick_stage(scheme, lim, stage, corr, mode);
    unsigned lds = 0;
    const hipError_t e = stage_dynamic_lds(fn, seam_bytes, lds_reserve, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(fn, grid, block, lds, s, desc, ctl, tasks, ntasks, part, ndone, compact, seam_rows);
    return hipGetLastError();
}

hipError_t stage_occupancy_fast(int scheme, int lim, bool corr, int waves_per_block, int mode, int row_waves,
                                unsigned seam_bytes, unsigned lds_reserve, int* workgroups) {
    const StageFn fn = pick_stage(scheme, lim, 2, corr, mode);
    unsigned lds = 0;
    const hipError_t e = stage_dynamic_lds(fn, seam_bytes, lds_reserve, lds);
    if (e != hipSuccess) return e;
    const int threads = (row_waves > 0 ? row_waves : waves_per_block) * kWave;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(workgroups, reinterpret_cast<const void*>(fn), threads, lds);
}

}  // namespace npde
