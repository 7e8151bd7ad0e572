// HIP kernels (gfx950) for the 2D Euler finite-volume step.
//
//  k_stage<SCHEME, LIM, STAGE, CORR>   one SSP-RK2 stage for every task of every stepping instance
//                                      (stage_kernel.hpp; exact arithmetic here, the fast-arithmetic
//                                      build in euler_kernels_fast.hip)
//  k_wavespeed                         max wave speeds + first non-physical cell of U (no update)
//  k_frame                             block-averaged float32 frame of U (trajectory output)
//  k_finish                            per-instance reduction of the partials and run-state decision
//  k_prep_host_step                    host-supplied dt for the next step
//  k_begin_run                         run state of every instance at the start of a device-resident run
//
// All arithmetic is binary64 with the reference's operation order (euler_numerics.hpp); the file is
// compiled with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include "run_state.hpp"
#include "stage_kernel.hpp"

namespace npde {

hipError_t launch_stage_fast(int scheme, int lim, int stage, bool corr, int waves_per_block, int mode, int row_waves,
                             int compact, int seam_rows, unsigned seam_bytes, unsigned lds_reserve,
                             const InstDesc* desc, Ctl* ctl, const Task* tasks, int ntasks, Partial* part, int* ndone,
                             hipStream_t s);
hipError_t stage_occupancy_fast(int scheme, int lim, bool corr, int waves_per_block, int mode, int row_waves,
                                unsigned seam_bytes, unsigned lds_reserve, int* workgroups);


// Reduce an instance's per-task wave-speed partials (one wavefront; every lane gets the result).  Runs in
// a launch after the kernel that wrote the partials, so the kernel boundary orders the partials before
// their reduction: no device-scope fence (which on this chip writes back and invalidates a whole L2) is
// needed inside the stage kernels.
__device__ void reduce_partials(const InstDesc& D, const Partial* __restrict__ part, int lane, double& sx,
                                double& sy, int& bad) {
    sx = 0.0;
    sy = 0.0;
    bad = kNoBad;
    for (int t = lane; t < D.ntasks; t += kWave) {
        const Partial p = part[D.task0 + t];
        sx = tmax(sx, p.sx);
        sy = tmax(sy, p.sy);
        bad = p.bad < bad ? p.bad : bad;
    }
    sx = wave_max(sx);
    sy = wave_max(sy);
    bad = wave_min(bad);
}

// Complete the step in flight and decide the next move (one wavefront).
__device__ void finish_instance(const InstDesc& D, Ctl* c, const Partial* __restrict__ part, int lane,
                                int* ndone) {
    double sx, sy;
    int bad;
    reduce_partials(D, part, lane, sx, sy, bad);
    if (lane == 0 && finish_step(D, c, sx, sy, bad)) atomicAdd(ndone + kCountDone, 1);
}

enum FinishMode : int { kAfterStep = 0, kAfterWavespeed = 1 };

// One wavefront per instance: after a stage-2 launch (instances that stepped) or after a wave-speed
// launch (instances not done).
__global__ __launch_bounds__(64) void k_finish(const InstDesc* __restrict__ desc, Ctl* __restrict__ ctl,
                                               const Partial* __restrict__ part, int n, int mode,
                                               int* __restrict__ ndone) {
    const int inst = blockIdx.x;
    if (inst >= n) return;
    Ctl* c = ctl + inst;
    const InstDesc& D = desc[inst];
    const int lane = threadIdx.x;
    const bool fresh = mode == kAfterStep ? c->active != 0 : c->done == 0;
    if (!fresh) return;
    finish_instance(D, c, part, lane, ndone);
}

// Max wave speeds and first non-physical cell of U for every instance that is not done.  Same task
// list, column mapping and partial slots as the stage kernels (strips, row groups or seams).
template <int MODE>
__global__ __launch_bounds__(kMaxRowWaves * kWave) void k_wavespeed(const InstDesc* __restrict__ desc,
                                                                    Ctl* __restrict__ ctl,
                                                                    const Task* __restrict__ tasks, int ntasks,
                                                                    Partial* __restrict__ part,
                                                                    int* __restrict__ ndone) {
    constexpr bool ROWS = MODE == kRowGroups;
    constexpr bool SEAMS = MODE == kSeams;
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int nw = blockDim.x / kWave;
    const int blk = xcd_block(int(blockIdx.x), int(gridDim.x));
    const int task = ROWS ? blk : __builtin_amdgcn_readfirstlane(blk * nw + wv);
    const int tid = ROWS ? __builtin_amdgcn_readfirstlane(blk * nw + wv) : task;
    if (task >= ntasks) return;
    const Task tk = tasks[task];
    Ctl* c = ctl + tk.inst;
    if (c->done) return;
    const InstDesc& D = desc[tk.inst];
    const int nx = D.nx;
    const long long P = D.plane;
    const int i = ROWS ? wv * kWave + lane : (SEAMS ? tk.i0 + lane : tk.i0 - kHalo + lane);
    const int ci = (ROWS || SEAMS) ? i : remap(i, nx, D.bc_x);
    const bool owner = ROWS || SEAMS || (lane >= kHalo && lane < kWave - kHalo && i < nx);
    double sx = 0.0, sy = 0.0;
    int bad = kNoBad;
    const int ny = D.ny;
    Cons un = load4(D.U, (long long)tk.j0 * nx + ci, P);  // one row in flight ahead of the arithmetic
    for (int f = tk.j0; f < tk.j1; ++f) {
        const Cons u = un;
        un = load4(D.U, (long long)(f + 1 < ny ? f + 1 : f) * nx + ci, P);
        const double rd = recip(u.d);
        const Prim w = prim_of(u.d, u.m, u.n, u.e, rd);
        const bool badc = !(w.d > 0.0) || !(w.p > 0.0);
        const double a = sound(w.d, w.p, rd);
        if (owner) {
            if (badc) {
                const int idx = f * nx + i;
                bad = idx < bad ? idx : bad;
            } else {
                sx = tmax(sx, dabs(w.u) + a);
                sy = tmax(sy, dabs(w.v) + a);
            }
        }
    }
    publish_partial(part, tid, lane, sx, sy, bad);
}

// One launch per iteration of a recording run, in place of the previous iteration's k_finish: with
// `complete` (every iteration but a launch sequence's first, whose predecessor ended with a k_finish),
// every frame workgroup of an instance first completes the step the previous iteration's stage kernels
// took (Ctl::active) -- each computes the same reduction and run-state decision from the same inputs into
// its own copy -- then, if that (or an earlier) decision landed on a frame, writes its share of the block
// averages (trajectory.hpp:39-48: sum over the
// f x f block, rows then columns, from 0.0; times 1/(f*f); to float).  The last of the instance's frame
// workgroups to finish stores the new run state and completes the landing (complete_landing; hold: a
// flux-network run).  Every workgroup read the run state before it counted itself in Ctl::fdone, so that
// store cannot overtake a reader.
__global__ __launch_bounds__(256) void k_frame(const InstDesc* __restrict__ desc, Ctl* __restrict__ ctl,
                                               const Partial* __restrict__ part, const FTask* __restrict__ ftasks,
                                               int nftasks, int complete, int hold, int* __restrict__ ndone) {
    const int bid = blockIdx.x;
    if (bid >= nftasks) return;
    const FTask ft = ftasks[bid];
    Ctl* c = ctl + ft.inst;
    const InstDesc& D = desc[ft.inst];
    __shared__ Ctl s_c;       // the instance's run state after this launch
    __shared__ int s_newly;   // 1 when the step completion finished the instance
    if (threadIdx.x < kWave) {
        const int lane = threadIdx.x;
        Ctl cl = *c;
        int newly = 0;
        if (complete && cl.active) {
            double sx, sy;
            int bad;
            reduce_partials(D, part, lane, sx, sy, bad);
            newly = finish_step(D, &cl, sx, sy, bad);
        }
        if (lane == 0) {
            s_c = cl;
            s_newly = newly;
        }
    }
    __syncthreads();
    const int k = s_c.pending;
    if (k >= 0) {
        const int lv = ft.level, m = D.lvl_m[lv], f = D.lvl_f[lv];
        const int nyc = D.ny / f;
        const int cell = ft.c0 + threadIdx.x;
        if (cell < m * nyc) {
            const int jc = cell / m, ic = cell - jc * m;
            const long long P = D.plane;
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
            for (int j = jc * f; j < (jc + 1) * f; ++j) {
                const long long rowoff = (long long)j * D.nx;
                for (int ii = ic * f; ii < (ic + 1) * f; ++ii) {
                    const long long o = rowoff + ii;
                    s0 += D.U[o];
                    s1 += D.U[o + P];
                    s2 += D.U[o + 2 * P];
                    s3 += D.U[o + 3 * P];
                }
            }
            const double inv = 1.0 / (double(f) * f);
            float* out = D.traj[lv] + (size_t)k * 4 * nyc * m;
            const size_t pl = (size_t)nyc * m;
            out[cell] = float(s0 * inv);
            out[pl + cell] = float(s1 * inv);
            out[2 * pl + cell] = float(s2 * inv);
            out[3 * pl + cell] = float(s3 * inv);
        }
    }
    if (threadIdx.x == 0 && atomicAdd(&c->fdone, 1) == D.nftasks - 1) {
        Ctl cl = s_c;
        cl.fdone = 0;
        *c = cl;
        if (s_newly) atomicAdd(ndone + kCountDone, 1);
        if (cl.pending >= 0) complete_landing(D, c, hold, ndone);
    }
}

// Block average of U into a plain (4, ny/f, m) float32 buffer (host-requested, any instance)
__global__ __launch_bounds__(256) void k_block_average(const InstDesc* __restrict__ desc, int inst, int m,
                                                       float* __restrict__ out) {
    const InstDesc& D = desc[inst];
    const int f = D.nx / m, nyc = D.ny / f;
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= m * nyc) return;
    const int jc = cell / m, ic = cell - jc * m;
    const long long P = D.plane;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (int j = jc * f; j < (jc + 1) * f; ++j) {
        const long long rowoff = (long long)j * D.nx;
        for (int ii = ic * f; ii < (ic + 1) * f; ++ii) {
            const long long o = rowoff + ii;
            s0 += D.U[o];
            s1 += D.U[o + P];
            s2 += D.U[o + 2 * P];
            s3 += D.U[o + 3 * P];
        }
    }
    const double inv = 1.0 / (double(f) * f);
    const size_t pl = (size_t)nyc * m;
    out[cell] = float(s0 * inv);
    out[pl + cell] = float(s1 * inv);
    out[2 * pl + cell] = float(s2 * inv);
    out[3 * pl + cell] = float(s3 * inv);
}

// Host-driven stepping: the next step of instance n uses dts[n] (negative: instance does not step).
__global__ void k_prep_host_step(Ctl* __restrict__ ctl, int n, const double* __restrict__ dts) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double dt = dts[k];
    ctl[k].mode = kModeHost;
    ctl[k].dt = dt;
    ctl[k].active = dt >= 0.0 ? 1 : 0;
}

// Start of a device-resident run (npde_advance_to / npde_run_trajectory): every instance's run state is
// set on the device, so the host needs no copy of it before the run.  max_new_steps < 0: no step budget.
// Thread 0 also clears the finished-instance counter.
__global__ void k_begin_run(Ctl* __restrict__ ctl, int n, double t_start, double t_end, int nout, int record,
                            long long max_new_steps, int* __restrict__ ndone) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) {
        ndone[kCountDone] = 0;
        ndone[kCountHeld] = 0;
        ndone[kCountNetGroups] = 0;
    }
    if (k >= n) return;
    Ctl& c = ctl[k];
    c.step_limit = max_new_steps < 0 ? (long long)(~0ull >> 1) : c.steps + max_new_steps;
    c.t_start = t_start;
    c.t_end = t_end;
    c.nout = nout;
    c.k = 0;
    c.mode = kModeAuto;
    c.record = record;
    c.pending = -1;
    c.done = 0;
    c.error = 0;
    c.active = 0;
    c.hold = 0;
    c.net_off = 0;
    c.rate0 = 0.0;
}

// Arithmetic self-test, element-wise: the shared-reciprocal quotient and the plain IEEE quotient,
// and the square root as the kernels evaluate it.
__global__ void k_selftest_arith(const double* __restrict__ x, const double* __restrict__ y, long long n,
                                 double* __restrict__ q_fast, double* __restrict__ q_ieee,
                                 double* __restrict__ root) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double a = x[k], b = y[k];
    q_fast[k] = div_r(a, b, recip(b));
    q_ieee[k] = a / b;
    root[k] = sqrt_core(a);
}

hipError_t launch_selftest_arith(const double* x, const double* y, long long n, double* qf, double* qi, double* rt,
                                 hipStream_t s) {
    hipLaunchKernelGGL(k_selftest_arith, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, x, y, n, qf, qi, rt);
    return hipGetLastError();
}

// Launch shape: strips put waves_per_block tasks (one per wavefront) in a workgroup; row groups
// (row_waves = W > 0) give each task (band) its own workgroup of W wavefronts.
static inline void launch_dims(int ntasks, int waves_per_block, int row_waves, dim3& grid, dim3& block) {
    if (row_waves > 0) {
        grid = dim3(unsigned(ntasks));
        block = dim3(unsigned(row_waves * kWave));
    } else {
        grid = dim3(unsigned((ntasks + waves_per_block - 1) / waves_per_block));
        block = dim3(unsigned(waves_per_block * kWave));
    }
}

// mode: StageMode; row_waves: wavefronts per workgroup in row-group mode (1 .. kMaxRowWaves)
static inline bool valid_mode(int mode, int row_waves) {
    if (mode == kRowGroups) return row_waves >= 1 && row_waves <= kMaxRowWaves;
    return (mode == kStrips || mode == kSeams) && row_waves == 0;
}

// seam mode: compact = read the stage input's edge columns from the edge arrays (the previous stage
// kernel of this launch sequence wrote them); seam_rows = the band height (at most kMaxSeamRows), which
// sizes each wavefront's LDS rows
// lds_reserve: LDS bytes each workgroup is to hold in total (an occupancy cap: a CU then holds only as many
// workgroups as fit), or 0
hipError_t launch_stage(int scheme, int lim, int stage, bool corr, bool fast, int waves_per_block, int mode,
                        int row_waves, int compact, int seam_rows, unsigned lds_reserve, const InstDesc* desc,
                        Ctl* ctl, const Task* tasks, int ntasks, Partial* part, int* ndone, hipStream_t s) {
    if (!valid_mode(mode, row_waves)) return hipErrorInvalidValue;
    if (mode == kSeams && (seam_rows < 1 || seam_rows > kMaxSeamRows)) return hipErrorInvalidValue;
    const unsigned seam = seam_lds_bytes(mode, waves_per_block, seam_rows);
    if (fast)
        return launch_stage_fast(scheme, lim, stage, corr, waves_per_block, mode, row_waves, compact, seam_rows, seam,
                                 lds_reserve, desc, ctl, tasks, ntasks, part, ndone, s);
    dim3 grid, block;
    launch_dims(ntasks, waves_per_block, row_waves, grid, block);
    const StageFn fn = pick_stage(scheme, lim, stage, corr, mode);
    unsigned lds = 0;
    const hipError_t e = stage_dynamic_lds(fn, seam, lds_reserve, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(fn, grid, block, lds, s, desc, ctl, tasks, ntasks, part, ndone, compact, seam_rows);
    return hipGetLastError();
}

// Workgroups of the stage-2 kernel one CU holds at once with the launch configuration launch_stage uses
// (HIP's occupancy calculator: registers, LDS including the reservation, wavefront slots).
hipError_t stage_occupancy(int scheme, int lim, bool corr, bool fast, int waves_per_block, int mode, int row_waves,
                           int seam_rows, unsigned lds_reserve, int* workgroups) {
    if (!valid_mode(mode, row_waves)) return hipErrorInvalidValue;
    const unsigned seam = seam_lds_bytes(mode, waves_per_block, seam_rows);
    if (fast)
        return stage_occupancy_fast(scheme, lim, corr, waves_per_block, mode, row_waves, seam, lds_reserve,
                                    workgroups);
    const StageFn fn = pick_stage(scheme, lim, 2, corr, mode);
    unsigned lds = 0;
    const hipError_t e = stage_dynamic_lds(fn, seam, lds_reserve, lds);
    if (e != hipSuccess) return e;
    const int threads = (row_waves > 0 ? row_waves : waves_per_block) * kWave;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(workgroups, reinterpret_cast<const void*>(fn), threads, lds);
}

hipError_t launch_wavespeed(int waves_per_block, int mode, int row_waves, const InstDesc* desc, Ctl* ctl,
                            const Task* tasks, int ntasks, Partial* part, int* ndone, hipStream_t s) {
    if (!valid_mode(mode, row_waves)) return hipErrorInvalidValue;
    dim3 grid, block;
    launch_dims(ntasks, waves_per_block, row_waves, grid, block);
    if (mode == kRowGroups)
        hipLaunchKernelGGL(k_wavespeed<kRowGroups>, grid, block, 0, s, desc, ctl, tasks, ntasks, part, ndone);
    else if (mode == kSeams)
        hipLaunchKernelGGL(k_wavespeed<kSeams>, grid, block, 0, s, desc, ctl, tasks, ntasks, part, ndone);
    else
        hipLaunchKernelGGL(k_wavespeed<kStrips>, grid, block, 0, s, desc, ctl, tasks, ntasks, part, ndone);
    return hipGetLastError();
}

hipError_t launch_frame(const InstDesc* desc, Ctl* ctl, const Partial* part, const FTask* ftasks, int nftasks,
                        int complete, int hold, int* ndone, hipStream_t s) {
    if (nftasks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_frame, dim3(nftasks), dim3(256), 0, s, desc, ctl, part, ftasks, nftasks, complete, hold,
                       ndone);
    return hipGetLastError();
}

hipError_t launch_finish(const InstDesc* desc, Ctl* ctl, const Partial* part, int n, int mode, int* ndone,
                         hipStream_t s) {
    hipLaunchKernelGGL(k_finish, dim3(n), dim3(64), 0, s, desc, ctl, part, n, mode, ndone);
    return hipGetLastError();
}

hipError_t launch_block_average(const InstDesc* desc, int inst, int m, int nyc, float* out, hipStream_t s) {
    const int cells = m * nyc;
    hipLaunchKernelGGL(k_block_average, dim3((cells + 255) / 256), dim3(256), 0, s, desc, inst, m, out);
    return hipGetLastError();
}


hipError_t launch_begin_run(Ctl* ctl, int n, double t_start, double t_end, int nout, int record, long long max_new_steps,
                            int* ndone, hipStream_t s) {
    hipLaunchKernelGGL(k_begin_run, dim3((n + 255) / 256), dim3(256), 0, s, ctl, n, t_start, t_end, nout, record,
                       max_new_steps, ndone);
    return hipGetLastError();
}

hipError_t launch_prep_host_step(Ctl* ctl, int n, const double* dts, hipStream_t s) {
    hipLaunchKernelGGL(k_prep_host_step, dim3((n + 255) / 256), dim3(256), 0, s, ctl, n, dts);
    return hipGetLastError();
}

}  // namespace npde
