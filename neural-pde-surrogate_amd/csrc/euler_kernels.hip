// HIP kernels (gfx950) for the 2D Euler finite-volume step.
//
//  k_stage<SCHEME, LIM, STAGE, CORR>   one SSP-RK2 stage for every task of every stepping instance
//  k_wavespeed                         max wave speeds + first non-physical cell of U (no update)
//  k_frame                             block-averaged float32 frame of U (trajectory output)
//  k_finish                            per-instance reduction of the partials and run-state decision
//  k_prep_host_step                    host-supplied dt for the next step
//
// Stage kernel design (one wavefront = one task = 60 interior columns x a band of rows):
//   * each lane owns one column; lanes 0,1 and 62,63 are the 2-column halo of the strip;
//   * the wavefront marches up its band one row per iteration, keeping a 3-row window of primitives
//     in registers, so every cell's primitives, limited slopes and four face-side states
//     (W +- dW/2 -> conserved -> primitive + sound speed) are computed exactly once per stage;
//   * x neighbours come from the adjacent lanes by DPP wavefront shifts; each lane solves the
//     Riemann problem at its left x face, and receives its right-face flux from the next lane;
//   * each y face is solved once, by the lane below/above it, carried to the next row in registers;
//   * loads of one row are four fully coalesced 512-byte reads (one per conserved plane);
//   * ghost cells are virtual: stencil indices outside the grid are remapped (periodic wrap or
//     transmissive clamp) before the load;
//   * stage 2 fuses the next step's dt reduction into its epilogue (to_primitive, positivity check,
//     max(|u|+a), max(|v|+a) of U^{n+1}) and writes one partial per task; k_finish (one wavefront per
//     instance) reduces them and advances that instance's run state.
//
// All arithmetic is binary64 with the reference's operation order (euler_numerics.hpp); the file is
// compiled with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include "euler_numerics.hpp"
#include "npde_device.hpp"

namespace npde {

// ---------------------------------------------------------------------------------------------
// cross-lane moves: lane l receives lane l-1 (from_prev) or lane l+1 (from_next)
__device__ __forceinline__ int dpp_prev(int x) {
    return __builtin_amdgcn_update_dpp(0, x, 0x138 /* wave_shr:1 */, 0xF, 0xF, false);
}
__device__ __forceinline__ int dpp_next(int x) {
    return __builtin_amdgcn_update_dpp(0, x, 0x130 /* wave_shl:1 */, 0xF, 0xF, false);
}
__device__ __forceinline__ double from_prev(double x) {
    return __hiloint2double(dpp_prev(__double2hiint(x)), dpp_prev(__double2loint(x)));
}
__device__ __forceinline__ double from_next(double x) {
    return __hiloint2double(dpp_next(__double2hiint(x)), dpp_next(__double2loint(x)));
}
__device__ __forceinline__ Prim from_prev(const Prim& w) {
    return {from_prev(w.d), from_prev(w.u), from_prev(w.v), from_prev(w.p)};
}
__device__ __forceinline__ Prim from_next(const Prim& w) {
    return {from_next(w.d), from_next(w.u), from_next(w.v), from_next(w.p)};
}
__device__ __forceinline__ Side from_prev(const Side& s) {
    return {from_prev(s.d),  from_prev(s.mn), from_prev(s.mt), from_prev(s.e),
            from_prev(s.un), from_prev(s.ut), from_prev(s.p),  from_prev(s.a),
            from_prev(s.rd)};
}
__device__ __forceinline__ Cons from_next(const Cons& c) {
    return {from_next(c.d), from_next(c.m), from_next(c.n), from_next(c.e)};
}

__device__ __forceinline__ double wave_max(double x) {
    for (int o = 32; o > 0; o >>= 1) x = tmax(x, __shfl_xor(x, o, 64));
    return x;
}
__device__ __forceinline__ int wave_min(int x) {
    for (int o = 32; o > 0; o >>= 1) {
        const int y = __shfl_xor(x, o, 64);
        x = y < x ? y : x;
    }
    return x;
}

// State pointers live in the instance descriptor, so the compiler cannot see that they address
// global memory; casting to the global address space gives global_load/global_store (scalar base +
// lane offset) instead of flat accesses.
typedef __attribute__((address_space(1))) const double gcdouble;
typedef __attribute__((address_space(1))) double gdouble;

__device__ __forceinline__ Cons load4(const double* B, long long off, long long P) {
    const gcdouble* g = (const gcdouble*)B;
    return {g[off], g[off + P], g[off + 2 * P], g[off + 3 * P]};
}

__device__ __forceinline__ void store4(double* B, long long off, long long P, const Cons& c) {
    gdouble* g = (gdouble*)B;
    g[off] = c.d;
    g[off + P] = c.m;
    g[off + 2 * P] = c.n;
    g[off + 3 * P] = c.e;
}

__device__ __forceinline__ Prim prim_of(const Cons& c) { return prim_of(c.d, c.m, c.n, c.e); }

// ---------------------------------------------------------------------------------------------
// Run-state machine of one instance (executed by one lane).  Mirrors the driver loop
// (main2d.cpp:224-246): targets T(0) = t_start, T(k) = t_start + (t_end - t_start) * k / nout; while
// t < T(k) take a step with dt = min(cfl / (sx/dx + sy/dy), T(k) - t); on reaching T(k) write frame k
// (if recording) and move to k + 1.  A non-physical state stops the instance when a step is needed
// (the reference checks in compute_dt, solver2d.hpp:75-80).
__device__ __forceinline__ double target_time(const Ctl* c, int k) {
    if (k == 0) return c->t_start;
    return c->t_start + (c->t_end - c->t_start) * double(k) / double(c->nout);
}

__device__ void decide_next(const InstDesc& D, Ctl* c, int* ndone) {
    if (c->done) return;
    for (;;) {
        if (c->k > c->nout) {
            c->done = 1;
            atomicAdd(ndone, 1);
            return;
        }
        const double T = target_time(c, c->k);
        if (c->t < T) {
            if (c->steps >= c->step_limit) {  // step budget of npde_advance_steps spent
                c->done = 1;
                atomicAdd(ndone, 1);
                return;
            }
            if (c->bad != kNoBad) {
                c->error = 1;
                c->done = 1;
                atomicAdd(ndone, 1);
                return;
            }
            const double dt = D.cfl / (c->sx / D.dx + c->sy / D.dy);
            c->dt = tmin(dt, T - c->t);
            c->active = 1;
            return;
        }
        if (c->record) {
            c->pending = c->k;
            return;
        }
        c->k += 1;
    }
}

// Reduce an instance's per-task wave-speed partials (one wavefront), complete the step in flight and
// decide the next one.  Runs in its own launch after the kernel that wrote the partials, so the
// kernel boundary orders the partials before their reduction: no device-scope fence (which on this
// chip writes back and invalidates a whole L2) is needed inside the stage kernels.
__device__ void finish_instance(const InstDesc& D, Ctl* c, const Partial* __restrict__ part, int lane,
                                int* ndone) {
    double sx = 0.0, sy = 0.0;
    int bad = kNoBad;
    for (int t = lane; t < D.ntasks; t += kWave) {
        const Partial p = part[D.task0 + t];
        sx = tmax(sx, p.sx);
        sy = tmax(sy, p.sy);
        bad = p.bad < bad ? p.bad : bad;
    }
    sx = wave_max(sx);
    sy = wave_max(sy);
    bad = wave_min(bad);
    if (lane == 0) {
        c->sx = sx;
        c->sy = sy;
        c->bad = bad;
        if (c->active) {
            c->t = c->t + c->dt;
            c->steps += 1;
            c->active = 0;
        }
        if (c->mode == kModeAuto) decide_next(D, c, ndone);
    }
}

// This task's maxima (all lanes) -> its partial slot.
__device__ __forceinline__ void publish_partial(Partial* __restrict__ part, int tid, int lane, double sx,
                                                double sy, int bad) {
    sx = wave_max(sx);
    sy = wave_max(sy);
    bad = wave_min(bad);
    if (lane == 0) {
        part[tid].sx = sx;
        part[tid].sy = sy;
        part[tid].bad = bad;
    }
}

enum FinishMode : int { kAfterStep = 0, kAfterWavespeed = 1, kAfterFrame = 2 };

// One wavefront per instance: after a stage-2 launch (instances that stepped), after a wave-speed
// launch (instances not done), or after a frame launch (instances whose frame was just written).
__global__ __launch_bounds__(64) void k_finish(const InstDesc* __restrict__ desc, Ctl* __restrict__ ctl,
                                               const Partial* __restrict__ part, int n, int mode,
                                               int* __restrict__ ndone) {
    const int inst = blockIdx.x;
    if (inst >= n) return;
    Ctl* c = ctl + inst;
    const InstDesc& D = desc[inst];
    const int lane = threadIdx.x;
    if (mode == kAfterFrame) {
        if (c->pending < 0 || lane != 0) return;
        const int k = c->pending;
        if (D.ftimes) D.ftimes[k] = c->t;
        c->pending = -1;
        c->k = k + 1;
        decide_next(D, c, ndone);
        return;
    }
    const bool fresh = mode == kAfterStep ? c->active != 0 : c->done == 0;
    if (!fresh) return;
    finish_instance(D, c, part, lane, ndone);
}

// ---------------------------------------------------------------------------------------------
template <int SCHEME, int LIM, int STAGE, bool CORR>
__global__ __launch_bounds__(256) void k_stage(const InstDesc* __restrict__ desc, Ctl* __restrict__ ctl,
                                               const Task* __restrict__ tasks, int ntasks,
                                               Partial* __restrict__ part, int* __restrict__ ndone) {
    const int lane = threadIdx.x & (kWave - 1);
    const int tid = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave);
    if (tid >= ntasks) return;
    const Task tk = tasks[tid];
    Ctl* c = ctl + tk.inst;
    if (!c->active) return;
    const double dt = c->dt;
    const InstDesc& D = desc[tk.inst];
    const int nx = D.nx, ny = D.ny, bcy = D.bc_y;
    const long long P = D.plane;
    const double cx = D.cx, cy = D.cy;

    const int i = tk.i0 - kHalo + lane;          // this lane's column
    const int ci = remap(i, nx, D.bc_x);          // where it is loaded from (in range for every lane)
    const bool owner = lane >= kHalo && lane < kWave - kHalo && i < nx;

    const double* __restrict__ V = (STAGE == 1) ? D.U : D.U1;  // stencil input
    double* O = (STAGE == 1) ? D.U1 : D.U;                     // output (stage 2: in place on U)

    auto row = [&](int j) -> long long { return (long long)remap(j, ny, bcy) * nx + ci; };

    // learned flux increments (coarse solver only): face left of column i, face below row r
    int cxi = -1;
    // an instance of a corrected batch without increments of its own reads none (null pointers)
    const bool has_corr = CORR && D.corr_x != nullptr && D.corr_y != nullptr;
    if constexpr (CORR) cxi = (i >= 0 && i < nx) ? i : (D.bc_x == kPeriodic ? remap(i, nx, kPeriodic) : -1);
    if (!has_corr) cxi = -1;
    auto corr_x = [&](int r) -> Cons {
        if (cxi < 0) return {0.0, 0.0, 0.0, 0.0};
        return load4(D.corr_x, (long long)r * nx + cxi, P);
    };
    auto corr_y = [&](int r) -> Cons {  // r in [0, ny]
        int rr = r < ny ? r : (D.bc_y == kPeriodic ? 0 : -1);
        if (rr < 0 || !has_corr) return {0.0, 0.0, 0.0, 0.0};
        return load4(D.corr_y, (long long)rr * nx + ci, P);
    };

    const int j0 = tk.j0, j1 = tk.j1;

    // prologue: rows j0-2 .. j0+1, the y face below row j0, and the top side of row j0
    Cons va = load4(V, row(j0), P);
    Cons vb = load4(V, row(j0 + 1), P);
    const Prim w0 = prim_of(load4(V, row(j0 - 2), P));
    const Prim w1 = prim_of(load4(V, row(j0 - 1), P));
    Prim wa = prim_of(va);
    Prim wb = prim_of(vb);
    Side top = edge_side(w1, slope<LIM>(w0, w1, wa), +0.5);
    Cons glo;
    {
        const Prim dy = slope<LIM>(w1, wa, wb);
        glo = flux_y<SCHEME>(top, edge_side(wa, dy, -0.5));
        if constexpr (CORR) {
            const Cons k = corr_y(j0);
            glo = {glo.d + k.d, glo.m + k.m, glo.n + k.n, glo.e + k.e};
        }
        top = edge_side(wa, dy, +0.5);
    }

    double sx = 0.0, sy = 0.0;
    int bad = kNoBad;

    // software pipeline: the rows an iteration consumes were loaded one iteration earlier, so HBM
    // latency overlaps a whole row of arithmetic instead of stalling the wavefront
    Cons vn = load4(V, row(j0 + 2), P);
    Cons ub{};
    if constexpr (STAGE == 2) ub = load4(O, (long long)j0 * nx + ci, P);  // U, point-wise

    for (int f = j0; f < j1; ++f) {
        const Cons vq = load4(V, row(f + 3), P);
        Cons uq{};
        if constexpr (STAGE == 2) uq = load4(O, (long long)(f + 1 < ny ? f + 1 : f) * nx + ci, P);
        const Prim wn = prim_of(vn);

        // y: slope of row f+1, its two y-face sides, and the face between rows f and f+1
        const Prim dyb = slope<LIM>(wa, wb, wn);
        Cons ghi = flux_y<SCHEME>(top, edge_side(wb, dyb, -0.5));
        if constexpr (CORR) {
            const Cons k = corr_y(f + 1);
            ghi = {ghi.d + k.d, ghi.m + k.m, ghi.n + k.n, ghi.e + k.e};
        }
        const Side topn = edge_side(wb, dyb, +0.5);

        // x: slope of this cell from the neighbouring lanes, left-face flux here, right-face flux from lane+1
        const Prim dxs = slope<LIM>(from_prev(wa), wa, from_next(wa));
        const Side lf = edge_side(wa, dxs, -0.5);
        const Side rt = edge_side(wa, dxs, +0.5);
        Cons fl = flux_x<SCHEME>(from_prev(rt), lf);
        if constexpr (CORR) {
            const Cons k = corr_x(f);
            fl = {fl.d + k.d, fl.m + k.m, fl.n + k.n, fl.e + k.e};
        }
        const Cons fr = from_next(fl);

        // L = -(F_right - F_left)/dx - (G_top - G_bottom)/dy, x first then y (solver2d.hpp:139-178)
        Cons L;
        L.d = (0.0 + cx * fl.d) - cx * fr.d;
        L.m = (0.0 + cx * fl.m) - cx * fr.m;
        L.n = (0.0 + cx * fl.n) - cx * fr.n;
        L.e = (0.0 + cx * fl.e) - cx * fr.e;
        L.d = (L.d + cy * glo.d) - cy * ghi.d;
        L.m = (L.m + cy * glo.m) - cy * ghi.m;
        L.n = (L.n + cy * glo.n) - cy * ghi.n;
        L.e = (L.e + cy * glo.e) - cy * ghi.e;

        Cons out;
        if constexpr (STAGE == 1) {  // U1 = U + dt L(U)
            out = {va.d + dt * L.d, va.m + dt * L.m, va.n + dt * L.n, va.e + dt * L.e};
        } else {  // U = 1/2 ((U + U1) + dt L(U1))
            out = {0.5 * (ub.d + va.d + dt * L.d), 0.5 * (ub.m + va.m + dt * L.m),
                   0.5 * (ub.n + va.n + dt * L.n), 0.5 * (ub.e + va.e + dt * L.e)};
        }
        if (owner) store4(O, (long long)f * nx + ci, P, out);

        if constexpr (STAGE == 2) {  // wave speeds of U^{n+1} for the next dt (solver2d.hpp:67-83)
            const double rd = recip(out.d);
            const Prim w = prim_of(out.d, out.m, out.n, out.e, rd);
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

        glo = ghi;
        top = topn;
        va = vb;
        vb = vn;
        wa = wb;
        wb = wn;
        vn = vq;
        if constexpr (STAGE == 2) ub = uq;
    }

    if constexpr (STAGE == 2) {
        publish_partial(part, tid, lane, sx, sy, bad);
    }
}

// Max wave speeds and first non-physical cell of U for every instance that is not done.
__global__ __launch_bounds__(256) void k_wavespeed(const InstDesc* __restrict__ desc, Ctl* __restrict__ ctl,
                                                   const Task* __restrict__ tasks, int ntasks,
                                                   Partial* __restrict__ part, int* __restrict__ ndone) {
    const int lane = threadIdx.x & (kWave - 1);
    const int tid = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave);
    if (tid >= ntasks) return;
    const Task tk = tasks[tid];
    Ctl* c = ctl + tk.inst;
    if (c->done) return;
    const InstDesc& D = desc[tk.inst];
    const int nx = D.nx;
    const long long P = D.plane;
    const int i = tk.i0 - kHalo + lane;
    const int ci = remap(i, nx, D.bc_x);
    const bool owner = lane >= kHalo && lane < kWave - kHalo && i < nx;
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

// Block averages of the conserved variables for instances with a pending frame
// (trajectory.hpp:39-48: sum over the f x f block, rows then columns, from 0.0; times 1/(f*f); to float).
__global__ __launch_bounds__(256) void k_frame(const InstDesc* __restrict__ desc, Ctl* __restrict__ ctl,
                                               const FTask* __restrict__ ftasks, int nftasks,
                                               int* __restrict__ ndone) {
    const int bid = blockIdx.x;
    if (bid >= nftasks) return;
    const FTask ft = ftasks[bid];
    Ctl* c = ctl + ft.inst;
    const int k = c->pending;
    if (k < 0) return;
    const InstDesc& D = desc[ft.inst];
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

// ---------------------------------------------------------------------------------------------
// launch table
using StageFn = void (*)(const InstDesc*, Ctl*, const Task*, int, Partial*, int*);

template <int S, int L>
constexpr StageFn stage_fn(int stage, bool corr) {
    return stage == 1 ? (corr ? k_stage<S, L, 1, true> : k_stage<S, L, 1, false>)
                      : (corr ? k_stage<S, L, 2, true> : k_stage<S, L, 2, false>);
}

StageFn pick_stage(int scheme, int lim, int stage, bool corr) {
    if (scheme == kRusanov) {
        switch (lim) {
            case kLimNone: return stage_fn<kRusanov, kLimNone>(stage, corr);
            case kLimUnlimited: return stage_fn<kRusanov, kLimUnlimited>(stage, corr);
            case kLimMinmod: return stage_fn<kRusanov, kLimMinmod>(stage, corr);
            default: return stage_fn<kRusanov, kLimVanLeer>(stage, corr);
        }
    }
    switch (lim) {
        case kLimNone: return stage_fn<kHllc, kLimNone>(stage, corr);
        case kLimUnlimited: return stage_fn<kHllc, kLimUnlimited>(stage, corr);
        case kLimMinmod: return stage_fn<kHllc, kLimMinmod>(stage, corr);
        default: return stage_fn<kHllc, kLimVanLeer>(stage, corr);
    }
}

hipError_t launch_stage(int scheme, int lim, int stage, bool corr, int waves_per_block, const InstDesc* desc,
                        Ctl* ctl, const Task* tasks, int ntasks, Partial* part, int* ndone, hipStream_t s) {
    const int blocks = (ntasks + waves_per_block - 1) / waves_per_block;
    hipLaunchKernelGGL(pick_stage(scheme, lim, stage, corr), dim3(blocks), dim3(waves_per_block * kWave), 0, s,
                       desc, ctl, tasks, ntasks, part, ndone);
    return hipGetLastError();
}

hipError_t launch_wavespeed(int waves_per_block, const InstDesc* desc, Ctl* ctl, const Task* tasks, int ntasks,
                            Partial* part, int* ndone, hipStream_t s) {
    const int blocks = (ntasks + waves_per_block - 1) / waves_per_block;
    hipLaunchKernelGGL(k_wavespeed, dim3(blocks), dim3(waves_per_block * kWave), 0, s, desc, ctl, tasks, ntasks,
                       part, ndone);
    return hipGetLastError();
}

hipError_t launch_frame(const InstDesc* desc, Ctl* ctl, const FTask* ftasks, int nftasks, int* ndone,
                        hipStream_t s) {
    if (nftasks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_frame, dim3(nftasks), dim3(256), 0, s, desc, ctl, ftasks, nftasks, ndone);
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

hipError_t launch_prep_host_step(Ctl* ctl, int n, const double* dts, hipStream_t s) {
    hipLaunchKernelGGL(k_prep_host_step, dim3((n + 255) / 256), dim3(256), 0, s, ctl, n, dts);
    return hipGetLastError();
}

}  // namespace npde
