// Run-state machine of a device-resident run, shared by the kernels that move an instance on: k_finish and
// k_frame (euler_kernels.hip) and, in flux-network runs, k_flux_net (flux_net.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "euler_numerics.hpp"
#include "npde_device.hpp"

namespace npde {

// Run-state machine of one instance (executed by one lane).  Mirrors the driver loop
// (main2d.cpp:224-246): targets T(0) = t_start, T(k) = t_start + (t_end - t_start) * k / nout; while
// t < T(k) take a step with dt = min(cfl / (sx/dx + sy/dy), T(k) - t); on reaching T(k) write frame k
// (if recording) and move to k + 1.  A non-physical state stops the instance when a step is needed
// (the reference checks in compute_dt, solver2d.hpp:75-80).
__device__ __forceinline__ double target_time(const Ctl* c, int k) {
    if (k == 0) return c->t_start;
    return c->t_start + (c->t_end - c->t_start) * double(k) / double(c->nout);
}

// The instance's next move; returns 1 when this call finished it (the caller counts it in
// counters[kCountDone]).
__device__ inline int decide_next(const InstDesc& D, Ctl* c) {
    if (c->done) return 0;
    for (;;) {
        if (c->k > c->nout) {
            c->done = 1;
            return 1;
        }
        const double T = target_time(c, c->k);
        if (c->t < T) {
            if (c->steps >= c->step_limit) {  // step budget of npde_advance_steps spent
                c->done = 1;
                return 1;
            }
            if (c->bad != kNoBad) {
                c->error = 1;
                c->done = 1;
                return 1;
            }
            const double dt = D.cfl / (c->sx / D.dx + c->sy / D.dy);
            c->dt = tmin(dt, T - c->t);
            c->active = 1;
            return 0;
        }
        if (c->record) {
            c->pending = c->k;
            return 0;
        }
        c->k += 1;
    }
}

// Completion of the step in flight (c->active) from the reduced wave-speed partials of the new state,
// then (device-decided runs) the next move; returns 1 when it finished the instance.
__device__ inline int finish_step(const InstDesc& D, Ctl* c, double sx, double sy, int bad) {
    c->sx = sx;
    c->sy = sy;
    c->bad = bad;
    if (c->active) {
        c->t = c->t + c->dt;
        c->steps += 1;
        c->active = 0;
    }
    return c->mode == kModeAuto ? decide_next(D, c) : 0;
}

// A frame landing is complete (its frame written): record its time and move to the next target.  hold: a
// flux-network run, where the instance then waits for the batch's network call at this frame (unless it
// was the last frame) and is counted in counters[kCountHeld]; otherwise it decides its next step at once.
// In a flux-network run the landing also watches the instance's CFL signal-speed sum sx/dx + sy/dy: the
// first frame records it, and a later frame at which it exceeds D.net_rate_limit times that value switches
// the network off for the instance (its factors are zero from this frame on: the plain coarse solver).  A
// corrected flow that drains a cell toward vacuum at finite pressure shows there first -- the sound speed,
// and with it every step of the whole frame-synchronous batch, grows without bound.
__device__ inline void complete_landing(const InstDesc& D, Ctl* c, int hold, int* counters) {
    const int k = c->pending;
    if (D.ftimes) D.ftimes[k] = c->t;
    if (hold) {
        const double rate = c->sx / D.dx + c->sy / D.dy;
        if (k == 0) c->rate0 = rate;
        else if (D.net_rate_limit > 0.0 && rate > D.net_rate_limit * c->rate0) c->net_off = 1;
    }
    c->pending = -1;
    c->k = k + 1;
    if (hold && c->k <= c->nout) {
        c->hold = 1;
        atomicAdd(counters + kCountHeld, 1);
        return;
    }
    if (decide_next(D, c)) atomicAdd(counters + kCountDone, 1);
}

}  // namespace npde
