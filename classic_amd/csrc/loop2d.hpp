// The decision of the claw2 time loop after one attempt of a step
// (src/2d/claw2.f:695-753), evaluated on the device so that a run of steps
// needs no host round trip per step.  Every wave of the next attempt's
// kernel evaluates it on the same inputs and gets the same state.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace clawgpu {

// s: state before attempt k-1; cfl: that attempt's Courant number;
// o: state before attempt k.  The arithmetic is the host loop's, operation
// for operation (no contraction), so a device run takes the same dt
// sequence as the reference.
__host__ __device__ inline void loop_next(const LoopArgs& l, const LoopState& s, double cfl,
                                          LoopState& o)
{
#pragma clang fp contract(off)
    o = s;
    if (s.done) return;
    o.cfl = cfl;
    if (!(cfl <= 1.7976931348623157e308)) {  // NaN or infinite wave speed
        o.done = 1;
        o.info = -5;
        return;
    }
    const double t = s.told + s.dt;
    if (cfl <= l.cflv1) {  // accept (claw2.f:695-697, :746-753)
        o.cflmax = fmax(cfl, s.cflmax);
        o.cur = (s.cur + 1) % 3;
        o.t = t;
        o.n = s.n + 1;
        double dt = s.dt;
        if (l.vardt) {
            dt = cfl == 0.0 ? l.dtv2 : fmin(l.dtv2, s.dt * l.cflv2 / cfl);
            o.dtmin = fmin(dt, s.dtmin);
            o.dtmax = fmax(dt, s.dtmax);
        }
        o.dt = dt;
        if (t >= l.tend || o.n >= l.maxn) {
            o.done = 1;
            return;
        }
        o.told = t;  // next step: dt clamped to hit tend (claw2.f:601)
        if (t + dt > l.tend && l.tstart < l.tend) o.dt = l.tend - t;
    } else {  // reject: q is still in buffer cur (claw2.f:699-714)
        o.t = s.told;
        if (l.vardt) {
            o.dt = fmin(l.dtv2, s.dt * l.cflv2 / cfl);
        } else {
            o.cflmax = fmax(cfl, s.cflmax);
            o.done = 1;
        }
    }
}

__device__ inline double loop_cfl(const LoopArgs& l, int k)
{
    return __longlong_as_double((long long)l.cflb[k % 3]);
}

}  // namespace clawgpu
