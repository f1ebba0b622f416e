// Arguments of the general 2D step (step2g.hip): solvers with aux arrays and
// the capacity function.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace clawgpu {
namespace g2 {

struct GenArgs {
    const double* __restrict__ q;    // meqn planes
    double* __restrict__ qn;
    const double* __restrict__ aux;  // maux planes (stride afs)
    double* __restrict__ rec;        // 10 planes: x then y records (qadd, fadd lo, fadd hi, g1, g2)
    Layout2 L;
    long afs;
    double dtdx, dtdy;
    int capa, order, trans, fwave, lim0;
    unsigned long long* cfl_bits;
    double* cfl_part;  // per-wave Courant numbers, reduced into cfl_bits after each kernel
};

// one sweep of the dimensionally split step (step2ds)
struct SplitArgs {
    const double* __restrict__ qin;
    double* __restrict__ qout;
    const double* __restrict__ aux;
    Layout2 L;
    long afs;
    double dtd;  // dt/dx (x sweep) or dt/dy (y sweep) of this sweep
    int capa, order, fwave, lim0;
    unsigned long long* cfl_bits;
    double* cfl_part;  // per-wave Courant numbers, reduced into cfl_bits after each kernel
};

}  // namespace g2

// variable-coefficient advection, meqn = mwaves = 1
int launch_step2_general(const g2::GenArgs& a, hipStream_t s);
// dir 1: x sweep, 2: y sweep
int launch_split_sweep_general(int dir, const g2::SplitArgs& a, hipStream_t s);

}  // namespace clawgpu
