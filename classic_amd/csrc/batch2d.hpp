// Batched 2D time stepping: many independent grids of different sizes and
// solvers advanced together, each by its own claw2 loop
// (src/2d/claw2.f:494-784) whose state lives on the device.
//
// One batched iteration is a short chain of kernels (batch2d.hip), captured
// once in a HIP graph and replayed:
//   prepare  per problem: at the start of a step (not a retake) clamp dt so
//            the step ends at tend (claw2.f:601), pick the output buffer,
//            clear the Courant accumulator;
//   step     a persistent kernel per solver: the (problem, 60-column strip,
//            row segment) tiles of all problems are dealt to its wavefronts
//            in turn and each runs the fused step of its tiles (step2_wave,
//            the code of the single-grid kernel), so grids of any size share
//            the GPU without a launch per grid; ghost cells are read through
//            each problem's bc2 index maps, so no ghost-fill pass runs;
//   accept   per problem: accept iff cfl <= cflv(1) (claw2.f:695), else keep
//            the old buffer; with variable dt the next dt is
//            min(dtv(2), dt*cflv(2)/cfl) after either outcome
//            (claw2.f:699-714, :746-750); count accepted steps, finish at
//            tend or after maxn steps.
// Every decision stays on the device, so a replay needs no host round trip,
// and each problem goes through exactly the dt, Courant numbers and states
// of a single-grid claw2 run of it.
//
// A "dry" iteration (autotuning) runs the step but changes no state: the
// step writes the spare buffer and prepare/accept leave t, dt and the
// buffer choice alone.
#pragma once

#include <hip/hip_runtime.h>

#include "bc2d.hpp"
#include "common.hpp"
#include "step2d_kernel.hpp"

namespace clawgpu {
namespace batch {

constexpr int kEuler = 2;  // CLAWGPU_SOLVER_EULER; every other id is acoustics here

struct Prob {
    double* q[3];  // three buffers of meqn planes each: q, step output, spare
    Layout2 L;
    SolverParams P;
    int solver, meqn;
    int mthbc[4];
    double dx, dy;
    int nstrips;
    // claw2 state
    int cur, res, nv, maxn, retake, done, info, nrej;
    double t, tstart, told, tend, dt, dtmin, dtmax;
    double cflmax, cfl;
    unsigned long long cfl_bits;
};

// Tiles are numbered per solver group (0: Euler, 1: acoustics), each group
// with its own persistent kernel, so each kernel keeps the register budget
// (and occupancy) of its own solver.
constexpr int kGroups = 2;

__host__ __device__ inline int group_of(int solver) { return solver == kEuler ? 0 : 1; }

struct Params {
    Prob* prob;
    int nprob;
    const int* rank_prob[kGroups];    // problem taking its tiles at rank r of a group's order
    const long* rank_tile0[kGroups];  // first tile of rank r (ascending in r)
    int nrank[kGroups];               // problems in the group
    long ntiles[kGroups];             // tiles of the group
    int walls[kGroups];               // some problem of the group has a wall side
    int* ndone;  // finished problems
    int order, trans, fwave, seglen;
    int lim[kMaxWaves];
    int vardt, dry;
    double dtv2, cflv1, cflv2;  // dt_max, cfl_max, cfl_desired
};

// largest r with start[r] <= t (start ascending, start[0] = 0)
__device__ inline int find_range(const long* start, int n, long t)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The tiles of group G; ghost cells are read through each problem's
// boundary maps (BC = 1: extrapolation and periodic sides; BC = 2: walls as
// well), so no ghost-fill pass runs between iterations.
template <class S, class Cfg, int G, int BC>
__device__ __forceinline__ void step_loop(const Params& b)
{
    const int lane = threadIdx.x & 63;
    const long wave = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
    const long nwaves = (long)gridDim.x * (blockDim.x >> 6);
    const long ntiles = b.ntiles[G];
    for (long tile = wave; tile < ntiles; tile += nwaves) {
        const int r = find_range(b.rank_tile0[G], b.nrank[G], tile);
        const int k = __builtin_amdgcn_readfirstlane(b.rank_prob[G][r]);
        const Prob& p = b.prob[k];
        if (__builtin_amdgcn_readfirstlane(p.done)) continue;
        const long local = tile - b.rank_tile0[G][r];
        StepArgs a;
        a.qold = p.q[p.cur];
        a.qnew = p.q[p.res];
        a.L = p.L;
        a.dtdx = p.dt / p.dx;
        a.dtdy = p.dt / p.dy;
        a.order = b.order;
        a.trans = b.trans;
        a.use_fwave = b.fwave;
        for (int w = 0; w < kMaxWaves; w++) a.lim[w] = b.lim[w];
        a.nstrips = p.nstrips;
        a.seglen = b.seglen;
        a.nsegs = (p.L.my + b.seglen - 1) / b.seglen;
        a.cfl_bits = &b.prob[k].cfl_bits;
        a.P = p.P;
        a.phix = a.phiy = nullptr;
        a.bcmap = 1;
        for (int e = 0; e < 4; e++) a.mthbc[e] = p.mthbc[e];
        step2_wave<S, Cfg, BC>(a, (int)(local % p.nstrips), (int)(local / p.nstrips), lane);
    }
}

// The compiler's register budget; step_kernel_w2 holds the Euler kernel
// to two waves per SIMD (256 registers) as the single-grid kernel reaches
// with fast arithmetic, step_kernel_w3 the acoustics kernel to three.
template <class S, class Cfg, int G, int BC>
__global__ __launch_bounds__(256, CLAWGPU_MIN_WAVES) void step_kernel(Params b)
{
    step_loop<S, Cfg, G, BC>(b);
}

template <class S, class Cfg, int G, int BC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void step_kernel_w2(Params b)
{
    step_loop<S, Cfg, G, BC>(b);
}

template <class S, class Cfg, int G, int BC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void step_kernel_w3(Params b)
{
    step_loop<S, Cfg, G, BC>(b);
}

inline bool mc_config(const Params& a)
{
    bool mc = a.order == 2 && a.trans == 2 && !a.fwave;
    for (int w = 0; w < 3; w++) mc = mc && a.lim[w] == 4;
    return mc;
}

template <class S, int G, int BC>
inline void launch_group_bc(const Params& a, bool mc, int nblk, bool w3, hipStream_t s)
{
    if constexpr (G == 1) {
        if (w3) {
            if (mc) hipLaunchKernelGGL((step_kernel_w3<S, CfgFix<2, 2, 4>, G, BC>), dim3(nblk), dim3(256), 0, s, a);
            else hipLaunchKernelGGL((step_kernel_w3<S, CfgAny, G, BC>), dim3(nblk), dim3(256), 0, s, a);
            return;
        }
    }
    if constexpr (G == 0 && S::Math::fast) {
        if (mc) hipLaunchKernelGGL((step_kernel_w2<S, CfgFix<2, 2, 4>, G, BC>), dim3(nblk), dim3(256), 0, s, a);
        else hipLaunchKernelGGL((step_kernel_w2<S, CfgAny, G, BC>), dim3(nblk), dim3(256), 0, s, a);
        return;
    }
    if (mc) hipLaunchKernelGGL((step_kernel<S, CfgFix<2, 2, 4>, G, BC>), dim3(nblk), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((step_kernel<S, CfgAny, G, BC>), dim3(nblk), dim3(256), 0, s, a);
}

template <class S, int G>
inline void launch_group(const Params& a, bool mc, int nblk, bool w3, hipStream_t s)
{
    if (a.walls[G]) launch_group_bc<S, G, 2>(a, mc, nblk, w3, s);
    else launch_group_bc<S, G, 1>(a, mc, nblk, w3, s);
}

// the step kernels of both solver groups (nblk blocks each)
template <class Mth>
inline int launch_steps_impl(const Params& a, int nblk, bool w3, hipStream_t s)
{
    const bool mc = mc_config(a);
    if (a.nrank[0] > 0) launch_group<Euler2<Mth>, 0>(a, mc, nblk, false, s);
    if (a.nrank[1] > 0) launch_group<Acoustics2<Mth>, 1>(a, mc, nblk, w3, s);
    return (int)hipGetLastError();
}

}  // namespace batch

// batch2d_exact.hip / batch2d_fast.hip (the latter with FMA contraction, as
// the single-grid step2d_fast.hip, so that both paths round alike)
int launch_batch_steps_exact(const batch::Params& a, int nblk, bool w3, hipStream_t s);
int launch_batch_steps_fast(const batch::Params& a, int nblk, bool w3, hipStream_t s);

}  // namespace clawgpu
