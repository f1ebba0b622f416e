// Fast-arithmetic instantiation of the fused step kernel: the same algorithm
// with multiply-adds fused (FMA contraction on for this translation unit)
// and reciprocals reused instead of repeated divisions.  Results differ from
// the reference arithmetic by a few ulps per operation; the tests hold it to
// 1e-12 relative per field over whole runs.
#pragma clang fp contract(fast)
#include "step2d_kernel.hpp"
#include "split2d_kernel.hpp"

namespace clawgpu {
int launch_step2_fast(int solver, const StepArgs& a, int nblocks, hipStream_t s)
{
    return launch_step2_impl<FastMath>(solver, a, nblocks, s);
}
int launch_loop2_fast(int solver, const StepArgs& a, const LoopArgs& l, int nblocks, hipStream_t s)
{
    return launch_loop2_impl<FastMath>(solver, a, l, nblocks, s);
}
int step2_waves_per_cu_fast(int solver, const StepArgs& a)
{
    return step2_waves_per_cu_impl<FastMath>(solver, a);
}
int launch_probe_fast(int solver, int n, int ixy, const SolverParams& P, const double* ql,
                      const double* qr, const double* asdq, double* wave, double* s, double* amdq,
                      double* apdq, double* bm, double* bp, hipStream_t st)
{
    return launch_probe_impl<FastMath>(solver, n, ixy, P, ql, qr, asdq, wave, s, amdq, apdq, bm, bp, st);
}
int launch_sweep_fast(int solver, int dir, const SweepArgs& a, int nblocks, hipStream_t s)
{
    return launch_sweep_impl<FastMath>(solver, dir, a, nblocks, s);
}
}  // namespace clawgpu
