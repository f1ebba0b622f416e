// Exact-arithmetic instantiation of the fused step kernel: compiled with FMA
// contraction off and correctly rounded division, so every value is the one
// the reference's Fortran arithmetic produces (src/2d/step2.f90, flux2.f90).
#include "step2d_kernel.hpp"
#include "split2d_kernel.hpp"

namespace clawgpu {
int launch_step2_exact(int solver, const StepArgs& a, int nblocks, hipStream_t s)
{
    return launch_step2_impl<ExactMath>(solver, a, nblocks, s);
}
int launch_loop2_exact(int solver, const StepArgs& a, const LoopArgs& l, int nblocks, hipStream_t s)
{
    return launch_loop2_impl<ExactMath>(solver, a, l, nblocks, s);
}
int step2_waves_per_cu_exact(int solver, const StepArgs& a)
{
    return step2_waves_per_cu_impl<ExactMath>(solver, a);
}
int launch_probe_exact(int solver, int n, int ixy, const SolverParams& P, const double* ql,
                      const double* qr, const double* asdq, double* wave, double* s, double* amdq,
                      double* apdq, double* bm, double* bp, hipStream_t st)
{
    return launch_probe_impl<ExactMath>(solver, n, ixy, P, ql, qr, asdq, wave, s, amdq, apdq, bm, bp, st);
}
int launch_sweep_exact(int solver, int dir, const SweepArgs& a, int nblocks, hipStream_t s)
{
    return launch_sweep_impl<ExactMath>(solver, dir, a, nblocks, s);
}
}  // namespace clawgpu
