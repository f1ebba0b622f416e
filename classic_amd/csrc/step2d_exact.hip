// Exact-arithmetic instantiation of the fused step kernel: compiled with FMA
// contraction off and correctly rounded division, so every value is the one
// the reference's Fortran arithmetic produces (src/2d/step2.f90, flux2.f90).
#include "step2d_kernel.hpp"

namespace clawgpu {
int launch_step2_exact(int solver, const StepArgs& a, int nblocks, hipStream_t s)
{
    return launch_step2_impl<ExactMath>(solver, a, nblocks, s);
}
}  // namespace clawgpu
