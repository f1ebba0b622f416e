// Fast-arithmetic instantiation of the batched step kernels, compiled like
// step2d_fast.hip (FMA contraction on for this translation unit) so that a
// batched problem rounds exactly as its single-grid fast run.
#pragma clang fp contract(fast)
#include "batch2d.hpp"

namespace clawgpu {
int launch_batch_steps_fast(const batch::Params& a, int nblk, bool w3, hipStream_t s)
{
    return batch::launch_steps_impl<FastMath>(a, nblk, w3, s);
}
}  // namespace clawgpu
