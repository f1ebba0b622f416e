// Exact-arithmetic instantiation of the batched step kernels (no FMA
// contraction, correctly rounded division): per problem, the same bits as
// the single-grid exact kernel and the reference.
#include "batch2d.hpp"

namespace clawgpu {
int launch_batch_steps_exact(const batch::Params& a, int nblk, bool w3, hipStream_t s)
{
    return batch::launch_steps_impl<ExactMath>(a, nblk, w3, s);
}
}  // namespace clawgpu
