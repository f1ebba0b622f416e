// Unsplit 3D step, FMA-contracted arithmetic (held to 1e-12 of the reference).
#pragma clang fp contract(fast)
#include "step3d.hpp"

namespace clawgpu {
namespace s3 {

int launch_step3_tile(const Args3& a, hipStream_t s)
{
    const long nblocks = (long)a.ntx * a.nty * a.ntz;
    if (nblocks <= 0) return 0;
    hipLaunchKernelGGL((step3_tile_kernel<kTileY, kTileZ>), dim3((unsigned)nblocks), dim3(64), 0,
                       s, a);
    return (int)hipGetLastError();
}

}  // namespace s3
}  // namespace clawgpu
