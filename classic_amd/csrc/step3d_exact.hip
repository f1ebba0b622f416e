// Split 3D sweeps with the reference's rounding (no contraction, IEEE
// division): bit for bit with step3ds.
#include "step3d.hpp"

namespace clawgpu {
namespace s3 {

int launch_sweep3(int dir, const SweepArgs3& a, hipStream_t s)
{
    const int wpb = 4;  // waves per 256-thread block
    long nwaves;
    if (dir == 0)
        nwaves = (long)a.nstrips * (a.L.my + 2 * a.L.mbc) * (a.L.mz + 2 * a.L.mbc);
    else
        nwaves = (long)a.nstrips * (dir == 1 ? a.L.mz + 2 * a.L.mbc : a.L.my + 2 * a.L.mbc) * a.nsegs;
    const unsigned nb = (unsigned)((nwaves + wpb - 1) / wpb);
    if (dir == 0)
        hipLaunchKernelGGL(sweep3_x_kernel<ExactMath>, dim3(nb), dim3(64 * wpb), 0, s, a);
    else
        hipLaunchKernelGGL(sweep3_yz_kernel<ExactMath>, dim3(nb), dim3(64 * wpb), 0, s, a, dir);
    return (int)hipGetLastError();
}

}  // namespace s3
}  // namespace clawgpu
