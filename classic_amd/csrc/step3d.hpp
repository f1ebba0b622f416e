// Host-side launchers of the 3D kernels (step3d_kernel.hpp).
#pragma once

#include <hip/hip_runtime.h>

#include "step3d_kernel.hpp"

namespace clawgpu {
namespace s3 {

// unsplit step, FMA-contracted arithmetic (step3d_fast.hip); tile selects
// the (TY, TZ) shape: 0 = 4x4, 1 = 4x2, 2 = 2x2, 3 = 8x2, 4 = 2x4
int launch_step3_tile(const Args3& a, int tile, hipStream_t s);
// split sweeps, the reference's rounding (step3d_exact.hip); dir 0, 1, 2
int launch_sweep3(int dir, const SweepArgs3& a, hipStream_t s);

}  // namespace s3
}  // namespace clawgpu
