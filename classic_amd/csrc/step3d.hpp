// Host-side launchers of the 3D kernels (step3d_kernel.hpp).
#pragma once

#include <hip/hip_runtime.h>

#include "step3d_kernel.hpp"

namespace clawgpu {
namespace s3 {

constexpr int kTileY = 4, kTileZ = 4;

// unsplit step, FMA-contracted arithmetic (step3d_fast.hip)
int launch_step3_tile(const Args3& a, hipStream_t s);
// split sweeps, the reference's rounding (step3d_exact.hip); dir 0, 1, 2
int launch_sweep3(int dir, const SweepArgs3& a, hipStream_t s);

}  // namespace s3
}  // namespace clawgpu
