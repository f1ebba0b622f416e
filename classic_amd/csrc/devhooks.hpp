// Device-side user routines: the loaded code object and the launch of one
// routine on the device q (devhooks.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "clawgpu.h"

struct clawgpu_dev_hooks {
    hipModule_t mod = nullptr;
    hipFunction_t bc2 = nullptr, b4step2 = nullptr, src2 = nullptr;
    double* user = nullptr;  // nuser doubles on the device
    int nuser = 0;
};

namespace clawgpu {

// One launch of routine f on q (one thread per cell, ghost cells included).
int launch_dev_hook(clawgpu_ctx* c, hipFunction_t f, const clawgpu_dev_hooks* h, double* q,
                    double xlower, double ylower, double dx, double dy, double t, double dt,
                    const int* mthbc);

}  // namespace clawgpu
