// User Riemann solvers: the loaded code object (kernels exported by
// include/clawgpu_solver.h) and the launches that replace the built-in
// solvers' kernels for a context created with clawgpu_create_user.
#pragma once

#include <hip/hip_runtime.h>

#include "clawgpu.h"
#include "common.hpp"

namespace clawgpu {
struct SweepArgs;
namespace s3u {
struct Sweep3U;
}
}

struct clawgpu_solver2 {
    hipModule_t mod = nullptr;
    hipFunction_t step = nullptr, loop1 = nullptr, loop2 = nullptr;
    hipFunction_t sweep_x = nullptr, sweep_y = nullptr, probe = nullptr;
    int meqn = 0, mwaves = 0, fast = 0;
};

// A 3D user solver (split step only): the x and y/z sweep kernels.
struct clawgpu_solver3 {
    hipModule_t mod = nullptr;
    hipFunction_t sweep_x = nullptr, sweep_yz = nullptr;
    int meqn = 0, mwaves = 0, maux = 0, fast = 0;
};

namespace clawgpu {

// One sweep of a 3D user solver (dir 0 x, 1 y, 2 z); the HIP error code.
int usolver3_sweep(const clawgpu_solver3* s, int dir, const s3u::Sweep3U& a, int nblocks,
                   hipStream_t st);

// Waves of the step kernel one compute unit holds (0 if unknown).
int usolver_waves_per_cu(const clawgpu_solver2* s);
// The launches; each returns the HIP error code of the launch.
int usolver_step(const clawgpu_solver2* s, const StepArgs& a, int nblocks, hipStream_t st);
int usolver_loop(const clawgpu_solver2* s, const StepArgs& a, const LoopArgs& l, int nblocks,
                 hipStream_t st);
int usolver_sweep(const clawgpu_solver2* s, int dir, const SweepArgs& a, int nblocks,
                  hipStream_t st);
int usolver_probe(const clawgpu_solver2* s, int n, int ixy, const SolverParams& P,
                  const double* ql, const double* qr, const double* asdq, double* wave,
                  double* sp, double* amdq, double* apdq, double* bm, double* bp, hipStream_t st);

}  // namespace clawgpu
