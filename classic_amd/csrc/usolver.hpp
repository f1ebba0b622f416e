// User Riemann solvers: the loaded code object (kernels exported by
// include/clawgpu_solver.h) and the launches that replace the built-in
// solvers' kernels for a context created with clawgpu_create_user.
#pragma once

#include <hip/hip_runtime.h>

#include "clawgpu.h"
#include "common.hpp"

namespace clawgpu {
struct SweepArgs;
}

struct clawgpu_solver2 {
    hipModule_t mod = nullptr;
    hipFunction_t step = nullptr, loop1 = nullptr, loop2 = nullptr;
    hipFunction_t sweep_x = nullptr, sweep_y = nullptr, probe = nullptr;
    int meqn = 0, mwaves = 0, fast = 0;
};

namespace clawgpu {

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
