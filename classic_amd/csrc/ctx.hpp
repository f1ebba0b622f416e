// Device context shared by the host-side translation units (api.hip,
// learned.hip): the grid, its device buffers and stream, and the error
// reporting helpers of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#include "common.hpp"
#include "../../include/clawgpu.h"

using namespace clawgpu;

struct clawgpu_solver2;

struct clawgpu_ctx {
    int meqn, mwaves, maux, mbc, mx, my, solver;
    SolverParams P;
    Layout2 L;
    double* dq[3];  // q and two work copies: qold/qnew of a step, sweeps of a split step
    int cur;
    double* staging;
    unsigned long long* dcfl;
    double* hcfl;  // pinned
    hipStream_t stream;
    int seglen;       // rows per wavefront segment of the unsplit step; 0: chosen per launch
    int ncu = 0;      // compute units of the context's device
    int auto_key[3] = {-1, -1, -1};  // (arith, fixed config, rows) the cached choice was made for
    int auto_rows = 32;
    int arith;
    double* daux = nullptr;   // aux planes (solvers with aux arrays)
    std::vector<double> haux;  // host copy of aux (reference layout) for user hooks
    double* grec = nullptr;   // records of the general step (step2g.hip)
    // learned limiter state (allocated on first use)
    double *phix = nullptr, *phiy = nullptr;  // phi tables read by the step kernel
    void* learned = nullptr;                   // LearnedState (learned.hip)
    // device-side claw2 loop: states before attempts of each parity, the
    // Courant slots of three consecutive attempts, pinned read-back copies
    LoopState* dloop = nullptr;
    unsigned long long* dcflb = nullptr;
    LoopState* hloop = nullptr;  // 3 entries
    double* dsave = nullptr;     // q before a Strang source half step (device user routines)
    const clawgpu_solver2* us = nullptr;  // user solver (CLAWGPU_SOLVER_USER), not owned
    int last_hip_error;
};

extern thread_local char g_errbuf[256];

inline int set_err(int code, const char* msg)
{
    snprintf(g_errbuf, sizeof g_errbuf, "%s", msg);
    return code;
}

// Rows per wavefront segment of the unsplit step for a grid of nstrips
// strips and my rows on ncu compute units holding waves_per_cu waves each.
int segment_rows_for(int ncu, int waves_per_cu, int nstrips, int my);

// SolverParams of a 2D solver from the (rho, bulk) / (gamma, efix) argument
// pair of clawgpu_create; CLAWGPU_OK or an error code with the message set.
int solver_params(int solver_id, const double* params, int nparams, SolverParams* P);

#define HIPCHK(ctx, call)                                                              \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            snprintf(g_errbuf, sizeof g_errbuf, "%s: %s", #call, hipGetErrorString(e_)); \
            if (ctx) (ctx)->last_hip_error = (int)e_;                                  \
            return CLAWGPU_ERR_HIP;                                                    \
        }                                                                              \
    } while (0)

