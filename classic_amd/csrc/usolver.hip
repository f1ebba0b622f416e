// User Riemann solvers (the rpn2/rpt2 a caller links into claw2,
// src/2d/claw2.f:312-402): a gfx950 code object whose kernels are the
// library's step, loop, sweep and probe kernel templates instantiated with
// the user's solver policy (include/clawgpu_solver.h).  Loading checks the
// argument layout the object was built against; a context created with
// clawgpu_create_user launches these kernels where a built-in solver's
// context launches its own.
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>

#include "ctx.hpp"
#include "split2d_kernel.hpp"
#include "step3d_user.hpp"
#include "usolver.hpp"

extern "C" int clawgpu_solver2_load(clawgpu_solver2** out, const char* path, const char* name)
{
    if (!out || !path || !name || strlen(name) > 200) return set_err(CLAWGPU_ERR_ARG, "bad argument");
    *out = nullptr;
    clawgpu_solver2* s = new (std::nothrow) clawgpu_solver2();
    if (!s) return set_err(CLAWGPU_ERR_ARG, "out of host memory");
    hipError_t e = hipModuleLoad(&s->mod, path);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // reported here: later launch checks must not see it
        snprintf(g_errbuf, sizeof g_errbuf, "cannot load code object %s: %s", path,
                 hipGetErrorString(e));
        delete s;
        return CLAWGPU_ERR_ARG;
    }
    char sym[256];
    int info[8] = {0};
    hipDeviceptr_t dinfo = nullptr;
    size_t bytes = 0;
    snprintf(sym, sizeof sym, "%s_info", name);
    e = hipModuleGetGlobal(&dinfo, &bytes, s->mod, sym);
    if (e == hipSuccess && bytes == sizeof info) e = hipMemcpy(info, dinfo, sizeof info, hipMemcpyDeviceToHost);
    if (e != hipSuccess || bytes != sizeof info) {
        (void)hipGetLastError();
        snprintf(g_errbuf, sizeof g_errbuf, "solver %s not found in %s", name, path);
        clawgpu_solver2_destroy(s);
        return CLAWGPU_ERR_ARG;
    }
    if (info[0] != 1 || info[4] != (int)sizeof(StepArgs) || info[5] != (int)sizeof(LoopArgs) ||
        info[6] != (int)sizeof(SweepArgs) || info[7] != (int)sizeof(SolverParams)) {
        clawgpu_solver2_destroy(s);
        return set_err(CLAWGPU_ERR_ARG, "the code object was built against other kernel arguments; rebuild it");
    }
    s->meqn = info[1];
    s->mwaves = info[2];
    s->fast = info[3];
    const char* suffix[6] = {"step", "loop1", "loop2", "sweep_x", "sweep_y", "probe"};
    hipFunction_t* fns[6] = {&s->step, &s->loop1, &s->loop2, &s->sweep_x, &s->sweep_y, &s->probe};
    for (int k = 0; k < 6; k++) {
        snprintf(sym, sizeof sym, "%s_%s", name, suffix[k]);
        if (hipModuleGetFunction(fns[k], s->mod, sym) != hipSuccess) {
            (void)hipGetLastError();
            snprintf(g_errbuf, sizeof g_errbuf, "kernel %s not found in %s", sym, path);
            clawgpu_solver2_destroy(s);
            return CLAWGPU_ERR_ARG;
        }
    }
    *out = s;
    return CLAWGPU_OK;
}

extern "C" int clawgpu_solver2_destroy(clawgpu_solver2* s)
{
    if (!s) return CLAWGPU_OK;
    if (s->mod) (void)hipModuleUnload(s->mod);
    delete s;
    return CLAWGPU_OK;
}

extern "C" int clawgpu_solver2_shape(const clawgpu_solver2* s, int* meqn, int* mwaves, int* arith)
{
    if (!s) return set_err(CLAWGPU_ERR_ARG, "null solver");
    if (meqn) *meqn = s->meqn;
    if (mwaves) *mwaves = s->mwaves;
    if (arith) *arith = s->fast ? CLAWGPU_ARITH_FAST : CLAWGPU_ARITH_EXACT;
    return CLAWGPU_OK;
}

// 3D solvers (split step): NAME_info3, NAME_sweep3_x, NAME_sweep3_yz.
extern "C" int clawgpu_solver3_load(clawgpu_solver3** out, const char* path, const char* name)
{
    if (!out || !path || !name || strlen(name) > 200) return set_err(CLAWGPU_ERR_ARG, "bad argument");
    *out = nullptr;
    clawgpu_solver3* s = new (std::nothrow) clawgpu_solver3();
    if (!s) return set_err(CLAWGPU_ERR_ARG, "out of host memory");
    hipError_t e = hipModuleLoad(&s->mod, path);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        snprintf(g_errbuf, sizeof g_errbuf, "cannot load code object %s: %s", path,
                 hipGetErrorString(e));
        delete s;
        return CLAWGPU_ERR_ARG;
    }
    char sym[256];
    int info[8] = {0};
    hipDeviceptr_t dinfo = nullptr;
    size_t bytes = 0;
    snprintf(sym, sizeof sym, "%s_info3", name);
    e = hipModuleGetGlobal(&dinfo, &bytes, s->mod, sym);
    if (e == hipSuccess && bytes == sizeof info) e = hipMemcpy(info, dinfo, sizeof info, hipMemcpyDeviceToHost);
    if (e != hipSuccess || bytes != sizeof info) {
        (void)hipGetLastError();
        snprintf(g_errbuf, sizeof g_errbuf, "3D solver %s not found in %s", name, path);
        clawgpu_solver3_destroy(s);
        return CLAWGPU_ERR_ARG;
    }
    if (info[0] != s3u::kAbi || info[5] != (int)sizeof(s3u::Sweep3U) ||
        info[6] != (int)sizeof(SolverParams)) {
        clawgpu_solver3_destroy(s);
        return set_err(CLAWGPU_ERR_ARG, "the code object was built against other kernel arguments; rebuild it");
    }
    if (info[1] < 1 || info[1] > kMaxWaves || info[2] < 1 || info[2] > kMaxWaves || info[3] < 0) {
        clawgpu_solver3_destroy(s);
        return set_err(CLAWGPU_ERR_ARG, "a 3D solver needs 1..8 equations and waves");
    }
    s->meqn = info[1];
    s->mwaves = info[2];
    s->maux = info[3];
    s->fast = info[4];
    const char* suffix[2] = {"sweep3_x", "sweep3_yz"};
    hipFunction_t* fns[2] = {&s->sweep_x, &s->sweep_yz};
    for (int k = 0; k < 2; k++) {
        snprintf(sym, sizeof sym, "%s_%s", name, suffix[k]);
        if (hipModuleGetFunction(fns[k], s->mod, sym) != hipSuccess) {
            (void)hipGetLastError();
            snprintf(g_errbuf, sizeof g_errbuf, "kernel %s not found in %s", sym, path);
            clawgpu_solver3_destroy(s);
            return CLAWGPU_ERR_ARG;
        }
    }
    *out = s;
    return CLAWGPU_OK;
}

extern "C" int clawgpu_solver3_destroy(clawgpu_solver3* s)
{
    if (!s) return CLAWGPU_OK;
    if (s->mod) (void)hipModuleUnload(s->mod);
    delete s;
    return CLAWGPU_OK;
}

extern "C" int clawgpu_solver3_shape(const clawgpu_solver3* s, int* meqn, int* mwaves, int* maux,
                                     int* arith)
{
    if (!s) return set_err(CLAWGPU_ERR_ARG, "null solver");
    if (meqn) *meqn = s->meqn;
    if (mwaves) *mwaves = s->mwaves;
    if (maux) *maux = s->maux;
    if (arith) *arith = s->fast ? CLAWGPU_ARITH_FAST : CLAWGPU_ARITH_EXACT;
    return CLAWGPU_OK;
}

namespace clawgpu {

int usolver3_sweep(const clawgpu_solver3* s, int dir, const s3u::Sweep3U& a, int nblocks,
                   hipStream_t st)
{
    s3u::Sweep3U x = a;
    int d = dir;
    void* args_x[] = {&x};
    void* args_yz[] = {&x, &d};
    const hipError_t e =
        hipModuleLaunchKernel(dir == 0 ? s->sweep_x : s->sweep_yz, (unsigned)nblocks, 1, 1, 256, 1,
                              1, 0, st, dir == 0 ? args_x : args_yz, nullptr);
    if (e != hipSuccess) (void)hipGetLastError();
    return (int)e;
}

int usolver_waves_per_cu(const clawgpu_solver2* s)
{
    int nb = 0;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nb, s->step, 256, 0) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return nb * kWavesPerBlock;
}

static int launch(hipFunction_t f, int nblocks, int block, hipStream_t st, void** args)
{
    const hipError_t e =
        hipModuleLaunchKernel(f, (unsigned)nblocks, 1, 1, (unsigned)block, 1, 1, 0, st, args, nullptr);
    if (e != hipSuccess) (void)hipGetLastError();  // returned to the caller, which reports it
    return (int)e;
}

int usolver_step(const clawgpu_solver2* s, const StepArgs& a, int nblocks, hipStream_t st)
{
    StepArgs x = a;
    void* args[] = {&x};
    return launch(s->step, nblocks, 256, st, args);
}

int usolver_loop(const clawgpu_solver2* s, const StepArgs& a, const LoopArgs& l, int nblocks,
                 hipStream_t st)
{
    bool walls = false;
    for (int e = 0; e < 4; e++) walls = walls || a.mthbc[e] == 3;
    StepArgs x = a;
    LoopArgs y = l;
    void* args[] = {&x, &y};
    return launch(walls ? s->loop2 : s->loop1, nblocks, 256, st, args);
}

int usolver_sweep(const clawgpu_solver2* s, int dir, const SweepArgs& a, int nblocks, hipStream_t st)
{
    SweepArgs x = a;
    void* args[] = {&x};
    return launch(dir == 1 ? s->sweep_x : s->sweep_y, nblocks, 256, st, args);
}

int usolver_probe(const clawgpu_solver2* s, int n, int ixy, const SolverParams& P,
                  const double* ql, const double* qr, const double* asdq, double* wave,
                  double* sp, double* amdq, double* apdq, double* bm, double* bp, hipStream_t st)
{
    SolverParams p = P;
    void* args[] = {&n, &ixy, &p, &ql, &qr, &asdq, &wave, &sp, &amdq, &apdq, &bm, &bp};
    return launch(s->probe, (n + 127) / 128, 128, st, args);
}

}  // namespace clawgpu
