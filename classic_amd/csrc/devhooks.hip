// Device-side user routines (bc2, b4step2, src2 of the plug-in contract,
// src/2d/claw2.f:620-737) loaded from a user's gfx950 code object and
// launched on the context's stream on the device q: the claw2 loop of
// clawgpu_claw2_dev_hooks calls them where the reference calls the Fortran
// routines, with no host copy of q.
#include <hip/hip_runtime.h>

#include <new>

#include "ctx.hpp"
#include "devhooks.hpp"

extern "C" int clawgpu_dev_hooks_load(clawgpu_dev_hooks** out, const char* path, const char* bc2,
                                      const char* b4step2, const char* src2, int nuser)
{
    if (!out || !path || nuser < 0) return set_err(CLAWGPU_ERR_ARG, "bad argument");
    *out = nullptr;
    clawgpu_dev_hooks* h = new (std::nothrow) clawgpu_dev_hooks();
    if (!h) return set_err(CLAWGPU_ERR_ARG, "out of host memory");
    hipError_t e = hipModuleLoad(&h->mod, path);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // reported here: later launch checks must not see it
        snprintf(g_errbuf, sizeof g_errbuf, "cannot load code object %s: %s", path,
                 hipGetErrorString(e));
        delete h;
        return CLAWGPU_ERR_ARG;
    }
    const char* names[3] = {bc2, b4step2, src2};
    hipFunction_t* fns[3] = {&h->bc2, &h->b4step2, &h->src2};
    for (int k = 0; k < 3; k++) {
        if (!names[k]) continue;
        e = hipModuleGetFunction(fns[k], h->mod, names[k]);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            snprintf(g_errbuf, sizeof g_errbuf, "kernel %s not found in %s", names[k], path);
            clawgpu_dev_hooks_destroy(h);
            return CLAWGPU_ERR_ARG;
        }
    }
    h->nuser = nuser;
    if (nuser > 0) {
        e = hipMalloc((void**)&h->user, (size_t)nuser * sizeof(double));
        if (e == hipSuccess) e = hipMemset(h->user, 0, (size_t)nuser * sizeof(double));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            clawgpu_dev_hooks_destroy(h);
            return set_err(CLAWGPU_ERR_HIP, "user array allocation failed");
        }
    }
    *out = h;
    return CLAWGPU_OK;
}

extern "C" int clawgpu_dev_hooks_destroy(clawgpu_dev_hooks* h)
{
    if (!h) return CLAWGPU_OK;
    (void)hipDeviceSynchronize();
    if (h->user) (void)hipFree(h->user);
    if (h->mod) (void)hipModuleUnload(h->mod);
    delete h;
    return CLAWGPU_OK;
}

extern "C" int clawgpu_dev_hooks_user(clawgpu_dev_hooks* h, const double* in, double* out, int n)
{
    if (!h || n < 0 || n > h->nuser) return set_err(CLAWGPU_ERR_ARG, "bad argument");
    if (n == 0) return CLAWGPU_OK;
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess && in) e = hipMemcpy(h->user, in, (size_t)n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess && out) e = hipMemcpy(out, h->user, (size_t)n * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return set_err(CLAWGPU_ERR_HIP, hipGetErrorString(e));
    return CLAWGPU_OK;
}

namespace clawgpu {

int launch_dev_hook(clawgpu_ctx* c, hipFunction_t f, const clawgpu_dev_hooks* h, double* q,
                    double xlower, double ylower, double dx, double dy, double t, double dt,
                    const int* mthbc)
{
    clawgpu_dev_grid g;
    g.q = q;
    g.aux = c->maux > 0 ? c->daux : nullptr;
    g.user = h->user;
    g.fs = c->L.fs;
    g.pitch = c->L.pitch;
    g.meqn = c->meqn;
    g.maux = c->maux;
    g.mbc = c->mbc;
    g.mx = c->mx;
    g.my = c->my;
    for (int k = 0; k < 4; k++) g.mthbc[k] = mthbc[k];
    g.xlower = xlower;
    g.ylower = ylower;
    g.dx = dx;
    g.dy = dy;
    g.t = t;
    g.dt = dt;
    const long n = (long)(c->mx + 2 * c->mbc) * (c->my + 2 * c->mbc);
    void* args[] = {&g};
    const hipError_t e = hipModuleLaunchKernel(f, (unsigned)((n + 255) / 256), 1, 1, 256, 1, 1, 0,
                                               c->stream, args, nullptr);
    if (e != hipSuccess) return set_err(CLAWGPU_ERR_HIP, "user routine launch failed");
    return CLAWGPU_OK;
}

}  // namespace clawgpu
