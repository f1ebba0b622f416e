// Host side of libclawgpu.so: device contexts, the C-ABI of
// include/clawgpu.h, the ghost-cell fill (bc2) and layout conversion kernels.
// The fused step kernels live in step2d_kernel.hpp (instantiated by
// step2d_exact.hip and step2d_fast.hip).

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>

#include "common.hpp"
#include "bc2d.hpp"
#include "loop2d.hpp"
#include "layout2d.hpp"
#include "split2d_kernel.hpp"
#include "../../include/clawgpu.h"

namespace clawgpu {

// ------------------------------------------------------------------ bc2
// Every ghost cell is written from the interior in one pass: the row index
// is mapped first (bottom/top), then the column (left/right), with the wall
// sign flips of q(3) and q(2).  This equals the reference's sequential fill
// (left, right, bottom, top; corners from the bottom/top pass over the
// already-filled ghost columns) because each side's rule is an index map
// plus a sign.  Thread 0 also clears the Courant accumulator for the step.
__global__ void bc2_kernel(BcArgs a)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0 && a.cfl_bits)
        for (int k = 0; k < 4; k++) a.cfl_bits[k] = 0ull;
    bc2_cell(a.q, a.L, a.meqn, a.mthbc, t);
}

// The decision of the last attempt of a device-side claw2 loop (attempt
// l.k - 1), stored in l.st[l.k & 1].
__global__ void loop2_finish_kernel(LoopArgs l)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    LoopState s;
    loop_next(l, l.st[(l.k - 1) & 1], loop_cfl(l, l.k - 1), s);
    l.st[l.k & 1] = s;
}

int launch_loop2_finish(const LoopArgs& l, hipStream_t s)
{
    hipLaunchKernelGGL(loop2_finish_kernel, dim3(1), dim3(64), 0, s, l);
    return (int)hipGetLastError();
}

}  // namespace clawgpu

// =====================================================================
// Host side: context, C-ABI.
// =====================================================================
using namespace clawgpu;

#include "ctx.hpp"
#include "devhooks.hpp"
#include "step2g.hpp"
#include "usolver.hpp"

thread_local char g_errbuf[256];

// learned limiter (learned.hip)
void learned_release(clawgpu_ctx* c);
int learned_prepare(clawgpu_ctx* c, double dx, double dy, double dt);
bool learned_ready(clawgpu_ctx* c);

extern "C" const char* clawgpu_last_error(void) { return g_errbuf; }

extern "C" int clawgpu_abi_version(void) { return CLAWGPU_ABI_VERSION; }

int solver_params(int solver_id, const double* params, int nparams, SolverParams* P)
{
    *P = SolverParams{};
    if (solver_id == CLAWGPU_SOLVER_ACOUSTICS) {
        if (nparams < 2 || !params) return set_err(CLAWGPU_ERR_ARG, "acoustics needs (rho, bulk)");
        P->rho = params[0];
        P->bulk = params[1];
        if (!(P->rho > 0.0) || !(P->bulk > 0.0)) return set_err(CLAWGPU_ERR_ARG, "rho and bulk must be positive");
        P->cc = sqrt(P->bulk / P->rho);
        P->zz = P->rho * P->cc;
    } else if (solver_id == CLAWGPU_SOLVER_EULER) {
        if (nparams < 2 || !params) return set_err(CLAWGPU_ERR_ARG, "euler needs (gamma, efix)");
        P->gamma = params[0];
        P->gamma1 = params[0] - 1.0;
        P->efix = params[1] != 0.0;
        if (!(P->gamma > 1.0)) return set_err(CLAWGPU_ERR_ARG, "gamma must exceed 1");
    }
    return CLAWGPU_OK;
}

static int solver_shape(int solver, int* meqn, int* mwaves)
{
    switch (solver) {
    case CLAWGPU_SOLVER_ACOUSTICS:
        *meqn = 3;
        *mwaves = 2;
        return 0;
    case CLAWGPU_SOLVER_EULER:
        *meqn = 4;
        *mwaves = 3;
        return 0;
    case CLAWGPU_SOLVER_VC_ADVECTION:
        *meqn = 1;
        *mwaves = 1;
        return 0;
    default: return -1;
    }
}

static int create_ctx(clawgpu_ctx** out, int meqn, int mwaves, int maux, int mbc, int mx, int my,
                      int solver_id, const SolverParams& P, const clawgpu_solver2* us);

extern "C" int clawgpu_create(clawgpu_ctx** out, int meqn, int mwaves, int maux, int mbc, int mx,
                              int my, int solver_id, const double* params, int nparams)
{
    if (!out) return set_err(CLAWGPU_ERR_ARG, "null output pointer");
    *out = nullptr;
    int sm, sw;
    if (solver_shape(solver_id, &sm, &sw) != 0) return set_err(CLAWGPU_ERR_UNSUPPORTED, "unknown solver id");
    if (meqn != sm || mwaves != sw)
        return set_err(CLAWGPU_ERR_ARG, "meqn/mwaves do not match the solver");
    const bool general = solver_id == CLAWGPU_SOLVER_VC_ADVECTION;
    if (!general && maux != 0)
        return set_err(CLAWGPU_ERR_UNSUPPORTED, "aux arrays are not supported by this solver");
    if (general && maux < 2)
        return set_err(CLAWGPU_ERR_ARG, "vc advection needs maux >= 2 (edge velocities u, v)");
    if (mbc < 2) return set_err(CLAWGPU_ERR_ARG, "mbc must be >= 2 (claw2 info=1)");
    if (mx < 1 || my < 1) return set_err(CLAWGPU_ERR_ARG, "grid must have at least one cell");
    SolverParams P{};
    const int prc = solver_params(solver_id, params, nparams, &P);
    if (prc) return prc;
    return create_ctx(out, meqn, mwaves, maux, mbc, mx, my, solver_id, P, nullptr);
}

// The context of a validated grid and solver: buffers, stream, layout.
static int create_ctx(clawgpu_ctx** out, int meqn, int mwaves, int maux, int mbc, int mx, int my,
                      int solver_id, const SolverParams& P, const clawgpu_solver2* us)
{
    const bool general = solver_id == CLAWGPU_SOLVER_VC_ADVECTION;
    clawgpu_ctx* c = new (std::nothrow) clawgpu_ctx();
    if (!c) return set_err(CLAWGPU_ERR_ARG, "out of host memory");
    c->meqn = meqn;
    c->mwaves = mwaves;
    c->maux = maux;
    c->mbc = mbc;
    c->mx = mx;
    c->my = my;
    c->solver = solver_id;
    c->us = us;
    c->P = P;
    c->L.mx = mx;
    c->L.my = my;
    c->L.mbc = mbc;
    c->L.pitch = ((mx + 2 * mbc) + 15) / 16 * 16;
    c->L.fs = ((long)c->L.pitch * (my + 2 * mbc) + 31) / 32 * 32;
    c->cur = 0;
    c->seglen = 0;
    c->arith = us && us->fast ? ARITH_FAST : ARITH_EXACT;
    {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            c->ncu = 0;
    }
    size_t qbytes = (size_t)c->L.fs * meqn * sizeof(double);
    size_t abytes = (size_t)(mx + 2 * mbc) * (my + 2 * mbc) * (meqn > maux ? meqn : maux) * sizeof(double);
    hipError_t e = hipSuccess;
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipMalloc((void**)&c->dq[0], qbytes)) != hipSuccess ||
        (e = hipMalloc((void**)&c->dq[1], qbytes)) != hipSuccess ||
        (e = hipMalloc((void**)&c->dq[2], qbytes)) != hipSuccess ||
        (e = hipMalloc((void**)&c->staging, abytes)) != hipSuccess ||
        (e = hipMalloc((void**)&c->dcfl, 4 * sizeof(unsigned long long))) != hipSuccess ||
        (e = hipHostMalloc((void**)&c->hcfl, 4 * sizeof(double), hipHostMallocDefault)) != hipSuccess ||
        (e = hipMalloc((void**)&c->dloop, 2 * sizeof(LoopState))) != hipSuccess ||
        (e = hipMalloc((void**)&c->dcflb, 3 * sizeof(unsigned long long))) != hipSuccess ||
        (e = hipHostMalloc((void**)&c->hloop, 3 * sizeof(LoopState), hipHostMallocDefault)) != hipSuccess ||
        (general && (e = hipMalloc((void**)&c->daux, (size_t)c->L.fs * maux * sizeof(double))) != hipSuccess) ||
        (general && (e = hipMalloc((void**)&c->grec, (size_t)c->L.fs * 10 * sizeof(double))) != hipSuccess)) {
        snprintf(g_errbuf, sizeof g_errbuf, "allocation failed: %s", hipGetErrorString(e));
        clawgpu_destroy(c);
        return CLAWGPU_ERR_HIP;
    }
    (void)hipMemsetAsync(c->dq[0], 0, qbytes, c->stream);
    (void)hipMemsetAsync(c->dq[1], 0, qbytes, c->stream);
    (void)hipMemsetAsync(c->dq[2], 0, qbytes, c->stream);
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) {
        snprintf(g_errbuf, sizeof g_errbuf, "init failed: %s", hipGetErrorString(e));
        clawgpu_destroy(c);
        return CLAWGPU_ERR_HIP;
    }
    *out = c;
    return CLAWGPU_OK;
}

extern "C" int clawgpu_create_user(clawgpu_ctx** out, const clawgpu_solver2* us, int mbc, int mx,
                                   int my, const double* params, int nparams)
{
    if (!out) return set_err(CLAWGPU_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (!us) return set_err(CLAWGPU_ERR_ARG, "null solver");
    if (nparams < 0 || nparams > kUserParams || (nparams > 0 && !params))
        return set_err(CLAWGPU_ERR_ARG, "a user solver takes 0..16 parameters");
    if (mbc < 2) return set_err(CLAWGPU_ERR_ARG, "mbc must be >= 2 (claw2 info=1)");
    if (mx < 1 || my < 1) return set_err(CLAWGPU_ERR_ARG, "grid must have at least one cell");
    SolverParams P{};
    P.nuser = nparams;
    for (int k = 0; k < nparams; k++) P.user[k] = params[k];
    return create_ctx(out, us->meqn, us->mwaves, 0, mbc, mx, my, CLAWGPU_SOLVER_USER, P, us);
}

extern "C" int clawgpu_destroy(clawgpu_ctx* c)
{
    if (!c) return CLAWGPU_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    learned_release(c);
    if (c->dq[0]) (void)hipFree(c->dq[0]);
    if (c->dq[1]) (void)hipFree(c->dq[1]);
    if (c->dq[2]) (void)hipFree(c->dq[2]);
    if (c->staging) (void)hipFree(c->staging);
    if (c->dcfl) (void)hipFree(c->dcfl);
    if (c->hcfl) (void)hipHostFree(c->hcfl);
    if (c->daux) (void)hipFree(c->daux);
    if (c->grec) (void)hipFree(c->grec);
    if (c->dloop) (void)hipFree(c->dloop);
    if (c->dcflb) (void)hipFree(c->dcflb);
    if (c->dsave) (void)hipFree(c->dsave);
    if (c->hloop) (void)hipHostFree(c->hloop);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return CLAWGPU_OK;
}

static long ncells_ghost(const clawgpu_ctx* c)
{
    return (long)(c->mx + 2 * c->mbc) * (c->my + 2 * c->mbc);
}

extern "C" int clawgpu_set_q(clawgpu_ctx* c, const double* q)
{
    if (!c || !q) return set_err(CLAWGPU_ERR_ARG, "null argument");
    long n = ncells_ghost(c);
    HIPCHK(c, hipMemcpyAsync(c->staging, q, (size_t)n * c->meqn * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
    int nb = (int)((n + 255) / 256);
    hipLaunchKernelGGL(aos_to_planes, dim3(nb), dim3(256), 0, c->stream, c->staging, c->dq[c->cur],
                       c->L, c->meqn);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLAWGPU_OK;
}

extern "C" int clawgpu_set_aux(clawgpu_ctx* c, const double* aux)
{
    if (!c || !aux) return set_err(CLAWGPU_ERR_ARG, "null argument");
    if (!c->daux) return set_err(CLAWGPU_ERR_ARG, "context has no aux arrays (maux = 0)");
    long n = ncells_ghost(c);
    HIPCHK(c, hipMemcpyAsync(c->staging, aux, (size_t)n * c->maux * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
    int nb = (int)((n + 255) / 256);
    hipLaunchKernelGGL(aos_to_planes, dim3(nb), dim3(256), 0, c->stream, c->staging, c->daux, c->L,
                       c->maux);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->haux.assign(aux, aux + (size_t)n * c->maux);
    return CLAWGPU_OK;
}

extern "C" int clawgpu_get_q(clawgpu_ctx* c, double* q)
{
    if (!c || !q) return set_err(CLAWGPU_ERR_ARG, "null argument");
    long n = ncells_ghost(c);
    int nb = (int)((n + 255) / 256);
    hipLaunchKernelGGL(planes_to_aos, dim3(nb), dim3(256), 0, c->stream, c->dq[c->cur], c->staging,
                       c->L, c->meqn);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(q, c->staging, (size_t)n * c->meqn * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLAWGPU_OK;
}

static int check_bc(const int* mthbc)
{
    for (int k = 0; k < 4; k++)
        if (mthbc[k] < 1 || mthbc[k] > 3) return -1;
    if ((mthbc[0] == 2) != (mthbc[1] == 2)) return -1;
    if ((mthbc[2] == 2) != (mthbc[3] == 2)) return -1;
    return 0;
}

// Walls negate components 2 and 3 (1-based), as bc2 does: a user solver
// with fewer equations has no such components.
static int check_walls(const clawgpu_ctx* c, const int* mthbc)
{
    if (c->us && c->meqn < 3)
        for (int k = 0; k < 4; k++)
            if (mthbc[k] == 3) return set_err(CLAWGPU_ERR_ARG, "wall boundaries (mthbc = 3) need meqn >= 3");
    return CLAWGPU_OK;
}

static int enqueue_bc(clawgpu_ctx* c, double* q, const int* mthbc, bool clear_cfl)
{
    BcArgs b;
    b.q = q;
    b.L = c->L;
    b.meqn = c->meqn;
    for (int k = 0; k < 4; k++) b.mthbc[k] = mthbc[k];
    b.cfl_bits = clear_cfl ? c->dcfl : nullptr;
    int ng = 2 * c->mbc * (c->mx + 2 * c->mbc) + 2 * c->mbc * c->my;
    hipLaunchKernelGGL(bc2_kernel, dim3((ng + 255) / 256), dim3(256), 0, c->stream, b);
    HIPCHK(c, hipGetLastError());
    return CLAWGPU_OK;
}

extern "C" int clawgpu_bc2(clawgpu_ctx* c, const int* mthbc)
{
    if (!c || !mthbc) return set_err(CLAWGPU_ERR_ARG, "null argument");
    if (check_bc(mthbc) != 0)
        return set_err(CLAWGPU_ERR_UNSUPPORTED, "mthbc must be 1, 2 or 3 (periodic on both sides)");
    if (const int wrc = check_walls(c, mthbc)) return wrc;
    int rc = enqueue_bc(c, c->dq[c->cur], mthbc, false);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLAWGPU_OK;
}

static bool general_path(const clawgpu_ctx* c) { return c->solver == CLAWGPU_SOLVER_VC_ADVECTION; }

static int check_method(const clawgpu_ctx* c, const int* method, const int* mthlim)
{
    if (method[1] != 1 && method[1] != 2) return set_err(CLAWGPU_ERR_ARG, "method(2) must be 1 or 2");
    if (method[2] < 0 || method[2] > 2)
        return set_err(CLAWGPU_ERR_UNSUPPORTED, "method(3) must be 0, 1 or 2 for the unsplit step");
    if (method[5] != 0 && !general_path(c))
        return set_err(CLAWGPU_ERR_UNSUPPORTED, "capacity function is not supported by this solver");
    if (method[5] < 0 || method[5] > c->maux)
        return set_err(CLAWGPU_ERR_ARG, "method(6) must be 0..maux");
    if (general_path(c) && mthlim[0] == kLimTable)
        return set_err(CLAWGPU_ERR_UNSUPPORTED, "the learned limiter runs with the built-in solvers only");
    if (c->mwaves > kMaxWaves) return set_err(CLAWGPU_ERR_UNSUPPORTED, "too many waves");
    for (int mw = 0; mw < c->mwaves; mw++) {
        if (mthlim[mw] == kLimTable && !learned_ready(const_cast<clawgpu_ctx*>(c)))
            return set_err(CLAWGPU_ERR_ARG, "limiter 6 (learned) needs clawgpu_learned_set_params");
        if (mthlim[mw] < 0 || mthlim[mw] > kLimTable)
            return set_err(CLAWGPU_ERR_ARG, "invalid limiter method");
    }
    return CLAWGPU_OK;
}

static bool uses_table(const clawgpu_ctx* c, const int* mthlim)
{
    for (int mw = 0; mw < c->mwaves; mw++)
        if (mthlim[mw] == kLimTable) return true;
    return false;
}

// Rows per wavefront segment of the unsplit step when none was set (also used
// by clawgpu_step2_noalloc, noalloc.hip).  Every
// wave of the kernel does the same work, a segment of s rows costing s + 3
// rows (halo rows, padded to an even count), and the chip runs the waves in
// rounds of (CUs x waves one CU holds).  The choice minimises rounds x rows
// per wave over s in [8, 72] (longer segments measured slower than this
// count predicts); 4096^2 Euler gets 70 rows: 4 rounds of 74 rows instead
// of 9 rounds of 36 with 32-row segments.
int segment_rows_for(int ncu, int waves_per_cu, int nstrips, int my)
{
    const long slots = (long)(ncu > 0 ? ncu : 1) * (waves_per_cu > 0 ? waves_per_cu : kWavesPerBlock);
    int best = 32;
    long best_cost = -1;
    for (int rows = 8; rows <= 72; rows++) {
        const long waves = (long)nstrips * ((my + rows - 1) / rows);
        const long rounds = (waves + slots - 1) / slots;
        const long cost = rounds * (rows + 3 + ((rows + 3) & 1));
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            best = rows;
        }
    }
    return best;
}

static int step_rows(clawgpu_ctx* c, const StepArgs& a)
{
    if (c->seglen > 0) return c->seglen;
    bool mc = a.order == 2 && a.trans == 2 && !a.use_fwave;
    for (int k = 0; k < c->mwaves; k++) mc = mc && a.lim[k] == 4;
    if (c->auto_key[0] == c->arith && c->auto_key[1] == (int)mc && c->auto_key[2] == c->my)
        return c->auto_rows;
    const int wpc = c->us                    ? usolver_waves_per_cu(c->us)
                    : c->arith == ARITH_FAST ? step2_waves_per_cu_fast(c->solver, a)
                                             : step2_waves_per_cu_exact(c->solver, a);
    const int best = segment_rows_for(c->ncu, wpc, a.nstrips, c->my);
    c->auto_key[0] = c->arith;
    c->auto_key[1] = (int)mc;
    c->auto_key[2] = c->my;
    c->auto_rows = best;
    return best;
}

// Arguments of the unsplit step kernel for this grid and method (buffers,
// dt ratios and the Courant slot left for the caller).
static StepArgs step_args(clawgpu_ctx* c, const int* method, const int* mthlim, int use_fwave)
{
    StepArgs a{};
    a.L = c->L;
    a.order = method[1];
    a.trans = method[2];
    a.use_fwave = use_fwave ? 1 : 0;
    for (int k = 0; k < kMaxWaves; k++) a.lim[k] = k < c->mwaves ? mthlim[k] : 0;
    a.nstrips = (c->mx + kStripOut - 1) / kStripOut;
    a.seglen = step_rows(c, a);
    a.nsegs = (c->my + a.seglen - 1) / a.seglen;
    a.P = c->P;
    a.phix = c->phix;
    a.phiy = c->phiy;
    a.bcmap = 0;
    return a;
}

// Enqueue one unsplit step from dq[cur] into dq[1-cur]; the Courant
// accumulator must have been cleared (by the bc pass or explicitly).
static int enqueue_step(clawgpu_ctx* c, double dx, double dy, double dt, const int* method,
                        const int* mthlim, int use_fwave)
{
    if (general_path(c)) {
        g2::GenArgs g;
        g.q = c->dq[c->cur];
        g.qn = c->dq[(c->cur + 1) % 3];
        g.aux = c->daux;
        g.rec = c->grec;
        g.L = c->L;
        g.afs = c->L.fs;
        g.dtdx = dt / dx;
        g.dtdy = dt / dy;
        g.capa = method[5];
        g.order = method[1];
        g.trans = method[2];
        g.fwave = use_fwave ? 1 : 0;
        g.lim0 = mthlim[0];
        g.cfl_bits = c->dcfl;
        g.cfl_part = c->staging;  // free during a step
        HIPCHK(c, (hipError_t)launch_step2_general(g, c->stream));
        return CLAWGPU_OK;
    }
    StepArgs a = step_args(c, method, mthlim, use_fwave);
    a.qold = c->dq[c->cur];
    a.qnew = c->dq[(c->cur + 1) % 3];
    a.dtdx = dt / dx;
    a.dtdy = dt / dy;
    a.cfl_bits = c->dcfl;
    long waves = (long)a.nstrips * a.nsegs;
    int nb = (int)((waves + kWavesPerBlock - 1) / kWavesPerBlock);
    int e = c->us                    ? usolver_step(c->us, a, nb, c->stream)
            : c->arith == ARITH_FAST ? launch_step2_fast(c->solver, a, nb, c->stream)
                                     : launch_step2_exact(c->solver, a, nb, c->stream);
    HIPCHK(c, (hipError_t)e);
    return CLAWGPU_OK;
}

static int read_cfl(clawgpu_ctx* c, double* cfl)
{
    HIPCHK(c, hipMemcpyAsync(c->hcfl, c->dcfl, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *cfl = *c->hcfl;
    return CLAWGPU_OK;
}

extern "C" int clawgpu_step2(clawgpu_ctx* c, double dx, double dy, double dt, const int* method,
                             const int* mthlim, int use_fwave, double* cfl)
{
    if (!c || !method || !mthlim || !cfl) return set_err(CLAWGPU_ERR_ARG, "null argument");
    int rc = check_method(c, method, mthlim);
    if (rc) return rc;
    HIPCHK(c, hipMemsetAsync(c->dcfl, 0, sizeof(unsigned long long), c->stream));
    if (uses_table(c, mthlim)) {
        rc = learned_prepare(c, dx, dy, dt);
        if (rc) return rc;
    }
    rc = enqueue_step(c, dx, dy, dt, method, mthlim, use_fwave);
    if (rc) return rc;
    rc = read_cfl(c, cfl);
    if (rc) return rc;
    c->cur = (c->cur + 1) % 3;
    return CLAWGPU_OK;
}

// One dimensionally split step (dimsp2, src/2d/dimsp2.f:4-98) from dq[cur]:
// Godunov x(dt) -> y(dt), or Strang x(dt/2) -> y(dt) -> x(dt/2).  Every
// sweep runs; each writes its own Courant number (slots 0, 1, 2) and
// split_cfl() applies dimsp2's early-abort rule to them, which gives the
// reference's cfl for the step (a step the reference abandons half way is
// rejected, so the extra sweep's result is never used).  *result receives
// the buffer holding the new q.
static int enqueue_sweep(clawgpu_ctx* c, int dir, int from, int to, double dtd, const int* method,
                         const int* mthlim, int use_fwave, int slot)
{
    if (general_path(c)) {
        g2::SplitArgs g;
        g.qin = c->dq[from];
        g.qout = c->dq[to];
        g.aux = c->daux;
        g.L = c->L;
        g.afs = c->L.fs;
        g.dtd = dtd;
        g.capa = method[5];
        g.order = method[1];
        g.fwave = use_fwave ? 1 : 0;
        g.lim0 = mthlim[0];
        g.cfl_bits = c->dcfl + slot;
        g.cfl_part = c->staging;  // free during a step
        HIPCHK(c, (hipError_t)launch_split_sweep_general(dir, g, c->stream));
        return CLAWGPU_OK;
    }
    SweepArgs a;
    a.qin = c->dq[from];
    a.qout = c->dq[to];
    a.L = c->L;
    a.dtd = dtd;
    a.order = method[1];
    a.use_fwave = use_fwave ? 1 : 0;
    for (int k = 0; k < kMaxWaves; k++) a.lim[k] = k < c->mwaves ? mthlim[k] : 0;
    a.seglen = c->seglen > 0 ? c->seglen : 32;
    if (dir == 1) {
        a.nstrips = (c->mx + kStripOut - 1) / kStripOut;
        a.nsegs = (c->my + 2 * c->mbc + a.seglen - 1) / a.seglen;
    } else {
        a.nstrips = (c->mx + 2 * c->mbc + 63) / 64;
        a.nsegs = (c->my + a.seglen - 1) / a.seglen;
    }
    a.cfl_bits = c->dcfl + slot;
    a.P = c->P;
    long waves = (long)a.nstrips * a.nsegs;
    int nb = (int)((waves + kWavesPerBlock - 1) / kWavesPerBlock);
    int e = c->us                    ? usolver_sweep(c->us, dir, a, nb, c->stream)
            : c->arith == ARITH_FAST ? launch_sweep_fast(c->solver, dir, a, nb, c->stream)
                                     : launch_sweep_exact(c->solver, dir, a, nb, c->stream);
    HIPCHK(c, (hipError_t)e);
    return CLAWGPU_OK;
}

static int enqueue_split(clawgpu_ctx* c, double dx, double dy, double dt, const int* method,
                         const int* mthlim, int use_fwave, int* result)
{
    const int A = c->cur, B = (c->cur + 1) % 3, C = (c->cur + 2) % 3;
    const bool strang = method[2] == -2;
    const double dtx = strang ? dt / 2.0 : dt;
    int rc = enqueue_sweep(c, 1, A, B, dtx / dx, method, mthlim, use_fwave, 0);
    if (!rc) rc = enqueue_sweep(c, 2, B, C, dt / dy, method, mthlim, use_fwave, 1);
    if (!rc && strang) rc = enqueue_sweep(c, 1, C, B, dtx / dx, method, mthlim, use_fwave, 2);
    *result = strang ? B : C;
    return rc;
}

static double split_cfl(const double* v, bool strang, double cflmax)
{
    double cflx = v[0], cfly = v[1];
    if (cflx > cflmax) return cflx;
    double cfl = fmax(cflx, cfly);
    if (strang) {
        if (cfly > cflmax) return cfly;
        cfl = fmax(cfl, v[2]);
    }
    return cfl;
}

static int read_cfl4(clawgpu_ctx* c, double* v)
{
    HIPCHK(c, hipMemcpyAsync(c->hcfl, c->dcfl, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < 4; k++) v[k] = c->hcfl[k];
    return CLAWGPU_OK;
}

extern "C" int clawgpu_dimsp2(clawgpu_ctx* c, double dx, double dy, double dt, const int* method,
                              const int* mthlim, const double* cflv, int use_fwave, double* cfl)
{
    if (!c || !method || !mthlim || !cflv || !cfl) return set_err(CLAWGPU_ERR_ARG, "null argument");
    if (method[2] != -1 && method[2] != -2) return set_err(CLAWGPU_ERR_ARG, "method(3) must be -1 or -2");
    if (method[1] != 1 && method[1] != 2) return set_err(CLAWGPU_ERR_ARG, "method(2) must be 1 or 2");
    if (method[5] != 0 && !general_path(c))
        return set_err(CLAWGPU_ERR_UNSUPPORTED, "capacity function is not supported by this solver");
    if (method[5] < 0 || method[5] > c->maux) return set_err(CLAWGPU_ERR_ARG, "method(6) must be 0..maux");
    for (int mw = 0; mw < c->mwaves; mw++)
        if (mthlim[mw] < 0 || mthlim[mw] > 5) return set_err(CLAWGPU_ERR_ARG, "invalid limiter method");
    HIPCHK(c, hipMemsetAsync(c->dcfl, 0, 4 * sizeof(unsigned long long), c->stream));
    int res;
    int rc = enqueue_split(c, dx, dy, dt, method, mthlim, use_fwave, &res);
    if (rc) return rc;
    double v[4];
    rc = read_cfl4(c, v);
    if (rc) return rc;
    *cfl = split_cfl(v, method[2] == -2, cflv[0]);
    c->cur = res;
    return CLAWGPU_OK;
}

// The claw2 loop without user routines, decided on the device: each
// attempt of a step is one launch of the loop kernel, which decides the
// previous attempt (accept: its output buffer becomes q and dt follows the
// Courant number; reject: dt shrinks and q stays, claw2.f:695-753) and
// reads the ghost cells through the boundary maps, so nothing returns to
// the host between steps.  The host keeps at most two chunks of attempts in
// flight and reads the state back once per chunk: attempts launched after
// the loop has finished are no-ops.  *out receives the final state.
static int claw2_device(clawgpu_ctx* c, double dx, double dy, double tstart, double tend,
                        const double* dtv, const double* cflv, int maxn, const int* method,
                        const int* mthlim, const int* mthbc, int use_fwave, LoopState* out)
{
    StepArgs a = step_args(c, method, mthlim, use_fwave);
    a.bcmap = 1;
    for (int e = 0; e < 4; e++) a.mthbc[e] = mthbc[e];
    LoopArgs l{};
    for (int b = 0; b < 3; b++) l.q[b] = c->dq[b];
    l.dx = dx;
    l.dy = dy;
    l.tstart = tstart;
    l.tend = tend;
    l.dtv2 = dtv[1];
    l.cflv1 = cflv[0];
    l.cflv2 = cflv[1];
    l.maxn = maxn;
    l.vardt = method[0] == 1;
    l.st = c->dloop;
    l.cflb = c->dcflb;
    // state before the first attempt (claw2.f:596-601)
    LoopState& s0 = c->hloop[2];
    s0 = LoopState{};
    s0.t = s0.told = tstart;
    s0.dt = dtv[0];
    if (tstart + s0.dt > tend && tstart < tend) s0.dt = tend - tstart;
    s0.dtmin = s0.dtmax = dtv[0];
    s0.cur = c->cur;
    HIPCHK(c, hipMemcpyAsync(c->dloop, &s0, sizeof(LoopState), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(c->dcflb, 0, 3 * sizeof(unsigned long long), c->stream));
    struct Evs {
        hipEvent_t e[2] = {nullptr, nullptr};
        ~Evs()
        {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } ev;
    for (int b = 0; b < 2; b++) HIPCHK(c, hipEventCreateWithFlags(&ev.e[b], hipEventDisableTiming));
    const long waves = (long)a.nstrips * a.nsegs;
    const int nb = (int)((waves + kWavesPerBlock - 1) / kWavesPerBlock);
    const bool fast = c->arith == ARITH_FAST;
    int k = 0, snap_k[2] = {0, 0}, qh = 0, qn = 0, n_seen = 0, k_known = 0;
    for (;;) {
        // attempts beyond those already launched that can still be needed,
        // if none of the undecided ones is rejected
        const long est = (long)maxn - n_seen - (k - k_known);
        if (qn < 2 && (est > 0 || qn == 0)) {
            const int m = est > 0 ? (int)(est < 32 ? est : 32) : 1;
            for (int j = 0; j < m; j++) {
                l.k = k++;
                HIPCHK(c, (hipError_t)(c->us ? usolver_loop(c->us, a, l, nb, c->stream)
                                       : fast  ? launch_loop2_fast(c->solver, a, l, nb, c->stream)
                                               : launch_loop2_exact(c->solver, a, l, nb, c->stream)));
            }
            const int slot = (qh + qn) % 2;
            snap_k[slot] = k;
            HIPCHK(c, hipMemcpyAsync(&c->hloop[slot], c->dloop + ((k - 1) & 1), sizeof(LoopState),
                                     hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipEventRecord(ev.e[slot], c->stream));
            qn++;
            continue;
        }
        HIPCHK(c, hipEventSynchronize(ev.e[qh]));
        const LoopState r = c->hloop[qh];  // state before attempt snap_k - 1
        const int ks = snap_k[qh];
        qh = (qh + 1) % 2;
        qn--;
        if (r.done) break;
        n_seen = r.n;
        k_known = ks - 1;
    }
    l.k = k;
    HIPCHK(c, (hipError_t)launch_loop2_finish(l, c->stream));
    HIPCHK(c, hipMemcpyAsync(&c->hloop[2], c->dloop + (k & 1), sizeof(LoopState),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *out = c->hloop[2];
    return CLAWGPU_OK;
}

// The claw2 time loop (src/2d/claw2.f:494-784) on device-resident q: per
// step bc2 -> fused step into the other buffer -> Courant number back to the
// host -> accept (swap buffers) or reject (keep the old buffer, which is the
// reference's copyq2 restore, and retake with a smaller dt).
// The claw2 loop (claw2.f:596-784).  With hooks, the user's bc2, b4step2
// and src2 are called at the reference's points: bc2 and b4step2 before
// every attempt (retakes included), src2 over dt/2 before the step and dt/2
// after it (method(5) = 2, the pre-source state kept for a retake) or over
// dt after it (method(5) = 1).  Host routines (h) get a host copy of q;
// device routines (dh) run on the device q.
static int claw2_loop(clawgpu_ctx* c, double dx, double dy, double xlower, double ylower,
                      double tstart, double* tend, double* dtv, double* cflv, int* nv,
                      const int* method, const int* mthlim, const int* mthbc, int use_fwave,
                      const clawgpu_hooks* h, const clawgpu_dev_hooks* dh, int* info_out)
{
    if (!c || !tend || !dtv || !cflv || !nv || !method || !mthlim || !mthbc || !info_out)
        return set_err(CLAWGPU_ERR_ARG, "null argument");
    const bool dev_bc = dh && dh->bc2;
    const bool user_bc = (h && h->bc2) || dev_bc;
    const bool hooks = h || dh;
    const bool src = (h && h->src2) || (dh && dh->src2);
    auto dev_call = [&](hipFunction_t f, double tt, double dtt) -> int {
        return launch_dev_hook(c, f, dh, c->dq[c->cur], xlower, ylower, dx, dy, tt, dtt, mthbc);
    };
    const size_t qbytes = (size_t)c->L.fs * c->meqn * sizeof(double);
    const bool split = method[2] < 0;
    int rc = split ? CLAWGPU_OK : check_method(c, method, mthlim);
    if (rc) return rc;
    if (split && method[2] != -1 && method[2] != -2)
        return set_err(CLAWGPU_ERR_ARG, "method(3) must be -2, -1, 0, 1 or 2");
    if (split && uses_table(c, mthlim))
        return set_err(CLAWGPU_ERR_UNSUPPORTED, "the learned limiter runs with the unsplit step only");
    if (method[5] != 0 && !general_path(c))
        return set_err(CLAWGPU_ERR_UNSUPPORTED, "capacity function is not supported by this solver");
    if (split && (method[5] < 0 || method[5] > c->maux))
        return set_err(CLAWGPU_ERR_ARG, "method(6) must be 0..maux");
    if (split)
        for (int mw = 0; mw < c->mwaves; mw++)
            if (mthlim[mw] < 0 || mthlim[mw] > 5) return set_err(CLAWGPU_ERR_ARG, "invalid limiter method");
    if (!user_bc && check_bc(mthbc) != 0)
        return set_err(CLAWGPU_ERR_UNSUPPORTED, "mthbc must be 1, 2 or 3 (periodic on both sides)");
    if (!user_bc) {
        if (const int wrc = check_walls(c, mthbc)) return wrc;
    }
    std::vector<double> hq, saved;  // host copies of q for the hooks
    double* haux = c->maux > 0 && !c->haux.empty() ? c->haux.data() : nullptr;
    auto pull = [&]() -> int {
        hq.resize((size_t)ncells_ghost(c) * c->meqn);
        return clawgpu_get_q(c, hq.data());
    };
    auto push = [&]() -> int { return clawgpu_set_q(c, hq.data()); };
    int info = 0;
    // Without user routines the ghost fill of the next step is enqueued
    // right behind the Courant-number read-back (into the buffer that becomes
    // q if the step is accepted), so it runs while the host decides; a
    // rejected step refills the unchanged q as usual.
    const bool overlap_bc = !hooks && !split;
    bool bc_ready = false;
    struct Ev {
        hipEvent_t e = nullptr;
        ~Ev() { if (e) (void)hipEventDestroy(e); }
    } ev;
    if (overlap_bc) HIPCHK(c, hipEventCreateWithFlags(&ev.e, hipEventDisableTiming));
    double t = tstart;
    int maxn = nv[0];
    double dt = dtv[0];
    double cflmax = 0.0, dtmin = dt, dtmax = dt, cfl = 0.0;
    nv[1] = 0;
    if (method[0] == 0) {
        if (*tend < tstart) {
            maxn = 1;
        } else {
            maxn = (int)lround((*tend - tstart) / dt);
            if (fabs(maxn * dt - (*tend - tstart)) > 1e-5 * (*tend - tstart)) {
                info = 2;
                goto done;
            }
        }
    }
    if (method[0] == 1 && cflv[1] > cflv[0]) { info = 3; goto done; }
    if (method[5] > method[6]) { info = 5; goto done; }
    if (method[1] < method[2]) { info = 6; goto done; }
    if (maxn == 0) goto done;
    if (!hooks && !split && !uses_table(c, mthlim) && !general_path(c)) {
        LoopState f;
        rc = claw2_device(c, dx, dy, tstart, *tend, dtv, cflv, maxn, method, mthlim, mthbc,
                          use_fwave, &f);
        if (rc) return rc;
        if (f.info == -5) {
            *info_out = -5;
            return set_err(CLAWGPU_ERR_NONFINITE, "non-finite wave speed (Courant number)");
        }
        c->cur = f.cur;
        t = f.t;
        cflmax = f.cflmax;
        cfl = f.cfl;
        dtmin = f.dtmin;
        dtmax = f.dtmax;
        dt = f.dt;
        nv[1] = f.n;
        goto done;
    }
    for (int n = 1; n <= maxn; n++) {
        double told = t;
        if (told + dt > *tend && tstart < *tend) dt = *tend - told;
        for (;;) {
            if (bc_ready) {
                bc_ready = false;  // filled while the host waited for the last Courant number
            } else if (dev_bc) {
                if ((rc = dev_call(dh->bc2, told, dt))) return rc;
                HIPCHK(c, hipMemsetAsync(c->dcfl, 0, 4 * sizeof(unsigned long long), c->stream));
            } else if (user_bc) {
                if ((rc = pull())) return rc;
                h->bc2(h->user, c->meqn, c->mbc, c->mx, c->my, xlower, ylower, dx, dy, hq.data(),
                       c->maux, haux, told, dt, mthbc);
                if ((rc = push())) return rc;
                HIPCHK(c, hipMemsetAsync(c->dcfl, 0, 4 * sizeof(unsigned long long), c->stream));
            } else {
                rc = enqueue_bc(c, c->dq[c->cur], mthbc, true);
                if (rc) return rc;
            }
            if (dh && dh->b4step2) {
                if ((rc = dev_call(dh->b4step2, told, dt))) return rc;
            } else if (h && h->b4step2) {
                if ((rc = pull())) return rc;
                h->b4step2(h->user, c->mbc, c->mx, c->my, c->meqn, hq.data(), xlower, ylower, dx, dy,
                           told, dt, c->maux, haux);
                if ((rc = push())) return rc;
                if (haux && (rc = clawgpu_set_aux(c, c->haux.data()))) return rc;
            }
            const bool strang_src = method[4] == 2 && src;
            if (strang_src && dh) {
                if (!c->dsave) HIPCHK(c, hipMalloc((void**)&c->dsave, qbytes));
                HIPCHK(c, hipMemcpyAsync(c->dsave, c->dq[c->cur], qbytes, hipMemcpyDeviceToDevice,
                                         c->stream));
                if ((rc = dev_call(dh->src2, told, dt / 2.0))) return rc;
            } else if (strang_src) {
                if ((rc = pull())) return rc;
                saved = hq;
                h->src2(h->user, c->meqn, c->mbc, c->mx, c->my, xlower, ylower, dx, dy, hq.data(),
                        c->maux, haux, told, dt / 2.0);
                if ((rc = push())) return rc;
            }
            int res = (c->cur + 1) % 3;
            if (split) {
                rc = enqueue_split(c, dx, dy, dt, method, mthlim, use_fwave, &res);
                if (rc) return rc;
                double v[4];
                rc = read_cfl4(c, v);
                if (rc) return rc;
                cfl = split_cfl(v, method[2] == -2, cflv[0]);
            } else {
                if (uses_table(c, mthlim)) {
                    rc = learned_prepare(c, dx, dy, dt);
                    if (rc) return rc;
                }
                rc = enqueue_step(c, dx, dy, dt, method, mthlim, use_fwave);
                if (rc) return rc;
                if (overlap_bc && n < maxn && told + dt < *tend) {
                    // another step follows if this one is accepted
                    HIPCHK(c, hipMemcpyAsync(c->hcfl, c->dcfl, sizeof(double), hipMemcpyDeviceToHost,
                                             c->stream));
                    HIPCHK(c, hipEventRecord(ev.e, c->stream));
                    rc = enqueue_bc(c, c->dq[res], mthbc, true);
                    if (rc) return rc;
                    HIPCHK(c, hipEventSynchronize(ev.e));
                    cfl = *c->hcfl;
                    bc_ready = true;
                } else {
                    rc = read_cfl(c, &cfl);
                    if (rc) return rc;
                }
            }
            if (!std::isfinite(cfl)) {
                *info_out = -5;
                return set_err(CLAWGPU_ERR_NONFINITE, "non-finite wave speed (Courant number)");
            }
            t = told + dt;
            if (cfl <= cflv[0]) {
                cflmax = fmax(cfl, cflmax);
                c->cur = res;  // accept: the new buffer becomes q
                break;
            }
            t = told;  // reject: dq[cur] still holds the step's initial data
            bc_ready = false;
            if (strang_src && dh) {  // back to the state before the source half step
                HIPCHK(c, hipMemcpyAsync(c->dq[c->cur], c->dsave, qbytes, hipMemcpyDeviceToDevice,
                                         c->stream));
            } else if (strang_src) {
                hq = saved;
                if ((rc = push())) return rc;
            }
            if (method[0] == 1) {
                dt = fmin(dtv[1], dt * cflv[1] / cfl);
                continue;
            }
            cflmax = fmax(cfl, cflmax);
            goto done;
        }
        if (dh && dh->src2 && (method[4] == 1 || method[4] == 2)) {
            if ((rc = dev_call(dh->src2, t, method[4] == 2 ? dt / 2.0 : dt))) return rc;
        } else if (h && h->src2 && (method[4] == 1 || method[4] == 2)) {
            if ((rc = pull())) return rc;
            h->src2(h->user, c->meqn, c->mbc, c->mx, c->my, xlower, ylower, dx, dy, hq.data(),
                    c->maux, haux, t, method[4] == 2 ? dt / 2.0 : dt);
            if ((rc = push())) return rc;
        }
        if (method[0] == 1) {
            if (cfl == 0.0) dt = dtv[1];
            else dt = fmin(dtv[1], dt * cflv[1] / cfl);
            dtmin = fmin(dt, dtmin);
            dtmax = fmax(dt, dtmax);
        }
        nv[1] = nv[1] + 1;
        if (t >= *tend) break;
    }
done:
    if (method[0] == 1 && t < *tend && nv[1] == maxn) info = 11;
    if (method[0] == 0 && cflmax > cflv[0]) info = 12;
    *tend = t;
    cflv[2] = cflmax;
    cflv[3] = cfl;
    dtv[2] = dtmin;
    dtv[3] = dtmax;
    dtv[4] = dt;
    *info_out = info;
    return CLAWGPU_OK;
}

extern "C" int clawgpu_claw2(clawgpu_ctx* c, double dx, double dy, double tstart, double* tend,
                             double* dtv, double* cflv, int* nv, const int* method,
                             const int* mthlim, const int* mthbc, int use_fwave, int* info_out)
{
    return claw2_loop(c, dx, dy, 0.0, 0.0, tstart, tend, dtv, cflv, nv, method, mthlim, mthbc,
                      use_fwave, nullptr, nullptr, info_out);
}

extern "C" int clawgpu_claw2_hooks(clawgpu_ctx* c, double dx, double dy, double xlower,
                                   double ylower, double tstart, double* tend, double* dtv,
                                   double* cflv, int* nv, const int* method, const int* mthlim,
                                   const int* mthbc, int use_fwave, const clawgpu_hooks* hooks,
                                   int* info_out)
{
    if (!hooks) return set_err(CLAWGPU_ERR_ARG, "null hooks (use clawgpu_claw2)");
    if (c && c->maux > 0 && c->haux.empty() && (hooks->b4step2 || hooks->src2 || hooks->bc2))
        return set_err(CLAWGPU_ERR_ARG, "aux has not been set (clawgpu_set_aux)");
    return claw2_loop(c, dx, dy, xlower, ylower, tstart, tend, dtv, cflv, nv, method, mthlim,
                      mthbc, use_fwave, hooks, nullptr, info_out);
}

extern "C" int clawgpu_claw2_dev_hooks(clawgpu_ctx* c, double dx, double dy, double xlower,
                                       double ylower, double tstart, double* tend, double* dtv,
                                       double* cflv, int* nv, const int* method,
                                       const int* mthlim, const int* mthbc, int use_fwave,
                                       const clawgpu_dev_hooks* hooks, int* info_out)
{
    if (!hooks) return set_err(CLAWGPU_ERR_ARG, "null hooks (use clawgpu_claw2)");
    return claw2_loop(c, dx, dy, xlower, ylower, tstart, tend, dtv, cflv, nv, method, mthlim,
                      mthbc, use_fwave, nullptr, hooks, info_out);
}

namespace {
__global__ void calib_fill(double* p, long n)
{
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
        p[e] = (double)(e & 1023);
}
// a plain stream copy, 16 bytes per lane per access
__global__ void calib_copy(const double2* __restrict__ s, double2* __restrict__ d, long n2)
{
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n2; e += (long)gridDim.x * blockDim.x)
        d[e] = s[e];
}
// a copy with U 16-byte loads in flight per lane (block-contiguous chunks of
// U * blockDim pairs), nontemporal when NT
template <int U, bool NT>
__global__ void copy_unrolled(const double2* __restrict__ s, double2* __restrict__ d, long n2)
{
    const long step = (long)gridDim.x * blockDim.x * U;
    for (long e = (long)blockIdx.x * blockDim.x * U + threadIdx.x; e < n2; e += step) {
        double2 v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const long k = e + (long)u * blockDim.x;
            if (k < n2) {
                if constexpr (NT) {
                    v[u].x = __builtin_nontemporal_load(&s[k].x);
                    v[u].y = __builtin_nontemporal_load(&s[k].y);
                } else {
                    v[u] = s[k];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const long k = e + (long)u * blockDim.x;
            if (k < n2) {
                if constexpr (NT) {
                    __builtin_nontemporal_store(v[u].x, &d[k].x);
                    __builtin_nontemporal_store(v[u].y, &d[k].y);
                } else {
                    d[k] = v[u];
                }
            }
        }
    }
}

// eight independent FMA chains per lane; the sum is stored so that nothing
// is optimised away
__global__ void fp64_fma_chains(double* out, int iters, double a, double b)
{
    double x[8];
#pragma unroll
    for (int c = 0; c < 8; c++) x[c] = 1e-3 * (threadIdx.x + c);
    for (int k = 0; k < iters; k++) {
#pragma unroll
        for (int c = 0; c < 8; c++) x[c] = fma(x[c], a, b);
    }
    double t = 0.0;
#pragma unroll
    for (int c = 0; c < 8; c++) t += x[c];
    if (t == -1.0) out[blockIdx.x] = t;  // never true for these inputs
}
}  // namespace

extern "C" int clawgpu_copy_bandwidth(long ndoubles, int variant, int reps, float* ms_avg)
{
    if (ndoubles < 2 || reps < 1 || !ms_avg || variant < 0 || variant > 3)
        return set_err(CLAWGPU_ERR_ARG, "bad argument");
    if (variant == 0) return clawgpu_copy_calibration(ndoubles, reps, ms_avg);
    ndoubles &= ~1L;
    double *a = nullptr, *b = nullptr;
    clawgpu_ctx* c = nullptr;
    HIPCHK(c, hipMalloc((void**)&a, ndoubles * sizeof(double)));
    if (hipMalloc((void**)&b, ndoubles * sizeof(double)) != hipSuccess) {
        (void)hipFree(a);
        return set_err(CLAWGPU_ERR_HIP, "allocation failed");
    }
    hipLaunchKernelGGL(calib_fill, dim3(4096), dim3(256), 0, 0, a, ndoubles);
    const long n2 = ndoubles / 2;
    auto launch = [&]() {
        if (variant == 1)
            hipLaunchKernelGGL((copy_unrolled<4, false>), dim3(4096), dim3(256), 0, 0, (const double2*)a, (double2*)b, n2);
        else if (variant == 2)
            hipLaunchKernelGGL((copy_unrolled<4, true>), dim3(4096), dim3(256), 0, 0, (const double2*)a, (double2*)b, n2);
        else
            hipLaunchKernelGGL((copy_unrolled<8, true>), dim3(2048), dim3(256), 0, 0, (const double2*)a, (double2*)b, n2);
    };
    launch();  // warm
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    HIPCHK(c, hipEventRecord(e0, 0));
    for (int r = 0; r < reps; r++) launch();
    HIPCHK(c, hipEventRecord(e1, 0));
    HIPCHK(c, hipEventSynchronize(e1));
    float ms = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
    *ms_avg = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return CLAWGPU_OK;
}

extern "C" int clawgpu_fp64_fma_peak(int iters, float* ms, double* tflops)
{
    if (iters < 1 || !ms || !tflops) return set_err(CLAWGPU_ERR_ARG, "bad argument");
    clawgpu_ctx* c = nullptr;
    int dev = 0, ncu = 0;
    HIPCHK(c, hipGetDevice(&dev));
    HIPCHK(c, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int nblocks = ncu * 8;  // 8 waves per SIMD of 256-lane blocks
    double* out = nullptr;
    HIPCHK(c, hipMalloc((void**)&out, nblocks * sizeof(double)));
    hipLaunchKernelGGL(fp64_fma_chains, dim3(nblocks), dim3(256), 0, 0, out, iters / 10 + 1,
                       0.999999, 1e-7);  // warm
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    HIPCHK(c, hipEventRecord(e0, 0));
    hipLaunchKernelGGL(fp64_fma_chains, dim3(nblocks), dim3(256), 0, 0, out, iters, 0.999999, 1e-7);
    HIPCHK(c, hipEventRecord(e1, 0));
    HIPCHK(c, hipEventSynchronize(e1));
    float t = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&t, e0, e1));
    *ms = t;
    *tflops = 2.0 * 8.0 * (double)iters * nblocks * 256.0 / (t * 1e-3) / 1e12;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(out);
    return CLAWGPU_OK;
}

extern "C" int clawgpu_copy_calibration(long ndoubles, int reps, float* ms_avg)
{
    // Benchmark helper (no reference counterpart): reps copies of ndoubles
    // doubles between two fresh device buffers, timed with HIP events; the
    // known byte count calibrates the FETCH_SIZE / WRITE_SIZE counters.
    if (ndoubles < 2 || reps < 1 || !ms_avg) return set_err(CLAWGPU_ERR_ARG, "bad argument");
    ndoubles &= ~1L;
    double *a = nullptr, *b = nullptr;
    hipEvent_t e0, e1;
    clawgpu_ctx* c = nullptr;
    HIPCHK(c, hipMalloc((void**)&a, ndoubles * sizeof(double)));
    if (hipMalloc((void**)&b, ndoubles * sizeof(double)) != hipSuccess) {
        (void)hipFree(a);
        return set_err(CLAWGPU_ERR_HIP, "allocation failed");
    }
    hipLaunchKernelGGL(calib_fill, dim3(4096), dim3(256), 0, 0, a, ndoubles);
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    HIPCHK(c, hipEventRecord(e0, 0));
    for (int r = 0; r < reps; r++)
        hipLaunchKernelGGL(calib_copy, dim3(8192), dim3(256), 0, 0, (const double2*)a, (double2*)b,
                           ndoubles / 2);
    HIPCHK(c, hipEventRecord(e1, 0));
    HIPCHK(c, hipEventSynchronize(e1));
    float ms = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
    *ms_avg = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return CLAWGPU_OK;
}

extern "C" int clawgpu_set_segment_rows(clawgpu_ctx* c, int rows)
{
    if (!c || rows < 0) return set_err(CLAWGPU_ERR_ARG, "rows must be >= 0 (0: chosen per grid)");
    c->seglen = rows;
    return CLAWGPU_OK;
}

extern "C" int clawgpu_get_segment_rows(clawgpu_ctx* c, int* rows)
{
    if (!c || !rows) return set_err(CLAWGPU_ERR_ARG, "null argument");
    *rows = c->seglen > 0 ? c->seglen : (c->auto_key[0] >= 0 ? c->auto_rows : 0);
    return CLAWGPU_OK;
}

extern "C" int clawgpu_set_arith(clawgpu_ctx* c, int mode)
{
    if (!c || (mode != CLAWGPU_ARITH_EXACT && mode != CLAWGPU_ARITH_FAST))
        return set_err(CLAWGPU_ERR_ARG, "mode must be CLAWGPU_ARITH_EXACT or CLAWGPU_ARITH_FAST");
    if (c->us && mode != c->arith)
        return set_err(CLAWGPU_ERR_UNSUPPORTED, "a user solver runs in the arithmetic its code object was built with");
    c->arith = mode;
    return CLAWGPU_OK;
}

extern "C" int clawgpu_time_steps(clawgpu_ctx* c, double dx, double dy, double dt, const int* method,
                                  const int* mthlim, const int* mthbc, int use_fwave, int nsteps,
                                  float* total_ms, float* step_kernel_ms, double* cfl_max_out)
{
    // Fixed-dt stepping timed with HIP events on the library's stream, with
    // no per-step host synchronisation: total_ms spans all launches (ghost
    // fill, the learned limiter's pre-pass with limiter 6, step kernel per
    // step), step_kernel_ms sums event pairs around each step-kernel launch
    // only.  Used by the benchmark.
    if (!c || !method || !mthlim || !mthbc || !total_ms || nsteps < 1)
        return set_err(CLAWGPU_ERR_ARG, "bad argument");
    int rc = check_method(c, method, mthlim);
    if (rc) return rc;
    if (check_bc(mthbc) != 0) return set_err(CLAWGPU_ERR_UNSUPPORTED, "bad mthbc");
    if (const int wrc = check_walls(c, mthbc)) return wrc;
    hipEvent_t* ev = new (std::nothrow) hipEvent_t[2 * nsteps + 2];
    if (!ev) return set_err(CLAWGPU_ERR_ARG, "out of host memory");
    for (int k = 0; k < 2 * nsteps + 2; k++) HIPCHK(c, hipEventCreate(&ev[k]));
    HIPCHK(c, hipEventRecord(ev[0], c->stream));
    for (int n = 0; n < nsteps; n++) {
        rc = enqueue_bc(c, c->dq[c->cur], mthbc, n == 0);
        if (rc) return rc;
        if (uses_table(c, mthlim)) {  // the learned limiter's pre-pass, outside the step-kernel pair
            rc = learned_prepare(c, dx, dy, dt);
            if (rc) return rc;
        }
        HIPCHK(c, hipEventRecord(ev[2 + 2 * n], c->stream));
        rc = enqueue_step(c, dx, dy, dt, method, mthlim, use_fwave);
        if (rc) return rc;
        HIPCHK(c, hipEventRecord(ev[3 + 2 * n], c->stream));
        c->cur = (c->cur + 1) % 3;
    }
    HIPCHK(c, hipEventRecord(ev[1], c->stream));
    HIPCHK(c, hipEventSynchronize(ev[1]));
    HIPCHK(c, hipEventElapsedTime(total_ms, ev[0], ev[1]));
    float ksum = 0.0f;
    for (int n = 0; n < nsteps; n++) {
        float k = 0.0f;
        HIPCHK(c, hipEventElapsedTime(&k, ev[2 + 2 * n], ev[3 + 2 * n]));
        ksum += k;
    }
    if (step_kernel_ms) *step_kernel_ms = ksum;
    for (int k = 0; k < 2 * nsteps + 2; k++) (void)hipEventDestroy(ev[k]);
    delete[] ev;
    if (cfl_max_out) {
        rc = read_cfl(c, cfl_max_out);
        if (rc) return rc;
    }
    return CLAWGPU_OK;
}

// Solver-level evaluation for tests (no reference counterpart): n interfaces
// with host arrays ql, qr, asdq (n*meqn) -> wave (n*mwaves*meqn), s
// (n*mwaves), amdq, apdq, bmasdq, bpasdq (n*meqn), all in the reference's
// per-interface layout.
extern "C" int clawgpu_solver_probe(clawgpu_ctx* c, int ixy, int n, const double* ql,
                                    const double* qr, const double* asdq, double* wave, double* s,
                                    double* amdq, double* apdq, double* bmasdq, double* bpasdq)
{
    if (!c || n < 1 || !ql || !qr || !asdq || !wave || !s || !amdq || !apdq || !bmasdq || !bpasdq)
        return set_err(CLAWGPU_ERR_ARG, "bad argument");
    if (ixy != 1 && ixy != 2) return set_err(CLAWGPU_ERR_ARG, "ixy must be 1 or 2");
    const int M = c->meqn, W = c->mwaves;
    size_t nq = (size_t)n * M, nw = (size_t)n * W;
    double* d = nullptr;
    size_t tot = 3 * nq + nw * M + nw + 4 * nq;
    HIPCHK(c, hipMalloc((void**)&d, tot * sizeof(double)));
    double *dql = d, *dqr = d + nq, *das = d + 2 * nq, *dw = d + 3 * nq, *ds = dw + nw * M;
    double *dam = ds + nw, *dap = dam + nq, *dbm = dap + nq, *dbp = dbm + nq;
    int rc = CLAWGPU_OK;
    hipError_t e;
    if ((e = hipMemcpy(dql, ql, nq * 8, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(dqr, qr, nq * 8, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(das, asdq, nq * 8, hipMemcpyHostToDevice)) != hipSuccess) {
        rc = set_err(CLAWGPU_ERR_HIP, hipGetErrorString(e));
    } else {
        int le = c->us ? usolver_probe(c->us, n, ixy, c->P, dql, dqr, das, dw, ds, dam, dap, dbm, dbp, c->stream)
                 : c->arith == ARITH_FAST
                     ? launch_probe_fast(c->solver, n, ixy, c->P, dql, dqr, das, dw, ds, dam, dap, dbm, dbp, c->stream)
                     : launch_probe_exact(c->solver, n, ixy, c->P, dql, dqr, das, dw, ds, dam, dap, dbm, dbp, c->stream);
        if (le != 0 || (e = hipStreamSynchronize(c->stream)) != hipSuccess ||
            (e = hipMemcpy(wave, dw, nw * M * 8, hipMemcpyDeviceToHost)) != hipSuccess ||
            (e = hipMemcpy(s, ds, nw * 8, hipMemcpyDeviceToHost)) != hipSuccess ||
            (e = hipMemcpy(amdq, dam, nq * 8, hipMemcpyDeviceToHost)) != hipSuccess ||
            (e = hipMemcpy(apdq, dap, nq * 8, hipMemcpyDeviceToHost)) != hipSuccess ||
            (e = hipMemcpy(bmasdq, dbm, nq * 8, hipMemcpyDeviceToHost)) != hipSuccess ||
            (e = hipMemcpy(bpasdq, dbp, nq * 8, hipMemcpyDeviceToHost)) != hipSuccess)
            rc = set_err(CLAWGPU_ERR_HIP, "solver probe failed");
    }
    (void)hipFree(d);
    return rc;
}

// entry points used by the learned-limiter training (learned.hip)
int enqueue_step_public(clawgpu_ctx* c, double dx, double dy, double dt, const int* method,
                        const int* mthlim, int use_fwave)
{
    return enqueue_step(c, dx, dy, dt, method, mthlim, use_fwave);
}

int enqueue_bc_public(clawgpu_ctx* c, const int* mthbc)
{
    if (check_bc(mthbc) != 0) return set_err(CLAWGPU_ERR_UNSUPPORTED, "bad mthbc");
    return enqueue_bc(c, c->dq[c->cur], mthbc, true);
}
