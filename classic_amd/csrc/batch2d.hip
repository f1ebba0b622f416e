// Host side of the batched 2D path (clawgpu_batch_* in include/clawgpu.h):
// device storage of all problems, the captured HIP graph of one batched
// iteration, and the schedule autotuner.  The kernels are in batch2d.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "batch2d.hpp"
#include "layout2d.hpp"
#include "ctx.hpp"

namespace {

struct Schedule {
    int seglen = 32;     // rows per tile
    int nblocks = 0;     // blocks of the persistent step kernel (4 waves each)
    int largest_first = 1;  // tile order: problems by decreasing cell count
    int w3 = 0;  // acoustics kernel held to three waves per SIMD (a few spills)
};

struct Plan {  // device tables of a schedule
    int* rank_prob = nullptr;    // group 0 ranks, then group 1 ranks
    long* rank_tile0 = nullptr;  // likewise
    int nrank[batch::kGroups] = {0, 0};
    long ntiles[batch::kGroups] = {0, 0};
};

}  // namespace

struct clawgpu_batch {
    int nprob = 0, mbc = 2, arith = ARITH_EXACT;
    std::vector<batch::Prob> hp;  // host copy of the problem table
    batch::Prob* dprob = nullptr;
    double* pool = nullptr;
    double* staging = nullptr;
    int* ndone = nullptr;
    int* hndone = nullptr;  // pinned
    int walls[batch::kGroups] = {0, 0};  // some problem of the group has a wall side
    int ncu = 256;
    hipStream_t stream = nullptr;
    Schedule sched;
    Plan plan;
    batch::Params prm{};
    bool started = false;
    hipGraphExec_t exec = nullptr;
    int graph_iters = 0;
    int last_hip_error = 0;
};

namespace clawgpu {
namespace batch {

__global__ void prepare_kernel(Params b)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= b.nprob) return;
    Prob& p = b.prob[k];
    p.cfl_bits = 0ull;
    if (p.done) return;
    if (b.dry) {
        p.res = (p.cur + 2) % 3;  // the spare buffer
        return;
    }
    p.res = (p.cur + 1) % 3;
    if (!p.retake) {
        p.told = p.t;
        if (p.told + p.dt > p.tend && p.tstart < p.tend) p.dt = p.tend - p.told;
    }
}

// The tiles of group G are dealt to the waves of the grid in turn (wave w
// takes tiles w, w + nwaves, ...).  Every value that steers the loop (tile,
// problem, its done flag) is the same in all 64 lanes and is moved to a
// scalar register with readfirstlane, so the loop stays uniform: no lane
// depends on another lane's private value (such as an atomic's result held
// by lane 0 alone), which the compiler may not keep converged across a
// loop with early exits.
__global__ void accept_kernel(Params b)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= b.nprob || b.dry) return;
    Prob& p = b.prob[k];
    if (p.done) return;
    const double cfl = __longlong_as_double((long long)p.cfl_bits);
    p.cfl = cfl;
    auto finish = [&](int info) {
        p.done = 1;
        p.info = info;
        atomicAdd(b.ndone, 1);
    };
    if (!(cfl < __builtin_huge_val())) {  // a NaN wave speed is published as +inf
        finish(-5);
        return;
    }
    p.t = p.told + p.dt;
    if (cfl <= b.cflv1) {
        p.cflmax = fmax(cfl, p.cflmax);
        p.cur = p.res;
        p.retake = 0;
        if (b.vardt) {
            p.dt = cfl == 0.0 ? b.dtv2 : fmin(b.dtv2, p.dt * b.cflv2 / cfl);
            p.dtmin = fmin(p.dt, p.dtmin);
            p.dtmax = fmax(p.dt, p.dtmax);
        }
        p.nv = p.nv + 1;
        if (p.t >= p.tend) finish(0);
        else if (p.nv >= p.maxn) finish(b.vardt ? 11 : 0);
    } else {
        p.t = p.told;
        if (b.vardt) {
            p.retake = 1;
            p.nrej = p.nrej + 1;
            p.dt = fmin(b.dtv2, p.dt * b.cflv2 / cfl);
        } else {
            p.cflmax = fmax(cfl, p.cflmax);
            finish(12);
        }
    }
}

}  // namespace batch
}  // namespace clawgpu

using namespace clawgpu;

static void free_plan(Plan& p)
{
    if (p.rank_prob) (void)hipFree(p.rank_prob);
    if (p.rank_tile0) (void)hipFree(p.rank_tile0);
    p = Plan{};
}

static void drop_graph(clawgpu_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
    b->graph_iters = 0;
}

static long ntiles_of(const batch::Prob& p, int seglen)
{
    return (long)p.nstrips * ((p.L.my + seglen - 1) / seglen);
}

// device tables of the tile order of schedule s: per solver group, the
// group's problems (by decreasing cell count if largest_first) and the
// first tile of each
static int make_plan(clawgpu_batch* b, const Schedule& s, Plan* out)
{
    const int n = b->nprob;
    std::vector<int> order;
    std::vector<long> t0;
    Plan p;
    for (int g = 0; g < batch::kGroups; g++) {
        std::vector<int> grp;
        for (int k = 0; k < n; k++)
            if (batch::group_of(b->hp[k].solver) == g) grp.push_back(k);
        if (s.largest_first)
            std::stable_sort(grp.begin(), grp.end(), [&](int x, int y) {
                return (long)b->hp[x].L.mx * b->hp[x].L.my > (long)b->hp[y].L.mx * b->hp[y].L.my;
            });
        long acc = 0;
        for (int k : grp) {
            order.push_back(k);
            t0.push_back(acc);
            acc += ntiles_of(b->hp[k], s.seglen);
        }
        p.nrank[g] = (int)grp.size();
        p.ntiles[g] = acc;
    }
    HIPCHK(b, hipMalloc((void**)&p.rank_prob, n * sizeof(int)));
    hipError_t e = hipMalloc((void**)&p.rank_tile0, n * sizeof(long));
    if (e == hipSuccess) e = hipMemcpyAsync(p.rank_prob, order.data(), n * sizeof(int), hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(p.rank_tile0, t0.data(), n * sizeof(long), hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) {
        free_plan(p);
        HIPCHK(b, e);
    }
    *out = p;
    return CLAWGPU_OK;
}

static int set_schedule(clawgpu_batch* b, const Schedule& s)
{
    Plan p;
    int rc = make_plan(b, s, &p);
    if (rc) return rc;
    free_plan(b->plan);
    b->plan = p;
    b->sched = s;
    drop_graph(b);
    return CLAWGPU_OK;
}

static batch::Params params_of(const clawgpu_batch* b, int dry)
{
    batch::Params a = b->prm;
    a.prob = b->dprob;
    a.nprob = b->nprob;
    int off = 0;
    for (int g = 0; g < batch::kGroups; g++) {
        a.rank_prob[g] = b->plan.rank_prob + off;
        a.rank_tile0[g] = b->plan.rank_tile0 + off;
        a.nrank[g] = b->plan.nrank[g];
        a.ntiles[g] = b->plan.ntiles[g];
        off += b->plan.nrank[g];
    }
    for (int g = 0; g < batch::kGroups; g++) a.walls[g] = b->walls[g];
    a.ndone = b->ndone;
    a.seglen = b->sched.seglen;
    a.dry = dry;
    return a;
}

// the kernels of one batched iteration; ev (optional) brackets the step
// kernels
static hipError_t enqueue_iteration(clawgpu_batch* b, int dry, hipEvent_t* ev = nullptr)
{
    const batch::Params a = params_of(b, dry);
    const int pb = (b->nprob + 127) / 128;
    hipLaunchKernelGGL(batch::prepare_kernel, dim3(pb), dim3(128), 0, b->stream, a);
    if (ev) (void)hipEventRecord(ev[0], b->stream);
    const int nblk = b->sched.nblocks > 0 ? b->sched.nblocks : 2 * b->ncu;
    const bool w3 = b->sched.w3 != 0;
    if (b->arith == ARITH_FAST) (void)launch_batch_steps_fast(a, nblk, w3, b->stream);
    else (void)launch_batch_steps_exact(a, nblk, w3, b->stream);
    if (ev) (void)hipEventRecord(ev[1], b->stream);
    hipLaunchKernelGGL(batch::accept_kernel, dim3(pb), dim3(128), 0, b->stream, a);
    return hipGetLastError();
}

static int upload_table(clawgpu_batch* b)
{
    HIPCHK(b, hipMemcpyAsync(b->dprob, b->hp.data(), b->nprob * sizeof(batch::Prob), hipMemcpyHostToDevice, b->stream));
    HIPCHK(b, hipStreamSynchronize(b->stream));
    return CLAWGPU_OK;
}

static int download_table(clawgpu_batch* b)
{
    HIPCHK(b, hipMemcpyAsync(b->hp.data(), b->dprob, b->nprob * sizeof(batch::Prob), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(b, hipStreamSynchronize(b->stream));
    return CLAWGPU_OK;
}

extern "C" int clawgpu_batch_destroy(clawgpu_batch* b)
{
    if (!b) return CLAWGPU_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    free_plan(b->plan);
    if (b->dprob) (void)hipFree(b->dprob);
    if (b->pool) (void)hipFree(b->pool);
    if (b->staging) (void)hipFree(b->staging);
    if (b->ndone) (void)hipFree(b->ndone);
    if (b->hndone) (void)hipHostFree(b->hndone);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return CLAWGPU_OK;
}

extern "C" int clawgpu_batch_create(clawgpu_batch** out, int nprob, int mbc, const int* solver,
                                    const int* mx, const int* my, const double* params,
                                    const int* mthbc, const double* dx, const double* dy)
{
    if (!out) return set_err(CLAWGPU_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (nprob < 1 || !solver || !mx || !my || !params || !mthbc || !dx || !dy)
        return set_err(CLAWGPU_ERR_ARG, "null argument or empty batch");
    if (mbc < 2) return set_err(CLAWGPU_ERR_ARG, "mbc must be >= 2 (claw2 info=1)");
    clawgpu_batch* b = new (std::nothrow) clawgpu_batch();
    if (!b) return set_err(CLAWGPU_ERR_ARG, "out of host memory");
    b->nprob = nprob;
    b->mbc = mbc;
    b->hp.resize(nprob);
    size_t pool = 0, stage = 0;
    for (int k = 0; k < nprob; k++) {
        batch::Prob& p = b->hp[k];
        p = batch::Prob{};
        if (solver[k] != CLAWGPU_SOLVER_ACOUSTICS && solver[k] != CLAWGPU_SOLVER_EULER) {
            delete b;
            return set_err(CLAWGPU_ERR_UNSUPPORTED, "batched problems use the acoustics or Euler solver");
        }
        if (mx[k] < 1 || my[k] < 1) {
            delete b;
            return set_err(CLAWGPU_ERR_ARG, "grid must have at least one cell");
        }
        const int rc = solver_params(solver[k], params + 2 * k, 2, &p.P);
        if (rc) {
            delete b;
            return rc;
        }
        for (int s = 0; s < 4; s++) {
            const int m = mthbc[4 * k + s];
            if (m != 1 && m != 2 && m != 3) {
                delete b;
                return set_err(CLAWGPU_ERR_UNSUPPORTED, "mthbc must be 1, 2 or 3");
            }
            p.mthbc[s] = m;
        }
        if ((p.mthbc[0] == 2) != (p.mthbc[1] == 2) || (p.mthbc[2] == 2) != (p.mthbc[3] == 2)) {
            delete b;
            return set_err(CLAWGPU_ERR_UNSUPPORTED, "periodic boundaries must be set on both sides");
        }
        p.solver = solver[k];
        p.meqn = solver[k] == CLAWGPU_SOLVER_EULER ? 4 : 3;
        p.dx = dx[k];
        p.dy = dy[k];
        p.L.mx = mx[k];
        p.L.my = my[k];
        p.L.mbc = mbc;
        p.L.pitch = ((mx[k] + 2 * mbc) + 15) / 16 * 16;
        p.L.fs = ((long)p.L.pitch * (my[k] + 2 * mbc) + 31) / 32 * 32;
        p.nstrips = (mx[k] + kStripOut - 1) / kStripOut;
        p.done = 1;  // until clawgpu_batch_start
        pool += 3 * (size_t)p.L.fs * p.meqn;
        stage = std::max(stage, (size_t)(mx[k] + 2 * mbc) * (my[k] + 2 * mbc) * p.meqn);
        for (int s = 0; s < 4; s++)
            if (p.mthbc[s] == 3) b->walls[batch::group_of(p.solver)] = 1;
    }
    int dev = 0;
    hipDeviceProp_t prop;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, dev);
    if (e == hipSuccess) b->ncu = prop.multiProcessorCount;
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void**)&b->pool, pool * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&b->staging, stage * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&b->dprob, nprob * sizeof(batch::Prob));
    if (e == hipSuccess) e = hipMalloc((void**)&b->ndone, sizeof(int));
    if (e == hipSuccess) e = hipHostMalloc((void**)&b->hndone, sizeof(int), hipHostMallocDefault);
    if (e != hipSuccess) {
        snprintf(g_errbuf, sizeof g_errbuf, "allocation failed: %s", hipGetErrorString(e));
        clawgpu_batch_destroy(b);
        return CLAWGPU_ERR_HIP;
    }
    size_t off = 0;
    for (int k = 0; k < nprob; k++) {
        batch::Prob& p = b->hp[k];
        for (int s = 0; s < 3; s++) {
            p.q[s] = b->pool + off;
            off += (size_t)p.L.fs * p.meqn;
        }
    }
    e = hipMemsetAsync(b->pool, 0, pool * sizeof(double), b->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b->ndone, 0, sizeof(int), b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) {
        snprintf(g_errbuf, sizeof g_errbuf, "init failed: %s", hipGetErrorString(e));
        clawgpu_batch_destroy(b);
        return CLAWGPU_ERR_HIP;
    }
    int rc = upload_table(b);
    if (!rc) rc = set_schedule(b, Schedule{});
    if (rc) {
        clawgpu_batch_destroy(b);
        return rc;
    }
    *out = b;
    return CLAWGPU_OK;
}

static int check_k(const clawgpu_batch* b, int k)
{
    if (!b) return set_err(CLAWGPU_ERR_ARG, "null batch");
    if (k < 0 || k >= b->nprob) return set_err(CLAWGPU_ERR_ARG, "problem index out of range");
    return CLAWGPU_OK;
}

extern "C" int clawgpu_batch_set_q(clawgpu_batch* b, int k, const double* q)
{
    int rc = check_k(b, k);
    if (rc) return rc;
    if (!q) return set_err(CLAWGPU_ERR_ARG, "null argument");
    rc = download_table(b);
    if (rc) return rc;
    const batch::Prob& p = b->hp[k];
    const long n = (long)(p.L.mx + 2 * p.L.mbc) * (p.L.my + 2 * p.L.mbc);
    HIPCHK(b, hipMemcpyAsync(b->staging, q, n * p.meqn * sizeof(double), hipMemcpyHostToDevice, b->stream));
    hipLaunchKernelGGL(aos_to_planes, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b->stream, b->staging,
                       p.q[p.cur], p.L, p.meqn);
    HIPCHK(b, hipGetLastError());
    HIPCHK(b, hipStreamSynchronize(b->stream));
    return CLAWGPU_OK;
}

extern "C" int clawgpu_batch_get_q(clawgpu_batch* b, int k, double* q)
{
    int rc = check_k(b, k);
    if (rc) return rc;
    if (!q) return set_err(CLAWGPU_ERR_ARG, "null argument");
    rc = download_table(b);
    if (rc) return rc;
    const batch::Prob& p = b->hp[k];
    const long n = (long)(p.L.mx + 2 * p.L.mbc) * (p.L.my + 2 * p.L.mbc);
    hipLaunchKernelGGL(planes_to_aos, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b->stream, p.q[p.cur],
                       b->staging, p.L, p.meqn);
    HIPCHK(b, hipGetLastError());
    HIPCHK(b, hipMemcpyAsync(q, b->staging, n * p.meqn * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(b, hipStreamSynchronize(b->stream));
    return CLAWGPU_OK;
}

extern "C" int clawgpu_batch_set_arith(clawgpu_batch* b, int arith)
{
    if (!b) return set_err(CLAWGPU_ERR_ARG, "null batch");
    if (arith != ARITH_EXACT && arith != ARITH_FAST) return set_err(CLAWGPU_ERR_ARG, "unknown arithmetic mode");
    if (arith != b->arith) drop_graph(b);
    b->arith = arith;
    return CLAWGPU_OK;
}

extern "C" int clawgpu_batch_start(clawgpu_batch* b, double tstart, const double* tend, const double* dtv,
                                   const double* cflv, int maxn, const int* method, const int* mthlim,
                                   int use_fwave)
{
    if (!b || !tend || !dtv || !cflv || !method || !mthlim) return set_err(CLAWGPU_ERR_ARG, "null argument");
    if (method[0] != 0 && method[0] != 1) return set_err(CLAWGPU_ERR_ARG, "method(1) must be 0 or 1");
    if (method[1] != 1 && method[1] != 2) return set_err(CLAWGPU_ERR_ARG, "method(2) must be 1 or 2");
    if (method[2] < 0 || method[2] > 2)
        return set_err(CLAWGPU_ERR_UNSUPPORTED, "batched problems take the unsplit step (method(3) = 0, 1, 2)");
    if (method[5] != 0) return set_err(CLAWGPU_ERR_UNSUPPORTED, "capacity function is not supported");
    if (method[1] < method[2]) return set_err(CLAWGPU_ERR_ARG, "method(2) < method(3) (claw2 info=6)");
    if (method[0] == 1 && cflv[1] > cflv[0]) return set_err(CLAWGPU_ERR_ARG, "cflv(2) > cflv(1) (claw2 info=3)");
    for (int w = 0; w < 3; w++)
        if (mthlim[w] < 0 || mthlim[w] > 5)
            return set_err(CLAWGPU_ERR_ARG, "mthlim must be 0..5 (the learned limiter runs on single grids)");
    if (maxn < 0) return set_err(CLAWGPU_ERR_ARG, "maxn must be >= 0");
    int rc = download_table(b);
    if (rc) return rc;
    int ndone = 0;
    for (int k = 0; k < b->nprob; k++) {
        batch::Prob& p = b->hp[k];
        p.t = tstart;
        p.tstart = tstart;
        p.told = tstart;
        p.tend = tend[k];
        p.dt = dtv[0];
        p.dtmin = p.dtmax = dtv[0];
        p.cflmax = 0.0;
        p.cfl = 0.0;
        p.nv = 0;
        p.nrej = 0;
        p.retake = 0;
        p.info = 0;
        p.done = 0;
        p.maxn = maxn;
        if (method[0] == 0) {  // fixed dt: the step count follows from tend (claw2.f:541-553)
            if (tend[k] < tstart) {
                p.maxn = 1;
            } else {
                p.maxn = (int)lround((tend[k] - tstart) / dtv[0]);
                if (fabs(p.maxn * dtv[0] - (tend[k] - tstart)) > 1e-5 * (tend[k] - tstart)) {
                    p.info = 2;
                    p.done = 1;
                }
            }
        }
        if (p.maxn == 0) p.done = 1;
        ndone += p.done;
    }
    batch::Params& a = b->prm;
    a.order = method[1];
    a.trans = method[2];
    a.fwave = use_fwave ? 1 : 0;
    for (int w = 0; w < kMaxWaves; w++) a.lim[w] = w < 3 ? mthlim[w] : 0;
    a.vardt = method[0];
    a.dtv2 = dtv[1];
    a.cflv1 = cflv[0];
    a.cflv2 = cflv[1];
    drop_graph(b);
    rc = upload_table(b);
    if (rc) return rc;
    *b->hndone = ndone;
    HIPCHK(b, hipMemcpyAsync(b->ndone, b->hndone, sizeof(int), hipMemcpyHostToDevice, b->stream));
    HIPCHK(b, hipStreamSynchronize(b->stream));
    b->started = true;
    return CLAWGPU_OK;
}

static int ensure_graph(clawgpu_batch* b, int iters)
{
    if (b->exec && b->graph_iters == iters) return CLAWGPU_OK;
    drop_graph(b);
    hipGraph_t g = nullptr;
    HIPCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int i = 0; i < iters && e == hipSuccess; i++) e = enqueue_iteration(b, 0);
    hipError_t e2 = hipStreamEndCapture(b->stream, &g);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        b->exec = nullptr;
        HIPCHK(b, e);
    }
    b->graph_iters = iters;
    return CLAWGPU_OK;
}

extern "C" int clawgpu_batch_advance(clawgpu_batch* b, int max_iters, int use_graph, int* nleft)
{
    if (!b || !nleft) return set_err(CLAWGPU_ERR_ARG, "null argument");
    if (!b->started) return set_err(CLAWGPU_ERR_ARG, "clawgpu_batch_start has not been called");
    if (max_iters < 0) return set_err(CLAWGPU_ERR_ARG, "max_iters must be >= 0");
    // Iterations are issued in chunks (one graph replay each); the number of
    // finished problems is read back after every few chunks so a run to
    // completion stops soon after the last problem reaches tend.
    const int chunk = 8, chunks_per_check = 4;
    int done_iters = 0;
    while (done_iters < max_iters) {
        for (int c = 0; c < chunks_per_check && done_iters < max_iters; c++) {
            const int n = std::min(chunk, max_iters - done_iters);
            if (use_graph && n == chunk) {
                int rc = ensure_graph(b, chunk);
                if (rc) return rc;
                HIPCHK(b, hipGraphLaunch(b->exec, b->stream));
            } else {
                for (int i = 0; i < n; i++) HIPCHK(b, enqueue_iteration(b, 0));
            }
            done_iters += n;
        }
        HIPCHK(b, hipMemcpyAsync(b->hndone, b->ndone, sizeof(int), hipMemcpyDeviceToHost, b->stream));
        HIPCHK(b, hipStreamSynchronize(b->stream));
        if (*b->hndone >= b->nprob) break;
    }
    HIPCHK(b, hipMemcpyAsync(b->hndone, b->ndone, sizeof(int), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(b, hipStreamSynchronize(b->stream));
    *nleft = b->nprob - *b->hndone;
    return CLAWGPU_OK;
}

extern "C" int clawgpu_batch_result(clawgpu_batch* b, int k, double* t, double* dtv, double* cflv, int* nv,
                                    int* info)
{
    int rc = check_k(b, k);
    if (rc) return rc;
    if (!t || !dtv || !cflv || !nv || !info) return set_err(CLAWGPU_ERR_ARG, "null argument");
    rc = download_table(b);
    if (rc) return rc;
    const batch::Prob& p = b->hp[k];
    *t = p.t;
    dtv[2] = p.dtmin;
    dtv[3] = p.dtmax;
    dtv[4] = p.dt;
    cflv[2] = p.cflmax;
    cflv[3] = p.cfl;
    nv[0] = p.maxn;
    nv[1] = p.nv;
    *info = p.info;
    return CLAWGPU_OK;
}

extern "C" int clawgpu_batch_set_schedule(clawgpu_batch* b, const int* sched)
{
    if (!b || !sched) return set_err(CLAWGPU_ERR_ARG, "null argument");
    if (sched[0] < 1 || sched[0] > 4096) return set_err(CLAWGPU_ERR_ARG, "seglen must be 1..4096");
    if (sched[1] < 0 || sched[1] > (1 << 20))
        return set_err(CLAWGPU_ERR_ARG, "nblocks must be 0..2^20 (0: two per CU)");
    return set_schedule(b, Schedule{sched[0], sched[1], sched[2] ? 1 : 0, sched[3] ? 1 : 0});
}

extern "C" int clawgpu_batch_get_schedule(clawgpu_batch* b, int* sched)
{
    if (!b || !sched) return set_err(CLAWGPU_ERR_ARG, "null argument");
    sched[0] = b->sched.seglen;
    sched[1] = b->sched.nblocks > 0 ? b->sched.nblocks : 2 * b->ncu;
    sched[2] = b->sched.largest_first;
    sched[3] = b->sched.w3;
    return CLAWGPU_OK;
}

// time `reps` dry iterations under the current schedule (ms per iteration)
static int time_dry(clawgpu_batch* b, int reps, float* ms)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = enqueue_iteration(b, 1);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, b->stream);
    for (int i = 0; i < reps && e == hipSuccess; i++) e = enqueue_iteration(b, 1);
    if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    HIPCHK(b, e);
    *ms = t / reps;
    return CLAWGPU_OK;
}

extern "C" int clawgpu_batch_tune(clawgpu_batch* b, int reps, int* ncand, int* cand_out, float* ms_out,
                                  int max_cand)
{
    // Autotuner of the persistent step kernels' schedule: tile height (rows
    // per tile), grid size, tile order and the acoustics kernel's register
    // budget are measured on dry iterations of the current batch (which
    // change no state) and the fastest is kept.  cand_out receives the four
    // schedule values per candidate and ms_out its time per iteration, up to
    // max_cand entries.
    if (!b || !ncand) return set_err(CLAWGPU_ERR_ARG, "null argument");
    if (!b->started) return set_err(CLAWGPU_ERR_ARG, "clawgpu_batch_start has not been called");
    if (reps < 1) return set_err(CLAWGPU_ERR_ARG, "reps must be >= 1");
    const int seglens[] = {16, 24, 32, 48, 64, 96, 128};
    const int waves_per_cu[] = {4, 8, 12, 16};
    Schedule best = b->sched;
    float best_ms = 0.f;
    int n = 0;
    for (int w3 = 0; w3 <= 1; w3++)
        for (int order = 1; order >= 0; order--)
            for (int sl : seglens)
                for (int w : waves_per_cu) {
                    Schedule s{sl, w * b->ncu / 4, order, w3};
                    int rc = set_schedule(b, s);
                    if (rc) return rc;
                    float ms = 0.f;
                    rc = time_dry(b, reps, &ms);
                    if (rc) return rc;
                    if (n < max_cand && cand_out && ms_out) {
                        cand_out[4 * n] = s.seglen;
                        cand_out[4 * n + 1] = s.nblocks;
                        cand_out[4 * n + 2] = s.largest_first;
                        cand_out[4 * n + 3] = s.w3;
                        ms_out[n] = ms;
                    }
                    if (n == 0 || ms < best_ms) {
                        best_ms = ms;
                        best = s;
                    }
                    n++;
                }
    *ncand = n;
    return set_schedule(b, best);
}

extern "C" int clawgpu_batch_time(clawgpu_batch* b, int iters, float* total_ms, float* step_kernel_ms)
{
    // Measurement helper: iters real batched iterations launched one kernel
    // at a time (no graph), with HIP events around every step-kernel launch
    // on the batch's stream.  *total_ms spans all launches, *step_kernel_ms
    // sums the step-kernel launches alone.
    if (!b || !total_ms || !step_kernel_ms) return set_err(CLAWGPU_ERR_ARG, "null argument");
    if (!b->started) return set_err(CLAWGPU_ERR_ARG, "clawgpu_batch_start has not been called");
    if (iters < 1) return set_err(CLAWGPU_ERR_ARG, "iters must be >= 1");
    std::vector<hipEvent_t> ev(2 * iters + 2, nullptr);
    hipError_t e = hipSuccess;
    for (size_t i = 0; i < ev.size() && e == hipSuccess; i++) e = hipEventCreate(&ev[i]);
    if (e == hipSuccess) e = hipEventRecord(ev[2 * iters], b->stream);
    for (int i = 0; i < iters && e == hipSuccess; i++) e = enqueue_iteration(b, 0, &ev[2 * i]);
    if (e == hipSuccess) e = hipEventRecord(ev[2 * iters + 1], b->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ev[2 * iters + 1]);
    float tot = 0.f, ker = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&tot, ev[2 * iters], ev[2 * iters + 1]);
    for (int i = 0; i < iters && e == hipSuccess; i++) {
        float t = 0.f;
        e = hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1]);
        ker += t;
    }
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    HIPCHK(b, e);
    *total_ms = tot;
    *step_kernel_ms = ker;
    return CLAWGPU_OK;
}
