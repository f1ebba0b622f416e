// clawgpu_step2_noalloc: one unsplit step (the contract of clawgpu_step2,
// src/2d/step2.f90:2-251) on caller-owned device buffers, for callers that
// cannot allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).  The only device memory beyond the
// caller's q buffers and Courant output is the step kernel's Courant-number
// accumulator, which lives in the file-scope array `scratch` below instead
// of a context allocation.
//
// Host code never passes `scratch` itself to a kernel: its name on the host
// is a stand-in, not a device address.  The address is looked up once per
// device with hipGetSymbolAddress (each GPU holds its own copy of the array)
// and that pointer is what the kernels receive.  Every call on a GPU shares
// that GPU's array, so calls that overlap in time on one GPU (two host
// threads, or two streams with no ordering between them) race on the
// accumulator; calls on one stream are ordered and safe.
#include <climits>
#include <mutex>

#include "ctx.hpp"

using namespace clawgpu;

namespace {

// 256 bytes: word 0 is the Courant accumulator of the step in flight (the
// maximum Courant number as the raw bits of a non-negative double, which
// order like the values); the rest is reserved padding to one 256-byte line.
constexpr int kScratchWords = 32;
static __device__ __align__(256) unsigned long long scratch[kScratchWords];

constexpr int kMaxDevices = 64;
// Largest mx or my accepted: the kernel's strip/segment indices and the
// int-sized row pitch stay far inside their ranges.
constexpr int kMaxCellsPerSide = 1 << 20;

struct DeviceScratch {
    unsigned long long* ptr;  // device address of this GPU's `scratch`
    int ncu;                  // compute units, for the segment-length choice
};

std::mutex g_mu;
DeviceScratch g_dev[kMaxDevices];

// The current device's scratch address and CU count, looked up on the first
// call on that device (with that device current).
int device_scratch(DeviceScratch* out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        snprintf(g_errbuf, sizeof g_errbuf, "hipGetDevice: %s", hipGetErrorString(e));
        return CLAWGPU_ERR_HIP;
    }
    if (dev < 0 || dev >= kMaxDevices) return set_err(CLAWGPU_ERR_UNSUPPORTED, "device index beyond 63");
    std::lock_guard<std::mutex> lk(g_mu);
    DeviceScratch& d = g_dev[dev];
    if (!d.ptr) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            snprintf(g_errbuf, sizeof g_errbuf, "hipGetSymbolAddress(scratch): %s",
                     hipGetErrorString(e));
            return CLAWGPU_ERR_HIP;
        }
        int ncu = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            ncu = 0;
        d.ncu = ncu;
        d.ptr = static_cast<unsigned long long*>(p);
    }
    *out = d;
    return CLAWGPU_OK;
}

// A fill kernel rather than hipMemsetAsync, so that the call can be
// captured into a HIP graph.
__global__ void noalloc_clear(unsigned long long* acc)
{
    if (threadIdx.x == 0) acc[0] = 0ull;
}

__global__ void noalloc_emit(const unsigned long long* acc, double* cfl)
{
    if (threadIdx.x == 0) *cfl = __longlong_as_double((long long)acc[0]);
}

}  // namespace

extern "C" int clawgpu_step2_noalloc(int solver_id, const double* params, int nparams, int arith,
                                     int mx, int my, int mbc, long pitch, long plane,
                                     const double* qold, double* qnew, double dx, double dy,
                                     double dt, const int* method, const int* mthlim,
                                     int use_fwave, double* cfl, void* stream)
{
    if (!qold || !qnew || !cfl || !method || !mthlim) return set_err(CLAWGPU_ERR_ARG, "null argument");
    if (solver_id == CLAWGPU_SOLVER_VC_ADVECTION)
        return set_err(CLAWGPU_ERR_UNSUPPORTED, "solvers with aux arrays need a context");
    if (solver_id != CLAWGPU_SOLVER_ACOUSTICS && solver_id != CLAWGPU_SOLVER_EULER)
        return set_err(CLAWGPU_ERR_UNSUPPORTED, "unknown solver id");
    if (arith != ARITH_EXACT && arith != ARITH_FAST) return set_err(CLAWGPU_ERR_ARG, "invalid arithmetic mode");
    if (mbc < 2) return set_err(CLAWGPU_ERR_ARG, "mbc must be >= 2 (claw2 info=1)");
    if (mx < 1 || my < 1) return set_err(CLAWGPU_ERR_ARG, "grid must have at least one cell");
    if (mx > kMaxCellsPerSide || my > kMaxCellsPerSide)
        return set_err(CLAWGPU_ERR_ARG, "mx and my must be at most 1048576");
    if (mbc > 64) return set_err(CLAWGPU_ERR_ARG, "mbc must be at most 64");
    if (pitch < mx + 2 * mbc || pitch > INT_MAX)
        return set_err(CLAWGPU_ERR_ARG, "pitch must be at least mx + 2*mbc (and fit an int)");
    if (plane < pitch * (my + 2 * mbc)) return set_err(CLAWGPU_ERR_ARG, "plane must be at least pitch*(my + 2*mbc)");
    if (qold == qnew) return set_err(CLAWGPU_ERR_ARG, "qold and qnew must be different buffers");
    if (method[1] != 1 && method[1] != 2) return set_err(CLAWGPU_ERR_ARG, "method(2) must be 1 or 2");
    if (method[2] < 0 || method[2] > 2)
        return set_err(CLAWGPU_ERR_UNSUPPORTED, "method(3) must be 0, 1 or 2 for the unsplit step");
    if (method[5] != 0) return set_err(CLAWGPU_ERR_UNSUPPORTED, "capacity function is not supported by this solver");
    SolverParams P{};
    int rc = solver_params(solver_id, params, nparams, &P);
    if (rc) return rc;
    const int mwaves = solver_id == CLAWGPU_SOLVER_ACOUSTICS ? 2 : 3;
    for (int mw = 0; mw < mwaves; mw++) {
        if (mthlim[mw] == kLimTable)
            return set_err(CLAWGPU_ERR_UNSUPPORTED, "limiter 6 (learned) reads tables held by a context");
        if (mthlim[mw] < 0 || mthlim[mw] > kLimTable) return set_err(CLAWGPU_ERR_ARG, "invalid limiter method");
    }
    DeviceScratch d{};
    rc = device_scratch(&d);
    if (rc) return rc;

    StepArgs a{};
    a.qold = qold;
    a.qnew = qnew;
    a.L.mx = mx;
    a.L.my = my;
    a.L.mbc = mbc;
    a.L.pitch = (int)pitch;
    a.L.fs = plane;
    a.dtdx = dt / dx;
    a.dtdy = dt / dy;
    a.order = method[1];
    a.trans = method[2];
    a.use_fwave = use_fwave ? 1 : 0;
    for (int k = 0; k < kMaxWaves; k++) a.lim[k] = k < mwaves ? mthlim[k] : 0;
    a.nstrips = (mx + kStripOut - 1) / kStripOut;
    const int wpc = arith == ARITH_FAST ? step2_waves_per_cu_fast(solver_id, a)
                                        : step2_waves_per_cu_exact(solver_id, a);
    a.seglen = segment_rows_for(d.ncu, wpc, a.nstrips, my);
    a.nsegs = (my + a.seglen - 1) / a.seglen;
    a.cfl_bits = d.ptr;
    a.P = P;
    a.phix = a.phiy = nullptr;
    a.bcmap = 0;

    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(noalloc_clear, dim3(1), dim3(64), 0, s, d.ptr);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        const long waves = (long)a.nstrips * a.nsegs;
        const int nb = (int)((waves + kWavesPerBlock - 1) / kWavesPerBlock);
        e = (hipError_t)(arith == ARITH_FAST ? launch_step2_fast(solver_id, a, nb, s)
                                             : launch_step2_exact(solver_id, a, nb, s));
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(noalloc_emit, dim3(1), dim3(64), 0, s, d.ptr, cfl);
        e = hipGetLastError();
    }
    if (e != hipSuccess) {
        snprintf(g_errbuf, sizeof g_errbuf, "step launch: %s", hipGetErrorString(e));
        return CLAWGPU_ERR_HIP;
    }
    return CLAWGPU_OK;
}
