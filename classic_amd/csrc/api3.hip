// C-ABI of the 3D path: context, boundary conditions, step3 / dimsp3 and the
// claw3 time loop on the device (declarations and reference citations in
// include/clawgpu.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <cstdint>
#include <new>
#include <utility>
#include <vector>

#include "ctx.hpp"
#include "devhooks.hpp"
#include "step3d.hpp"
#include "step3d_user.hpp"
#include "step3g.hpp"
#include "usolver.hpp"

using namespace clawgpu::s3;

struct clawgpu3_ctx {
    int meqn, mwaves, maux, mbc, mx, my, mz, solver;
    Layout3 L;
    double* dq[3];    // q (field planes) and two work buffers
    int cur;
    double* daux;     // aux planes (Z, c, then the rest of aux)
    double* staging;  // reference-layout copy for uploads and downloads
    double cmax;      // max |c| over the cells whose interfaces the sweeps visit
    double* exy;      // piston: exp(-500 r^2) per (i, j) column, from the host's libm
    double exy_geom[4];
    hipStream_t stream;
    int seg;          // cells per segment of the y/z split sweeps
    int tile;         // tile shape of the unsplit kernel (launch_step3_tile)
    int xcd;          // tile order bits: 1 XCD-contiguous, 2 z fastest
    unsigned char* dmat = nullptr;  // material index per cell (nmat > 0)
    double* dlut = nullptr;         // (Z, c) of each material
    int nmat = 0;                   // distinct (Z, c) pairs found by set_aux, 0 if too many
    int use_table = 1;              // read the material table instead of aux when nmat > 0
    int last_hip_error;
    std::vector<double> haux;       // host copy of aux for the user routines
    int exact = 0;                  // unsplit step with the reference's rounding and order
    double* drec = nullptr;         // records of the exact / capacity unsplit step
    double* dsave = nullptr;        // q before a Strang source half step (device user routines)
    const clawgpu_solver3* us = nullptr;  // user solver (split step only), not owned
    SolverParams P{};                     // its parameters
};

namespace {

__global__ void aos_to_planes3(const double* __restrict__ src, double* __restrict__ dst, long n,
                               int ncomp, int take, long fs)
{
    long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    for (int m = 0; m < take; m++) dst[m * fs + c] = src[c * ncomp + m];
}

__global__ void planes_to_aos3(const double* __restrict__ src, double* __restrict__ dst, long n,
                               int ncomp, long fs)
{
    long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    for (int m = 0; m < ncomp; m++) dst[c * ncomp + m] = src[m * fs + c];
}

struct Bc3 {
    int code[6];
    int meqn;                // fields (at most kMaxWaves)
    double t, et;            // time and exp(-500 t^2) (host libm)
    const double* exy;       // exp(-500 (x^2 + y^2)) per column (host libm)
};

// bc3.f: faces x-lo, x-hi, y-lo, y-hi, z-lo, z-hi in that order, each over
// the full extent of the other indices, so a cell outside in several
// directions ends with the value of the last face: the source cell has every
// outside index mapped (extrapolation, periodic, mirror), and each wall
// negates its normal velocity.  Code 0 on z-hi is the piston wall of the
// heterogeneous-acoustics example (examples/acoustics_3d_heterogeneous/bc3.f
// labels 600-606): w = 2 s(x, y, t) - w(mirror).
// One thread per ghost cell (0-based padded indices ii, jj, kk): the ghost
// planes of the z faces (whole planes), then the ghost rows of the y faces
// in the interior planes, then the ghost cells of the x faces in the
// interior rows, each run along its contiguous index.  Every source cell is
// an interior cell, so the order of the writes does not matter.
__host__ __device__ inline long bc3_ghosts(const Layout3& L)
{
    const long m2 = 2L * L.mbc;
    return m2 * L.nx * L.ny + m2 * L.nx * L.mz + m2 * L.my * L.mz;
}

__global__ void bc3_kernel(double* __restrict__ q, Layout3 L, Bc3 b)
{
    const int mb = L.mbc;
    const long nzf = 2L * mb * L.nx * L.ny, nyf = 2L * mb * L.nx * L.mz;
    long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int ii, jj, kk;
    if (e < nzf) {
        ii = (int)(e % L.nx);
        jj = (int)((e / L.nx) % L.ny);
        const int r = (int)(e / ((long)L.nx * L.ny));  // 0 .. 2mbc-1
        kk = r < mb ? r : L.nz - 2 * mb + r;
    } else if ((e -= nzf) < nyf) {
        ii = (int)(e % L.nx);
        const long t = e / L.nx;
        const int r = (int)(t % (2 * mb));
        kk = mb + (int)(t / (2 * mb));
        jj = r < mb ? r : L.ny - 2 * mb + r;
    } else if ((e -= nyf) < 2L * mb * L.my * L.mz) {
        const int r = (int)(e % (2 * mb));
        const long t = e / (2 * mb);
        jj = mb + (int)(t % L.my);
        kk = mb + (int)(t / L.my);
        ii = r < mb ? r : L.nx - 2 * mb + r;
    } else {
        return;
    }
    const int g[3] = {ii + 1 - L.mbc, jj + 1 - L.mbc, kk + 1 - L.mbc};
    const int nn[3] = {L.mx, L.my, L.mz};
    int s[3];
    bool flip[3] = {false, false, false}, piston = false, ghost = false;
    for (int d = 0; d < 3; d++) {
        s[d] = g[d];
        if (g[d] >= 1 && g[d] <= nn[d]) continue;
        ghost = true;
        const bool hi = g[d] > nn[d];
        const int code = b.code[2 * d + (hi ? 1 : 0)];
        if (code == 1) s[d] = hi ? nn[d] : 1;
        else if (code == 2) s[d] = hi ? g[d] - nn[d] : g[d] + nn[d];
        else {
            s[d] = hi ? 2 * nn[d] + 1 - g[d] : 1 - g[d];
            if (code == 3) flip[d] = true;
            else piston = true;  // code 0, z-hi only (checked on the host)
        }
    }
    if (!ghost) return;  // not reached: every enumerated cell is a ghost cell
    const long src = L.off(s[0], s[1], s[2]);
    double v[kMaxWaves];
    for (int m = 0; m < b.meqn; m++) v[m] = q[m * L.fs + src];
    for (int d = 0; d < 3; d++)
        if (flip[d] && 1 + d < b.meqn) v[1 + d] = -v[1 + d];
    if (piston) {
        double sp = 0.0;
        if (b.t < 0.1) {
            // s = -exp(-500 r^2) * exp(-500 t^2), both exponentials taken on
            // the host so that they round as the reference's do
            sp = -(b.exy[(long)jj * L.nx + ii] * b.et);
            if (fabs(sp) < 1e-6) sp = 0.0;
        }
        v[3] = 2.0 * sp - v[3];
    }
    const long dst = L.off(g[0], g[1], g[2]);
    for (int m = 0; m < b.meqn; m++) q[m * L.fs + dst] = v[m];
}

// max |c| over cells with all indices in [0, n+1] (the cells adjacent to
// the interfaces 1..n+1 of the slices 0..n+1 that step3/step3ds visit);
// NaN wins.  Result as raw bits in out[0] (values are >= 0) and a NaN flag
// in out[1].
__global__ void cmax_kernel(const double* __restrict__ c, Layout3 L, unsigned long long* out)
{
    const long n = (long)(L.mx + 2) * (L.my + 2) * (L.mz + 2);
    double m = 0.0;
    bool nan = false;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (long)gridDim.x * blockDim.x) {
        const int i = (int)(e % (L.mx + 2)), j = (int)((e / (L.mx + 2)) % (L.my + 2)),
                  k = (int)(e / ((long)(L.mx + 2) * (L.my + 2)));
        const double v = fabs(c[L.off(i, j, k)]);
        if (v != v) nan = true;
        else m = fmax(m, v);
    }
    for (int d = 32; d >= 1; d >>= 1) {
        m = fmax(m, __shfl_xor(m, d));
        nan = nan || __shfl_xor((int)nan, d);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&out[0], (unsigned long long)__double_as_longlong(m));
        if (nan) atomicOr(&out[1], 1ull);
    }
}

// Whether every cell's (Z, c) still equals its material's table entry, bit
// for bit (out[0] set when some cell differs).
__global__ void table_check_kernel(const double* __restrict__ Z, const double* __restrict__ cc,
                                   const unsigned char* __restrict__ mat,
                                   const double* __restrict__ lut, long n, unsigned long long* out)
{
    bool diff = false;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (long)gridDim.x * blockDim.x) {
        const int id = mat[e];
        diff = diff || __double_as_longlong(Z[e]) != __double_as_longlong(lut[2 * id]) ||
               __double_as_longlong(cc[e]) != __double_as_longlong(lut[2 * id + 1]);
    }
    if (__any(diff) && (threadIdx.x & 63) == 0) atomicOr(&out[0], 1ull);
}

// max of n non-negative per-wave Courant numbers, as raw bits in *out
// (one atomic per wave of this kernel)
__global__ void wave_max_kernel(const double* __restrict__ part, long n, unsigned long long* out)
{
    clawgpu::wave_max_reduce(part, n, out);
}

int check_method3(const int* method)
{
    int info = 0;
    if (method[0] != 0 && method[0] != 1) info = 6;
    if (method[1] != 1 && method[1] != 2) info = 6;
    if (method[1] == 1 && method[2] >= 0 && method[2] != 0 && method[2] != 10 && method[2] != 11)
        info = 6;
    if (method[1] == 2 && method[2] >= 0 && method[2] != 0 && method[2] != 10 && method[2] != 11 &&
        method[2] != 20 && method[2] != 21 && method[2] != 22)
        info = 6;
    return info;
}

int check_args3(clawgpu3_ctx* c, const int* method, const int* mthlim)
{
    if (method[2] < 0 && method[2] != -1)
        return set_err(CLAWGPU_ERR_ARG, "method(3) < 0 must be -1 (dimsp3 is Godunov splitting only)");
    if (c->us && method[2] != -1)
        return set_err(CLAWGPU_ERR_UNSUPPORTED, "a 3D user solver runs the split step (method(3) = -1) only");
    if (c->us && method[5] != 0)
        return set_err(CLAWGPU_ERR_UNSUPPORTED, "a 3D user solver has no capacity function");
    if (check_method3(method) != 0) return set_err(CLAWGPU_ERR_ARG, "invalid method (chkmth info=6)");
    if (method[5] < 0 || method[5] > c->maux)
        return set_err(CLAWGPU_ERR_ARG, "method(6) must be 0 or an aux index (capacity)");
    for (int w = 0; w < c->mwaves; w++)
        if (mthlim[w] < 0 || mthlim[w] > 5) return set_err(CLAWGPU_ERR_ARG, "mthlim must be 0..5");
    return CLAWGPU_OK;
}

// Walls negate components 2, 3, 4 (1-based) as bc3 does, and the piston
// (code 0) is the acoustics example's: a user solver needs those components
// and takes no piston.
static int check_walls3(const clawgpu3_ctx* c, const int* mthbc)
{
    if (!c->us) return CLAWGPU_OK;
    if (mthbc[5] == 0) return set_err(CLAWGPU_ERR_UNSUPPORTED, "the piston boundary (code 0) is the built-in acoustics solver's");
    for (int d = 0; d < 3; d++)
        if ((mthbc[2 * d] == 3 || mthbc[2 * d + 1] == 3) && 1 + d >= c->meqn)
            return set_err(CLAWGPU_ERR_ARG, "a wall in this direction needs the normal velocity component (meqn too small)");
    return CLAWGPU_OK;
}

int check_bc3(const int* mthbc)
{
    for (int f = 0; f < 6; f++) {
        const int lo = f == 5 ? 0 : 1;
        if (mthbc[f] < lo || mthbc[f] > 3) return -1;
    }
    for (int d = 0; d < 3; d++)
        if ((mthbc[2 * d] == 2) != (mthbc[2 * d + 1] == 2)) return -1;
    return 0;
}

// exp(-500 (x^2 + y^2)) of the piston at every (i, j) column, computed with
// the host's exp (the reference's dexp) whenever the geometry changes
int piston_table(clawgpu3_ctx* c, const double* geom)
{
    const double g4[4] = {geom[0], geom[1], geom[3], geom[4]};
    if (c->exy && std::memcmp(g4, c->exy_geom, sizeof g4) == 0) return CLAWGPU_OK;
    const int nx = c->L.nx, ny = c->L.ny;
    double* h = new (std::nothrow) double[(size_t)nx * ny];
    if (!h) return set_err(CLAWGPU_ERR_ARG, "out of host memory");
    for (int jj = 0; jj < ny; jj++)
        for (int ii = 0; ii < nx; ii++) {
            const double xi = geom[0] + ((ii + 1 - c->mbc) - 0.5) * geom[3];
            const double yj = geom[1] + ((jj + 1 - c->mbc) - 0.5) * geom[4];
            const double r2 = (xi * xi + yj * yj);
            h[(size_t)jj * nx + ii] = std::exp(-500.0 * r2);
        }
    hipError_t e = hipSuccess;
    if (!c->exy) e = hipMalloc((void**)&c->exy, (size_t)nx * ny * sizeof(double));
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->exy, h, (size_t)nx * ny * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    delete[] h;
    HIPCHK(c, e);
    std::memcpy(c->exy_geom, g4, sizeof g4);
    return CLAWGPU_OK;
}

int enqueue_bc3(clawgpu3_ctx* c, double* q, const double* geom, double t, const int* mthbc)
{
    int rc0 = check_walls3(c, mthbc);
    if (rc0) return rc0;
    Bc3 b;
    for (int f = 0; f < 6; f++) b.code[f] = mthbc[f];
    b.meqn = c->meqn;
    b.t = t;
    b.et = 0.0;
    b.exy = nullptr;
    if (mthbc[5] == 0) {
        int rc = piston_table(c, geom);
        if (rc) return rc;
        b.exy = c->exy;
        b.et = std::exp(-(500.0 * (t * t)));
    }
    const long n = bc3_ghosts(c->L);
    hipLaunchKernelGGL(bc3_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, q,
                       c->L, b);
    HIPCHK(c, hipGetLastError());
    return CLAWGPU_OK;
}

Args3 step_args(clawgpu3_ctx* c, const double* q, double* qn, double dx, double dy, double dz,
                double dt, const int* method, const int* mthlim, int use_fwave)
{
    Args3 a;
    a.q = q;
    a.qn = qn;
    a.aux = c->daux;
    a.mat = c->dmat;
    a.lut = c->dlut;
    a.nmat = c->use_table ? c->nmat : 0;
    a.L = c->L;
    a.afs = c->L.fs;
    a.dtdx = dt / dx;
    a.dtdy = dt / dy;
    a.dtdz = dt / dz;
    a.m2 = method[1];
    a.m3 = method[2] / 10;
    a.m4 = method[2] - 10 * a.m3;
    a.fwave = use_fwave ? 1 : 0;
    a.lim0 = mthlim[0];
    a.lim1 = mthlim[1];
    a.ntx = (c->mx + kOut - 1) / kOut;
    a.nty = a.ntz = 0;  // set by the launcher for its tile shape
    a.xcd = c->xcd;
    return a;
}

const double* capa_plane(const clawgpu3_ctx* c, const int* method)
{
    return method[5] > 0 ? c->daux + (long)(method[5] - 1) * c->L.fs : nullptr;
}

// The unsplit step: with a capacity function or in exact arithmetic the
// records path (step3g.hip, the reference's rounding and order), else the
// z-marching / tile kernels.
int enqueue_step3(clawgpu3_ctx* c, int res, double dx, double dy, double dz, double dt,
                  const int* method, const int* mthlim, int use_fwave)
{
    if (method[5] > 0 || c->exact) {
        if (c->L.mx + 2 > 8190 || c->L.my + 2 > 8190 || c->L.mz + 2 > 8190)
            return set_err(CLAWGPU_ERR_UNSUPPORTED, "exact 3D step: sides up to 8188");
        if (!c->drec) HIPCHK(c, hipMalloc((void**)&c->drec, rec3_doubles(c->L) * sizeof(double)));
        Args3g g;
        g.q = c->dq[c->cur];
        g.qn = c->dq[res];
        g.aux = c->daux;
        g.capa = capa_plane(c, method);
        g.rec = c->drec;
        g.L = c->L;
        g.dt3[0] = dt / dx;
        g.dt3[1] = dt / dy;
        g.dt3[2] = dt / dz;
        g.m2 = method[1];
        g.m3 = method[2] / 10;
        g.m4 = method[2] - 10 * g.m3;
        g.fwave = use_fwave ? 1 : 0;
        g.lim0 = mthlim[0];
        g.lim1 = mthlim[1];
        g.dir = 0;
        if (launch_step3_records(g, c->stream)) return set_err(CLAWGPU_ERR_HIP, "step3 launch failed");
        return CLAWGPU_OK;
    }
    Args3 a = step_args(c, c->dq[c->cur], c->dq[res], dx, dy, dz, dt, method, mthlim, use_fwave);
    const int e = launch_step3_tile(a, c->tile, c->stream);
    if (e) return set_err(CLAWGPU_ERR_HIP, "step3 kernel launch failed");
    return CLAWGPU_OK;
}

// The Courant numbers of the x, y and z sweeps at this dt: dt/d * max|c|,
// or with a capacity function the max over the cells of (dt/d / kappa) * c
// (flux3's dtdx1d(i) s), computed on the device.  A NaN sound speed makes
// all three NaN.
int courant3(clawgpu3_ctx* c, const int* method, double dx, double dy, double dz, double dt,
             double* cf)
{
    if (method[5] <= 0) {
        cf[0] = dt / dx * c->cmax;
        cf[1] = dt / dy * c->cmax;
        cf[2] = dt / dz * c->cmax;
        return CLAWGPU_OK;
    }
    unsigned long long* bits = reinterpret_cast<unsigned long long*>(c->staging);
    HIPCHK(c, hipMemsetAsync(bits, 0, 4 * sizeof(unsigned long long), c->stream));
    const double d3[3] = {dt / dx, dt / dy, dt / dz};
    if (launch_capa_cfl(c->daux + c->L.fs, capa_plane(c, method), c->L, d3, bits, c->stream))
        return set_err(CLAWGPU_ERR_HIP, "Courant kernel launch failed");
    unsigned long long h[4] = {0, 0, 0, 0};
    HIPCHK(c, hipMemcpyAsync(h, bits, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int r = 0; r < 3; r++) {
        std::memcpy(&cf[r], &h[r], sizeof(double));
        if (h[3]) cf[r] = NAN;
    }
    return CLAWGPU_OK;
}

// One sweep reading field m from buffer from[m] and writing it to buffer
// to[m] (-1: not written; see SweepArgs3).
int enqueue_sweep3(clawgpu3_ctx* c, int dir, const int from[4], const int to[4], double d,
                   const int* method, const int* mthlim, int use_fwave)
{
    SweepArgs3 a;
    for (int m = 0; m < 4; m++) {
        a.qf[m] = c->dq[from[m]] + (size_t)m * c->L.fs;
        a.qo[m] = to[m] >= 0 ? c->dq[to[m]] + (size_t)m * c->L.fs : nullptr;
    }
    a.aux = c->daux;
    const bool table = c->use_table && c->nmat > 0;
    a.mat = table ? c->dmat : nullptr;
    a.lut = table ? c->dlut : nullptr;
    a.capa = capa_plane(c, method);
    a.L = c->L;
    a.afs = c->L.fs;
    a.d = d;
    a.m2 = method[1];
    a.fwave = use_fwave ? 1 : 0;
    a.lim0 = mthlim[0];
    a.lim1 = mthlim[1];
    const int n = dir == 1 ? c->my : c->mz;
    a.seg = c->seg;
    a.nsegs = (n + c->seg - 1) / c->seg;
    a.nstrips = dir == 0 ? (c->mx + 2 + kOut - 1) / kOut : (c->L.nx + 63) / 64;
    const int e = launch_sweep3(dir, a, c->stream);
    if (e) return set_err(CLAWGPU_ERR_HIP, "sweep kernel launch failed");
    return CLAWGPU_OK;
}

// dimsp3.f:5-85 on the device: x, y, z sweeps out of place.  The Courant
// numbers are known before the sweeps (dt/dx * max|c|), so the early returns
// of dimsp3 become a count of sweeps to run.  A full step (buffers A = cur,
// B, C) moves each field only where a sweep changes it: x writes (p, u) to
// B; y reads p from B and v from A and writes (p, v) to C; z reads p and v
// from C and w from A and writes (p, w) to B with v copied there, so B holds
// the new state (u from the x sweep).  A step that stops after its first or
// second sweep writes every field in each sweep, so its last buffer is
// complete.
int enqueue_dimsp3(clawgpu3_ctx* c, double dx, double dy, double dz, double dt, const int* method,
                   const int* mthlim, const double* cflv, int use_fwave, int* res, double* cfl)
{
    double cf[3];
    int rc0 = courant3(c, method, dx, dy, dz, dt, cf);
    if (rc0) return rc0;
    const double cx = cf[0], cy = cf[1], cz = cf[2];
    const int A = c->cur, B = (c->cur + 1) % 3, C = (c->cur + 2) % 3;
    if (cx > cflv[0] || cy > cflv[0]) {
        const int fa[4] = {A, A, A, A}, tb[4] = {B, B, B, B};
        int rc = enqueue_sweep3(c, 0, fa, tb, dt / dx, method, mthlim, use_fwave);
        if (rc) return rc;
        *res = B;
        if (cx > cflv[0]) { *cfl = cx; return CLAWGPU_OK; }
        const int tc[4] = {C, C, C, C};
        rc = enqueue_sweep3(c, 1, tb, tc, dt / dy, method, mthlim, use_fwave);
        if (rc) return rc;
        *res = C;
        *cfl = cy;
        return CLAWGPU_OK;
    }
    const int fx[4] = {A, A, A, A}, tx[4] = {B, B, -1, -1};
    const int fy[4] = {B, B, A, A}, ty[4] = {C, -1, C, -1};
    const int fz[4] = {C, B, C, A}, tz[4] = {B, -1, B, B};
    int rc = enqueue_sweep3(c, 0, fx, tx, dt / dx, method, mthlim, use_fwave);
    if (!rc) rc = enqueue_sweep3(c, 1, fy, ty, dt / dy, method, mthlim, use_fwave);
    if (!rc) rc = enqueue_sweep3(c, 2, fz, tz, dt / dz, method, mthlim, use_fwave);
    if (rc) return rc;
    *res = B;
    *cfl = fmax(cx, fmax(cy, cz));
    return CLAWGPU_OK;
}

// dimsp3 with a user solver: x (A -> B), y (B -> C), z (C -> B) sweeps of
// every field, each writing its Courant number to its own slot; the numbers
// come back to the host after the sweeps (the wave speeds are known only
// there), so every sweep runs and a rejected step's buffers are not used.
// With stop (cflv(1)) the numbers come back after each sweep and the step
// ends after the first sweep above it, as dimsp3's early returns.
int enqueue_dimsp3_user(clawgpu3_ctx* c, double dx, double dy, double dz, double dt,
                        const int* method, const int* mthlim, int use_fwave, const double* stop,
                        int* res, double cf[3])
{
    const int A = c->cur, B = (c->cur + 1) % 3, C = (c->cur + 2) % 3;
    unsigned long long* bits = reinterpret_cast<unsigned long long*>(c->staging);
    double* part = c->staging + 8;  // per-wave Courant numbers of the running sweep
    const long room = (long)(c->meqn > c->maux ? c->meqn : (c->maux > 0 ? c->maux : 1)) * c->L.fs - 8;
    HIPCHK(c, hipMemsetAsync(bits, 0, 3 * sizeof(unsigned long long), c->stream));
    const int from[3] = {A, B, C}, to[3] = {B, C, B};
    const double dd[3] = {dt / dx, dt / dy, dt / dz};
    for (int dir = 0; dir < 3; dir++) {
        s3u::Sweep3U a;
        a.q = c->dq[from[dir]];
        a.qn = c->dq[to[dir]];
        a.aux = c->daux;
        a.L = c->L;
        a.d = dd[dir];
        a.m2 = method[1];
        a.fwave = use_fwave ? 1 : 0;
        for (int w = 0; w < kMaxWaves; w++) a.lim[w] = w < c->mwaves ? mthlim[w] : 0;
        a.seg = c->seg;
        const int n = dir == 1 ? c->my : c->mz;
        a.nsegs = (n + c->seg - 1) / c->seg;
        a.nstrips = dir == 0 ? (c->mx + 2 + kOut - 1) / kOut : (c->L.nx + 63) / 64;
        a.cfl_part = part;
        a.P = c->P;
        const long nwaves =
            dir == 0 ? (long)a.nstrips * (c->L.ny) * (c->L.nz)
                     : (long)a.nstrips * (dir == 1 ? c->L.nz : c->L.ny) * a.nsegs;
        if (nwaves > room) return set_err(CLAWGPU_ERR_ARG, "grid too large for the Courant partials");
        if (usolver3_sweep(c->us, dir, a, (int)((nwaves + 3) / 4), c->stream))
            return set_err(CLAWGPU_ERR_HIP, "user sweep kernel launch failed");
        const long nb = (nwaves + 255) / 256;
        hipLaunchKernelGGL(wave_max_kernel, dim3((unsigned)(nb < 512 ? nb : 512)), dim3(256), 0,
                           c->stream, part, nwaves, bits + dir);
        HIPCHK(c, hipGetLastError());
        *res = to[dir];
        if (stop && dir < 2) {
            unsigned long long h;
            HIPCHK(c, hipMemcpyAsync(&h, bits + dir, sizeof h, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            std::memcpy(&cf[dir], &h, sizeof(double));
            if (cf[dir] > *stop) {
                for (int e = dir + 1; e < 3; e++) cf[e] = 0.0;
                return CLAWGPU_OK;
            }
        }
    }
    unsigned long long h[3];
    HIPCHK(c, hipMemcpyAsync(h, bits, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int d = 0; d < 3; d++) std::memcpy(&cf[d], &h[d], sizeof(double));
    return CLAWGPU_OK;
}

// dimsp3's Courant number from the three sweeps' (claw3 rejects above cflv(1))
double split_cfl(const double cf[3], double cfl1)
{
    return cf[0] > cfl1 ? cf[0] : cf[1] > cfl1 ? cf[1] : fmax(cf[0], fmax(cf[1], cf[2]));
}

// the Courant number of an unsplit step (the max over all slices); NaN
// when a sound speed is NaN
int unsplit_cfl(clawgpu3_ctx* c, const int* method, double dx, double dy, double dz, double dt,
                double* cfl)
{
    double cf[3];
    int rc = courant3(c, method, dx, dy, dz, dt, cf);
    if (rc) return rc;
    *cfl = (cf[0] != cf[0]) ? cf[0] : fmax(cf[0], fmax(cf[1], cf[2]));
    return CLAWGPU_OK;
}

}  // namespace

static int create3_ctx(clawgpu3_ctx** out, int meqn, int mwaves, int maux, int mx, int my, int mz,
                       int solver_id, const clawgpu_solver3* us)
{
    const int mbc = 2;
    if (mx < 1 || my < 1 || mz < 1) return set_err(CLAWGPU_ERR_ARG, "grid sizes must be >= 1");
    clawgpu3_ctx* c = new (std::nothrow) clawgpu3_ctx();
    if (!c) return set_err(CLAWGPU_ERR_ARG, "out of host memory");
    c->meqn = meqn; c->mwaves = mwaves; c->maux = maux; c->mbc = mbc;
    c->mx = mx; c->my = my; c->mz = mz; c->solver = solver_id;
    c->us = us;
    c->L.mx = mx; c->L.my = my; c->L.mz = mz; c->L.mbc = mbc;
    c->L.nx = mx + 2 * mbc; c->L.ny = my + 2 * mbc; c->L.nz = mz + 2 * mbc;
    c->L.fs = (long)c->L.nx * c->L.ny * c->L.nz;
    c->seg = 64;
    // z-marching blocks of 60x6 columns, each XCD on a contiguous range of
    // blocks (8.2 ms per 512^3 layered step against 13.2 ms for the best
    // tile kernel); without a material table the tile kernel 60x2x4 runs,
    // its tiles numbered z fastest (more halo re-reads served by the caches)
    c->tile = 6;
    c->xcd = 3;
    c->cmax = 0.0;
    const size_t planes = (size_t)c->L.fs * sizeof(double);
    const int naux = maux > 0 ? maux : 1;
    int rc = CLAWGPU_OK;
    auto fail = [&](hipError_t e) {
        snprintf(g_errbuf, sizeof g_errbuf, "device allocation failed: %s", hipGetErrorString(e));
        rc = CLAWGPU_ERR_HIP;
    };
    hipError_t e;
    for (int k = 0; k < 3 && rc == CLAWGPU_OK; k++)
        if ((e = hipMalloc((void**)&c->dq[k], (size_t)meqn * planes)) != hipSuccess) fail(e);
    if (rc == CLAWGPU_OK && (e = hipMalloc((void**)&c->daux, (size_t)naux * planes)) != hipSuccess) fail(e);
    if (rc == CLAWGPU_OK &&
        (e = hipMalloc((void**)&c->staging, (size_t)(meqn > naux ? meqn : naux) * planes)) != hipSuccess)
        fail(e);
    if (rc == CLAWGPU_OK && (e = hipStreamCreate(&c->stream)) != hipSuccess) fail(e);
    if (rc == CLAWGPU_OK && (e = hipMemsetAsync(c->daux, 0, (size_t)naux * planes, c->stream)) != hipSuccess) fail(e);
    if (rc == CLAWGPU_OK && (e = hipStreamSynchronize(c->stream)) != hipSuccess) fail(e);
    if (rc != CLAWGPU_OK) {
        for (int k = 0; k < 3; k++) (void)hipFree(c->dq[k]);
        (void)hipFree(c->daux);
        (void)hipFree(c->staging);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    *out = c;
    return CLAWGPU_OK;
}

extern "C" int clawgpu3_create(clawgpu3_ctx** out, int meqn, int mwaves, int maux, int mbc, int mx,
                               int my, int mz, int solver_id)
{
    if (!out) return set_err(CLAWGPU_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (solver_id != CLAWGPU_SOLVER_VC_ACOUSTICS)
        return set_err(CLAWGPU_ERR_UNSUPPORTED, "the 3D path has the variable-coefficient acoustics solver only (or a user solver: clawgpu3_create_user)");
    if (meqn != 4 || mwaves != 2 || maux < 2)
        return set_err(CLAWGPU_ERR_ARG, "vc acoustics needs meqn=4, mwaves=2, maux>=2 (Z, c)");
    if (mbc != 2) return set_err(CLAWGPU_ERR_UNSUPPORTED, "the 3D path needs mbc = 2");
    return create3_ctx(out, meqn, mwaves, maux, mx, my, mz, solver_id, nullptr);
}

extern "C" int clawgpu3_create_user(clawgpu3_ctx** out, const clawgpu_solver3* s, int maux, int mx,
                                    int my, int mz, const double* params, int nparams)
{
    if (!out) return set_err(CLAWGPU_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (!s) return set_err(CLAWGPU_ERR_ARG, "null solver");
    if (maux < s->maux) return set_err(CLAWGPU_ERR_ARG, "maux is below the solver's aux count");
    if (nparams < 0 || nparams > clawgpu::kUserParams || (nparams > 0 && !params))
        return set_err(CLAWGPU_ERR_ARG, "at most 16 solver parameters");
    clawgpu3_ctx* c = nullptr;
    int rc = create3_ctx(&c, s->meqn, s->mwaves, maux, mx, my, mz, CLAWGPU_SOLVER_USER, s);
    if (rc) return rc;
    c->P.nuser = nparams;
    for (int k = 0; k < nparams; k++) c->P.user[k] = params[k];
    c->exact = s->fast ? 0 : 1;
    *out = c;
    return CLAWGPU_OK;
}

extern "C" int clawgpu3_destroy(clawgpu3_ctx* c)
{
    if (!c) return CLAWGPU_OK;
    (void)hipStreamSynchronize(c->stream);
    for (int k = 0; k < 3; k++) (void)hipFree(c->dq[k]);
    (void)hipFree(c->daux);
    (void)hipFree(c->staging);
    (void)hipFree(c->exy);
    if (c->dmat) (void)hipFree(c->dmat);
    if (c->dlut) (void)hipFree(c->dlut);
    if (c->drec) (void)hipFree(c->drec);
    if (c->dsave) (void)hipFree(c->dsave);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return CLAWGPU_OK;
}

extern "C" int clawgpu3_set_q(clawgpu3_ctx* c, const double* q)
{
    if (!c || !q) return set_err(CLAWGPU_ERR_ARG, "null argument");
    const long n = c->L.fs;
    HIPCHK(c, hipMemcpyAsync(c->staging, q, (size_t)n * c->meqn * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(aos_to_planes3, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                       c->staging, c->dq[c->cur], n, c->meqn, c->meqn, n);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLAWGPU_OK;
}

extern "C" int clawgpu3_get_q(clawgpu3_ctx* c, double* q)
{
    if (!c || !q) return set_err(CLAWGPU_ERR_ARG, "null argument");
    const long n = c->L.fs;
    hipLaunchKernelGGL(planes_to_aos3, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                       c->dq[c->cur], c->staging, n, c->meqn, n);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(q, c->staging, (size_t)n * c->meqn * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLAWGPU_OK;
}

// The distinct (Z, c) pairs of aux (compared as bit patterns, ghost cells
// included) and, when there are at most kMaxMat of them, a byte index per
// cell in the device's cell order (the host order: no padding in 3D).
static int build_material_table(clawgpu3_ctx* c, const double* aux, long n)
{
    c->nmat = 0;
    std::vector<std::pair<uint64_t, uint64_t>> mats;
    std::vector<unsigned char> idx((size_t)n);
    std::pair<uint64_t, uint64_t> last{0, 0};
    int last_i = -1;
    for (long e = 0; e < n; e++) {
        std::pair<uint64_t, uint64_t> k;
        std::memcpy(&k.first, &aux[e * c->maux], 8);
        std::memcpy(&k.second, &aux[e * c->maux + 1], 8);
        if (last_i < 0 || k != last) {
            last_i = -1;
            for (size_t s = 0; s < mats.size(); s++)
                if (mats[s] == k) {
                    last_i = (int)s;
                    break;
                }
            if (last_i < 0) {
                if ((int)mats.size() == kMaxMat) return CLAWGPU_OK;  // too many: read aux
                mats.push_back(k);
                last_i = (int)mats.size() - 1;
            }
            last = k;
        }
        idx[(size_t)e] = (unsigned char)last_i;
    }
    if (!c->dmat) HIPCHK(c, hipMalloc((void**)&c->dmat, (size_t)n));
    if (!c->dlut) HIPCHK(c, hipMalloc((void**)&c->dlut, 2 * kMaxMat * sizeof(double)));
    std::vector<double> lut(2 * mats.size());
    for (size_t s = 0; s < mats.size(); s++) {
        std::memcpy(&lut[2 * s], &mats[s].first, 8);
        std::memcpy(&lut[2 * s + 1], &mats[s].second, 8);
    }
    HIPCHK(c, hipMemcpyAsync(c->dmat, idx.data(), (size_t)n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->dlut, lut.data(), lut.size() * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->nmat = (int)mats.size();
    return CLAWGPU_OK;
}

extern "C" int clawgpu3_set_aux(clawgpu3_ctx* c, const double* aux)
{
    if (!c || !aux) return set_err(CLAWGPU_ERR_ARG, "null argument");
    const long n = c->L.fs;
    if (aux != c->haux.data()) c->haux.assign(aux, aux + (size_t)n * c->maux);
    HIPCHK(c, hipMemcpyAsync(c->staging, aux, (size_t)n * c->maux * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(aos_to_planes3, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                       c->staging, c->daux, n, c->maux, c->maux, n);
    HIPCHK(c, hipGetLastError());
    if (c->us) {  // a user solver reads aux itself: no sound speeds or material table
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return CLAWGPU_OK;
    }
    unsigned long long* bits = reinterpret_cast<unsigned long long*>(c->staging);
    HIPCHK(c, hipMemsetAsync(bits, 0, 2 * sizeof(unsigned long long), c->stream));
    hipLaunchKernelGGL(cmax_kernel, dim3(1024), dim3(256), 0, c->stream, c->daux + n, c->L, bits);
    HIPCHK(c, hipGetLastError());
    unsigned long long h[2] = {0, 0};
    HIPCHK(c, hipMemcpyAsync(h, bits, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double m;
    std::memcpy(&m, &h[0], sizeof m);
    c->cmax = h[1] ? NAN : m;
    return build_material_table(c, aux, n);
}

extern "C" int clawgpu3_set_aux_table(clawgpu3_ctx* c, int enable, int* nmat)
{
    if (!c) return set_err(CLAWGPU_ERR_ARG, "null argument");
    c->use_table = enable ? 1 : 0;
    if (nmat) *nmat = c->nmat;
    return CLAWGPU_OK;
}

extern "C" int clawgpu3_bc3(clawgpu3_ctx* c, const double* geom, double t, const int* mthbc)
{
    if (!c || !geom || !mthbc) return set_err(CLAWGPU_ERR_ARG, "null argument");
    if (check_bc3(mthbc) != 0)
        return set_err(CLAWGPU_ERR_UNSUPPORTED, "mthbc must be 1, 2 or 3 (periodic on both sides), or 0 on z-hi (piston)");
    int rc = enqueue_bc3(c, c->dq[c->cur], geom, t, mthbc);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLAWGPU_OK;
}

extern "C" int clawgpu3_step3(clawgpu3_ctx* c, double dx, double dy, double dz, double dt,
                              const int* method, const int* mthlim, int use_fwave, double* cfl)
{
    if (!c || !method || !mthlim || !cfl) return set_err(CLAWGPU_ERR_ARG, "null argument");
    if (method[2] < 0) return set_err(CLAWGPU_ERR_ARG, "method(3) < 0: use clawgpu3_dimsp3");
    int rc = check_args3(c, method, mthlim);
    if (rc) return rc;
    const int res = (c->cur + 1) % 3;
    rc = enqueue_step3(c, res, dx, dy, dz, dt, method, mthlim, use_fwave);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->cur = res;
    return unsplit_cfl(c, method, dx, dy, dz, dt, cfl);
}

extern "C" int clawgpu3_dimsp3(clawgpu3_ctx* c, double dx, double dy, double dz, double dt,
                               const int* method, const int* mthlim, const double* cflv,
                               int use_fwave, double* cfl)
{
    if (!c || !method || !mthlim || !cflv || !cfl) return set_err(CLAWGPU_ERR_ARG, "null argument");
    if (method[2] != -1) return set_err(CLAWGPU_ERR_ARG, "dimsp3 needs method(3) = -1");
    int rc = check_args3(c, method, mthlim);
    if (rc) return rc;
    int res = c->cur;
    if (c->us) {
        double cf[3] = {0.0, 0.0, 0.0};
        rc = enqueue_dimsp3_user(c, dx, dy, dz, dt, method, mthlim, use_fwave, cflv, &res, cf);
        if (rc == CLAWGPU_OK) *cfl = split_cfl(cf, cflv[0]);
    } else {
        rc = enqueue_dimsp3(c, dx, dy, dz, dt, method, mthlim, cflv, use_fwave, &res, cfl);
    }
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->cur = res;
    return CLAWGPU_OK;
}

// After aux changed on the device (a device b4step3): the sound-speed
// maximum again, and the material table kept only if every cell still
// matches it (otherwise the steps read aux).
static int refresh_aux_on_device(clawgpu3_ctx* c)
{
    if (c->us) return CLAWGPU_OK;
    const long n = c->L.fs;
    unsigned long long* bits = reinterpret_cast<unsigned long long*>(c->staging);
    HIPCHK(c, hipMemsetAsync(bits, 0, 3 * sizeof(unsigned long long), c->stream));
    hipLaunchKernelGGL(cmax_kernel, dim3(1024), dim3(256), 0, c->stream, c->daux + n, c->L, bits);
    if (c->nmat > 0)
        hipLaunchKernelGGL(table_check_kernel, dim3(1024), dim3(256), 0, c->stream, c->daux,
                           c->daux + n, c->dmat, c->dlut, n, bits + 2);
    HIPCHK(c, hipGetLastError());
    unsigned long long h[3] = {0, 0, 0};
    HIPCHK(c, hipMemcpyAsync(h, bits, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double m;
    std::memcpy(&m, &h[0], sizeof m);
    c->cmax = h[1] ? NAN : m;
    if (h[2]) c->nmat = 0;
    return CLAWGPU_OK;
}

// One launch of device routine f on q (one thread per cell, ghost cells
// included, blocks of 256).
static int launch_dev_hook3(clawgpu3_ctx* c, hipFunction_t f, const clawgpu_dev_hooks* h, double* q,
                            const double* geom, double t, double dt, const int* mthbc)
{
    clawgpu_dev_grid3 g;
    g.q = q;
    g.aux = c->maux > 0 ? c->daux : nullptr;
    g.user = h->user;
    g.fs = c->L.fs;
    g.meqn = c->meqn;
    g.maux = c->maux;
    g.mbc = c->mbc;
    g.mx = c->mx;
    g.my = c->my;
    g.mz = c->mz;
    for (int k = 0; k < 6; k++) g.mthbc[k] = mthbc[k];
    g.xlower = geom[0];
    g.ylower = geom[1];
    g.zlower = geom[2];
    g.dx = geom[3];
    g.dy = geom[4];
    g.dz = geom[5];
    g.t = t;
    g.dt = dt;
    void* args[] = {&g};
    const hipError_t e = hipModuleLaunchKernel(f, (unsigned)((c->L.fs + 255) / 256), 1, 1, 256, 1, 1,
                                               0, c->stream, args, nullptr);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return set_err(CLAWGPU_ERR_HIP, "user routine launch failed");
    }
    return CLAWGPU_OK;
}

// The claw3 loop (src/3d/claw3.f:735-904).  The Courant number of this
// solver is dt/d * max|c|, known before the step, so an attempt that would be
// rejected is not run (the reference runs it, then discards the result and
// restores q).  With hooks, the user's bc3, b4step3 and src3 are called on a
// host copy of q at the reference's points (claw3.f:766-885): bc3 and b4step3
// before every attempt, src3 over dt/2 before the step (its input kept for a
// retake) and over dt/2 after it (method(5) = 2), or over dt after it
// (method(5) = 1).  Device routines (dh) run on the device q at the same
// points; the state before a Strang half step is kept on the device.
static int claw3_loop(clawgpu3_ctx* c, const double* geom, double tstart, double* tend,
                      double* dtv, double* cflv, int* nv, const int* method, const int* mthlim,
                      const int* mthbc, int use_fwave, const clawgpu3_hooks* h,
                      const clawgpu_dev_hooks* dh, int* info_out)
{
    if (!c || !geom || !tend || !dtv || !cflv || !nv || !method || !mthlim || !mthbc || !info_out)
        return set_err(CLAWGPU_ERR_ARG, "null argument");
    const bool dev_bc = dh && dh->bc2;
    const bool user_bc = (h && h->bc3) || dev_bc;
    const bool dev_src = dh && dh->src2;
    const size_t qbytes = (size_t)c->L.fs * c->meqn * sizeof(double);
    auto dev_call = [&](hipFunction_t f, double tt, double dtt) -> int {
        return launch_dev_hook3(c, f, dh, c->dq[c->cur], geom, tt, dtt, mthbc);
    };
    if (!user_bc && check_bc3(mthbc) != 0)
        return set_err(CLAWGPU_ERR_UNSUPPORTED, "mthbc must be 1, 2 or 3 (periodic on both sides), or 0 on z-hi (piston)");
    const double xl = geom[0], yl = geom[1], zl = geom[2];
    const double dx = geom[3], dy = geom[4], dz = geom[5];
    std::vector<double> hq, saved;  // host copies of q for the hooks
    double* haux = c->maux > 0 && !c->haux.empty() ? c->haux.data() : nullptr;
    auto pull = [&]() -> int {
        hq.resize((size_t)c->L.fs * c->meqn);
        return clawgpu3_get_q(c, hq.data());
    };
    auto push = [&]() -> int { return clawgpu3_set_q(c, hq.data()); };
    int info = 0, rc;
    double t = tstart;
    int maxn = nv[0];
    double dt = dtv[0];
    double cflmax = 0.0, dtmin = dt, dtmax = dt, cfl = 0.0;
    nv[1] = 0;
    if (check_method3(method) == 6) { info = 6; goto done; }
    rc = check_args3(c, method, mthlim);
    if (rc) return rc;
    if (method[0] == 0) {
        if (*tend < tstart) {
            maxn = 1;
        } else {
            maxn = (int)lround((*tend - tstart) / dt);
            if (fabs(maxn * dt - (*tend - tstart)) > 1e-5 * (*tend - tstart)) { info = 2; goto done; }
        }
    }
    if (method[0] == 1 && cflv[1] > cflv[0]) { info = 3; goto done; }
    if (method[5] > method[6]) { info = 5; goto done; }
    if (maxn == 0) goto done;
    for (int n = 1; n <= maxn; n++) {
        double told = t;
        if (told + dt > *tend && tstart < *tend) dt = *tend - told;
        for (;;) {
            if (dev_bc) {
                if ((rc = dev_call(dh->bc2, told, dt))) return rc;
            } else if (user_bc) {
                if ((rc = pull())) return rc;
                h->bc3(h->user, c->meqn, c->mbc, c->mx, c->my, c->mz, xl, yl, zl, dx, dy, dz,
                       hq.data(), c->maux, haux, told, dt, mthbc);
                if ((rc = push())) return rc;
            } else {
                rc = enqueue_bc3(c, c->dq[c->cur], geom, told, mthbc);
                if (rc) return rc;
            }
            if (h && h->b4step3) {
                if ((rc = pull())) return rc;
                h->b4step3(h->user, c->mbc, c->mx, c->my, c->mz, c->meqn, hq.data(), xl, yl, zl,
                           dx, dy, dz, told, dt, c->maux, haux);
                if ((rc = push())) return rc;
                if (haux && (rc = clawgpu3_set_aux(c, c->haux.data()))) return rc;
            } else if (dh && dh->b4step2) {
                if ((rc = dev_call(dh->b4step2, told, dt))) return rc;
                if (c->maux > 0 && (rc = refresh_aux_on_device(c))) return rc;
            }
            const bool strang_src = method[4] == 2 && h && h->src3;
            const bool dev_strang = method[4] == 2 && dev_src;
            if (dev_strang) {
                if (!c->dsave) HIPCHK(c, hipMalloc((void**)&c->dsave, qbytes));
                HIPCHK(c, hipMemcpyAsync(c->dsave, c->dq[c->cur], qbytes, hipMemcpyDeviceToDevice,
                                         c->stream));
                if ((rc = dev_call(dh->src2, told, dt / 2.0))) return rc;
            }
            if (strang_src) {
                if ((rc = pull())) return rc;
                saved = hq;
                h->src3(h->user, c->meqn, c->mbc, c->mx, c->my, c->mz, xl, yl, zl, dx, dy, dz,
                        hq.data(), c->maux, haux, told, dt / 2.0);
                if ((rc = push())) return rc;
            }
            double cf[3] = {0.0, 0.0, 0.0};
            int ures = c->cur;  // a user solver's step is taken before its Courant number is known
            if (c->us) {
                if ((rc = enqueue_dimsp3_user(c, dx, dy, dz, dt, method, mthlim, use_fwave, nullptr,
                                              &ures, cf)))
                    return rc;
                cfl = split_cfl(cf, cflv[0]);
                if (!std::isfinite(cfl)) {
                    *info_out = -5;
                    return set_err(CLAWGPU_ERR_NONFINITE, "non-finite wave speed (Courant number)");
                }
            } else {
                if ((rc = courant3(c, method, dx, dy, dz, dt, cf))) return rc;
                if (!std::isfinite(c->cmax) || !std::isfinite(cf[0]) || !std::isfinite(cf[1]) ||
                    !std::isfinite(cf[2])) {
                    *info_out = -5;
                    return set_err(CLAWGPU_ERR_NONFINITE, "non-finite sound speed (Courant number)");
                }
                if (method[2] >= 0) cfl = fmax(cf[0], fmax(cf[1], cf[2]));
                else cfl = split_cfl(cf, cflv[0]);
            }
            t = told + dt;
            if (cfl <= cflv[0]) {
                int res = (c->cur + 1) % 3;
                if (c->us) res = ures;
                else if (method[2] >= 0) rc = enqueue_step3(c, res, dx, dy, dz, dt, method, mthlim, use_fwave);
                else {
                    double cfs;
                    rc = enqueue_dimsp3(c, dx, dy, dz, dt, method, mthlim, cflv, use_fwave, &res, &cfs);
                }
                if (rc) return rc;
                cflmax = fmax(cfl, cflmax);
                c->cur = res;
                break;
            }
            t = told;
            if (strang_src) {  // back to the state before the source half step
                hq = saved;
                if ((rc = push())) return rc;
            }
            if (dev_strang)
                HIPCHK(c, hipMemcpyAsync(c->dq[c->cur], c->dsave, qbytes, hipMemcpyDeviceToDevice,
                                         c->stream));
            if (method[0] == 1) {
                dt = fmin(dtv[1], dt * cflv[1] / cfl);
                continue;
            }
            cflmax = fmax(cfl, cflmax);
            goto done;
        }
        if (h && h->src3 && (method[4] == 1 || method[4] == 2)) {
            if ((rc = pull())) return rc;
            h->src3(h->user, c->meqn, c->mbc, c->mx, c->my, c->mz, xl, yl, zl, dx, dy, dz,
                    hq.data(), c->maux, haux, t, method[4] == 2 ? dt / 2.0 : dt);
            if ((rc = push())) return rc;
        } else if (dev_src && (method[4] == 1 || method[4] == 2)) {
            if ((rc = dev_call(dh->src2, t, method[4] == 2 ? dt / 2.0 : dt))) return rc;
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
    HIPCHK(c, hipStreamSynchronize(c->stream));
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

extern "C" int clawgpu3_claw3(clawgpu3_ctx* c, const double* geom, double tstart, double* tend,
                              double* dtv, double* cflv, int* nv, const int* method,
                              const int* mthlim, const int* mthbc, int use_fwave, int* info_out)
{
    return claw3_loop(c, geom, tstart, tend, dtv, cflv, nv, method, mthlim, mthbc, use_fwave,
                      nullptr, nullptr, info_out);
}

extern "C" int clawgpu3_claw3_dev_hooks(clawgpu3_ctx* c, const double* geom, double tstart,
                                        double* tend, double* dtv, double* cflv, int* nv,
                                        const int* method, const int* mthlim, const int* mthbc,
                                        int use_fwave, const clawgpu_dev_hooks* hooks,
                                        int* info_out)
{
    if (!hooks) return set_err(CLAWGPU_ERR_ARG, "null hooks (use clawgpu3_claw3)");
    return claw3_loop(c, geom, tstart, tend, dtv, cflv, nv, method, mthlim, mthbc, use_fwave,
                      nullptr, hooks, info_out);
}

extern "C" int clawgpu3_claw3_hooks(clawgpu3_ctx* c, const double* geom, double tstart,
                                    double* tend, double* dtv, double* cflv, int* nv,
                                    const int* method, const int* mthlim, const int* mthbc,
                                    int use_fwave, const clawgpu3_hooks* hooks, int* info_out)
{
    if (!hooks) return set_err(CLAWGPU_ERR_ARG, "null hooks (use clawgpu3_claw3)");
    if (c && c->maux > 0 && c->haux.empty() && (hooks->b4step3 || hooks->src3 || hooks->bc3))
        return set_err(CLAWGPU_ERR_ARG, "aux has not been set (clawgpu3_set_aux)");
    return claw3_loop(c, geom, tstart, tend, dtv, cflv, nv, method, mthlim, mthbc, use_fwave,
                      hooks, nullptr, info_out);
}

extern "C" int clawgpu3_set_segment(clawgpu3_ctx* c, int cells)
{
    if (!c || cells < 1) return set_err(CLAWGPU_ERR_ARG, "cells must be >= 1");
    c->seg = cells;
    return CLAWGPU_OK;
}

extern "C" int clawgpu3_set_tile(clawgpu3_ctx* c, int tile, int xcd_order)
{
    if (!c || tile < 0 || tile > 6) return set_err(CLAWGPU_ERR_ARG, "tile must be 0..6");
    if (xcd_order < 0 || xcd_order > 3) return set_err(CLAWGPU_ERR_ARG, "tile order must be 0..3");
    c->tile = tile;
    c->xcd = xcd_order;
    return CLAWGPU_OK;
}

extern "C" int clawgpu3_get_tile(const clawgpu3_ctx* c, int* tile, int* xcd_order)
{
    if (!c || !tile || !xcd_order) return set_err(CLAWGPU_ERR_ARG, "null argument");
    *tile = c->tile;
    *xcd_order = c->xcd;
    return CLAWGPU_OK;
}

extern "C" int clawgpu3_time_steps(clawgpu3_ctx* c, const double* geom, double dt,
                                   const int* method, const int* mthlim, const int* mthbc,
                                   int use_fwave, int nsteps, float* total_ms,
                                   float* step_kernel_ms, double* cfl_out)
{
    // Fixed-dt stepping (unsplit, or dimensionally split with method(3) =
    // -1, all three sweeps) timed with HIP events on the library's stream,
    // no host synchronisation between steps: total_ms spans all launches
    // (ghost fill + step kernels per step), step_kernel_ms sums the event
    // pairs around each step's kernels.  Used by the benchmark.
    if (!c || !geom || !method || !mthlim || !mthbc || !total_ms || nsteps < 1)
        return set_err(CLAWGPU_ERR_ARG, "bad argument");
    if (method[2] < -1) return set_err(CLAWGPU_ERR_ARG, "method(3) must be -1 or >= 0");
    if (c->us) return set_err(CLAWGPU_ERR_UNSUPPORTED, "timed steps run the built-in solver; time a user solver's claw3");
    int rc = check_args3(c, method, mthlim);
    if (rc) return rc;
    if (check_bc3(mthbc) != 0) return set_err(CLAWGPU_ERR_UNSUPPORTED, "bad mthbc");
    const double dx = geom[3], dy = geom[4], dz = geom[5];
    hipEvent_t* ev = new (std::nothrow) hipEvent_t[2 * nsteps + 2];
    if (!ev) return set_err(CLAWGPU_ERR_ARG, "out of host memory");
    for (int k = 0; k < 2 * nsteps + 2; k++) HIPCHK(c, hipEventCreate(&ev[k]));
    HIPCHK(c, hipEventRecord(ev[0], c->stream));
    for (int n = 0; n < nsteps; n++) {
        rc = enqueue_bc3(c, c->dq[c->cur], geom, n * dt, mthbc);
        if (rc) return rc;
        HIPCHK(c, hipEventRecord(ev[2 + 2 * n], c->stream));
        int res = (c->cur + 1) % 3;
        if (method[2] == -1) {
            const double all[2] = {1e300, 1e300};  // every sweep runs
            double cf;
            rc = enqueue_dimsp3(c, dx, dy, dz, dt, method, mthlim, all, use_fwave, &res, &cf);
        } else {
            rc = enqueue_step3(c, res, dx, dy, dz, dt, method, mthlim, use_fwave);
        }
        if (rc) return rc;
        HIPCHK(c, hipEventRecord(ev[3 + 2 * n], c->stream));
        c->cur = res;
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
    if (cfl_out) return unsplit_cfl(c, method, dx, dy, dz, dt, cfl_out);
    return CLAWGPU_OK;
}

extern "C" int clawgpu3_set_arith(clawgpu3_ctx* c, int mode)
{
    if (!c || (mode != 0 && mode != 1)) return set_err(CLAWGPU_ERR_ARG, "mode must be 0 or 1");
    if (c->us && mode != (c->us->fast ? 1 : 0))
        return set_err(CLAWGPU_ERR_UNSUPPORTED, "a user solver's arithmetic is fixed when its code object is built");
    c->exact = mode == 0 ? 1 : 0;
    return CLAWGPU_OK;
}
