// Fused 2D wave-propagation step for gfx950 (FP64).
//
// Replaces, for one unsplit time step, the reference's step2 -> flux2 ->
// rpn2/limiter/rpt2 path (src/2d/step2.f90:2-251, src/2d/flux2.f90:5-235,
// src/2d/inlinelimiter.f90:25-105) plus the bc2 ghost fill
// (src/2d/bc2.f:4-215), in one kernel launch per step.
//
// Work decomposition ("column sweep").  Every 64-lane wavefront owns a strip
// of 64 adjacent columns and a segment of rows.  Lane l holds column
// i = 61*strip + l; lanes 1..61 produce output cells, lanes 0, 62 and 63 are
// halo.  The wave walks down its segment one row J per iteration:
//   * x-interfaces of row J are solved one per lane; the limiter's
//     neighbouring waves and the contributions owed to the left neighbour
//     cell come from adjacent lanes through cross-lane shuffles;
//   * the y-interface between rows J-1 and J is solved per lane; the limiter
//     needs the interfaces above and below, which the lane still holds in
//     registers from the previous iteration, so the y-direction pipeline
//     lags one row and finishes cell (i, J-2) at iteration J.
// Every cell update is accumulated in registers in the reference's exact
// order (x-sweep rows j-1, j, j+1, then y-sweep columns i-1, i, i+1, with the
// parenthesisation of step2.f90:141-145 and :224-228), so with FMA
// contraction off the result equals the reference arithmetic bit for bit.
// q is read once from HBM (plus a 3-column / 3-row halo) and written once.
//
// Device layout: field planes (structure of arrays).  Element (m, i, j) of a
// grid with mbc ghost cells lives at  base + m*fs + (j+mbc-1)*pitch +
// (i+mbc-1), so the 64 lanes of a wave read 64 consecutive doubles of one
// row of one field (512 B, coalesced).

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>

#include "riemann2d.hpp"
#include "../../include/clawgpu.h"

namespace clawgpu {

constexpr int kLanes = 64;
constexpr int kOut = 61;  // output columns per wave strip
constexpr int kWavesPerBlock = 4;
constexpr int kMaxWaves = 8;

struct Layout2 {
    int mx, my, mbc, pitch;
    long fs;
    __host__ __device__ long off(int i, int j) const
    {
        return (long)(j + mbc - 1) * pitch + (i + mbc - 1);
    }
};

struct StepArgs {
    const double* qold;
    double* qnew;
    Layout2 L;
    double dtdx, dtdy;
    int order, trans, use_fwave;
    int lim[kMaxWaves];
    int nstrips, seglen, nsegs;
    unsigned long long* cfl_bits;  // max of the Courant number, as raw bits
    SolverParams P;
};

__device__ inline double shup(double v) { return __shfl_up(v, 1); }
__device__ inline double shdn(double v) { return __shfl_down(v, 1); }

// Max of a non-negative double over the wave, then one atomic per wave.
// A NaN speed is reported as +inf so that the host can refuse the step.
__device__ inline void publish_cfl(double cfl, unsigned long long* bits)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) cfl = fmax(cfl, __shfl_xor(cfl, d));
    if ((threadIdx.x & 63) == 0) {
        unsigned long long b = (unsigned long long)__double_as_longlong(cfl);
        atomicMax(bits, b);
    }
}

__device__ inline double cfl_term(double dtdx, double s)
{
    double t = fmax(dtdx * s, -(dtdx * s));
    return (t == t) ? t : INFINITY;
}

template <class S>
__global__ __launch_bounds__(256) void step2_unsplit_kernel(StepArgs a)
{
    constexpr int M = S::M, MW = S::MW;
    const int lane = threadIdx.x & 63;
    const int wid = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (wid >= a.nstrips * a.nsegs) return;  // uniform per wave
    const int strip = wid % a.nstrips;
    const int seg = wid / a.nstrips;
    const Layout2 L = a.L;
    const int mx = L.mx, my = L.my, mbc = L.mbc;
    const int i = kOut * strip + lane;
    const int ic = i < mx + mbc ? i : mx + mbc;          // clamped column for loads
    const int icl = (i - 1 < mx + mbc) ? i - 1 : mx + mbc;  // left neighbour (lane 0 only)
    const int r0 = 1 + seg * a.seglen;
    const int r1 = min(r0 + a.seglen, my + 1);  // output rows r0 .. r1-1
    const bool out_lane = lane >= 1 && lane <= kOut && i <= mx;
    const bool xcfl_lane = i >= 1 && i <= mx + 1;
    const bool ycfl_lane = i <= mx + 1;
    const double dtdx = a.dtdx, dtdy = a.dtdy;
    const bool second = a.order == 2;
    const bool tr2 = second && a.trans == 2;
    const double* __restrict__ qo = a.qold;
    double* __restrict__ qn = a.qnew;

    double qprev[M];
#pragma unroll
    for (int m = 0; m < M; m++) qprev[m] = qo[m * L.fs + L.off(ic, r0 - 2)];

    double vA[M], vB[M], g2prev[M];
    double wyp[MW][M], syp[MW], amyp[M], apyp[M], dotly[MW];
    double qa_part[M], fa_part[M], gp1[M], gp2[M];
    typename S::Cache cyp;
#pragma unroll
    for (int m = 0; m < M; m++) {
        vA[m] = vB[m] = g2prev[m] = 0.0;
        amyp[m] = apyp[m] = qa_part[m] = fa_part[m] = gp1[m] = gp2[m] = 0.0;
    }
#pragma unroll
    for (int mw = 0; mw < MW; mw++) {
        syp[mw] = dotly[mw] = 0.0;
#pragma unroll
        for (int m = 0; m < M; m++) wyp[mw][m] = 0.0;
    }
    double cfl = 0.0;

    for (int J = r0 - 1; J <= r1 + 1; ++J) {
        double qc[M];
#pragma unroll
        for (int m = 0; m < M; m++) qc[m] = qo[m * L.fs + L.off(ic, J)];
        double vC[M];

        // ------------------------------------------------ x-row J (flux2, ixy=1)
        if (J <= r1) {
            double ql[M];
#pragma unroll
            for (int m = 0; m < M; m++) {
                ql[m] = shup(qc[m]);
                if (lane == 0) ql[m] = qo[m * L.fs + L.off(icl, J)];
            }
            double wx[MW][M], sx[MW], amx[M], apx[M];
            typename S::Cache cx;
            S::template rpn<1, 2>(a.P, ql, qc, wx, sx, amx, apx, cx);
            if (xcfl_lane) {
#pragma unroll
                for (int mw = 0; mw < MW; mw++) cfl = fmax(cfl, cfl_term(dtdx, sx[mw]));
            }
            double cq[M], fa[M];
#pragma unroll
            for (int m = 0; m < M; m++) cq[m] = fa[m] = 0.0;
            if (second) {
                double dotr[MW], wn2[MW];
#pragma unroll
                for (int mw = 0; mw < MW; mw++) {
                    double dr = 0.0, nn = 0.0;
#pragma unroll
                    for (int m = 0; m < M; m++) {
                        double wr = shdn(wx[mw][m]);
                        nn = nn + wx[mw][m] * wx[mw][m];
                        dr = dr + wx[mw][m] * wr;
                    }
                    dotr[mw] = dr;
                    wn2[mw] = nn;
                }
#pragma unroll
                for (int mw = 0; mw < MW; mw++) {
                    double dl = shup(dotr[mw]);
                    if (a.lim[mw] != 0 && wn2[mw] != 0.0) {
                        double r = (sx[mw] > 0.0 ? dl : dotr[mw]) / wn2[mw];
                        double phi = philim(r, a.lim[mw]);
#pragma unroll
                        for (int m = 0; m < M; m++) wx[mw][m] = phi * wx[mw][m];
                    }
                }
#pragma unroll
                for (int mw = 0; mw < MW; mw++) {
                    double sa = fabs(sx[mw]);
                    double sw = a.use_fwave ? copysign(1.0, sx[mw]) : sa;
                    double coef = sw * (1.0 - sa * dtdx);
#pragma unroll
                    for (int m = 0; m < M; m++) cq[m] = cq[m] + coef * wx[mw][m];
                }
#pragma unroll
                for (int m = 0; m < M; m++) fa[m] = 0.5 * cq[m];
            }
            double G1[M], G2[M];
            if (a.trans > 0) {
                double amt[M], apt[M];
#pragma unroll
                for (int m = 0; m < M; m++) {
                    amt[m] = tr2 ? amx[m] + cq[m] : amx[m];
                    apt[m] = tr2 ? apx[m] - cq[m] : apx[m];
                }
                double bma[M], bpa[M], bmp[M], bpp[M];
                S::template rpt<1, 2>(a.P, cx, amt, bma, bpa);
                S::template rpt<1, 2>(a.P, cx, apt, bmp, bpp);
#pragma unroll
                for (int m = 0; m < M; m++) {
                    double g1n = shdn(-0.5 * dtdx * bma[m]);
                    double g2n = shdn(-0.5 * dtdx * bpa[m]);
                    G1[m] = g1n - 0.5 * dtdx * bmp[m];
                    G2[m] = g2n - 0.5 * dtdx * bpp[m];
                }
            } else {
#pragma unroll
                for (int m = 0; m < M; m++) G1[m] = G2[m] = 0.0;
            }
#pragma unroll
            for (int m = 0; m < M; m++) {
                double amn = shdn(amx[m]);
                double fan = shdn(fa[m]);
                double qadd = (0.0 - dtdx * apx[m]) - dtdx * amn;
                double v = qc[m] + dtdy * g2prev[m];
                v = v + qadd - dtdx * (fan - fa[m]) - dtdy * (G2[m] - G1[m]);
                vC[m] = v;
                vB[m] = vB[m] - dtdy * G1[m];
                g2prev[m] = G2[m];
            }
        } else {
#pragma unroll
            for (int m = 0; m < M; m++) vC[m] = 0.0;
        }

        // ------------------------------ y-interface J (flux2, ixy=2) on column i
        double wy[MW][M], sy[MW], amy[M], apy[M];
        typename S::Cache cy;
        S::template rpn<2, 1>(a.P, qprev, qc, wy, sy, amy, apy, cy);
        if (ycfl_lane && J >= 1 && J <= my + 1) {
#pragma unroll
            for (int mw = 0; mw < MW; mw++) cfl = fmax(cfl, cfl_term(dtdy, sy[mw]));
        }
        double dotry[MW];
#pragma unroll
        for (int mw = 0; mw < MW; mw++) {
            double dr = 0.0;
#pragma unroll
            for (int m = 0; m < M; m++) dr = dr + wyp[mw][m] * wy[mw][m];
            dotry[mw] = dr;
        }
        if (J - 1 >= r0) {
            // limit interface J-1 (held in wyp) with its neighbours J-2 and J
            double cq[M], fa[M];
#pragma unroll
            for (int m = 0; m < M; m++) cq[m] = fa[m] = 0.0;
            if (second) {
#pragma unroll
                for (int mw = 0; mw < MW; mw++) {
                    double nn = 0.0;
#pragma unroll
                    for (int m = 0; m < M; m++) nn = nn + wyp[mw][m] * wyp[mw][m];
                    if (a.lim[mw] != 0 && nn != 0.0) {
                        double r = (syp[mw] > 0.0 ? dotly[mw] : dotry[mw]) / nn;
                        double phi = philim(r, a.lim[mw]);
#pragma unroll
                        for (int m = 0; m < M; m++) wyp[mw][m] = phi * wyp[mw][m];
                    }
                }
#pragma unroll
                for (int mw = 0; mw < MW; mw++) {
                    double sa = fabs(syp[mw]);
                    double sw = a.use_fwave ? copysign(1.0, syp[mw]) : sa;
                    double coef = sw * (1.0 - sa * dtdy);
#pragma unroll
                    for (int m = 0; m < M; m++) cq[m] = cq[m] + coef * wyp[mw][m];
                }
#pragma unroll
                for (int m = 0; m < M; m++) fa[m] = 0.5 * cq[m];
            }
            double bma[M], bpa[M], bmp[M], bpp[M];
            if (a.trans > 0) {
                double amt[M], apt[M];
#pragma unroll
                for (int m = 0; m < M; m++) {
                    amt[m] = tr2 ? amyp[m] + cq[m] : amyp[m];
                    apt[m] = tr2 ? apyp[m] - cq[m] : apyp[m];
                }
                S::template rpt<2, 1>(a.P, cyp, amt, bma, bpa);
                S::template rpt<2, 1>(a.P, cyp, apt, bmp, bpp);
            } else {
#pragma unroll
                for (int m = 0; m < M; m++) bma[m] = bpa[m] = bmp[m] = bpp[m] = 0.0;
            }
            if (J - 2 >= r0) {
                // cell (i, J-2): its y-sweep terms are now complete
#pragma unroll
                for (int m = 0; m < M; m++) {
                    double qadd = qa_part[m] - dtdy * amyp[m];
                    double dF = fa[m] - fa_part[m];
                    double G1 = a.trans > 0 ? (-0.5 * dtdy * bma[m]) - gp1[m] : 0.0;
                    double G2 = a.trans > 0 ? (-0.5 * dtdy * bpa[m]) - gp2[m] : 0.0;
                    double g2l = shup(G2);
                    double g1r = shdn(G1);
                    double v = vA[m] + dtdx * g2l;
                    v = v + (qadd - dtdy * dF - dtdx * (G2 - G1));
                    v = v - dtdx * g1r;
                    if (out_lane) qn[m * L.fs + L.off(i, J - 2)] = v;
                }
            }
#pragma unroll
            for (int m = 0; m < M; m++) {
                qa_part[m] = 0.0 - dtdy * apyp[m];
                fa_part[m] = fa[m];
                gp1[m] = 0.5 * dtdy * bmp[m];
                gp2[m] = 0.5 * dtdy * bpp[m];
            }
        }
        // ------------------------------------------------------ shift the pipeline
#pragma unroll
        for (int mw = 0; mw < MW; mw++) {
            dotly[mw] = dotry[mw];
            syp[mw] = sy[mw];
#pragma unroll
            for (int m = 0; m < M; m++) wyp[mw][m] = wy[mw][m];
        }
#pragma unroll
        for (int m = 0; m < M; m++) {
            amyp[m] = amy[m];
            apyp[m] = apy[m];
            vA[m] = vB[m];
            vB[m] = vC[m];
            qprev[m] = qc[m];
        }
        cyp = cy;
    }
    publish_cfl(cfl, a.cfl_bits);
}

// ------------------------------------------------------------------ bc2
// Every ghost cell is written from the interior in one pass: the row index
// is mapped first (bottom/top), then the column (left/right), with the wall
// sign flips of q(3) and q(2).  This equals the reference's sequential fill
// (left, right, bottom, top; corners from the bottom/top pass over the
// already-filled ghost columns) because each side's rule is an index map
// plus a sign.  Thread 0 also clears the Courant accumulator for the step.
struct BcArgs {
    double* q;
    Layout2 L;
    int meqn;
    int mthbc[4];
    unsigned long long* cfl_bits;
};

__device__ inline int bc_map(int k, int n, int mbc, int lo, int hi, bool& flip)
{
    (void)mbc;
    flip = false;
    if (k < 1) {
        int g = 1 - k;  // ghost layer 1..mbc
        if (lo == 1) return 1;
        if (lo == 2) return n + 1 - g;
        flip = true;
        return g;
    }
    if (k > n) {
        int g = k - n;
        if (hi == 1) return n;
        if (hi == 2) return g;
        flip = true;
        return n + 1 - g;
    }
    return k;
}

__global__ void bc2_kernel(BcArgs a)
{
    const Layout2 L = a.L;
    const int nx = L.mx + 2 * L.mbc, ny = L.my + 2 * L.mbc;
    const int ng = 2 * L.mbc * nx + 2 * L.mbc * L.my;  // ghost rows + ghost columns of interior rows
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0 && a.cfl_bits) *a.cfl_bits = 0ull;
    if (t >= ng) return;
    int i, j;
    if (t < 2 * L.mbc * nx) {
        int r = t / nx;  // 0..2mbc-1
        i = 1 - L.mbc + t % nx;
        j = r < L.mbc ? 1 - L.mbc + r : L.my + 1 + (r - L.mbc);
    } else {
        int u = t - 2 * L.mbc * nx;
        int c = u / L.my;  // 0..2mbc-1
        j = 1 + u % L.my;
        i = c < L.mbc ? 1 - L.mbc + c : L.mx + 1 + (c - L.mbc);
    }
    (void)ny;
    bool fy, fx;
    int js = bc_map(j, L.my, L.mbc, a.mthbc[2], a.mthbc[3], fy);
    int is = bc_map(i, L.mx, L.mbc, a.mthbc[0], a.mthbc[1], fx);
    long src = L.off(is, js), dst = L.off(i, j);
    for (int m = 0; m < a.meqn; m++) {
        double v = a.q[m * L.fs + src];
        if ((m == 1 && fx) || (m == 2 && fy)) v = -v;
        a.q[m * L.fs + dst] = v;
    }
}

// ------------------------------------------------- layout conversion kernels
// Reference layout q(meqn, 1-mbc:mx+mbc, 1-mbc:my+mbc) (m fastest) <-> planes.
__global__ void aos_to_planes(const double* __restrict__ src, double* __restrict__ dst, Layout2 L,
                              int meqn)
{
    const int nx = L.mx + 2 * L.mbc, ny = L.my + 2 * L.mbc;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nx * ny) return;
    int ii = (int)(t % nx), jj = (int)(t / nx);
    for (int m = 0; m < meqn; m++) dst[m * L.fs + (long)jj * L.pitch + ii] = src[t * meqn + m];
}

__global__ void planes_to_aos(const double* __restrict__ src, double* __restrict__ dst, Layout2 L,
                              int meqn)
{
    const int nx = L.mx + 2 * L.mbc, ny = L.my + 2 * L.mbc;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nx * ny) return;
    int ii = (int)(t % nx), jj = (int)(t / nx);
    for (int m = 0; m < meqn; m++) dst[t * meqn + m] = src[m * L.fs + (long)jj * L.pitch + ii];
}

}  // namespace clawgpu

// =====================================================================
// Host side: context, C-ABI.
// =====================================================================
using namespace clawgpu;

struct clawgpu_ctx {
    int meqn, mwaves, maux, mbc, mx, my, solver;
    SolverParams P;
    Layout2 L;
    double* dq[2];
    int cur;
    double* staging;
    unsigned long long* dcfl;
    double* hcfl;  // pinned
    hipStream_t stream;
    int seglen;
    int last_hip_error;
};

static thread_local char g_errbuf[256];

static int set_err(int code, const char* msg)
{
    snprintf(g_errbuf, sizeof g_errbuf, "%s", msg);
    return code;
}

#define HIPCHK(ctx, call)                                                              \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            snprintf(g_errbuf, sizeof g_errbuf, "%s: %s", #call, hipGetErrorString(e_)); \
            if (ctx) (ctx)->last_hip_error = (int)e_;                                  \
            return CLAWGPU_ERR_HIP;                                                    \
        }                                                                              \
    } while (0)

extern "C" const char* clawgpu_last_error(void) { return g_errbuf; }

extern "C" int clawgpu_abi_version(void) { return CLAWGPU_ABI_VERSION; }

static int solver_shape(int solver, int* meqn, int* mwaves)
{
    switch (solver) {
    case CLAWGPU_SOLVER_ACOUSTICS:
        *meqn = Acoustics2::M;
        *mwaves = Acoustics2::MW;
        return 0;
    case CLAWGPU_SOLVER_EULER:
        *meqn = Euler2::M;
        *mwaves = Euler2::MW;
        return 0;
    default: return -1;
    }
}

extern "C" int clawgpu_create(clawgpu_ctx** out, int meqn, int mwaves, int maux, int mbc, int mx,
                              int my, int solver_id, const double* params, int nparams)
{
    if (!out) return set_err(CLAWGPU_ERR_ARG, "null output pointer");
    *out = nullptr;
    int sm, sw;
    if (solver_shape(solver_id, &sm, &sw) != 0) return set_err(CLAWGPU_ERR_UNSUPPORTED, "unknown solver id");
    if (meqn != sm || mwaves != sw)
        return set_err(CLAWGPU_ERR_ARG, "meqn/mwaves do not match the solver");
    if (maux != 0) return set_err(CLAWGPU_ERR_UNSUPPORTED, "aux arrays are not supported by this solver");
    if (mbc < 2) return set_err(CLAWGPU_ERR_ARG, "mbc must be >= 2 (claw2 info=1)");
    if (mx < 1 || my < 1) return set_err(CLAWGPU_ERR_ARG, "grid must have at least one cell");
    SolverParams P{};
    if (solver_id == CLAWGPU_SOLVER_ACOUSTICS) {
        if (nparams < 2 || !params) return set_err(CLAWGPU_ERR_ARG, "acoustics needs (rho, bulk)");
        P.rho = params[0];
        P.bulk = params[1];
        if (!(P.rho > 0.0) || !(P.bulk > 0.0)) return set_err(CLAWGPU_ERR_ARG, "rho and bulk must be positive");
        P.cc = sqrt(P.bulk / P.rho);
        P.zz = P.rho * P.cc;
    } else {
        if (nparams < 2 || !params) return set_err(CLAWGPU_ERR_ARG, "euler needs (gamma, efix)");
        P.gamma = params[0];
        P.gamma1 = params[0] - 1.0;
        P.efix = params[1] != 0.0;
        if (!(P.gamma > 1.0)) return set_err(CLAWGPU_ERR_ARG, "gamma must exceed 1");
    }
    clawgpu_ctx* c = new (std::nothrow) clawgpu_ctx();
    if (!c) return set_err(CLAWGPU_ERR_ARG, "out of host memory");
    c->meqn = meqn;
    c->mwaves = mwaves;
    c->maux = maux;
    c->mbc = mbc;
    c->mx = mx;
    c->my = my;
    c->solver = solver_id;
    c->P = P;
    c->L.mx = mx;
    c->L.my = my;
    c->L.mbc = mbc;
    c->L.pitch = ((mx + 2 * mbc) + 15) / 16 * 16;
    c->L.fs = ((long)c->L.pitch * (my + 2 * mbc) + 31) / 32 * 32;
    c->cur = 0;
    c->seglen = 64;
    size_t qbytes = (size_t)c->L.fs * meqn * sizeof(double);
    size_t abytes = (size_t)(mx + 2 * mbc) * (my + 2 * mbc) * meqn * sizeof(double);
    hipError_t e = hipSuccess;
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipMalloc((void**)&c->dq[0], qbytes)) != hipSuccess ||
        (e = hipMalloc((void**)&c->dq[1], qbytes)) != hipSuccess ||
        (e = hipMalloc((void**)&c->staging, abytes)) != hipSuccess ||
        (e = hipMalloc((void**)&c->dcfl, sizeof(unsigned long long))) != hipSuccess ||
        (e = hipHostMalloc((void**)&c->hcfl, sizeof(double), hipHostMallocDefault)) != hipSuccess) {
        snprintf(g_errbuf, sizeof g_errbuf, "allocation failed: %s", hipGetErrorString(e));
        clawgpu_destroy(c);
        return CLAWGPU_ERR_HIP;
    }
    (void)hipMemsetAsync(c->dq[0], 0, qbytes, c->stream);
    (void)hipMemsetAsync(c->dq[1], 0, qbytes, c->stream);
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) {
        snprintf(g_errbuf, sizeof g_errbuf, "init failed: %s", hipGetErrorString(e));
        clawgpu_destroy(c);
        return CLAWGPU_ERR_HIP;
    }
    *out = c;
    return CLAWGPU_OK;
}

extern "C" int clawgpu_destroy(clawgpu_ctx* c)
{
    if (!c) return CLAWGPU_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->dq[0]) (void)hipFree(c->dq[0]);
    if (c->dq[1]) (void)hipFree(c->dq[1]);
    if (c->staging) (void)hipFree(c->staging);
    if (c->dcfl) (void)hipFree(c->dcfl);
    if (c->hcfl) (void)hipHostFree(c->hcfl);
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
    int rc = enqueue_bc(c, c->dq[c->cur], mthbc, false);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLAWGPU_OK;
}

static int check_method(const clawgpu_ctx* c, const int* method, const int* mthlim)
{
    if (method[1] != 1 && method[1] != 2) return set_err(CLAWGPU_ERR_ARG, "method(2) must be 1 or 2");
    if (method[2] < 0 || method[2] > 2)
        return set_err(CLAWGPU_ERR_UNSUPPORTED, "method(3) must be 0, 1 or 2 for the unsplit step");
    if (method[5] != 0) return set_err(CLAWGPU_ERR_UNSUPPORTED, "capacity function is not supported by this solver");
    if (c->mwaves > kMaxWaves) return set_err(CLAWGPU_ERR_UNSUPPORTED, "too many waves");
    for (int mw = 0; mw < c->mwaves; mw++)
        if (mthlim[mw] < 0 || mthlim[mw] > 5) return set_err(CLAWGPU_ERR_ARG, "invalid limiter method");
    return CLAWGPU_OK;
}

// Enqueue one unsplit step from dq[cur] into dq[1-cur]; the Courant
// accumulator must have been cleared (by the bc pass or explicitly).
static int enqueue_step(clawgpu_ctx* c, double dx, double dy, double dt, const int* method,
                        const int* mthlim, int use_fwave)
{
    StepArgs a;
    a.qold = c->dq[c->cur];
    a.qnew = c->dq[1 - c->cur];
    a.L = c->L;
    a.dtdx = dt / dx;
    a.dtdy = dt / dy;
    a.order = method[1];
    a.trans = method[2];
    a.use_fwave = use_fwave ? 1 : 0;
    for (int k = 0; k < kMaxWaves; k++) a.lim[k] = k < c->mwaves ? mthlim[k] : 0;
    a.nstrips = (c->mx + kOut - 1) / kOut;
    a.seglen = c->seglen;
    a.nsegs = (c->my + a.seglen - 1) / a.seglen;
    a.cfl_bits = c->dcfl;
    a.P = c->P;
    long waves = (long)a.nstrips * a.nsegs;
    int nb = (int)((waves + kWavesPerBlock - 1) / kWavesPerBlock);
    if (c->solver == CLAWGPU_SOLVER_ACOUSTICS)
        hipLaunchKernelGGL(step2_unsplit_kernel<Acoustics2>, dim3(nb), dim3(256), 0, c->stream, a);
    else
        hipLaunchKernelGGL(step2_unsplit_kernel<Euler2>, dim3(nb), dim3(256), 0, c->stream, a);
    HIPCHK(c, hipGetLastError());
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
    rc = enqueue_step(c, dx, dy, dt, method, mthlim, use_fwave);
    if (rc) return rc;
    rc = read_cfl(c, cfl);
    if (rc) return rc;
    c->cur = 1 - c->cur;
    return CLAWGPU_OK;
}

// The claw2 time loop (src/2d/claw2.f:494-784) on device-resident q: per
// step bc2 -> fused step into the other buffer -> Courant number back to the
// host -> accept (swap buffers) or reject (keep the old buffer, which is the
// reference's copyq2 restore, and retake with a smaller dt).
extern "C" int clawgpu_claw2(clawgpu_ctx* c, double dx, double dy, double tstart, double* tend,
                             double* dtv, double* cflv, int* nv, const int* method,
                             const int* mthlim, const int* mthbc, int use_fwave, int* info_out)
{
    if (!c || !tend || !dtv || !cflv || !nv || !method || !mthlim || !mthbc || !info_out)
        return set_err(CLAWGPU_ERR_ARG, "null argument");
    int rc = check_method(c, method, mthlim);
    if (rc) return rc;
    if (check_bc(mthbc) != 0)
        return set_err(CLAWGPU_ERR_UNSUPPORTED, "mthbc must be 1, 2 or 3 (periodic on both sides)");
    int info = 0;
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
    for (int n = 1; n <= maxn; n++) {
        double told = t;
        if (told + dt > *tend && tstart < *tend) dt = *tend - told;
        for (;;) {
            rc = enqueue_bc(c, c->dq[c->cur], mthbc, true);
            if (rc) return rc;
            rc = enqueue_step(c, dx, dy, dt, method, mthlim, use_fwave);
            if (rc) return rc;
            rc = read_cfl(c, &cfl);
            if (rc) return rc;
            if (!std::isfinite(cfl)) {
                *info_out = -5;
                return set_err(CLAWGPU_ERR_NONFINITE, "non-finite wave speed (Courant number)");
            }
            t = told + dt;
            if (cfl <= cflv[0]) {
                cflmax = fmax(cfl, cflmax);
                c->cur = 1 - c->cur;  // accept: the new buffer becomes q
                break;
            }
            t = told;  // reject: dq[cur] still holds the step's initial data
            if (method[0] == 1) {
                dt = fmin(dtv[1], dt * cflv[1] / cfl);
                continue;
            }
            cflmax = fmax(cfl, cflmax);
            goto done;
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

extern "C" int clawgpu_set_segment_rows(clawgpu_ctx* c, int rows)
{
    if (!c || rows < 1) return set_err(CLAWGPU_ERR_ARG, "rows must be >= 1");
    c->seglen = rows;
    return CLAWGPU_OK;
}

extern "C" int clawgpu_time_steps(clawgpu_ctx* c, double dx, double dy, double dt, const int* method,
                                  const int* mthlim, const int* mthbc, int use_fwave, int nsteps,
                                  float* total_ms, float* step_kernel_ms, double* cfl_max_out)
{
    // Fixed-dt stepping timed with HIP events on the library's stream, with
    // no per-step host synchronisation: total_ms spans all launches (ghost
    // fill + step kernel per step), step_kernel_ms sums event pairs around
    // each step-kernel launch only.  Used by the benchmark.
    if (!c || !method || !mthlim || !mthbc || !total_ms || nsteps < 1)
        return set_err(CLAWGPU_ERR_ARG, "bad argument");
    int rc = check_method(c, method, mthlim);
    if (rc) return rc;
    if (check_bc(mthbc) != 0) return set_err(CLAWGPU_ERR_UNSUPPORTED, "bad mthbc");
    hipEvent_t* ev = new (std::nothrow) hipEvent_t[2 * nsteps + 2];
    if (!ev) return set_err(CLAWGPU_ERR_ARG, "out of host memory");
    for (int k = 0; k < 2 * nsteps + 2; k++) HIPCHK(c, hipEventCreate(&ev[k]));
    HIPCHK(c, hipEventRecord(ev[0], c->stream));
    for (int n = 0; n < nsteps; n++) {
        rc = enqueue_bc(c, c->dq[c->cur], mthbc, n == 0);
        if (rc) return rc;
        HIPCHK(c, hipEventRecord(ev[2 + 2 * n], c->stream));
        rc = enqueue_step(c, dx, dy, dt, method, mthlim, use_fwave);
        if (rc) return rc;
        HIPCHK(c, hipEventRecord(ev[3 + 2 * n], c->stream));
        c->cur = 1 - c->cur;
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
