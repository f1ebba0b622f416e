// 1D wave-propagation step for gfx950 (FP64) and its C-ABI.
//
// Replaces step1 (src/1d/step1.f90) and the claw1 time loop
// (src/1d/claw1.f) for the rp1 solvers of the reference's 1D examples:
// constant-coefficient acoustics, variable-coefficient acoustics
// (aux = (Z, c)) and scalar advection.  One 64-lane wavefront holds 64
// consecutive cells (lanes 2..61 are output cells); lane l solves the
// Riemann problem at the left face of its cell and gets its neighbours' data
// through DPP shifts.  The step is out of place, with the reference's
// expression order per cell,
//     q(i) = ((q(i) - dtdx(i)*apdq(i)) - dtdx(i)*amdq(i+1))
//                   - dtdx(i)*(f(i+1) - f(i)),
// which equals step1's in-place loops because every update reads only
// fluctuations and fluxes computed from the old state; compiled without
// contraction it agrees with the reference bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "ctx.hpp"
#include "lanes.hpp"
#include "riemann2d.hpp"

namespace {

using clawgpu::lane_dn;
using clawgpu::lane_up;

struct P1 {
    double zz, cc, u;
};

struct Args1 {
    const double* __restrict__ q;    // (mx+2mbc) x meqn, m fastest
    double* __restrict__ qn;
    const double* __restrict__ aux;  // (mx+2mbc) x maux
    int mx, mbc, maux, capa;         // capa: 1-based aux index of the capacity, or 0
    double dx, dt;
    int order, fwave, lim0, lim1;
    P1 P;
    unsigned long long* cfl_bits;
    double* cfl_part;  // per-strip Courant numbers, reduced into cfl_bits afterwards
};

// rp1 of one interface: q and aux of the left (L) and right (R) cell
template <int SOLVER>
struct Rp1;

template <>
struct Rp1<CLAWGPU_SOLVER_ACOUSTICS> {
    static constexpr int M = 2, W = 2;
    __device__ static void solve(const double* qL, const double* qR, const double*, const double*,
                                 const P1& P, double (&w)[W][M], double (&s)[W], double (&am)[M],
                                 double (&ap)[M])
    {
        const double delta1 = qR[0] - qL[0], delta2 = qR[1] - qL[1];
        const double a1 = (-delta1 + P.zz * delta2) / (2.0 * P.zz);
        const double a2 = (delta1 + P.zz * delta2) / (2.0 * P.zz);
        w[0][0] = -a1 * P.zz;
        w[0][1] = a1;
        s[0] = -P.cc;
        w[1][0] = a2 * P.zz;
        w[1][1] = a2;
        s[1] = P.cc;
        for (int m = 0; m < M; m++) {
            am[m] = s[0] * w[0][m];
            ap[m] = s[1] * w[1][m];
        }
    }
};

template <>
struct Rp1<CLAWGPU_SOLVER_VC_ACOUSTICS> {
    static constexpr int M = 2, W = 2;
    __device__ static void solve(const double* qL, const double* qR, const double* aL,
                                 const double* aR, const P1&, double (&w)[W][M], double (&s)[W],
                                 double (&am)[M], double (&ap)[M])
    {
        const double delta1 = qR[0] - qL[0], delta2 = qR[1] - qL[1];
        const double zi = aR[0], zim = aL[0];
        const double a1 = (-delta1 + zi * delta2) / (zim + zi);
        const double a2 = (delta1 + zim * delta2) / (zim + zi);
        w[0][0] = -a1 * zim;
        w[0][1] = a1;
        s[0] = -aL[1];
        w[1][0] = a2 * zi;
        w[1][1] = a2;
        s[1] = aR[1];
        for (int m = 0; m < M; m++) {
            am[m] = s[0] * w[0][m];
            ap[m] = s[1] * w[1][m];
        }
    }
};

template <>
struct Rp1<CLAWGPU_SOLVER_ADVECTION> {
    static constexpr int M = 1, W = 1;
    __device__ static void solve(const double* qL, const double* qR, const double*, const double*,
                                 const P1& P, double (&w)[W][M], double (&s)[W], double (&am)[M],
                                 double (&ap)[M])
    {
        w[0][0] = qR[0] - qL[0];
        s[0] = P.u;
        am[0] = fmin(P.u, 0.0) * w[0][0];
        ap[0] = fmax(P.u, 0.0) * w[0][0];
    }
};

__global__ void cfl1_reduce_kernel(const double* part, long n, unsigned long long* out)
{
    clawgpu::wave_max_reduce(part, n, out);
}

template <int SOLVER>
__global__ __launch_bounds__(256) void step1_kernel(Args1 a)
{
    using R = Rp1<SOLVER>;
    constexpr int M = R::M, W = R::W;
    using Mth = clawgpu::ExactMath;
    const int l = threadIdx.x & 63;
    const int strip = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (strip * 60 > a.mx + 2) return;  // uniform per wave
    const int i = 60 * strip + l - 1;
    const int ic = i < 1 - a.mbc ? 1 - a.mbc : (i > a.mx + a.mbc ? a.mx + a.mbc : i);
    const long c = ic + a.mbc - 1;
    double q[M], qL[M], ax[2] = {0.0, 0.0}, aL[2];
    for (int m = 0; m < M; m++) q[m] = a.q[c * M + m];
    if (a.maux > 0) {
        ax[0] = a.aux[c * a.maux];
        ax[1] = a.maux > 1 ? a.aux[c * a.maux + 1] : 0.0;
    }
    const double dtdx = a.capa > 0 ? a.dt / (a.dx * a.aux[c * a.maux + a.capa - 1]) : a.dt / a.dx;
    for (int m = 0; m < M; m++) qL[m] = lane_up(q[m]);
    aL[0] = lane_up(ax[0]);
    aL[1] = lane_up(ax[1]);
    const double dtdxL = lane_up(dtdx);
    double w[W][M], s[W], am[M], ap[M];
    R::solve(qL, q, aL, ax, a.P, w, s, am, ap);  // interface i (between i-1 and i)

    // Courant number over interfaces 1..mx+1 (step1.f90 cfl loop)
    double cfl = 0.0;
    if (l >= 1 && i >= 1 && i <= a.mx + 1) {
        for (int mw = 0; mw < W; mw++) {
            double t = fmax(dtdx * s[mw], -dtdxL * s[mw]);
            cfl = fmax(cfl, (t == t) ? t : __builtin_huge_val());
        }
    }
    clawgpu::store_wave_max(cfl, a.cfl_part, strip);

    // second-order flux f(i) at this interface from the limited waves
    double f[M];
    for (int m = 0; m < M; m++) f[m] = 0.0;
    if (a.order == 2) {
        const double dtdxave = 0.5 * (dtdxL + dtdx);
        for (int mw = 0; mw < W; mw++) {
            double wl[M], wr[M], wv[M];
            for (int m = 0; m < M; m++) {
                wv[m] = w[mw][m];
                wl[m] = lane_up(wv[m]);
                wr[m] = lane_dn(wv[m]);
            }
            const int lim = mw == 0 ? a.lim0 : a.lim1;
            if (lim > 0) {  // inlinelimiter.f90:51-100
                double wn2 = 0.0, dotl = 0.0, dotr = 0.0;
                for (int m = 0; m < M; m++) {
                    wn2 = wn2 + wv[m] * wv[m];
                    dotr = dotr + wv[m] * wr[m];
                    dotl = dotl + wl[m] * wv[m];
                }
                if (wn2 != 0.0) {
                    const double r = (s[mw] > 0.0 ? dotl : dotr) / wn2;
                    const double phi = clawgpu::philim<Mth>(r, lim);
                    for (int m = 0; m < M; m++) wv[m] = phi * wv[m];
                }
            }
            const double sa = fabs(s[mw]);
            const double lead = a.fwave ? 0.5 * copysign(1.0, s[mw]) : 0.5 * sa;
            for (int m = 0; m < M; m++) f[m] = f[m] + lead * (1.0 - sa * dtdxave) * wv[m];
        }
    }
    double amhi[M], fhi[M];
    for (int m = 0; m < M; m++) {
        amhi[m] = lane_dn(am[m]);
        fhi[m] = lane_dn(f[m]);
    }
    const bool own = (l >= 2 && l < 62) || (strip == 0 && l < 2);
    if (!own || i > a.mx + a.mbc || i < 1 - a.mbc) return;
    const long o = (long)(i + a.mbc - 1);
    for (int m = 0; m < M; m++) {
        double v = q[m];
        if (i >= 1 && i <= a.mx) {
            v = v - dtdx * ap[m];
            v = v - dtdx * amhi[m];
            if (a.order == 2) v = v - dtdx * (fhi[m] - f[m]);
        }
        a.qn[o * M + m] = v;
    }
}

__global__ void bc1_kernel(double* q, int meqn, int mbc, int mx, int lo, int hi)
{
    // bc1.f: left then right; codes 1 extrap, 2 periodic, 3 wall (q(2) negated)
    const int t = threadIdx.x;
    const int ibc = t + 1;
#define QA(m, i) q[(long)((i) + mbc - 1) * meqn + (m)]
    if (t < mbc) {
        for (int m = 0; m < meqn; m++) {
            const int src = lo == 1 ? 1 : lo == 2 ? mx + 1 - ibc : ibc;
            QA(m, 1 - ibc) = QA(m, src);
        }
        if (lo == 3 && meqn > 1) QA(1, 1 - ibc) = -QA(1, ibc);
    }
    __syncthreads();
    if (t < mbc) {
        for (int m = 0; m < meqn; m++) {
            const int src = hi == 1 ? mx : hi == 2 ? ibc : mx + 1 - ibc;
            QA(m, mx + ibc) = QA(m, src);
        }
        if (hi == 3 && meqn > 1) QA(1, mx + ibc) = -QA(1, mx + 1 - ibc);
    }
#undef QA
}

}  // namespace

struct clawgpu1_ctx {
    int meqn, mwaves, maux, mbc, mx, solver;
    P1 P;
    double* dq[2];
    int cur;
    double* daux;
    unsigned long long* dcfl;
    double* dpart;  // per-strip Courant numbers of a step
    hipStream_t stream;
    int last_hip_error;
    std::vector<double> haux;  // host copy of aux for the user routines
};

namespace {

int solver1_shape(int id, int* m, int* w)
{
    switch (id) {
    case CLAWGPU_SOLVER_ACOUSTICS: *m = 2; *w = 2; return 0;
    case CLAWGPU_SOLVER_VC_ACOUSTICS: *m = 2; *w = 2; return 0;
    case CLAWGPU_SOLVER_ADVECTION: *m = 1; *w = 1; return 0;
    default: return -1;
    }
}

int check1(clawgpu1_ctx* c, const int* method, const int* mthlim)
{
    if (method[1] != 1 && method[1] != 2) return set_err(CLAWGPU_ERR_ARG, "method(2) must be 1 or 2");
    if (method[5] < 0 || method[5] > c->maux) return set_err(CLAWGPU_ERR_ARG, "method(6) must be 0..maux");
    for (int w = 0; w < c->mwaves; w++)
        if (mthlim[w] < 0 || mthlim[w] > 5) return set_err(CLAWGPU_ERR_ARG, "mthlim must be 0..5");
    return CLAWGPU_OK;
}

int enqueue_step1(clawgpu1_ctx* c, double dx, double dt, const int* method, const int* mthlim,
                  int use_fwave)
{
    Args1 a;
    a.q = c->dq[c->cur];
    a.qn = c->dq[1 - c->cur];
    a.aux = c->daux;
    a.mx = c->mx;
    a.mbc = c->mbc;
    a.maux = c->maux;
    a.capa = method[5];
    a.dx = dx;
    a.dt = dt;
    a.order = method[1];
    a.fwave = use_fwave ? 1 : 0;
    a.lim0 = mthlim[0];
    a.lim1 = c->mwaves > 1 ? mthlim[1] : 0;
    a.P = c->P;
    a.cfl_bits = c->dcfl;
    a.cfl_part = c->dpart;
    HIPCHK(c, hipMemsetAsync(c->dcfl, 0, sizeof(unsigned long long), c->stream));
    const int nstrips = (c->mx + 2) / 60 + 1;  // strips 0..nstrips-1 store their Courant numbers
    const unsigned nb = (unsigned)((nstrips + 3) / 4);
    switch (c->solver) {
    case CLAWGPU_SOLVER_ACOUSTICS:
        hipLaunchKernelGGL(step1_kernel<CLAWGPU_SOLVER_ACOUSTICS>, dim3(nb), dim3(256), 0, c->stream, a);
        break;
    case CLAWGPU_SOLVER_VC_ACOUSTICS:
        hipLaunchKernelGGL(step1_kernel<CLAWGPU_SOLVER_VC_ACOUSTICS>, dim3(nb), dim3(256), 0, c->stream, a);
        break;
    default:
        hipLaunchKernelGGL(step1_kernel<CLAWGPU_SOLVER_ADVECTION>, dim3(nb), dim3(256), 0, c->stream, a);
        break;
    }
    const long rb = (nstrips + 255) / 256;
    hipLaunchKernelGGL(cfl1_reduce_kernel, dim3((unsigned)(rb < 512 ? rb : 512)), dim3(256), 0,
                       c->stream, c->dpart, (long)nstrips, c->dcfl);
    HIPCHK(c, hipGetLastError());
    return CLAWGPU_OK;
}

int read_cfl1(clawgpu1_ctx* c, double* cfl)
{
    unsigned long long bits = 0;
    HIPCHK(c, hipMemcpyAsync(&bits, c->dcfl, sizeof bits, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::memcpy(cfl, &bits, sizeof *cfl);
    return CLAWGPU_OK;
}

int enqueue_bc1(clawgpu1_ctx* c, const int* mthbc)
{
    for (int s = 0; s < 2; s++)
        if (mthbc[s] < 1 || mthbc[s] > 3)
            return set_err(CLAWGPU_ERR_UNSUPPORTED, "mthbc must be 1, 2 or 3");
    if ((mthbc[0] == 2) != (mthbc[1] == 2))
        return set_err(CLAWGPU_ERR_UNSUPPORTED, "periodic boundaries come in pairs");
    hipLaunchKernelGGL(bc1_kernel, dim3(1), dim3(64), 0, c->stream, c->dq[c->cur], c->meqn, c->mbc,
                       c->mx, mthbc[0], mthbc[1]);
    HIPCHK(c, hipGetLastError());
    return CLAWGPU_OK;
}

}  // namespace

extern "C" int clawgpu1_create(clawgpu1_ctx** out, int meqn, int mwaves, int maux, int mbc, int mx,
                               int solver_id, const double* params, int nparams)
{
    if (!out) return set_err(CLAWGPU_ERR_ARG, "null output pointer");
    *out = nullptr;
    int sm, sw;
    if (solver1_shape(solver_id, &sm, &sw) != 0)
        return set_err(CLAWGPU_ERR_UNSUPPORTED, "1D solvers: acoustics, vc acoustics, advection");
    if (meqn != sm || mwaves != sw) return set_err(CLAWGPU_ERR_ARG, "meqn/mwaves do not match the solver");
    if (solver_id == CLAWGPU_SOLVER_VC_ACOUSTICS && maux < 2)
        return set_err(CLAWGPU_ERR_ARG, "vc acoustics needs maux >= 2 (Z, c)");
    if (mbc != 2) return set_err(CLAWGPU_ERR_UNSUPPORTED, "the 1D path needs mbc = 2");
    if (mx < 1 || maux < 0) return set_err(CLAWGPU_ERR_ARG, "bad sizes");
    P1 P = {0.0, 0.0, 0.0};
    if (solver_id == CLAWGPU_SOLVER_ACOUSTICS) {
        if (!params || nparams < 2 || !(params[0] > 0) || !(params[1] > 0))
            return set_err(CLAWGPU_ERR_ARG, "acoustics params: rho > 0, bulk > 0");
        P.cc = std::sqrt(params[1] / params[0]);
        P.zz = P.cc * params[0];
    } else if (solver_id == CLAWGPU_SOLVER_ADVECTION) {
        if (!params || nparams < 1) return set_err(CLAWGPU_ERR_ARG, "advection params: u");
        P.u = params[0];
    }
    clawgpu1_ctx* c = new (std::nothrow) clawgpu1_ctx();
    if (!c) return set_err(CLAWGPU_ERR_ARG, "out of host memory");
    c->meqn = meqn; c->mwaves = mwaves; c->maux = maux; c->mbc = mbc; c->mx = mx;
    c->solver = solver_id;
    c->P = P;
    const size_t n = (size_t)(mx + 2 * mbc);
    hipError_t e = hipMalloc((void**)&c->dq[0], n * meqn * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&c->dq[1], n * meqn * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&c->daux, n * (maux > 0 ? maux : 1) * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&c->dcfl, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc((void**)&c->dpart, (size_t)((mx + 2) / 60 + 1) * sizeof(double));
    if (e == hipSuccess) e = hipStreamCreate(&c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->dq[1], 0, n * meqn * sizeof(double), c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        snprintf(g_errbuf, sizeof g_errbuf, "device allocation failed: %s", hipGetErrorString(e));
        (void)hipFree(c->dq[0]); (void)hipFree(c->dq[1]); (void)hipFree(c->daux); (void)hipFree(c->dcfl);
        (void)hipFree(c->dpart);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return CLAWGPU_ERR_HIP;
    }
    *out = c;
    return CLAWGPU_OK;
}

extern "C" int clawgpu1_destroy(clawgpu1_ctx* c)
{
    if (!c) return CLAWGPU_OK;
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(c->dq[0]); (void)hipFree(c->dq[1]); (void)hipFree(c->daux); (void)hipFree(c->dcfl);
    (void)hipFree(c->dpart);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return CLAWGPU_OK;
}

extern "C" int clawgpu1_set_q(clawgpu1_ctx* c, const double* q)
{
    if (!c || !q) return set_err(CLAWGPU_ERR_ARG, "null argument");
    const size_t n = (size_t)(c->mx + 2 * c->mbc) * c->meqn;
    HIPCHK(c, hipMemcpyAsync(c->dq[c->cur], q, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLAWGPU_OK;
}

extern "C" int clawgpu1_get_q(clawgpu1_ctx* c, double* q)
{
    if (!c || !q) return set_err(CLAWGPU_ERR_ARG, "null argument");
    const size_t n = (size_t)(c->mx + 2 * c->mbc) * c->meqn;
    HIPCHK(c, hipMemcpyAsync(q, c->dq[c->cur], n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLAWGPU_OK;
}

extern "C" int clawgpu1_set_aux(clawgpu1_ctx* c, const double* aux)
{
    if (!c || !aux) return set_err(CLAWGPU_ERR_ARG, "null argument");
    if (c->maux == 0) return set_err(CLAWGPU_ERR_ARG, "context has maux = 0");
    const size_t n = (size_t)(c->mx + 2 * c->mbc) * c->maux;
    if (aux != c->haux.data()) c->haux.assign(aux, aux + n);
    HIPCHK(c, hipMemcpyAsync(c->daux, aux, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLAWGPU_OK;
}

extern "C" int clawgpu1_bc1(clawgpu1_ctx* c, const int* mthbc)
{
    if (!c || !mthbc) return set_err(CLAWGPU_ERR_ARG, "null argument");
    int rc = enqueue_bc1(c, mthbc);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLAWGPU_OK;
}

extern "C" int clawgpu1_step1(clawgpu1_ctx* c, double dx, double dt, const int* method,
                              const int* mthlim, int use_fwave, double* cfl)
{
    if (!c || !method || !mthlim || !cfl) return set_err(CLAWGPU_ERR_ARG, "null argument");
    int rc = check1(c, method, mthlim);
    if (rc) return rc;
    rc = enqueue_step1(c, dx, dt, method, mthlim, use_fwave);
    if (rc) return rc;
    rc = read_cfl1(c, cfl);
    if (rc) return rc;
    c->cur = 1 - c->cur;
    if (!std::isfinite(*cfl)) return set_err(CLAWGPU_ERR_NONFINITE, "non-finite wave speed");
    return CLAWGPU_OK;
}

// The claw1 time loop (src/1d/claw1.f:448-560).  With hooks, the user's
// bc1, b4step1 and src1 are called on a host copy of q where claw1 calls
// them (claw1.f:478-510): bc1 and b4step1 before every attempt, src1 over
// dt/2 before the step and over dt/2 after it (method(5) = 2) or over dt
// after it (method(5) = 1).  The post-step source comes before the Courant
// check, so a rejected attempt calls it too; a rejection then restores q as
// it was at the start of the step (claw1.f's copyq1 of qwork).
static int claw1_loop(clawgpu1_ctx* c, double dx, double xlower, double tstart, double* tend,
                      double* dtv, double* cflv, int* nv, const int* method, const int* mthlim,
                      const int* mthbc, int use_fwave, const clawgpu1_hooks* h, int* info_out)
{
    if (!c || !tend || !dtv || !cflv || !nv || !method || !mthlim || !mthbc || !info_out)
        return set_err(CLAWGPU_ERR_ARG, "null argument");
    int rc = check1(c, method, mthlim);
    if (rc) return rc;
    const bool user_bc = h && h->bc1;
    const bool hooked = h && (h->bc1 || h->b4step1 || h->src1);
    const bool src = h && h->src1 && (method[4] == 1 || method[4] == 2);
    std::vector<double> hq, qwork;  // host copies of q for the hooks
    double* haux = c->maux > 0 && !c->haux.empty() ? c->haux.data() : nullptr;
    const size_t nq = (size_t)(c->mx + 2 * c->mbc) * c->meqn;
    auto pull = [&]() -> int {
        hq.resize(nq);
        return clawgpu1_get_q(c, hq.data());
    };
    auto push = [&]() -> int { return clawgpu1_set_q(c, hq.data()); };
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
            if (fabs(maxn * dt - (*tend - tstart)) > 1e-5 * (*tend - tstart)) { info = 2; goto done; }
        }
    }
    if (method[0] == 1 && cflv[1] > cflv[0]) { info = 3; goto done; }
    if (maxn == 0) goto done;
    for (int n = 1; n <= maxn; n++) {
        const double told = t;
        if (told + dt > *tend && tstart < *tend) dt = *tend - told;
        if (hooked) {  // claw1.f's qwork, restored on a rejection
            if ((rc = pull())) return rc;
            qwork = hq;
        }
        for (;;) {  // claw1.f label 40
            const double dt2 = dt / 2.0;
            const double thalf = told + dt2;
            t = told + dt;
            if (user_bc) {
                if ((rc = pull())) return rc;
                h->bc1(h->user, c->meqn, c->mbc, c->mx, xlower, dx, hq.data(), c->maux, haux, told,
                       dt, mthbc);
                if ((rc = push())) return rc;
            } else {
                rc = enqueue_bc1(c, mthbc);
                if (rc) return rc;
            }
            if (h && h->b4step1) {
                if ((rc = pull())) return rc;
                h->b4step1(h->user, c->mbc, c->mx, c->meqn, hq.data(), xlower, dx, told, dt,
                           c->maux, haux);
                if ((rc = push())) return rc;
                if (haux && (rc = clawgpu1_set_aux(c, c->haux.data()))) return rc;
            }
            if (src && method[4] == 2) {
                if ((rc = pull())) return rc;
                h->src1(h->user, c->meqn, c->mbc, c->mx, xlower, dx, hq.data(), c->maux, haux, told,
                        dt2);
                if ((rc = push())) return rc;
            }
            rc = enqueue_step1(c, dx, dt, method, mthlim, use_fwave);
            if (rc) return rc;
            rc = read_cfl1(c, &cfl);
            if (rc) return rc;
            if (!std::isfinite(cfl)) {
                *info_out = -5;
                return set_err(CLAWGPU_ERR_NONFINITE, "non-finite wave speed (Courant number)");
            }
            c->cur = 1 - c->cur;  // the new buffer holds the step's result
            if (src) {
                if ((rc = pull())) return rc;
                h->src1(h->user, c->meqn, c->mbc, c->mx, xlower, dx, hq.data(), c->maux, haux,
                        method[4] == 2 ? thalf : t, method[4] == 2 ? dt2 : dt);
                if ((rc = push())) return rc;
            }
            // the new dt is chosen before the Courant check (claw1.f)
            if (method[0] == 1) {
                if (cfl > 0.0) {
                    dt = fmin(dtv[1], dt * cflv[1] / cfl);
                    dtmin = fmin(dt, dtmin);
                    dtmax = fmax(dt, dtmax);
                } else {
                    dt = dtv[1];
                }
            }
            if (cfl <= cflv[0]) {
                cflmax = fmax(cfl, cflmax);
                break;
            }
            // reject: back to the old buffer, which still holds q unless a
            // user routine changed it
            t = told;
            c->cur = 1 - c->cur;
            if (hooked) {
                hq = qwork;
                if ((rc = push())) return rc;
            }
            if (method[0] == 1) continue;
            cflmax = fmax(cfl, cflmax);
            goto done;
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

extern "C" int clawgpu1_claw1(clawgpu1_ctx* c, double dx, double tstart, double* tend,
                              double* dtv, double* cflv, int* nv, const int* method,
                              const int* mthlim, const int* mthbc, int use_fwave, int* info_out)
{
    return claw1_loop(c, dx, 0.0, tstart, tend, dtv, cflv, nv, method, mthlim, mthbc, use_fwave,
                      nullptr, info_out);
}

extern "C" int clawgpu1_claw1_hooks(clawgpu1_ctx* c, double dx, double xlower, double tstart,
                                    double* tend, double* dtv, double* cflv, int* nv,
                                    const int* method, const int* mthlim, const int* mthbc,
                                    int use_fwave, const clawgpu1_hooks* hooks, int* info_out)
{
    if (!hooks) return set_err(CLAWGPU_ERR_ARG, "null hooks (use clawgpu1_claw1)");
    if (c && c->maux > 0 && c->haux.empty() && (hooks->b4step1 || hooks->src1 || hooks->bc1))
        return set_err(CLAWGPU_ERR_ARG, "aux has not been set (clawgpu1_set_aux)");
    return claw1_loop(c, dx, xlower, tstart, tend, dtv, cflv, nv, method, mthlim, mthbc,
                      use_fwave, hooks, info_out);
}
