// General 2D unsplit step for gfx950 (FP64): solvers with aux arrays and the
// capacity function (method(6) > 0).
//
// Replaces step2/flux2 (src/2d/step2.f90:2-251, src/2d/flux2.f90:5-235) for
// the variable-coefficient advection solver of the reference's annulus
// example (edge velocities in aux, capacity = cell-area ratio).  Three
// kernels:
//   * x records: one thread per slice position (i, j), j = 0..my+1, solves
//     the four interfaces i-1..i+2 of its x-slice, limits i and i+1 and forms
//     the position's qadd, fadd(i), fadd(i+1), gadd(1,i), gadd(2,i) exactly
//     as flux2 does (dtdx1d = dtdx / capa per cell, cqxx with the average of
//     the neighbouring dtdx1d);
//   * y records: the same for the y-slices i = 0..mx+1;
//   * combine: one thread per cell applies, in step2's order, the updates of
//     x-slices j-1, j, j+1 and then of y-slices i-1, i, i+1, with step2's
//     expressions for capa = 0 and capa > 0.
// Recomputing the neighbouring interfaces instead of exchanging them keeps
// each thread independent; compiled without contraction the result equals
// the reference bit for bit.
#include <hip/hip_runtime.h>

#include "ctx.hpp"
#include "lanes.hpp"
#include "riemann2d.hpp"
#include "step2g.hpp"

namespace clawgpu {
namespace g2 {

__device__ inline double lim_phi(double r, int lim) { return philim<ExactMath>(r, lim); }

// One slice position of flux2 for the variable-coefficient advection
// solver: cells p-2..p+2 along the slice, the four interfaces p-1..p+2
// between them (t = 0..3: between local cells t and t+1), the Courant
// contribution of interface p (and of n+1 for the last position) and the
// second-order corrections of interfaces p (cq[1]) and p+1 (cq[2]).
struct Pos {
    double q[5], dtd[5], w[4], s[4], am[4], ap[4], cq[3];
    double cfl;
};

// off(pp) gives the offset of cell pp of the slice; dtb = dt/dx or dt/dy
// (divided by the capacity per cell when capa > 0, flux2.f90:38-45).
template <class Off>
__device__ inline void eval_position(const double* q, const double* aux, long afs, const double* vel,
                                     Off off, int p, int n, bool live, double dtb, int capa,
                                     int order, int lim0, int fwave, Pos& o)
{
    double un[5];
#pragma unroll
    for (int k = 0; k < 5; k++) {
        const long c = off(p - 2 + k);
        o.q[k] = q[c];
        un[k] = vel[c];
        o.dtd[k] = capa > 0 ? dtb / aux[(long)(capa - 1) * afs + c] : dtb;
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
        o.w[t] = o.q[t + 1] - o.q[t];
        o.s[t] = un[t + 1];
        o.am[t] = fmin(o.s[t], 0.0) * o.w[t];
        o.ap[t] = fmax(o.s[t], 0.0) * o.w[t];
    }
    // Courant number: interface p (t = 1) and, for the last position, n+1
    // (flux2.f90:123-132)
    o.cfl = 0.0;
    if (live) {
        const double* dtd = o.dtd;
        const double* sp = o.s;
        o.cfl = fmax(dtd[2] * sp[1], -dtd[1] * sp[1]);
        if (p == n) o.cfl = fmax(o.cfl, fmax(dtd[3] * sp[2], -dtd[2] * sp[2]));
        if (o.cfl != o.cfl || sp[1] != sp[1] || (p == n && sp[2] != sp[2])) o.cfl = __builtin_huge_val();
    }
    // second-order corrections of interfaces p (t = 1) and p+1 (t = 2)
    // (flux2.f90:134-183)
    o.cq[0] = o.cq[1] = o.cq[2] = 0.0;
    if (order == 2) {
#pragma unroll
        for (int t = 1; t <= 2; t++) {
            double wl = o.w[t];
            if (lim0 != 0) {
                const double wn2 = o.w[t] * o.w[t];
                if (wn2 != 0.0) {
                    const double dotl = o.w[t - 1] * o.w[t], dotr = o.w[t] * o.w[t + 1];
                    const double r = (o.s[t] > 0.0 ? dotl : dotr) / wn2;
                    wl = lim_phi(r, lim0) * o.w[t];
                }
            }
            const double dtdxave = 0.5 * (o.dtd[t] + o.dtd[t + 1]);
            const double sa = fabs(o.s[t]);
            const double sw = fwave ? copysign(1.0, o.s[t]) : sa;
            o.cq[t] = 0.0 + sw * (1.0 - sa * dtdxave) * wl;
        }
    }
}

// The x (DIR = 1) or y (DIR = 2) records of slice position (p, s): p runs
// along the slice, s is the slice index.
template <int DIR>
__global__ __launch_bounds__(256) void records_kernel(GenArgs a)
{
    const Layout2& L = a.L;
    const int n = DIR == 1 ? L.mx : L.my;        // positions 1..n
    const int ns = DIR == 1 ? L.my : L.mx;       // slices 0..ns+1
    const int p = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int sl = blockIdx.y * blockDim.y + threadIdx.y;
    const bool live = p <= n && sl <= ns + 1;
    const int pc = live ? p : 1, sc = live ? sl : 0;
    const double dtb = DIR == 1 ? a.dtdx : a.dtdy;
    auto off = [&](int pp, int ss) -> long { return DIR == 1 ? L.off(pp, ss) : L.off(ss, pp); };
    const double* vel = a.aux + (DIR == 1 ? 0 : a.afs);   // normal edge velocity (aux 1 or 2)
    const double* tvel = a.aux + (DIR == 1 ? a.afs : 0);  // transverse edge velocity
    Pos ps;
    eval_position(a.q, a.aux, a.afs, vel, [&](int pp) { return off(pp, sc); }, pc, n, live, dtb,
                  a.capa, a.order, a.lim0, a.fwave, ps);
    store_wave_max(ps.cfl, a.cfl_part, ((long)blockIdx.y * gridDim.x + blockIdx.x) * 4 + threadIdx.y);
    const double* dtd = ps.dtd;
    const double* am = ps.am;
    const double* ap = ps.ap;
    const double* cq = ps.cq;
    const double fl = 0.5 * cq[1], fh = 0.5 * cq[2];
    const double qadd = (0.0 - dtd[2] * ap[1]) - dtd[2] * am[2];
    double g1 = 0.0, g2 = 0.0;
    if (a.trans > 0) {
        const double amt = (a.order > 1 && a.trans == 2) ? am[2] + cq[2] : am[2];
        const double apt = (a.order > 1 && a.trans == 2) ? ap[1] - cq[1] : ap[1];
        // both splits move into this position's cell: down with its own
        // transverse edge velocity, up with the one of the next cell across
        const double vdn = tvel[off(pc, sc)];
        const double vup = tvel[DIR == 1 ? L.off(pc, sc + 1) : L.off(sc + 1, pc)];
        const double bm1 = fmin(vdn, 0.0) * amt, bp1 = fmax(vup, 0.0) * amt;
        const double bm2 = fmin(vdn, 0.0) * apt, bp2 = fmax(vup, 0.0) * apt;
        g1 = -0.5 * dtd[2] * bm1;
        g2 = -0.5 * dtd[2] * bp1;
        g1 = g1 - 0.5 * dtd[2] * bm2;
        g2 = g2 - 0.5 * dtd[2] * bp2;
    }
    if (!live) return;
    double* R = a.rec + (DIR == 1 ? 0 : 5 * a.L.fs);
    const long o = off(p, sl);
    R[o] = qadd;
    R[a.L.fs + o] = fl;
    R[2 * a.L.fs + o] = fh;
    R[3 * a.L.fs + o] = g1;
    R[4 * a.L.fs + o] = g2;
}

// step2's updates of cell (i, j), in the order its x-slices j-1, j, j+1 and
// then its y-slices i-1, i, i+1 apply them (step2.f90:131-163, :206-247)
__global__ __launch_bounds__(256) void combine_kernel(GenArgs a)
{
    const Layout2& L = a.L;
    const int nx = L.mx + 2 * L.mbc, ny = L.my + 2 * L.mbc;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)nx * ny) return;
    const int i = (int)(e % nx) + 1 - L.mbc, j = (int)(e / nx) + 1 - L.mbc;
    const long c = L.off(i, j);
    double v = a.q[c];
    if (i >= 1 && i <= L.mx && j >= 1 && j <= L.my) {
        const long fs = L.fs;
        const double* X = a.rec;
        const double* Y = a.rec + 5 * fs;
        const double dtdx = a.dtdx, dtdy = a.dtdy;
        const long cxm = L.off(i, j - 1), cxp = L.off(i, j + 1);
        const long cym = L.off(i - 1, j), cyp = L.off(i + 1, j);
        if (a.capa == 0) {
            v = v + dtdy * X[4 * fs + cxm];
            v = v + X[c] - dtdx * (X[2 * fs + c] - X[fs + c]) - dtdy * (X[4 * fs + c] - X[3 * fs + c]);
            v = v - dtdy * X[3 * fs + cxp];
            v = v + dtdx * Y[4 * fs + cym];
            v = v + (Y[c] - dtdy * (Y[2 * fs + c] - Y[fs + c]) - dtdx * (Y[4 * fs + c] - Y[3 * fs + c]));
            v = v - dtdx * Y[3 * fs + cyp];
        } else {
            const double k = a.aux[(long)(a.capa - 1) * a.afs + c];
            v = v + dtdy * X[4 * fs + cxm] / k;
            v = v + X[c] - (dtdx * (X[2 * fs + c] - X[fs + c]) + dtdy * (X[4 * fs + c] - X[3 * fs + c])) / k;
            v = v - dtdy * X[3 * fs + cxp] / k;
            v = v + dtdx * Y[4 * fs + cym] / k;
            v = v + Y[c] - (dtdy * (Y[2 * fs + c] - Y[fs + c]) + dtdx * (Y[4 * fs + c] - Y[3 * fs + c])) / k;
            v = v - dtdx * Y[3 * fs + cyp] / k;
        }
    }
    a.qn[c] = v;
}


// One sweep of the dimensionally split step (step2ds.f90:96-163 for x,
// :186-255 for y) for the variable-coefficient advection solver: every
// slice, ghost slices included, updates its cells 1..n by
//   q + qadd - dtd (fadd(p+1) - fadd(p))   [/ capa(p) with a capacity],
// with no transverse terms; the other cells are copied, so that ghost data
// flow to the next sweep as in the reference's single array.  Lanes run
// along x in both directions (coalesced rows).
template <int DIR>
__global__ __launch_bounds__(256) void split_sweep_kernel(SplitArgs a)
{
    const Layout2& L = a.L;
    const int mbc = L.mbc;
    const int n = DIR == 1 ? L.mx : L.my;
    const int ns = DIR == 1 ? L.my : L.mx;
    const int ix = blockIdx.x * blockDim.x + threadIdx.x + 1 - mbc;
    const int iy = blockIdx.y * blockDim.y + threadIdx.y + 1 - mbc;
    const int p = DIR == 1 ? ix : iy, sl = DIR == 1 ? iy : ix;
    const bool in_grid = p <= n + mbc && sl <= ns + mbc;
    const bool live = in_grid && p >= 1 && p <= n;
    const int pc = live ? p : 1, sc = in_grid ? sl : 1 - mbc;
    auto off = [&](int pp) -> long { return DIR == 1 ? L.off(pp, sc) : L.off(sc, pp); };
    const double* vel = a.aux + (DIR == 1 ? 0 : a.afs);
    Pos o;
    eval_position(a.qin, a.aux, a.afs, vel, off, pc, n, live, a.dtd, a.capa, a.order, a.lim0,
                  a.fwave, o);
    store_wave_max(o.cfl, a.cfl_part, ((long)blockIdx.y * gridDim.x + blockIdx.x) * 4 + threadIdx.y);
    if (!in_grid) return;
    if (!live) {
        const long c = off(p);
        a.qout[c] = a.qin[c];
        return;
    }
    // qadd(p): A^+ of interface p, then A^- of interface p+1 (flux2.f90:113-121)
    const double qadd = (0.0 - o.dtd[2] * o.ap[1]) - o.dtd[2] * o.am[2];
    const double fl = 0.5 * o.cq[1], fh = 0.5 * o.cq[2];
    const long c = off(p);
    double v;
    if (a.capa == 0)
        v = o.q[2] + qadd - a.dtd * (fh - fl);
    else
        v = o.q[2] + qadd - a.dtd * (fh - fl) / a.aux[(long)(a.capa - 1) * a.afs + c];
    a.qout[c] = v;
}

__global__ void cfl_reduce_kernel(const double* part, long n, unsigned long long* out)
{
    wave_max_reduce(part, n, out);
}

// the per-wave Courant numbers of a launch of grid g (blocks of 64 x 4)
// into *out
void reduce_cfl(const double* part, dim3 g, unsigned long long* out, hipStream_t s)
{
    const long n = (long)g.x * g.y * 4;
    const long nb = (n + 255) / 256;
    hipLaunchKernelGGL(cfl_reduce_kernel, dim3((unsigned)(nb < 512 ? nb : 512)), dim3(256), 0, s,
                       part, n, out);
}

}  // namespace g2

// The kernels' blocks are 64 x 4 threads, one wave per row of the block.
int launch_step2_general(const g2::GenArgs& a, hipStream_t s)
{
    const dim3 blk(64, 4);
    const dim3 gx((a.L.mx + 63) / 64, (a.L.my + 2 + 3) / 4);
    const dim3 gy((a.L.my + 63) / 64, (a.L.mx + 2 + 3) / 4);
    hipLaunchKernelGGL(g2::records_kernel<1>, gx, blk, 0, s, a);
    g2::reduce_cfl(a.cfl_part, gx, a.cfl_bits, s);
    hipLaunchKernelGGL(g2::records_kernel<2>, gy, blk, 0, s, a);
    g2::reduce_cfl(a.cfl_part, gy, a.cfl_bits, s);
    const long n = (long)(a.L.mx + 2 * a.L.mbc) * (a.L.my + 2 * a.L.mbc);
    hipLaunchKernelGGL(g2::combine_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
    return (int)hipGetLastError();
}

int launch_split_sweep_general(int dir, const g2::SplitArgs& a, hipStream_t s)
{
    const dim3 blk(64, 4);
    const int nx = a.L.mx + 2 * a.L.mbc, ny = a.L.my + 2 * a.L.mbc;
    const dim3 g((nx + 63) / 64, (ny + 3) / 4);
    if (dir == 1)
        hipLaunchKernelGGL(g2::split_sweep_kernel<1>, g, blk, 0, s, a);
    else
        hipLaunchKernelGGL(g2::split_sweep_kernel<2>, g, blk, 0, s, a);
    g2::reduce_cfl(a.cfl_part, g, a.cfl_bits, s);
    return (int)hipGetLastError();
}

}  // namespace clawgpu
