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
    double q[5], dtd[5];
    double un[5];
#pragma unroll
    for (int k = 0; k < 5; k++) {
        const long c = off(pc - 2 + k, sc);
        q[k] = a.q[c];
        un[k] = vel[c];
        dtd[k] = a.capa > 0 ? dtb / a.aux[(long)(a.capa - 1) * a.afs + c] : dtb;
    }
    // interfaces t = 0..3 (I = p-1+t) between local cells t and t+1
    double w[4], s[4], am[4], ap[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        w[t] = q[t + 1] - q[t];
        s[t] = un[t + 1];
        am[t] = fmin(s[t], 0.0) * w[t];
        ap[t] = fmax(s[t], 0.0) * w[t];
    }
    // Courant number: interface p (t = 1) and, for the last position, n+1
    double cfl = 0.0;
    if (live) {
        cfl = fmax(dtd[2] * s[1], -dtd[1] * s[1]);
        if (p == n) cfl = fmax(cfl, fmax(dtd[3] * s[2], -dtd[2] * s[2]));
        if (cfl != cfl || s[1] != s[1] || (p == n && s[2] != s[2])) cfl = __builtin_huge_val();
    }
    publish_cfl(cfl, a.cfl_bits);
    // second-order corrections of interfaces p (t = 1) and p+1 (t = 2)
    double cq[3] = {0.0, 0.0, 0.0};
    if (a.order == 2) {
#pragma unroll
        for (int t = 1; t <= 2; t++) {
            double wl = w[t];
            if (a.lim0 != 0) {
                const double wn2 = w[t] * w[t];
                if (wn2 != 0.0) {
                    const double dotl = w[t - 1] * w[t], dotr = w[t] * w[t + 1];
                    const double r = (s[t] > 0.0 ? dotl : dotr) / wn2;
                    wl = lim_phi(r, a.lim0) * w[t];
                }
            }
            const double dtdxave = 0.5 * (dtd[t] + dtd[t + 1]);
            const double sa = fabs(s[t]);
            const double sw = a.fwave ? copysign(1.0, s[t]) : sa;
            cq[t] = 0.0 + sw * (1.0 - sa * dtdxave) * wl;
        }
    }
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

}  // namespace g2

int launch_step2_general(const g2::GenArgs& a, hipStream_t s)
{
    const dim3 blk(64, 4);
    const dim3 gx((a.L.mx + 63) / 64, (a.L.my + 2 + 3) / 4);
    const dim3 gy((a.L.my + 63) / 64, (a.L.mx + 2 + 3) / 4);
    hipLaunchKernelGGL(g2::records_kernel<1>, gx, blk, 0, s, a);
    hipLaunchKernelGGL(g2::records_kernel<2>, gy, blk, 0, s, a);
    const long n = (long)(a.L.mx + 2 * a.L.mbc) * (a.L.my + 2 * a.L.mbc);
    hipLaunchKernelGGL(g2::combine_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
    return (int)hipGetLastError();
}

}  // namespace clawgpu
