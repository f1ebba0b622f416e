// Dimensionally split 2D step for gfx950 (FP64): one-direction sweeps.
//
// Replaces step2ds (src/2d/step2ds.f90:2-275) as called by dimsp2
// (src/2d/dimsp2.f:4-98): an x-sweep over every row including the ghost
// rows (step2ds.f90:96) and a y-sweep over every column including the ghost
// columns (:186), each  q <- q + qadd - dt/dx (fadd(i+1) - fadd(i))
// (:158-163, :250-255) with no transverse terms.
//
// Both sweeps read one buffer and write another (the reference's in-place y
// and Strang sweeps become out-of-place, which removes every ordering
// hazard between waves).  A sweep writes every cell of its output: the cells
// it updates get the update, the others are copied, so ghost data flow from
// one sweep to the next exactly as in the reference's single array.
//
// x-sweep: one wave per (64-column strip, group of rows); lanes are
// interfaces, the limiter's neighbours and amdq(i+1)/fadd(i+1) come through
// lane shifts.  y-sweep: one wave per (64-column strip, row segment); lane =
// column, walking down the rows with the same one-row-lag pipeline as the
// unsplit kernel.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "riemann2d.hpp"
#include "step2d_kernel.hpp"

namespace clawgpu {

struct SweepArgs {
    const double* qin;
    double* qout;
    Layout2 L;
    double dtd;  // dt/dx for x, dt/dy for y
    int order, use_fwave;
    int lim[kMaxWaves];
    int nstrips, seglen, nsegs;
    unsigned long long* cfl_bits;
    SolverParams P;
};

// Kernel bodies (shared with the kernels of a user solver's code object,
// include/clawgpu_solver.h); the kernels follow launch_sweep_impl's callers.
template <class S>
__device__ __forceinline__ void sweep_x_body(const SweepArgs& a)
{
    using Mth = typename S::Math;
    using Cell = typename S::Cell;
    using Cache = typename S::Cache;
    constexpr int M = S::M, MW = S::MW;
    const int lane = threadIdx.x & 63;
    const int wid = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (wid >= a.nstrips * a.nsegs) return;
    const int strip = wid % a.nstrips, seg = wid / a.nstrips;
    const Layout2 L = a.L;
    const int mx = L.mx, my = L.my, mbc = L.mbc;
    const int i = kStripOut * strip + lane - 1;
    const int ic = min(i, mx + mbc);
    const bool upd = lane >= 2 && lane <= kStripOut + 1 && i >= 1 && i <= mx;
    // ghost columns are copied by the strips at the two ends
    const bool copy_lo = strip == 0 && lane < 2;                       // i = -1, 0
    const bool copy_hi = strip == a.nstrips - 1 && i > mx && i <= mx + mbc;  // i = mx+1, mx+2
    const bool cfl_lane = lane >= 1 && i >= 1 && i <= mx + 1;
    const int j0 = 1 - mbc + seg * a.seglen;
    const int j1 = min(j0 + a.seglen, my + mbc + 1);
    const double dtdx = a.dtd;
    const bool second = a.order == 2;
    const SolverParams P = a.P;
    double cfl = 0.0;
    for (int j = j0; j < j1; ++j) {
        double qc[M];
#pragma unroll
        for (int m = 0; m < M; m++) qc[m] = a.qin[m * L.fs + L.off(ic, j)];
        const Cell cc = S::cell(P, qc);
        Cell cl;
        {
            const double* src = reinterpret_cast<const double*>(&cc);
            double* dst = reinterpret_cast<double*>(&cl);
#pragma unroll
            for (int k = 0; k < S::XSHIFT; k++) dst[k] = lane_up(src[k]);
        }
        double w[MW][M], s[MW], am[M], ap[M];
        Cache cx;
        S::template rpn<1, 2>(P, cl, cc, w, s, am, ap, cx);
        if (cfl_lane) {
#pragma unroll
            for (int mw = 0; mw < MW; mw++) cfl = fmax(cfl, cfl_term(dtdx, s[mw]));
        }
        double fa[M];
#pragma unroll
        for (int m = 0; m < M; m++) fa[m] = 0.0;
        if (second) {
            double dotr[MW], wn2[MW];
#pragma unroll
            for (int mw = 0; mw < MW; mw++) {
                double dr = 0.0, nn = 0.0;
#pragma unroll
                for (int m = 0; m < M; m++) {
                    double wr = lane_dn(w[mw][m]);
                    nn = nn + w[mw][m] * w[mw][m];
                    dr = dr + w[mw][m] * wr;
                }
                dotr[mw] = dr;
                wn2[mw] = nn;
            }
#pragma unroll
            for (int mw = 0; mw < MW; mw++) {
                double dl = lane_up(dotr[mw]);
                if (a.lim[mw] != 0 && wn2[mw] != 0.0) {
                    double r = Mth::div(s[mw] > 0.0 ? dl : dotr[mw], wn2[mw]);
                    double phi = philim<Mth>(r, a.lim[mw]);
#pragma unroll
                    for (int m = 0; m < M; m++) w[mw][m] = phi * w[mw][m];
                }
            }
            double cq[M];
#pragma unroll
            for (int m = 0; m < M; m++) cq[m] = 0.0;
#pragma unroll
            for (int mw = 0; mw < MW; mw++) {
                double sa = fabs(s[mw]);
                double sw = a.use_fwave ? copysign(1.0, s[mw]) : sa;
                double coef = sw * (1.0 - sa * dtdx);
#pragma unroll
                for (int m = 0; m < M; m++) cq[m] = cq[m] + coef * w[mw][m];
            }
#pragma unroll
            for (int m = 0; m < M; m++) fa[m] = 0.5 * cq[m];
        }
#pragma unroll
        for (int m = 0; m < M; m++) {
            double amn = lane_dn(am[m]);
            double fan = lane_dn(fa[m]);
            double qadd = (0.0 - dtdx * ap[m]) - dtdx * amn;
            double v = qc[m] + qadd - dtdx * (fan - fa[m]);
            if (upd)
                a.qout[m * L.fs + L.off(i, j)] = v;
            else if (copy_lo || copy_hi)
                a.qout[m * L.fs + L.off(i, j)] = qc[m];
        }
    }
    publish_cfl(cfl, a.cfl_bits);
}

template <class S>
__device__ __forceinline__ void sweep_y_body(const SweepArgs& a)
{
    using Mth = typename S::Math;
    using Cell = typename S::Cell;
    using Cache = typename S::Cache;
    constexpr int M = S::M, MW = S::MW;
    const int lane = threadIdx.x & 63;
    const int wid = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (wid >= a.nstrips * a.nsegs) return;
    const int strip = wid % a.nstrips, seg = wid / a.nstrips;
    const Layout2 L = a.L;
    const int mx = L.mx, my = L.my, mbc = L.mbc;
    const int i = 1 - mbc + 64 * strip + lane;  // every column, ghosts included
    const bool col_ok = i <= mx + mbc;
    const int ic = min(i, mx + mbc);
    const int r0 = 1 + seg * a.seglen;
    const int r1 = min(r0 + a.seglen, my + 1);
    const double dtdy = a.dtd;
    const bool second = a.order == 2;
    const SolverParams P = a.P;
    const double* __restrict__ qi = a.qin + L.off(ic, 0);
    double* __restrict__ qo = a.qout + L.off(ic, 0);
    const long pitch = L.pitch, fs = L.fs;

    // ghost rows are copied by the first and last segments
    if (col_ok) {
        if (seg == 0)
            for (int j = 1 - mbc; j <= 0; j++)
                for (int m = 0; m < M; m++) qo[m * fs + j * pitch] = qi[m * fs + j * pitch];
        if (r1 == my + 1)
            for (int j = my + 1; j <= my + mbc; j++)
                for (int m = 0; m < M; m++) qo[m * fs + j * pitch] = qi[m * fs + j * pitch];
    }

    Cell cprev;
    {
        double q0[M];
#pragma unroll
        for (int m = 0; m < M; m++) q0[m] = qi[m * fs + (r0 - 2) * pitch];
        cprev = S::cell(P, q0);
    }
    double wyp[MW][M], syp[MW], amyp[M], apyp[M], dotly[MW], qcp[M], qcpp[M];
    double qa_part[M], fa_part[M];
#pragma unroll
    for (int m = 0; m < M; m++) {
        amyp[m] = apyp[m] = qa_part[m] = fa_part[m] = qcp[m] = qcpp[m] = 0.0;
#pragma unroll
        for (int mw = 0; mw < MW; mw++) wyp[mw][m] = 0.0;
    }
#pragma unroll
    for (int mw = 0; mw < MW; mw++) syp[mw] = dotly[mw] = 0.0;
    double cfl = 0.0;
    for (int J = r0 - 1; J <= r1 + 1; ++J) {
        double qc[M];
#pragma unroll
        for (int m = 0; m < M; m++) qc[m] = qi[m * fs + (long)J * pitch];
        const Cell cc = S::cell(P, qc);
        double wy[MW][M], sy[MW], amy[M], apy[M];
        Cache cy;
        S::template rpn<2, 1>(P, cprev, cc, wy, sy, amy, apy, cy);
        if (col_ok && J >= 1 && J <= my + 1) {
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
            double fa[M];
#pragma unroll
            for (int m = 0; m < M; m++) fa[m] = 0.0;
            if (second) {
#pragma unroll
                for (int mw = 0; mw < MW; mw++) {
                    double nn = 0.0;
#pragma unroll
                    for (int m = 0; m < M; m++) nn = nn + wyp[mw][m] * wyp[mw][m];
                    if (a.lim[mw] != 0 && nn != 0.0) {
                        double r = Mth::div(syp[mw] > 0.0 ? dotly[mw] : dotry[mw], nn);
                        double phi = philim<Mth>(r, a.lim[mw]);
#pragma unroll
                        for (int m = 0; m < M; m++) wyp[mw][m] = phi * wyp[mw][m];
                    }
                }
                double cq[M];
#pragma unroll
                for (int m = 0; m < M; m++) cq[m] = 0.0;
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
            if (J - 2 >= r0 && col_ok) {
                // cell (i, J-2) = qcpp
#pragma unroll
                for (int m = 0; m < M; m++) {
                    double qadd = qa_part[m] - dtdy * amyp[m];
                    double v = qcpp[m] + qadd - dtdy * (fa[m] - fa_part[m]);
                    qo[m * fs + (long)(J - 2) * pitch] = v;
                }
            }
#pragma unroll
            for (int m = 0; m < M; m++) {
                qa_part[m] = 0.0 - dtdy * apyp[m];
                fa_part[m] = fa[m];
            }
        }
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
            qcpp[m] = qcp[m];
            qcp[m] = qc[m];
        }
        cprev = cc;
    }
    publish_cfl(cfl, a.cfl_bits);
}

template <class S>
__global__ __launch_bounds__(256) void sweep_x_kernel(SweepArgs a)
{
    sweep_x_body<S>(a);
}

template <class S>
__global__ __launch_bounds__(256) void sweep_y_kernel(SweepArgs a)
{
    sweep_y_body<S>(a);
}

template <class Mth>
static int launch_sweep_impl(int solver, int dir, const SweepArgs& a, int nblocks, hipStream_t s)
{
    if (solver == 1) {
        if (dir == 1)
            hipLaunchKernelGGL(sweep_x_kernel<Acoustics2<Mth>>, dim3(nblocks), dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL(sweep_y_kernel<Acoustics2<Mth>>, dim3(nblocks), dim3(256), 0, s, a);
    } else {
        if (dir == 1)
            hipLaunchKernelGGL(sweep_x_kernel<Euler2<Mth>>, dim3(nblocks), dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL(sweep_y_kernel<Euler2<Mth>>, dim3(nblocks), dim3(256), 0, s, a);
    }
    return (int)hipGetLastError();
}

}  // namespace clawgpu
