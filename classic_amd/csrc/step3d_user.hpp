// Dimensionally split 3D step for user solvers (step3ds.f90 driven by
// dimsp3.f with a user rpn3; method(3) = -1 has no transverse terms, so a
// solver needs only its normal Riemann solve).  The sweeps follow the
// built-in acoustics sweeps of step3d_kernel.hpp with a general number of
// equations M and waves MW:
//   * x: one wavefront per 60-cell strip of an x-line, the left cell's data
//     and the neighbouring interfaces' waves moved between lanes by DPP;
//   * y, z: each lane marches along its line, keeping three interfaces and
//     two limited cqxx in registers.
// Per interface the limiter works on the unlimited waves of the
// neighbouring interfaces (src/3d/limiter.f90:34-58), cqxx includes the
// factor 1/2 (flux3.f90:248), and a cell is updated as
// q + ((0 - d apdq(i)) - d amdq(i+1)) - d (cqxx(i+1) - cqxx(i))
// (step3ds.f90), in IEEE arithmetic without contraction when the solver
// uses ExactMath.  Each sweep also forms its Courant number d max|s| over
// the interfaces 1..n+1 of the slices step3ds visits: every wave stores its
// maximum in cfl_part[wave] and the library reduces those afterwards (one
// atomic per wave on one address would serialise at about 10 ns each, and
// the x sweep runs a wave per strip of every x-line).
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "lanes.hpp"
#include "riemann2d.hpp"
#include "step3d_kernel.hpp"

namespace clawgpu {
namespace s3u {

// version of the sweep-argument contract, checked when a code object loads
constexpr int kAbi = 2;

struct Sweep3U {
    const double* __restrict__ q;
    double* __restrict__ qn;
    const double* __restrict__ aux;  // aux planes (maux >= the solver's MAUX)
    s3::Layout3 L;
    double d;  // dt / (spacing along the sweep)
    int m2, fwave;
    int lim[kMaxWaves];
    int nstrips, seg, nsegs;
    double* cfl_part;  // per-wave Courant numbers, indexed by the wave number
    SolverParams P;
};

template <class S>
struct IfaceU {
    double w[S::MW][S::M], s[S::MW], am[S::M], ap[S::M];
};

template <class S>
__device__ inline typename S::Cell load_cell(const Sweep3U& a, long c, double (&qv)[S::M])
{
    constexpr int NA = S::MAUX > 0 ? S::MAUX : 1;
    double x[NA];
#pragma unroll
    for (int m = 0; m < S::M; m++) qv[m] = a.q[m * a.L.fs + c];
#pragma unroll
    for (int m = 0; m < NA; m++) x[m] = m < S::MAUX ? a.aux[m * a.L.fs + c] : 0.0;
    return S::cell(a.P, qv, x);
}

// Wave mw of f limited by the unlimited waves of the interfaces left (wl)
// and right (wr) of it (limiter.f90: r from dot products over the
// components, phi(r) from philim).
template <class S>
__device__ inline void limit_waves(IfaceU<S>& f, const double (&wl)[S::MW][S::M],
                                   const double (&wr)[S::MW][S::M], const int* lim)
{
    using Mth = typename S::Math;
#pragma unroll
    for (int mw = 0; mw < S::MW; mw++) {
        if (lim[mw] == 0) continue;
        double wn2 = f.w[mw][0] * f.w[mw][0];
#pragma unroll
        for (int m = 1; m < S::M; m++) wn2 = wn2 + f.w[mw][m] * f.w[mw][m];
        if (wn2 == 0.0) continue;
        double dot;
        if (f.s[mw] > 0.0) {
            dot = wl[mw][0] * f.w[mw][0];
#pragma unroll
            for (int m = 1; m < S::M; m++) dot = dot + wl[mw][m] * f.w[mw][m];
        } else {
            dot = f.w[mw][0] * wr[mw][0];
#pragma unroll
            for (int m = 1; m < S::M; m++) dot = dot + f.w[mw][m] * wr[mw][m];
        }
        const double phi = philim<Mth>(Mth::div(dot, wn2), lim[mw]);
#pragma unroll
        for (int m = 0; m < S::M; m++) f.w[mw][m] = phi * f.w[mw][m];
    }
}

// cqxx = sum over waves of |s|/2 (1 - |s| d) w (f-waves: sign(s)/2)
template <class S>
__device__ inline void cqxx_u(const IfaceU<S>& f, double d, int fwave, double (&cq)[S::M])
{
    double fac[S::MW];
#pragma unroll
    for (int mw = 0; mw < S::MW; mw++) {
        const double a = fabs(f.s[mw]);
        const double l = fwave ? 0.5 * copysign(1.0, f.s[mw]) : 0.5 * a;
        fac[mw] = l * (1.0 - a * d);
    }
#pragma unroll
    for (int m = 0; m < S::M; m++) {
        double v = fac[0] * f.w[0][m];
#pragma unroll
        for (int mw = 1; mw < S::MW; mw++) v = v + fac[mw] * f.w[mw][m];
        cq[m] = v;
    }
}

template <class S>
__device__ inline double courant_u(const IfaceU<S>& f, double d)
{
    double c = 0.0;
#pragma unroll
    for (int mw = 0; mw < S::MW; mw++) c = fmax(c, cfl_term(d, f.s[mw]));
    return c;
}

// x sweep: as s3::sweep3_x_kernel, one wave per 60-cell strip of one x-line
template <class S>
__device__ __forceinline__ void sweep3u_x_body(const Sweep3U& a)
{
    constexpr int M = S::M, MW = S::MW;
    using Cell = typename S::Cell;
    const int l = threadIdx.x & 63;
    const long w = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const s3::Layout3& L = a.L;
    const int nyt = L.my + 2 * L.mbc;
    const long nsl = (long)nyt * (L.mz + 2 * L.mbc);
    if (w >= nsl * a.nstrips) return;  // uniform per wave
    const int strip = (int)(w % a.nstrips);
    const long sl = w / a.nstrips;
    const int j = (int)(sl % nyt) + 1 - L.mbc;
    const int k = (int)(sl / nyt) + 1 - L.mbc;
    const int i = s3::kOut * strip + l - 1;
    const long c = L.offc(i, j, k);
    double qv[M];
    const Cell R = load_cell<S>(a, c, qv);
    Cell Lc;
    {
        const double* src = reinterpret_cast<const double*>(&R);
        double* dst = reinterpret_cast<double*>(&Lc);
#pragma unroll
        for (int e = 0; e < (int)(sizeof(Cell) / sizeof(double)); e++) dst[e] = lane_up(src[e]);
    }
    IfaceU<S> f;
    S::template rpn<1>(a.P, Lc, R, f.w, f.s, f.am, f.ap);
    const bool slice = j >= 0 && j <= L.my + 1 && k >= 0 && k <= L.mz + 1;
    const double cfl = (l >= 1 && i >= 1 && i <= L.mx + 1 && slice) ? courant_u<S>(f, a.d) : 0.0;
    store_wave_max(cfl, a.cfl_part, w);
    double cq[M];
#pragma unroll
    for (int m = 0; m < M; m++) cq[m] = 0.0;
    if (a.m2 == 2) {
        double wl[MW][M], wr[MW][M];
#pragma unroll
        for (int mw = 0; mw < MW; mw++)
#pragma unroll
            for (int m = 0; m < M; m++) {
                wl[mw][m] = lane_up(f.w[mw][m]);
                wr[mw][m] = lane_dn(f.w[mw][m]);
            }
        limit_waves<S>(f, wl, wr, a.lim);
        cqxx_u<S>(f, a.d, a.fwave, cq);
    }
    double amhi[M], cqhi[M];
#pragma unroll
    for (int m = 0; m < M; m++) {
        amhi[m] = lane_dn(f.am[m]);
        cqhi[m] = lane_dn(cq[m]);
    }
    const bool own = (l >= 2 && l < 62) || (strip == 0 && l < 2);
    if (!own || i > L.mx + L.mbc) return;
    const bool upd = i >= 1 && i <= L.mx && slice;
    const long o = L.off(i, j, k);
#pragma unroll
    for (int m = 0; m < M; m++) {
        double v = qv[m];
        if (upd) {
            const double qadd = (0.0 - a.d * f.ap[m]) - a.d * amhi[m];
            v = v + qadd - a.d * (cqhi[m] - cq[m]);
        }
        a.qn[m * L.fs + o] = v;
    }
}

// y (dir = 1) or z (dir = 2) sweep: as s3::sweep3_yz_kernel, lanes on 64
// consecutive x-cells, each marching along its line over one segment
template <class S>
__device__ __forceinline__ void sweep3u_yz_body(const Sweep3U& a, int dir)
{
    constexpr int M = S::M, MW = S::MW;
    using Cell = typename S::Cell;
    const int l = threadIdx.x & 63;
    const long w = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const s3::Layout3& L = a.L;
    const int nt = dir == 1 ? L.mz + 2 * L.mbc : L.my + 2 * L.mbc;  // other transverse extent
    if (w >= (long)a.nstrips * nt * a.nsegs) return;               // uniform per wave
    const int blk = (int)(w % a.nstrips);
    const long r = w / a.nstrips;
    const int seg = (int)(r % a.nsegs);
    const int o2 = (int)(r / a.nsegs) + 1 - L.mbc;  // k for dir 1, j for dir 2
    const int i = 64 * blk + l + 1 - L.mbc;
    const bool live = i <= L.mx + L.mbc;  // lanes beyond the array only join the reduction
    const int ic = live ? i : L.mx + L.mbc;
    const int n = dir == 1 ? L.my : L.mz;
    const int n2 = dir == 1 ? L.mz : L.my;
    const bool slice_ok = i >= 0 && i <= L.mx + 1 && o2 >= 0 && o2 <= n2 + 1;
    const int s0 = seg == 0 ? 1 - L.mbc : 1 + seg * a.seg;
    const int s1 = seg == a.nsegs - 1 ? n + L.mbc : (seg + 1) * a.seg;
    auto idx = [&](int s) -> long {
        s = s < 1 - L.mbc ? 1 - L.mbc : (s > n + L.mbc ? n + L.mbc : s);
        return dir == 1 ? L.off(ic, s, o2) : L.off(ic, o2, s);
    };
    auto solve = [&](const Cell& lo, const Cell& hi, IfaceU<S>& f) {
        if (dir == 1) S::template rpn<2>(a.P, lo, hi, f.w, f.s, f.am, f.ap);
        else S::template rpn<3>(a.P, lo, hi, f.w, f.s, f.am, f.ap);
    };
    auto lim_cq = [&](IfaceU<S> f, const IfaceU<S>& fl, const IfaceU<S>& fr, double (&cq)[M]) {
#pragma unroll
        for (int m = 0; m < M; m++) cq[m] = 0.0;
        if (a.m2 != 2) return;
        limit_waves<S>(f, fl.w, fr.w, a.lim);
        cqxx_u<S>(f, a.d, a.fwave, cq);
    };
    double qa[M], qb[M], qc[M], qd[M];  // q of cells s0-2 .. s0+1 (qd: scratch)
    Cell cm2 = load_cell<S>(a, idx(s0 - 2), qd);
    Cell cm1 = load_cell<S>(a, idx(s0 - 1), qa);
    Cell c0 = load_cell<S>(a, idx(s0), qb);
    Cell c1 = load_cell<S>(a, idx(s0 + 1), qc);
    IfaceU<S> fprev, fcur, fnext;
    solve(cm2, cm1, fprev);  // interface s0-1
    solve(cm1, c0, fcur);    // interface s0
    solve(c0, c1, fnext);    // interface s0+1
    double cqcur[M], cqnext[M];
    lim_cq(fcur, fprev, fnext, cqcur);
    double cfl = 0.0;
    for (int s = s0; s <= s1; s++) {
        double q2[M];
        const Cell c2 = load_cell<S>(a, idx(s + 2), q2);
        IfaceU<S> fnn;
        solve(c1, c2, fnn);  // interface s+2
        lim_cq(fnext, fcur, fnn, cqnext);
        if (slice_ok && s >= 1 && s <= n + 1) cfl = fmax(cfl, courant_u<S>(fcur, a.d));
        if (live) {
            const long o = idx(s);
            const bool upd = slice_ok && s >= 1 && s <= n;
#pragma unroll
            for (int m = 0; m < M; m++) {
                double v = qb[m];
                if (upd) {
                    const double qadd = (0.0 - a.d * fcur.ap[m]) - a.d * fnext.am[m];
                    v = v + qadd - a.d * (cqnext[m] - cqcur[m]);
                }
                a.qn[m * L.fs + o] = v;
            }
        }
        fcur = fnext;
        fnext = fnn;
#pragma unroll
        for (int m = 0; m < M; m++) {
            cqcur[m] = cqnext[m];
            qb[m] = qc[m];
            qc[m] = q2[m];
        }
        c1 = c2;
    }
    store_wave_max(cfl, a.cfl_part, w);
}

}  // namespace s3u
}  // namespace clawgpu
