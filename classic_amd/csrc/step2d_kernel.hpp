// Fused 2D wave-propagation step for gfx950 (FP64): kernel template.
//
// Replaces, for one unsplit time step, the reference's step2 -> flux2 ->
// rpn2/limiter/rpt2 path (src/2d/step2.f90:2-251, src/2d/flux2.f90:5-235,
// src/2d/inlinelimiter.f90:25-105) with one kernel launch.
//
// Work decomposition ("column sweep").  Every 64-lane wavefront owns a strip
// of 64 adjacent columns and a segment of rows.  Lane l holds column
// i = 60*strip + l - 1; lanes 2..61 produce output cells, lanes 0, 1, 62, 63
// are halo.  The wave walks down its segment one row J per iteration:
//   * x-interfaces of row J are solved one per lane (interface i between
//     cells i-1 and i); the left cell's data, the limiter's neighbouring
//     waves and the contributions owed to the left neighbour cell come from
//     adjacent lanes through cross-lane shifts;
//   * the y-interface between rows J-1 and J is solved per lane; the limiter
//     needs the interfaces above and below, which the lane still holds in
//     registers from the previous iteration, so the y pipeline lags one row
//     and finishes cell (i, J-2) at iteration J.
// Every cell update is accumulated in registers in the reference's order
// (x-sweep rows j-1, j, j+1, then y-sweep columns i-1, i, i+1, with the
// parenthesisation of step2.f90:141-145 and :224-228).  q is read once from
// HBM (plus a 4-column / 3-row halo) and written once; the next row is
// prefetched while the current one is processed, and the row loop is
// unrolled by two so that the y pipeline alternates between two register
// sets instead of copying one into the other.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "bc2d.hpp"
#include "common.hpp"
#include "loop2d.hpp"
#include "riemann2d.hpp"
#include "lanes.hpp"

#ifndef CLAWGPU_MIN_WAVES
#define CLAWGPU_MIN_WAVES 1
#endif

namespace clawgpu {

// Destination of the row stores of lanes that have no cell to write (halo
// lanes, rows outside the segment).  Every row iteration issues its four
// stores unconditionally, redirected here when there is nothing to store:
// with stores under a branch the compiler cannot count the outstanding
// vector-memory operations, and its wait for the prefetched next row would
// also wait for the previous row's stores to be acknowledged.
static __device__ double g_store_sink[kMaxWaves * 2 * 64];

// The reads of one lane's column of qold.  BC = 0: the ghost cells are
// filled; the column is clamped into the array and rows are read as stored.
// BC = 1, 2 (the device-side loop, a.bcmap): the ghost cells were not
// filled; a ghost column or row is read from the interior cell bc2 copies
// into it (src/2d/bc2.f:4-215; a corner from the cell both maps give, as
// bc2's bottom/top passes over the filled ghost columns do), and with
// BC = 2 (some side is a wall) the wall sign flips of component 2 (x-walls)
// and 3 (y-walls) are applied as sign-bit flips.
template <int BC>
struct ColumnReader {
    const double* col;
    long pitch, fs;
    int my;
    int xflip;  // sign bit flipped in component 2 of every row of this column
    // BC > 0: source rows of the ghost rows -1, 0, my+1, my+2 (a wave reads
    // no row outside [-1, my+2] that it uses; rows beyond are clamped) and
    // their sign bits for component 3, all uniform, so that the row loop
    // picks them with scalar selects
    int gsrc[4], gflip[4];

    __device__ ColumnReader(const StepArgs& a, int i)
    {
        const Layout2& L = a.L;
        pitch = L.pitch;
        fs = L.fs;
        my = L.my;
        int ic;
        xflip = 0;
        if constexpr (BC > 0) {
            bool f;
            const int i0 = min(max(i, 1 - L.mbc), L.mx + L.mbc);
            ic = bc_map(i0, L.mx, L.mbc, a.mthbc[0], a.mthbc[1], f);
            if constexpr (BC == 2) xflip = f ? (int)0x80000000u : 0;
            const int rows[4] = {-1, 0, L.my + 1, L.my + 2};
#pragma unroll
            for (int e = 0; e < 4; e++) {
                gsrc[e] = bc_map(rows[e], L.my, L.mbc, a.mthbc[2], a.mthbc[3], f);
                gflip[e] = f ? (int)0x80000000u : 0;
            }
        } else {
            ic = i < L.mx + L.mbc ? i : L.mx + L.mbc;
        }
        col = a.qold + L.off(ic, 0);
    }

    // components 0..M-1 of row j (uniform across the wave)
    template <int M>
    __device__ __forceinline__ void row(int j, double (&v)[M]) const
    {
        int jr = j, yflip = 0;
        if constexpr (BC > 0) {
            if (j < 1 || j > my) {
                const int e = j < 0 ? 0 : j == 0 ? 1 : j == my + 1 ? 2 : 3;
                jr = gsrc[e];
                yflip = gflip[e];
            }
        }
#pragma unroll
        for (int m = 0; m < M; m++) v[m] = col[m * fs + (long)jr * pitch];
        if constexpr (BC == 2 && M >= 3) {
            v[1] = __hiloint2double(__double2hiint(v[1]) ^ xflip, __double2loint(v[1]));
            v[2] = __hiloint2double(__double2hiint(v[2]) ^ yflip, __double2loint(v[2]));
        }
    }
};

// Method flags fixed at compile time (CfgFix) or read from the arguments
// (CfgAny).  The fixed variant serves the common case method(2)=2,
// method(3)=2, one limiter for every wave, no f-waves.
struct CfgAny {
    static constexpr bool fixed = false;
    static constexpr int order = 0, trans = 0, lim = 0;
};
template <int ORDER, int TRANS, int LIM>
struct CfgFix {
    static constexpr bool fixed = true;
    static constexpr int order = ORDER, trans = TRANS, lim = LIM;
};

// The same wavefront work for constant-coefficient acoustics in fast
// arithmetic and the fixed configuration (second order, transverse 2, one
// limiter): the solver's structure written out.  Both waves of an interface
// are multiples of fixed eigenvectors, (-Z, 1) and (Z, 1) in (p, u_n) at
// speeds -c and +c, so with alpha1, alpha2 their strengths
//   * the limiter ratio w(k')·w(k) / w(k)·w(k) is alpha(k') / alpha(k);
//   * cqxx = kappa (Z (b2 - b1), b1 + b2) with b = phi alpha and
//     kappa = c (1 - c dt/dx) for both waves;
//   * a fluctuation has no transverse velocity, so its transverse split
//     depends on its pressure a alone: B^-a = (-c/2 a, c/(2Z) a),
//     B^+a = (c/2 a, c/(2Z) a) in (p, v_t);
// and the Courant number of every interface is c dt/dx.  The sums are the
// generic kernel's (fast-arithmetic order); results agree with it and with
// the reference to rounding.
template <class Cfg, int BC>
__device__ __forceinline__ void acoustics_wave_fast(const StepArgs& a, const int strip,
                                                    const int seg, const int lane)
{
    using Mth = FastMath;
    const Layout2 L = a.L;
    const int mx = L.mx, my = L.my, mbc = L.mbc;
    const int i = kStripOut * strip + lane - 1;
    const int r0 = 1 + seg * a.seglen;
    const int r1 = min(r0 + a.seglen, my + 1);
    const bool out_lane = lane >= 2 && lane <= kStripOut + 1 && i <= mx;
    const bool xcfl_lane = lane >= 1 && i >= 1 && i <= mx + 1;
    const bool ycfl_lane = i >= 0 && i <= mx + 1;
    const double dtdx = a.dtdx, dtdy = a.dtdy;
    const ColumnReader<BC> rd(a, i);
    double* __restrict__ qn = a.qnew + L.off(i, 0);
    const long pitch = L.pitch, fs = L.fs;
    const int rowmax = my + mbc;
    const double zz = a.P.zz, cc = a.P.cc;
    const double h = 0.5 / zz;             // 1 / (2Z)
    const double hc = 0.5 * cc, gam = cc * h;  // B^+- coefficients c/2, c/(2Z)
    const double kx = cc * (1.0 - cc * dtdx), ky = cc * (1.0 - cc * dtdy);
    const double cz = cc * zz;
    const double cflx = cfl_term(dtdx, cc), cfly = cfl_term(dtdy, cc);

    // one y-interface: strengths of its two waves and of the 2-wave of the
    // interface below it (the limiter's upwind neighbour)
    struct YA {
        double a1, a2, a2lo;
    };
    YA ysA{0.0, 0.0, 0.0}, ysB{0.0, 0.0, 0.0};
    double qnx[3];
    rd.row(r0 - 2, qnx);
    double pp = qnx[0], vp = qnx[2];  // cell J-1
    rd.row(r0 - 1, qnx);
    double vA[3] = {0.0, 0.0, 0.0}, vB[3] = {0.0, 0.0, 0.0};
    double g2p = 0.0, g2v = 0.0;  // x-row J-1's G2 owed to row J (p, v)
    double cfl = 0.0;

    auto phi = [&](double num, double al) {
        const double r = Mth::ratio(num, al);
        return al != 0.0 ? philim<Mth>(r, Cfg::lim) : 0.0;
    };

    auto row = [&](YA& yp, YA& yc, const int J) {
        const double p = qnx[0], u = qnx[1], v = qnx[2];
        rd.row(min(J + 1, rowmax), qnx);
        double vC[3];
        // ---------------------------------------------- x-row J
        if (J <= r1) {
            if (xcfl_lane) cfl = fmax(cfl, cflx);
            const double dp = p - lane_up(p), du = u - lane_up(u);
            const double a1 = (-dp + zz * du) * h, a2 = (dp + zz * du) * h;
            const double b1 = phi(lane_dn(a1), a1) * a1;  // 1-wave: speed -c, upwind = right
            const double b2 = phi(lane_up(a2), a2) * a2;  // 2-wave: speed +c, upwind = left
            const double cqp = kx * (zz * (b2 - b1)), cqu = kx * (b1 + b2);
            const double fap = 0.5 * cqp, fau = 0.5 * cqu;
            const double amp = cz * a1, amu = -cc * a1;  // A^- dq = -c w1
            const double app = cz * a2, apu = cc * a2;   // A^+ dq = +c w2
            const double amt = amp + cqp, apt = app - cqp;  // pressures to split in y
            // G1 = B^- parts, G2 = B^+ parts (p, v), the A^- side shifted left
            const double hx = -0.5 * dtdx;
            const double g1p = lane_dn(hx * (-hc * amt)) + hx * (-hc * apt);
            const double g2pn = lane_dn(hx * (hc * amt)) + hx * (hc * apt);
            const double g1v = lane_dn(hx * (gam * amt)) + hx * (gam * apt);
            const double g2vn = g1v;  // B^- and B^+ have the same v part
            const double amnp = lane_dn(amp), amnu = lane_dn(amu);
            const double fanp = lane_dn(fap), fanu = lane_dn(fau);
            {
                const double qadd = (0.0 - dtdx * app) - dtdx * amnp;
                double w = p + dtdy * g2p;
                vC[0] = w + qadd - dtdx * (fanp - fap) - dtdy * (g2pn - g1p);
            }
            {
                const double qadd = (0.0 - dtdx * apu) - dtdx * amnu;
                vC[1] = u + qadd - dtdx * (fanu - fau);
            }
            vC[2] = v + dtdy * g2v - dtdy * (g2vn - g1v);
            vB[0] = vB[0] - dtdy * g1p;
            vB[2] = vB[2] - dtdy * g1v;
            g2p = g2pn;
            g2v = g2vn;
        } else {
            vC[0] = vC[1] = vC[2] = 0.0;
        }

        // ---------------------------------- y-interface J on column i
        if (ycfl_lane && J >= 1 && J <= my + 1) cfl = fmax(cfl, cfly);
        {
            const double dp = p - pp, dv = v - vp;
            yc.a1 = (-dp + zz * dv) * h;
            yc.a2 = (dp + zz * dv) * h;
            yc.a2lo = yp.a2;
        }
        pp = p;
        vp = v;
        double vout[3] = {0.0, 0.0, 0.0};
        bool store = false;
        if (J - 1 >= r0) {
            // interface J-1 (yp) limited with its neighbours J-2 and J
            const double b1 = phi(yc.a1, yp.a1) * yp.a1;
            const double b2 = phi(yp.a2lo, yp.a2) * yp.a2;
            const double cqp = ky * (zz * (b2 - b1)), cqv = ky * (b1 + b2);
            const double fap = 0.5 * cqp, fav = 0.5 * cqv;
            const double amp = cz * yp.a1, amv = -cc * yp.a1;
            const double app = cz * yp.a2, apv = cc * yp.a2;
            const double amt = amp + cqp, apt = app - cqp;  // pressures to split in x
            const double hy = -0.5 * dtdy;
            if (J - 2 >= r0) {
                // cell J-2: the upper-face terms complete it
                store = out_lane && J - 2 < r1;
                const double G1p = hy * (-hc * amt), G2p = hy * (hc * amt), Gu = hy * (gam * amt);
                vout[0] = vA[0] - dtdy * amp - dtdy * fap;
                vout[0] = vout[0] + dtdx * lane_up(G2p) - dtdx * lane_dn(G1p) - dtdx * (G2p - G1p);
                vout[1] = vA[1] + dtdx * lane_up(Gu) - dtdx * lane_dn(Gu);
                vout[2] = vA[2] - dtdy * amv - dtdy * fav;
            }
            // cell J-1: the lower-face terms
            const double G1p = hy * (-hc * apt), G2p = hy * (hc * apt), Gu = hy * (gam * apt);
            double w = vB[0] - dtdy * app + dtdy * fap;
            vB[0] = w + dtdx * lane_up(G2p) - dtdx * lane_dn(G1p) - dtdx * (G2p - G1p);
            vB[1] = vB[1] + dtdx * lane_up(Gu) - dtdx * lane_dn(Gu);
            vB[2] = vB[2] - dtdy * apv + dtdy * fav;
        }
        {
            double* dst = store ? qn + (long)(J - 2) * pitch : g_store_sink + lane;
            const long stride = store ? fs : 64;
#pragma unroll
            for (int m = 0; m < 3; m++) dst[m * stride] = vout[m];
        }
#pragma unroll
        for (int m = 0; m < 3; m++) {
            vA[m] = vB[m];
            vB[m] = vC[m];
        }
    };

    const int nrow = r1 - r0 + 3;
    const int jend = r0 - 1 + nrow + (nrow & 1);
#pragma unroll
    for (int m = 0; m < 3; m++) g_store_sink[m * 64 + lane] = 0.0;
    for (int J = r0 - 1; J < jend; J += 2) {
        row(ysA, ysB, J);
        row(ysB, ysA, J + 1);
    }
    publish_cfl(cfl, a.cfl_bits);
}

// FAST arithmetic in the fixed configuration (second order, transverse 2,
// limiter Cfg::lim on every wave, no f-waves): the sweep of step2_wave with
// the arithmetic rearranged to issue fewer instructions --
//  * the limiter factor phi scales the wave's second-order coefficient
//    |s|(1 - |s| dt/dx) instead of its components (a zero wave keeps
//    phi = 1, as the reference skips it);
//  * an interface's Courant number is dt/dx max(|s_1|, |s_mw|), which bounds
//    every wave speed for these solvers (S::kOuterSpeedsBound); a NaN speed
//    stays visible through a running sum and is published as +inf;
//  * the transverse splits move between lanes unscaled and enter the sums
//    through the one constant -dt/dx dt/dy / 2;
//  * the normal fluctuation and the correction flux are combined before
//    they move (amdq + cq/2 to the left cell, apdq - cq/2 to the right).
// Same terms as the reference, summed in another order: the tests hold it to
// the 1e-12 bar.
template <class S, class Cfg, int BC>
__device__ __forceinline__ void step2_wave_fast(const StepArgs& a, const int strip, const int seg,
                                                const int lane)
{
    using Mth = typename S::Math;
    using Cell = typename S::Cell;
    using Cache = typename S::Cache;
    constexpr int M = S::M, MW = S::MW;
    static_assert(Mth::fast && Cfg::fixed && S::kOuterSpeedsBound, "fast fixed path");
    const Layout2 L = a.L;
    const int mx = L.mx, my = L.my, mbc = L.mbc;
    const int i = kStripOut * strip + lane - 1;
    const int r0 = 1 + seg * a.seglen;
    const int r1 = min(r0 + a.seglen, my + 1);  // output rows r0 .. r1-1
    const bool out_lane = lane >= 2 && lane <= kStripOut + 1 && i <= mx;
    const bool xcfl_lane = lane >= 1 && i >= 1 && i <= mx + 1;
    const bool ycfl_lane = i >= 0 && i <= mx + 1;
    const double dtdx = a.dtdx, dtdy = a.dtdy;
    const double kt = -0.5 * dtdx * dtdy;  // transverse terms of either sweep
    const ColumnReader<BC> rd(a, i);
    double* __restrict__ qn_base = a.qnew + L.off(i, 0);
    const long pitch = L.pitch, fs = L.fs;
    const int rowmax = my + mbc;
    const SolverParams P = a.P;

    struct YS {
        Cell cell;  // cell (i, J)
        double w[MW][M], s[MW], am[M], ap[M], dotl[MW];
        Cache cache;
    };
    YS ysA, ysB;
    {
        double q0[M];
        rd.row(r0 - 2, q0);
        ysA.cell = S::cell(P, q0);
#pragma unroll
        for (int mw = 0; mw < MW; mw++) {
            ysA.s[mw] = ysA.dotl[mw] = 0.0;
#pragma unroll
            for (int m = 0; m < M; m++) ysA.w[mw][m] = 0.0;
        }
#pragma unroll
        for (int m = 0; m < M; m++) ysA.am[m] = ysA.ap[m] = 0.0;
    }
    double qX[M], qY[M];  // q of the current row and the next (loaded ahead)
    rd.row(r0 - 1, qX);
    double vA[M], vB[M], b2prev[M];
#pragma unroll
    for (int m = 0; m < M; m++) vA[m] = vB[m] = b2prev[m] = 0.0;
    double cfl = 0.0, cfl_sum = 0.0;

    // second-order coefficient of a wave of speed s, limiter factor included
    auto coef = [&](double s, double num, double wn2, double dtd) {
        double phi = philim<Mth>(Mth::ratio(num, wn2), Cfg::lim);
        // minmod, superbee and MC give a finite phi even for the 0/0 ratio of
        // a zero wave, whose components are then zero whatever phi is; van
        // Leer and Beam-Warming would give NaN there
        if constexpr (Cfg::lim == 3 || Cfg::lim == 5) phi = wn2 != 0.0 ? phi : 1.0;
        const double sa = fabs(s);
        return (sa * (1.0 - sa * dtd)) * phi;
    };
    auto courant = [&](const double* s, double dtd, bool lane_ok) {
        double t = dtd * fmax(fabs(s[0]), fabs(s[MW - 1]));
        t = lane_ok ? t : 0.0;
        cfl = fmax(cfl, t);
        cfl_sum = cfl_sum + t;
    };

    // One row of the sweep.  yp holds interface J-1 (and cell J-1), yc
    // receives interface J (and cell J); qc holds q of row J and qn receives
    // row J+1; va holds cell J-2, which the y-interface J-1 completes, and
    // then receives the new cell J of the x-row; vb holds cell J-1.  The
    // caller alternates the two sets of each pair, so nothing is copied
    // between iterations.
    auto row = [&](YS& yp, YS& yc, const int J, const double (&qc)[M], double (&qn)[M],
                   double (&va)[M], double (&vb)[M]) {
        rd.row(min(J + 1, rowmax), qn);
        yc.cell = S::cell(P, qc);
        const Cell& cc = yc.cell;

        double vout[M];  // cell (i, J-2) when it is complete and this lane owns it
        bool store = false;
#pragma unroll
        for (int m = 0; m < M; m++) vout[m] = 0.0;

        // ------------------------------ y-interface J (flux2, ixy=2) on column i
        S::template rpn<2, 1>(P, yp.cell, cc, yc.w, yc.s, yc.am, yc.ap, yc.cache);
        if (J >= 1 && J <= my + 1) courant(yc.s, dtdy, ycfl_lane);
#pragma unroll
        for (int mw = 0; mw < MW; mw++) {
            double dr = 0.0;
#pragma unroll
            for (int m = 0; m < M; m++) dr = dr + yp.w[mw][m] * yc.w[mw][m];
            yc.dotl[mw] = dr;  // dotr of interface J-1 = dotl of interface J
        }
        if (J - 1 >= r0) {
            // interface J-1 (held in yp), limited with its neighbours J-2 and J
            double cq[M];
#pragma unroll
            for (int m = 0; m < M; m++) cq[m] = 0.0;
#pragma unroll
            for (int mw = 0; mw < MW; mw++) {
                double nn = 0.0;
#pragma unroll
                for (int m = 0; m < M; m++) nn = nn + yp.w[mw][m] * yp.w[mw][m];
                const double c =
                    coef(yp.s[mw], yp.s[mw] > 0.0 ? yp.dotl[mw] : yc.dotl[mw], nn, dtdy);
#pragma unroll
                for (int m = 0; m < M; m++) cq[m] = cq[m] + c * yp.w[mw][m];
            }
            double amt[M], apt[M];
#pragma unroll
            for (int m = 0; m < M; m++) {
                amt[m] = yp.am[m] + cq[m];
                apt[m] = yp.ap[m] - cq[m];
            }
            double bma[M], bpa[M], bmp[M], bpp[M];
            S::template rpt<2, 1>(P, yp.cache, amt, bma, bpa);
            S::template rpt<2, 1>(P, yp.cache, apt, bmp, bpp);
            // B^- parts of the columns to the right, B^+ parts of those to the
            // left, two components at a time
            if (J - 2 >= r0) store = out_lane && J - 2 < r1;  // cell (i, J-2) completes
#pragma unroll
            for (int m = 0; m < M; m += 2) {
                double bmar[2], bmpr[2], bpal[2], bppl[2];
#pragma unroll
                for (int e = 0; e < 2; e++) {
                    bmar[e] = lane_dn(bma[m + e]);
                    bmpr[e] = lane_dn(bmp[m + e]);
                    bpal[e] = lane_up(bpa[m + e]);
                    bppl[e] = lane_up(bpp[m + e]);
                }
#pragma unroll
                for (int e = 0; e < 2; e++) {
                    const int c = m + e;
                    if (J - 2 >= r0) {
                        // cell (i, J-2): its upper face
                        const double t = (bpal[e] - bmar[e]) - (bpa[c] - bma[c]);
                        const double v2 = va[c] - dtdy * (yp.am[c] + 0.5 * cq[c]);
                        vout[c] = v2 + kt * t;
                    }
                    // cell (i, J-1): its lower face
                    const double t = (bppl[e] - bmpr[e]) - (bpp[c] - bmp[c]);
                    const double v2 = vb[c] - dtdy * (yp.ap[c] - 0.5 * cq[c]);
                    vb[c] = v2 + kt * t;
                }
            }
        }
        {
            double* dst = store ? qn_base + (long)(J - 2) * pitch : g_store_sink + lane;
            const long stride = store ? fs : 64;
#pragma unroll
            for (int m = 0; m < M; m++) dst[m * stride] = vout[m];
        }

        // ------------------------------------------------ x-row J (flux2, ixy=1)
        if (J <= r1) {
            Cell cl;
            {
                const double* src = reinterpret_cast<const double*>(&cc);
                double* dst = reinterpret_cast<double*>(&cl);
#pragma unroll
                for (int k = 0; k < S::XSHIFT; k++) dst[k] = lane_up(src[k]);
            }
            double wx[MW][M], sx[MW], amx[M], apx[M];
            Cache cx;
            S::template rpn<1, 2>(P, cl, cc, wx, sx, amx, apx, cx);
            courant(sx, dtdx, xcfl_lane);
            double cq[M];
#pragma unroll
            for (int m = 0; m < M; m++) cq[m] = 0.0;
            double dr[MW], nn[MW];
#pragma unroll
            for (int mw = 0; mw < MW; mw++) {
                // wave mw of the interface to the right (lane+1)
                double wr[M];
#pragma unroll
                for (int m = 0; m < M; m++) wr[m] = lane_dn(wx[mw][m]);
                double d = 0.0, n2 = 0.0;
#pragma unroll
                for (int m = 0; m < M; m++) {
                    n2 = n2 + wx[mw][m] * wx[mw][m];
                    d = d + wx[mw][m] * wr[m];
                }
                dr[mw] = d;
                nn[mw] = n2;
            }
            double dl[MW];  // dotr of the interface to the left (lane-1)
#pragma unroll
            for (int mw = 0; mw < MW; mw++) dl[mw] = lane_up(dr[mw]);
#pragma unroll
            for (int mw = 0; mw < MW; mw++) {
                const double c = coef(sx[mw], sx[mw] > 0.0 ? dl[mw] : dr[mw], nn[mw], dtdx);
#pragma unroll
                for (int m = 0; m < M; m++) cq[m] = cq[m] + c * wx[mw][m];
            }
            double amt[M], apt[M], lf[M];
#pragma unroll
            for (int m = 0; m < M; m++) {
                amt[m] = amx[m] + cq[m];
                apt[m] = apx[m] - cq[m];
                lf[m] = amx[m] + 0.5 * cq[m];   // to the left cell
                apx[m] = apx[m] - 0.5 * cq[m];  // to the right cell
            }
            double bma[M], bpa[M], bmp[M], bpp[M];
            S::template rpt<1, 2>(P, cx, amt, bma, bpa);
            S::template rpt<1, 2>(P, cx, apt, bmp, bpp);
            // what the interface to the right (lane+1) owes this cell,
            // exchanged two components at a time
#pragma unroll
            for (int m = 0; m < M; m += 2) {
                double lfr[2], bmar[2], bpar[2];
#pragma unroll
                for (int e = 0; e < 2; e++) {
                    lfr[e] = lane_dn(lf[m + e]);
                    bmar[e] = lane_dn(bma[m + e]);
                    bpar[e] = lane_dn(bpa[m + e]);
                }
#pragma unroll
                for (int e = 0; e < 2; e++) {
                    const int c = m + e;
                    const double b1 = bmar[e] + bmp[c];  // G1 / (-dt/2dx)
                    const double b2 = bpar[e] + bpp[c];  // G2 / (-dt/2dx)
                    double v2 = qc[c] + kt * b2prev[c];
                    v2 = v2 - dtdx * apx[c];
                    v2 = v2 - dtdx * lfr[e];
                    va[c] = v2 - kt * (b2 - b1);
                    vb[c] = vb[c] - kt * b1;
                    b2prev[c] = b2;
                }
            }
        } else {
#pragma unroll
            for (int m = 0; m < M; m++) va[m] = 0.0;
        }
    };

    const int nrow = r1 - r0 + 3;
    const int jend = r0 - 1 + nrow + (nrow & 1);
#pragma unroll
    for (int m = 0; m < M; m++) g_store_sink[m * 64 + lane] = 0.0;
    for (int J = r0 - 1; J < jend; J += 2) {
        row(ysA, ysB, J, qX, qY, vA, vB);
        row(ysB, ysA, J + 1, qY, qX, vB, vA);
    }
    publish_cfl(cfl_sum == cfl_sum ? cfl : __builtin_huge_val(), a.cfl_bits);
}

// The work of one wavefront: strip `strip` (64 columns) and row segment `seg`
// of the grid described by a; BC: how ghost cells are read (ColumnReader).
template <class S, class Cfg, int BC = 0>
__device__ __forceinline__ void step2_wave(const StepArgs& a, const int strip, const int seg,
                                           const int lane)
{
    if constexpr (std::is_same_v<S, Acoustics2<FastMath>> && Cfg::fixed) {
        acoustics_wave_fast<Cfg, BC>(a, strip, seg, lane);
        return;
    } else if constexpr (S::Math::fast && Cfg::fixed) {
        step2_wave_fast<S, Cfg, BC>(a, strip, seg, lane);
        return;
    }
    using Mth = typename S::Math;
    using Cell = typename S::Cell;
    using Cache = typename S::Cache;
    constexpr int M = S::M, MW = S::MW;
    const Layout2 L = a.L;
    const int mx = L.mx, my = L.my, mbc = L.mbc;
    const int i = kStripOut * strip + lane - 1;
    const int ic = i < mx + mbc ? i : mx + mbc;  // clamped column (learned-limiter tables)
    const int r0 = 1 + seg * a.seglen;
    const int r1 = min(r0 + a.seglen, my + 1);  // output rows r0 .. r1-1
    const bool out_lane = lane >= 2 && lane <= kStripOut + 1 && i <= mx;
    const bool xcfl_lane = lane >= 1 && i >= 1 && i <= mx + 1;
    const bool ycfl_lane = i >= 0 && i <= mx + 1;
    const double dtdx = a.dtdx, dtdy = a.dtdy;
    const bool second = Cfg::fixed ? Cfg::order == 2 : a.order == 2;
    const int trn = Cfg::fixed ? Cfg::trans : a.trans;
    const bool tr2 = second && trn == 2;
    const bool trans = trn > 0;
    const bool fwave = Cfg::fixed ? false : a.use_fwave != 0;
    const ColumnReader<BC> rd(a, i);
    double* __restrict__ qn = a.qnew + L.off(i, 0);
    const long pitch = L.pitch, fs = L.fs;
    const int rowmax = my + mbc;
    const SolverParams P = a.P;
    auto limw = [&](int mw) { return Cfg::fixed ? Cfg::lim : a.lim[mw]; };

    // y-direction pipeline state of one interface (between rows J-1, J)
    struct YS {
        Cell cell;  // cell (i, J)
        double w[MW][M], s[MW], am[M], ap[M], dotl[MW];
        Cache cache;
    };
    YS ysA, ysB;
    {
        double q0[M];
        rd.row(r0 - 2, q0);
        ysA.cell = S::cell(P, q0);
#pragma unroll
        for (int mw = 0; mw < MW; mw++) {
            ysA.s[mw] = ysA.dotl[mw] = 0.0;
#pragma unroll
            for (int m = 0; m < M; m++) ysA.w[mw][m] = 0.0;
        }
#pragma unroll
        for (int m = 0; m < M; m++) ysA.am[m] = ysA.ap[m] = 0.0;
    }
    double qnx[M];
    rd.row(r0 - 1, qnx);

    double vA[M], vB[M], g2prev[M];
    double qa_part[M], fa_part[M], gp1[M], gp2[M];
#pragma unroll
    for (int m = 0; m < M; m++) {
        vA[m] = vB[m] = g2prev[m] = 0.0;
        qa_part[m] = fa_part[m] = gp1[m] = gp2[m] = 0.0;
    }
    double cfl = 0.0;

    // One row of the sweep.  yp holds interface J-1 (and cell J-1), yc
    // receives interface J (and cell J); the caller alternates the two
    // states so that nothing is copied between iterations.
    auto row = [&](YS& yp, YS& yc, const int J) {
        double qc[M];
#pragma unroll
        for (int m = 0; m < M; m++) qc[m] = qnx[m];
        rd.row(min(J + 1, rowmax), qnx);
        yc.cell = S::cell(P, qc);
        const Cell& cc = yc.cell;
        double vC[M];

        // ------------------------------------------------ x-row J (flux2, ixy=1)
        if (J <= r1) {
            Cell cl;
            {
                const double* src = reinterpret_cast<const double*>(&cc);
                double* dst = reinterpret_cast<double*>(&cl);
#pragma unroll
                for (int k = 0; k < S::XSHIFT; k++) dst[k] = lane_up(src[k]);
            }
            double wx[MW][M], sx[MW], amx[M], apx[M];
            Cache cx;
            S::template rpn<1, 2>(P, cl, cc, wx, sx, amx, apx, cx);
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
                        double wr = lane_dn(wx[mw][m]);
                        nn = nn + wx[mw][m] * wx[mw][m];
                        dr = dr + wx[mw][m] * wr;
                    }
                    dotr[mw] = dr;
                    wn2[mw] = nn;
                }
#pragma unroll
                for (int mw = 0; mw < MW; mw++) {
                    double dl = lane_up(dotr[mw]);
                    if (limw(mw) != 0 && wn2[mw] != 0.0) {
                        double phi;
                        if (!Cfg::fixed && limw(mw) == kLimTable) {
                            phi = a.phix[phi_index(ic, J, mx, my, mw, MW)];
                        } else {
                            double r = Mth::ratio(sx[mw] > 0.0 ? dl : dotr[mw], wn2[mw]);
                            phi = philim<Mth>(r, limw(mw));
                        }
#pragma unroll
                        for (int m = 0; m < M; m++) wx[mw][m] = phi * wx[mw][m];
                    }
                }
#pragma unroll
                for (int mw = 0; mw < MW; mw++) {
                    double sa = fabs(sx[mw]);
                    double sw = fwave ? copysign(1.0, sx[mw]) : sa;
                    double coef = sw * (1.0 - sa * dtdx);
#pragma unroll
                    for (int m = 0; m < M; m++) cq[m] = cq[m] + coef * wx[mw][m];
                }
#pragma unroll
                for (int m = 0; m < M; m++) fa[m] = 0.5 * cq[m];
            }
            double G1[M], G2[M];
            if (trans) {
                double amt[M], apt[M];
#pragma unroll
                for (int m = 0; m < M; m++) {
                    amt[m] = tr2 ? amx[m] + cq[m] : amx[m];
                    apt[m] = tr2 ? apx[m] - cq[m] : apx[m];
                }
                double bma[M], bpa[M], bmp[M], bpp[M];
                S::template rpt<1, 2>(P, cx, amt, bma, bpa);
                S::template rpt<1, 2>(P, cx, apt, bmp, bpp);
#pragma unroll
                for (int m = 0; m < M; m++) {
                    double g1n = lane_dn(-0.5 * dtdx * bma[m]);
                    double g2n = lane_dn(-0.5 * dtdx * bpa[m]);
                    G1[m] = g1n - 0.5 * dtdx * bmp[m];
                    G2[m] = g2n - 0.5 * dtdx * bpp[m];
                }
            } else {
#pragma unroll
                for (int m = 0; m < M; m++) G1[m] = G2[m] = 0.0;
            }
#pragma unroll
            for (int m = 0; m < M; m++) {
                double amn = lane_dn(amx[m]);
                double fan = lane_dn(fa[m]);
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

        double vout[M];  // cell (i, J-2) when it is complete and this lane owns it
        bool store = false;
#pragma unroll
        for (int m = 0; m < M; m++) vout[m] = 0.0;

        // ------------------------------ y-interface J (flux2, ixy=2) on column i
        S::template rpn<2, 1>(P, yp.cell, cc, yc.w, yc.s, yc.am, yc.ap, yc.cache);
        if (ycfl_lane && J >= 1 && J <= my + 1) {
#pragma unroll
            for (int mw = 0; mw < MW; mw++) cfl = fmax(cfl, cfl_term(dtdy, yc.s[mw]));
        }
#pragma unroll
        for (int mw = 0; mw < MW; mw++) {
            double dr = 0.0;
#pragma unroll
            for (int m = 0; m < M; m++) dr = dr + yp.w[mw][m] * yc.w[mw][m];
            yc.dotl[mw] = dr;  // dotr of interface J-1 = dotl of interface J
        }
        if (J - 1 >= r0) {
            // limit interface J-1 (held in yp) with its neighbours J-2 and J
            double cq[M], fa[M];
#pragma unroll
            for (int m = 0; m < M; m++) cq[m] = fa[m] = 0.0;
            if (second) {
#pragma unroll
                for (int mw = 0; mw < MW; mw++) {
                    double nn = 0.0;
#pragma unroll
                    for (int m = 0; m < M; m++) nn = nn + yp.w[mw][m] * yp.w[mw][m];
                    if (limw(mw) != 0 && nn != 0.0) {
                        double phi;
                        if (!Cfg::fixed && limw(mw) == kLimTable) {
                            phi = a.phiy[phi_index(ic, J - 1, mx, my, mw, MW)];
                        } else {
                            double r = Mth::ratio(yp.s[mw] > 0.0 ? yp.dotl[mw] : yc.dotl[mw], nn);
                            phi = philim<Mth>(r, limw(mw));
                        }
#pragma unroll
                        for (int m = 0; m < M; m++) yp.w[mw][m] = phi * yp.w[mw][m];
                    }
                }
#pragma unroll
                for (int mw = 0; mw < MW; mw++) {
                    double sa = fabs(yp.s[mw]);
                    double sw = fwave ? copysign(1.0, yp.s[mw]) : sa;
                    double coef = sw * (1.0 - sa * dtdy);
#pragma unroll
                    for (int m = 0; m < M; m++) cq[m] = cq[m] + coef * yp.w[mw][m];
                }
#pragma unroll
                for (int m = 0; m < M; m++) fa[m] = 0.5 * cq[m];
            }
            double bma[M], bpa[M], bmp[M], bpp[M];
            if (trans) {
                double amt[M], apt[M];
#pragma unroll
                for (int m = 0; m < M; m++) {
                    amt[m] = tr2 ? yp.am[m] + cq[m] : yp.am[m];
                    apt[m] = tr2 ? yp.ap[m] - cq[m] : yp.ap[m];
                }
                S::template rpt<2, 1>(P, yp.cache, amt, bma, bpa);
                S::template rpt<2, 1>(P, yp.cache, apt, bmp, bpp);
            } else {
#pragma unroll
                for (int m = 0; m < M; m++) bma[m] = bpa[m] = bmp[m] = bpp[m] = 0.0;
            }
            if constexpr (Mth::fast) {
                // Fast arithmetic sums a cell's y terms in another order:
                // interface J-1 adds its upper-face terms to cell J-2 (now
                // complete) and its lower-face terms to cell J-1 at once,
                // so nothing of it is carried to the next row.
                if (J - 2 >= r0) {
                    store = out_lane && J - 2 < r1;
#pragma unroll
                    for (int m = 0; m < M; m++) {
                        double G1 = trans ? -0.5 * dtdy * bma[m] : 0.0;
                        double G2 = trans ? -0.5 * dtdy * bpa[m] : 0.0;
                        double v = vA[m] - dtdy * yp.am[m] - dtdy * fa[m];
                        v = v + dtdx * lane_up(G2) - dtdx * lane_dn(G1) - dtdx * (G2 - G1);
                        vout[m] = v;
                    }
                }
#pragma unroll
                for (int m = 0; m < M; m++) {
                    double G1 = trans ? -0.5 * dtdy * bmp[m] : 0.0;
                    double G2 = trans ? -0.5 * dtdy * bpp[m] : 0.0;
                    double v = vB[m] - dtdy * yp.ap[m] + dtdy * fa[m];
                    vB[m] = v + dtdx * lane_up(G2) - dtdx * lane_dn(G1) - dtdx * (G2 - G1);
                }
            } else {
                if (J - 2 >= r0) {
                    // cell (i, J-2): its y-sweep terms are now complete
                    store = out_lane && J - 2 < r1;
#pragma unroll
                    for (int m = 0; m < M; m++) {
                        double qadd = qa_part[m] - dtdy * yp.am[m];
                        double dF = fa[m] - fa_part[m];
                        double G1 = trans ? (-0.5 * dtdy * bma[m]) - gp1[m] : 0.0;
                        double G2 = trans ? (-0.5 * dtdy * bpa[m]) - gp2[m] : 0.0;
                        double g2l = lane_up(G2);
                        double g1r = lane_dn(G1);
                        double v = vA[m] + dtdx * g2l;
                        v = v + (qadd - dtdy * dF - dtdx * (G2 - G1));
                        v = v - dtdx * g1r;
                        vout[m] = v;
                    }
                }
#pragma unroll
                for (int m = 0; m < M; m++) {
                    qa_part[m] = 0.0 - dtdy * yp.ap[m];
                    fa_part[m] = fa[m];
                    gp1[m] = 0.5 * dtdy * bmp[m];
                    gp2[m] = 0.5 * dtdy * bpp[m];
                }
            }
        }
        {
            double* dst = store ? qn + (long)(J - 2) * pitch : g_store_sink + lane;
            const long stride = store ? fs : 64;
#pragma unroll
            for (int m = 0; m < M; m++) dst[m * stride] = vout[m];
        }
#pragma unroll
        for (int m = 0; m < M; m++) {
            vA[m] = vB[m];
            vB[m] = vC[m];
        }
    };

    // rows r0-1 .. r1+1, padded to an even count (a padded last row only
    // recomputes in-bounds data and stores nothing)
    const int nrow = r1 - r0 + 3;
    const int jend = r0 - 1 + nrow + (nrow & 1);
    // the loop is entered, like every iteration ends, with the row loads
    // followed by M stores outstanding
#pragma unroll
    for (int m = 0; m < M; m++) g_store_sink[m * 64 + lane] = 0.0;
    for (int J = r0 - 1; J < jend; J += 2) {
        row(ysA, ysB, J);
        row(ysB, ysA, J + 1);
    }
    publish_cfl(cfl, a.cfl_bits);
}

// Kernel bodies: the library's kernels below and the kernels a user solver's
// code object exports (include/clawgpu_solver.h) are these functions with
// the same launch bounds.
template <class S, class Cfg>
__device__ __forceinline__ void step2_unsplit_body(const StepArgs& a)
{
    const int lane = threadIdx.x & 63;
    const int wid = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (wid >= a.nstrips * a.nsegs) return;  // uniform per wave
    step2_wave<S, Cfg>(a, wid % a.nstrips, wid / a.nstrips, lane);
}

template <class S, class Cfg>
__global__ __launch_bounds__(256, CLAWGPU_MIN_WAVES) void step2_unsplit_kernel(StepArgs a)
{
    step2_unsplit_body<S, Cfg>(a);
}

// A value equal in all lanes, moved to scalar registers.
__device__ __forceinline__ double uniform_double(double x)
{
    return __hiloint2double(__builtin_amdgcn_readfirstlane(__double2hiint(x)),
                            __builtin_amdgcn_readfirstlane(__double2loint(x)));
}

// Attempt l.k of the device-side claw2 loop: every wave first decides the
// previous attempt (loop_next) from the state before it and its Courant
// number, which gives the buffers and dt of this attempt; wave 0 of block 0
// stores that state and clears the Courant slot of the next attempt.  A
// finished loop makes the attempt a no-op.  The ghost cells are read
// through the boundary maps (a.bcmap), so no ghost-fill pass runs between
// attempts.
// (FAST arithmetic: held to two waves per SIMD, the unsplit kernel's
// occupancy, which the loop's few extra live values would otherwise cost.)
template <class S, class Cfg, int BC>
__device__ __forceinline__ void step2_loop_body(StepArgs a, const LoopArgs& l)
{
    LoopState s;
    if (l.k == 0)
        s = l.st[0];
    else
        loop_next(l, l.st[(l.k - 1) & 1], loop_cfl(l, l.k - 1), s);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (l.k > 0) l.st[l.k & 1] = s;
        l.cflb[(l.k + 1) % 3] = 0ull;
    }
    // the decision is computed in vector registers; its results are the
    // same in every lane, and moving them to scalar registers keeps the
    // step's constants (buffers, dt ratios) out of the vector budget
    if (__builtin_amdgcn_readfirstlane(s.done)) return;
    const int lane = threadIdx.x & 63;
    const int wid = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (wid >= a.nstrips * a.nsegs) return;  // uniform per wave
    const int cur = __builtin_amdgcn_readfirstlane(s.cur);
    a.qold = l.q[cur];
    a.qnew = l.q[(cur + 1) % 3];
    a.dtdx = uniform_double(s.dt / l.dx);
    a.dtdy = uniform_double(s.dt / l.dy);
    a.cfl_bits = l.cflb + l.k % 3;
    step2_wave<S, Cfg, BC>(a, wid % a.nstrips, wid / a.nstrips, lane);
}

template <class S, class Cfg, int BC>
__global__ __launch_bounds__(256, S::Math::fast ? 2 : CLAWGPU_MIN_WAVES) void step2_loop_kernel(StepArgs a, LoopArgs l)
{
    step2_loop_body<S, Cfg, BC>(a, l);
}

// Solver-level evaluation for tests: interface k between states ql[k]
// (left, reference layout q(meqn)) and qr[k]; outputs wave(meqn, mwaves),
// s(mwaves), amdq, apdq, and the transverse split of asdq[k] (imp = 1 uses
// amdq-like data, the split does not depend on imp) into bmasdq, bpasdq.
template <class S>
__device__ __forceinline__ void solver_probe_body(int n, int ixy, const SolverParams& P,
                                                  const double* ql, const double* qr,
                                                  const double* asdq, double* wave, double* s,
                                                  double* amdq, double* apdq, double* bm,
                                                  double* bp)
{
    constexpr int M = S::M, MW = S::MW;
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    typename S::Cell L = S::cell(P, ql + k * M), R = S::cell(P, qr + k * M);
    double w[MW][M], sp[MW], am[M], ap[M], b0[M], b1[M];
    typename S::Cache c;
    if (ixy == 1) {
        S::template rpn<1, 2>(P, L, R, w, sp, am, ap, c);
        S::template rpt<1, 2>(P, c, asdq + k * M, b0, b1);
    } else {
        S::template rpn<2, 1>(P, L, R, w, sp, am, ap, c);
        S::template rpt<2, 1>(P, c, asdq + k * M, b0, b1);
    }
    for (int mw = 0; mw < MW; mw++) {
        s[k * MW + mw] = sp[mw];
        for (int m = 0; m < M; m++) wave[(k * MW + mw) * M + m] = w[mw][m];
    }
    for (int m = 0; m < M; m++) {
        amdq[k * M + m] = am[m];
        apdq[k * M + m] = ap[m];
        bm[k * M + m] = b0[m];
        bp[k * M + m] = b1[m];
    }
}

template <class S>
__global__ void solver_probe_kernel(int n, int ixy, SolverParams P, const double* ql,
                                    const double* qr, const double* asdq, double* wave,
                                    double* s, double* amdq, double* apdq, double* bm,
                                    double* bp)
{
    solver_probe_body<S>(n, ixy, P, ql, qr, asdq, wave, s, amdq, apdq, bm, bp);
}

template <class Mth>
static int launch_probe_impl(int solver, int n, int ixy, const SolverParams& P, const double* ql,
                             const double* qr, const double* asdq, double* wave, double* s,
                             double* amdq, double* apdq, double* bm, double* bp, hipStream_t st)
{
    int nb = (n + 127) / 128;
    if (solver == 1)
        hipLaunchKernelGGL(solver_probe_kernel<Acoustics2<Mth>>, dim3(nb), dim3(128), 0, st, n, ixy,
                           P, ql, qr, asdq, wave, s, amdq, apdq, bm, bp);
    else
        hipLaunchKernelGGL(solver_probe_kernel<Euler2<Mth>>, dim3(nb), dim3(128), 0, st, n, ixy, P,
                           ql, qr, asdq, wave, s, amdq, apdq, bm, bp);
    return (int)hipGetLastError();
}

// the common configuration (second order, transverse 2, MC on every wave,
// no f-waves) has its own instantiation with the flags fixed
template <class S>
static bool fixed_cfg(const StepArgs& a)
{
    bool mc = a.order == 2 && a.trans == 2 && !a.use_fwave;
    for (int k = 0; k < S::MW; k++) mc = mc && a.lim[k] == 4;
    return mc;
}

template <class S>
static int waves_per_cu_cfg(const StepArgs& a)
{
    int nb = 0;
    hipError_t e = fixed_cfg<S>(a)
        ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, step2_unsplit_kernel<S, CfgFix<2, 2, 4>>, 256, 0)
        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, step2_unsplit_kernel<S, CfgAny>, 256, 0);
    return e == hipSuccess ? nb * kWavesPerBlock : 0;
}

template <class Mth>
static int step2_waves_per_cu_impl(int solver, const StepArgs& a)
{
    return solver == 1 ? waves_per_cu_cfg<Acoustics2<Mth>>(a) : waves_per_cu_cfg<Euler2<Mth>>(a);
}

template <class S>
static void launch_cfg(const StepArgs& a, int nblocks, hipStream_t s)
{
    const bool mc = fixed_cfg<S>(a);
    if (mc)
        hipLaunchKernelGGL((step2_unsplit_kernel<S, CfgFix<2, 2, 4>>), dim3(nblocks), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((step2_unsplit_kernel<S, CfgAny>), dim3(nblocks), dim3(256), 0, s, a);
}

template <class S, int BC>
static void launch_loop_bc(const StepArgs& a, const LoopArgs& l, int nblocks, hipStream_t s)
{
    if (fixed_cfg<S>(a))
        hipLaunchKernelGGL((step2_loop_kernel<S, CfgFix<2, 2, 4>, BC>), dim3(nblocks), dim3(256), 0, s, a, l);
    else
        hipLaunchKernelGGL((step2_loop_kernel<S, CfgAny, BC>), dim3(nblocks), dim3(256), 0, s, a, l);
}

template <class S>
static void launch_loop_cfg(const StepArgs& a, const LoopArgs& l, int nblocks, hipStream_t s)
{
    bool walls = false;
    for (int e = 0; e < 4; e++) walls = walls || a.mthbc[e] == 3;
    if (walls)
        launch_loop_bc<S, 2>(a, l, nblocks, s);
    else
        launch_loop_bc<S, 1>(a, l, nblocks, s);
}

template <class Mth>
static int launch_loop2_impl(int solver, const StepArgs& a, const LoopArgs& l, int nblocks,
                             hipStream_t s)
{
    if (solver == 1)
        launch_loop_cfg<Acoustics2<Mth>>(a, l, nblocks, s);
    else
        launch_loop_cfg<Euler2<Mth>>(a, l, nblocks, s);
    return (int)hipGetLastError();
}

template <class Mth>
static int launch_step2_impl(int solver, const StepArgs& a, int nblocks, hipStream_t s)
{
    if (solver == 1)
        launch_cfg<Acoustics2<Mth>>(a, nblocks, s);
    else
        launch_cfg<Euler2<Mth>>(a, nblocks, s);
    return (int)hipGetLastError();
}

}  // namespace clawgpu
