// Device Riemann solvers for the 2D wave-propagation step (gfx950, FP64).
//
// These replace the rpn2/rpt2 plug-ins of the reference contract
// (src/2d/claw2.f:312-402; call sites src/2d/flux2.f90:110-111, :204-205,
// :220-221).  A GPU cannot call host Fortran procedures, so each solver is a
// compile-time policy selected by a solver id.  One call solves ONE interface
// between a left cell L (cell i-1) and a right cell R (cell i); the normal
// momentum sits in slot MU and the transverse one in slot MV (0-based), which
// is how the reference's ixy argument is expressed here.
//
// Quantities that depend on one cell only (sqrt(rho), pressure, sound speed,
// velocities, the Roe weights q/sqrt(rho)) are evaluated once per cell in
// cell() and shared by the cell's two interfaces in each direction.
//
// The arithmetic is written from LeVeque, "Finite Volume Methods for
// Hyperbolic Problems": constant-coefficient acoustics (ch. 18/21) and the
// Roe solver for the Euler equations with the Harten-Hyman entropy fix
// (sec. 15.3).  With ExactMath every operation is the one the CPU
// restatement performs, in the same order, so (compiled without FMA
// contraction) results agree bit for bit; FastMath reuses reciprocals and
// is compiled with contraction, agreeing to a few ulps per operation.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace clawgpu {

struct ExactMath {
    static constexpr bool fast = false;
    __device__ static inline double div(double a, double b) { return a / b; }
    __device__ static inline double rcp(double b) { return 1.0 / b; }
    __device__ static inline double ratio(double a, double b) { return a / b; }
    __device__ static inline double sqrt_(double a) { return sqrt(a); }
};

struct FastMath {
    static constexpr bool fast = true;
    // hardware reciprocal refined by two Newton steps (about 1 ulp)
    __device__ static inline double rcp(double b)
    {
        double r = __builtin_amdgcn_rcp(b);
        double e = fma(-b, r, 1.0);
        r = fma(r, e, r);
        e = fma(-b, r, 1.0);
        return fma(r, e, r);
    }
    __device__ static inline double div(double a, double b)
    {
        double r = rcp(b);
        double q = a * r;
        return fma(fma(-b, q, a), r, q);
    }
    // a/b where a few ulps do: the limiter's wave ratio, which only selects
    // phi(r).  Hardware reciprocal (relative error up to about 5e-8 on
    // gfx950) refined by one Newton step: within about 11 ulps.
    __device__ static inline double ratio(double a, double b)
    {
        double r = __builtin_amdgcn_rcp(b);
        r = fma(r, fma(-b, r, 1.0), r);
        return a * r;
    }
    // 1/sqrt(a) for a > 0: the hardware estimate refined by two Newton steps
    // y <- y + y (1 - a y^2) / 2 (about 1 ulp); sqrt(a) = a y and 1/a = y^2
    // then cost one multiplication each instead of a square root and a
    // reciprocal
    __device__ static inline double rsqrt(double a)
    {
        double y = __builtin_amdgcn_rsq(a);
        double e = fma(-a * y, y, 1.0);
        y = fma(0.5 * y, e, y);
        e = fma(-a * y, y, 1.0);
        return fma(0.5 * y, e, y);
    }
    // hardware reciprocal square root, one coupled Newton step for sqrt and
    // 1/(2 sqrt), then a residual correction (a > 0; about 1 ulp)
    __device__ static inline double sqrt_(double a)
    {
        double y = __builtin_amdgcn_rsq(a);
        double g = a * y;
        double h = 0.5 * y;
        double r = fma(-g, h, 0.5);
        g = fma(g, r, g);
        h = fma(h, r, h);
        double d = fma(-g, g, a);
        return fma(d, h, g);
    }
};

// Constant-coefficient acoustics, q = (p, u, v).  Waves r1 = (-Z, 1, 0) at
// speed -c and r2 = (Z, 1, 0) at speed +c along the normal direction.
template <class Mth>
struct Acoustics2 {
    using Math = Mth;
    static constexpr int M = 3;
    static constexpr int MW = 2;
    struct Cell {
        double q[M];
    };
    static constexpr int XSHIFT = 3;  // leading doubles of Cell an x-interface needs
    // the largest wave speed magnitude of an interface is that of its first
    // or its last wave (here both are c)
    static constexpr bool kOuterSpeedsBound = true;
    struct Cache {};

    __device__ static inline Cell cell(const SolverParams&, const double* q)
    {
        Cell c;
#pragma unroll
        for (int m = 0; m < M; m++) c.q[m] = q[m];
        return c;
    }

    template <int MU, int MV>
    __device__ static inline void rpn(const SolverParams& P, const Cell& L, const Cell& R,
                                      double w[MW][M], double s[MW], double am[M], double ap[M],
                                      Cache&)
    {
        const double zz = P.zz, cc = P.cc;
        double dp = R.q[0] - L.q[0];
        double du = R.q[MU] - L.q[MU];
        double a1, a2;
        if constexpr (Mth::fast) {
            double h = 0.5 / zz;
            a1 = (-dp + zz * du) * h;
            a2 = (dp + zz * du) * h;
        } else {
            a1 = (-dp + zz * du) / (2.0 * zz);
            a2 = (dp + zz * du) / (2.0 * zz);
        }
        w[0][0] = -a1 * zz;
        w[0][MU] = a1;
        w[0][MV] = 0.0;
        s[0] = -cc;
        w[1][0] = a2 * zz;
        w[1][MU] = a2;
        w[1][MV] = 0.0;
        s[1] = cc;
#pragma unroll
        for (int m = 0; m < M; m++) {
            am[m] = s[0] * w[0][m];
            ap[m] = s[1] * w[1][m];
        }
    }

    // Split asdq into down/up-going parts with r1 = (-Z, 0, 1) at -c and
    // r2 = (Z, 0, 1) at +c in the transverse direction.
    template <int MU, int MV>
    __device__ static inline void rpt(const SolverParams& P, const Cache&, const double* asdq,
                                      double bm[M], double bp[M])
    {
        const double zz = P.zz, cc = P.cc;
        double a1, a2;
        if constexpr (Mth::fast) {
            double h = 0.5 / zz;
            a1 = (-asdq[0] + zz * asdq[MV]) * h;
            a2 = (asdq[0] + zz * asdq[MV]) * h;
        } else {
            a1 = (-asdq[0] + zz * asdq[MV]) / (2.0 * zz);
            a2 = (asdq[0] + zz * asdq[MV]) / (2.0 * zz);
        }
        bm[0] = cc * a1 * zz;
        bm[MU] = 0.0;
        bm[MV] = -cc * a1;
        bp[0] = cc * a2 * zz;
        bp[MU] = 0.0;
        bp[MV] = cc * a2;
    }
};

// Euler equations, q = (rho, rho u, rho v, E), ideal gas.  Three waves: the
// u-a acoustic wave, the entropy and shear waves lumped at speed u, and the
// u+a acoustic wave.
template <class Mth>
struct Euler2 {
    using Math = Mth;
    static constexpr int M = 4;
    static constexpr int MW = 3;
    struct Cell {
        double q[M];
        double sr;     // sqrt(rho)
        double rw[2];  // q(2)/sqrt(rho), q(3)/sqrt(rho): Roe weights of the momenta
        double hw;     // (E + p)/sqrt(rho): Roe weight of the enthalpy
        double c;      // sound speed
        double vel[2]; // q(2)/rho, q(3)/rho
    };
    // an x-interface needs all but vel[1] of its left cell (Cell order)
    static constexpr int XSHIFT = 10;
    // |u| <= max(|u - a|, |u + a|) for a >= 0, also after rounding (fl is
    // monotonic): the middle wave never sets the Courant number
    static constexpr bool kOuterSpeedsBound = true;
    struct Cache {
        double u, v, enth, a, g1a2, euv, u2v2, h2a;  // h2a = 1/(2a) (fast math only)
    };

    __device__ static inline double pressure(const double* q, double g1)
    {
        if constexpr (Mth::fast)
            return g1 * (q[3] - 0.5 * (q[1] * q[1] + q[2] * q[2]) * Mth::rcp(q[0]));
        else
            return g1 * (q[3] - 0.5 * (q[1] * q[1] + q[2] * q[2]) / q[0]);
    }

    __device__ static inline Cell cell(const SolverParams& P, const double* q)
    {
        Cell c;
#pragma unroll
        for (int m = 0; m < M; m++) c.q[m] = q[m];
        const double g1 = P.gamma1;
        if constexpr (Mth::fast) {
            // one reciprocal square root gives sqrt(rho), 1/rho and 1/sqrt(rho)
            const double isr = Mth::rsqrt(q[0]);
            c.sr = q[0] * isr;
            const double ri = isr * isr;
            double p = g1 * (q[3] - 0.5 * (q[1] * q[1] + q[2] * q[2]) * ri);
            c.rw[0] = q[1] * isr;
            c.rw[1] = q[2] * isr;
            c.hw = (q[3] + p) * isr;
            c.c = Mth::sqrt_(P.gamma * p * ri);
            c.vel[0] = q[1] * ri;
            c.vel[1] = q[2] * ri;
        } else {
            c.sr = Mth::sqrt_(q[0]);
            double p = g1 * (q[3] - 0.5 * (q[1] * q[1] + q[2] * q[2]) / q[0]);
            c.rw[0] = q[1] / c.sr;
            c.rw[1] = q[2] / c.sr;
            c.hw = (q[3] + p) / c.sr;
            c.c = sqrt(P.gamma * p / q[0]);
            c.vel[0] = q[1] / q[0];
            c.vel[1] = q[2] / q[0];
        }
        return c;
    }

    template <int MU, int MV>
    __device__ static inline void rpn(const SolverParams& P, const Cell& L, const Cell& R,
                                      double w[MW][M], double s[MW], double am[M], double ap[M],
                                      Cache& r)
    {
        const double g1 = P.gamma1;
        const double* qa = L.q;
        const double* qb = R.q;
        // Roe averages (FVMHP 15.3.2)
        double rhsq2 = L.sr + R.sr;
        if constexpr (Mth::fast) {
            double inv = Mth::rcp(rhsq2);
            r.u = (L.rw[MU - 1] + R.rw[MU - 1]) * inv;
            r.v = (L.rw[MV - 1] + R.rw[MV - 1]) * inv;
            r.enth = (L.hw + R.hw) * inv;
        } else {
            r.u = (L.rw[MU - 1] + R.rw[MU - 1]) / rhsq2;
            r.v = (L.rw[MV - 1] + R.rw[MV - 1]) / rhsq2;
            r.enth = (L.hw + R.hw) / rhsq2;
        }
        r.u2v2 = r.u * r.u + r.v * r.v;
        double a2 = g1 * (r.enth - 0.5 * r.u2v2);
        if constexpr (Mth::fast) {
            // a = a2 / sqrt(a2), 1/a2 = (1/sqrt(a2))^2, 1/(2a) = 0.5/sqrt(a2)
            const double ia = Mth::rsqrt(a2);
            r.a = a2 * ia;
            r.g1a2 = g1 * (ia * ia);
            r.h2a = 0.5 * ia;
        } else {
            r.a = Mth::sqrt_(a2);
            r.g1a2 = g1 / a2;
            r.h2a = 0.0;
        }
        r.euv = r.enth - r.u2v2;

        double d1 = qb[0] - qa[0];
        double d2 = qb[MU] - qa[MU];
        double d3 = qb[MV] - qa[MV];
        double d4 = qb[3] - qa[3];
        double a3 = r.g1a2 * (r.euv * d1 + r.u * d2 + r.v * d3 - d4);
        double a2s = d3 - r.v * d1;
        double a4;
        if constexpr (Mth::fast)
            a4 = (d2 + (r.a - r.u) * d1 - r.a * a3) * r.h2a;
        else
            a4 = (d2 + (r.a - r.u) * d1 - r.a * a3) / (2.0 * r.a);
        double a1 = d1 - a3 - a4;
        w[0][0] = a1;
        w[0][MU] = a1 * (r.u - r.a);
        w[0][MV] = a1 * r.v;
        w[0][3] = a1 * (r.enth - r.u * r.a);
        s[0] = r.u - r.a;
        w[1][0] = a3;
        w[1][MU] = a3 * r.u;
        w[1][MV] = a3 * r.v + a2s;
        w[1][3] = a3 * 0.5 * r.u2v2 + a2s * r.v;
        s[1] = r.u;
        w[2][0] = a4;
        w[2][MU] = a4 * (r.u + r.a);
        w[2][MV] = a4 * r.v;
        w[2][3] = a4 * (r.enth + r.u * r.a);
        s[2] = r.u + r.a;

        if constexpr (Mth::fast) {
            if (P.efix) {
                // Harten-Hyman as below, with the transonic tests done on
                // signs: s1 = u_m - c_m > 0 iff rho u_m > 0 and
                // (rho u_m)^2 > gamma rho_m p_m, which needs no sqrt or
                // division; the sonic speeds are evaluated only when a
                // transonic rarefaction is present.
                const double gamma = P.gamma, gg1 = P.gamma * g1;
                const double s0 = L.vel[MU - 1] - L.c;
                const bool super = s0 >= 0.0 && s[0] > 0.0;
                double qm0 = qa[0] + w[0][0], qmu = qa[MU] + w[0][MU];
                double qmv = qa[MV] + w[0][MV], qme = qa[3] + w[0][3];
                double km = qmu * qmu;
                double tm = gg1 * (qme * qm0 - 0.5 * (km + qmv * qmv));
                bool tr1 = s0 < 0.0 && qmu > 0.0 && tm >= 0.0 && km > tm;
                // without a transonic 1-wave the left-going part is min(s1, 0):
                // a supersonic interface (s0 >= 0 and s1 > 0) gives 0 as well
                double f1 = fmin(s[0], 0.0);
                if (tr1) {
                    double ri = Mth::rcp(qm0);
                    double pm = g1 * (qme - 0.5 * (km + qmv * qmv) * ri);
                    double s1 = qmu * ri - Mth::sqrt_(gamma * pm * ri);
                    f1 = s0 * (s1 - s[0]) / (s1 - s0);
                }
                const bool left2 = !super && s[1] < 0.0;
                const double s3 = R.vel[MU - 1] + R.c;
                double qn0 = qb[0] - w[2][0], qnu = qb[MU] - w[2][MU];
                double qnv = qb[MV] - w[2][MV], qne = qb[3] - w[2][3];
                double kn = qnu * qnu;
                double tn = gg1 * (qne * qn0 - 0.5 * (kn + qnv * qnv));
                bool tr3 = left2 && s3 > 0.0 && qnu < 0.0 && tn >= 0.0 && kn > tn;
                // s1 <= s2 <= s3 after rounding (u - a <= u <= u + a, a >= 0):
                // where the 2-wave is not left-going (or the interface is
                // supersonic) min(s, 0) of the 2- and 3-waves is already 0
                double f3 = fmin(s[2], 0.0);
                if (tr3) {
                    double ri = Mth::rcp(qn0);
                    double pn = g1 * (qne - 0.5 * (kn + qnv * qnv) * ri);
                    double s2 = qnu * ri + Mth::sqrt_(gamma * pn * ri);
                    f3 = s2 * (s3 - s[2]) / (s3 - s2);
                }
                const double f2 = fmin(s[1], 0.0);
                const double g1c = s[0] - f1, g2c = s[1] - f2, g3c = s[2] - f3;
#pragma unroll
                for (int m = 0; m < M; m++) {
                    am[m] = f1 * w[0][m] + f2 * w[1][m] + f3 * w[2][m];
                    ap[m] = g1c * w[0][m] + g2c * w[1][m] + g3c * w[2][m];
                }
                return;
            }
        }

        if (!P.efix) {
#pragma unroll
            for (int m = 0; m < M; m++) {
                double amm = 0.0, apm = 0.0;
#pragma unroll
                for (int mw = 0; mw < MW; mw++) {
                    if (s[mw] < 0.0)
                        amm = amm + s[mw] * w[mw][m];
                    else
                        apm = apm + s[mw] * w[mw][m];
                }
                am[m] = amm;
                ap[m] = apm;
            }
            return;
        }

        // Harten-Hyman: left-going flux difference with a transonic split of
        // the 1- or 3-wave; the right-going part is the rest of sum(s*W).
        const double gamma = P.gamma;
        double amv[M] = {0.0, 0.0, 0.0, 0.0};
        do {
            double s0 = L.vel[MU - 1] - L.c;  // u-c in the left cell
            if (s0 >= 0.0 && s[0] > 0.0) break;
            double qm[M];
#pragma unroll
            for (int m = 0; m < M; m++) qm[m] = qa[m] + w[0][m];
            double s1;
            if constexpr (Mth::fast) {
                double ri = Mth::rcp(qm[0]);
                double pm = g1 * (qm[3] - 0.5 * (qm[1] * qm[1] + qm[2] * qm[2]) * ri);
                s1 = qm[MU] * ri - Mth::sqrt_(gamma * pm * ri);
            } else {
                double pm = pressure(qm, g1);
                double cm = sqrt(gamma * pm / qm[0]);
                s1 = qm[MU] / qm[0] - cm;
            }
            double sfract;
            if (s0 < 0.0 && s1 > 0.0)
                sfract = s0 * (s1 - s[0]) / (s1 - s0);
            else if (s[0] < 0.0)
                sfract = s[0];
            else
                sfract = 0.0;
#pragma unroll
            for (int m = 0; m < M; m++) amv[m] = sfract * w[0][m];
            if (s[1] >= 0.0) break;
#pragma unroll
            for (int m = 0; m < M; m++) amv[m] = amv[m] + s[1] * w[1][m];
            double s3 = R.vel[MU - 1] + R.c;  // u+c in the right cell
            double qn[M];
#pragma unroll
            for (int m = 0; m < M; m++) qn[m] = qb[m] - w[2][m];
            double s2;
            if constexpr (Mth::fast) {
                double ri = Mth::rcp(qn[0]);
                double pn = g1 * (qn[3] - 0.5 * (qn[1] * qn[1] + qn[2] * qn[2]) * ri);
                s2 = qn[MU] * ri + Mth::sqrt_(gamma * pn * ri);
            } else {
                double pn = pressure(qn, g1);
                double cn = sqrt(gamma * pn / qn[0]);
                s2 = qn[MU] / qn[0] + cn;
            }
            if (s2 < 0.0 && s3 > 0.0)
                sfract = s2 * (s3 - s[2]) / (s3 - s2);
            else if (s[2] < 0.0)
                sfract = s[2];
            else
                break;
#pragma unroll
            for (int m = 0; m < M; m++) amv[m] = amv[m] + sfract * w[2][m];
        } while (0);
#pragma unroll
        for (int m = 0; m < M; m++) {
            double df = 0.0;
#pragma unroll
            for (int mw = 0; mw < MW; mw++) df = df + s[mw] * w[mw][m];
            am[m] = amv[m];
            ap[m] = df - amv[m];
        }
    }

    // Transverse split with the eigenvectors of the transverse Jacobian at
    // the same Roe state as the normal solve of this interface.
    template <int MU, int MV>
    __device__ static inline void rpt(const SolverParams&, const Cache& r, const double* asdq,
                                      double bm[M], double bp[M])
    {
        if constexpr (Mth::fast) {
            // the same split written on the eigen-coefficients: with
            // c_p = min(s_p, 0) * alpha_p (resp. max for bpasdq)
            double d1 = asdq[0], d2 = asdq[MU], d3 = asdq[MV], d4 = asdq[3];
            double a3 = r.g1a2 * (r.euv * d1 + r.u * d2 + r.v * d3 - d4);
            double a2 = d2 - r.u * d1;
            double a4 = (d3 + (r.a - r.v) * d1 - r.a * a3) * r.h2a;
            double a1 = d1 - a3 - a4;
            const double l1 = r.v - r.a, l2 = r.v, l3 = r.v + r.a;
            const double va = r.v * r.a, hk = 0.5 * r.u2v2;
            {
                double m1 = fmin(l1, 0.0) * a1, m2 = fmin(l2, 0.0), m3 = fmin(l3, 0.0) * a4;
                double S = m1 + m3, D = m3 - m1, e2 = m2 * a3, s2 = m2 * a2;
                double Se = S + e2;
                bm[0] = Se;
                bm[MU] = Se * r.u + s2;
                bm[MV] = Se * r.v + D * r.a;
                bm[3] = S * r.enth + D * va + e2 * hk + s2 * r.u;
            }
            {
                double m1 = fmax(l1, 0.0) * a1, m2 = fmax(l2, 0.0), m3 = fmax(l3, 0.0) * a4;
                double S = m1 + m3, D = m3 - m1, e2 = m2 * a3, s2 = m2 * a2;
                double Se = S + e2;
                bp[0] = Se;
                bp[MU] = Se * r.u + s2;
                bp[MV] = Se * r.v + D * r.a;
                bp[3] = S * r.enth + D * va + e2 * hk + s2 * r.u;
            }
            return;
        }
        double d1 = asdq[0], d2 = asdq[MU], d3 = asdq[MV], d4 = asdq[3];
        double a3 = r.g1a2 * (r.euv * d1 + r.u * d2 + r.v * d3 - d4);
        double a2 = d2 - r.u * d1;
        double a4;
        if constexpr (Mth::fast)
            a4 = (d3 + (r.a - r.v) * d1 - r.a * a3) * r.h2a;
        else
            a4 = (d3 + (r.a - r.v) * d1 - r.a * a3) / (2.0 * r.a);
        double a1 = d1 - a3 - a4;
        double wb[3][M], sb[3];
        wb[0][0] = a1;
        wb[0][MU] = a1 * r.u;
        wb[0][MV] = a1 * (r.v - r.a);
        wb[0][3] = a1 * (r.enth - r.v * r.a);
        sb[0] = r.v - r.a;
        wb[1][0] = a3;
        wb[1][MU] = a3 * r.u + a2;
        wb[1][MV] = a3 * r.v;
        wb[1][3] = a3 * 0.5 * r.u2v2 + a2 * r.u;
        sb[1] = r.v;
        wb[2][0] = a4;
        wb[2][MU] = a4 * r.u;
        wb[2][MV] = a4 * (r.v + r.a);
        wb[2][3] = a4 * (r.enth + r.v * r.a);
        sb[2] = r.v + r.a;
#pragma unroll
        for (int m = 0; m < M; m++) {
            double b0 = 0.0, b1 = 0.0;
#pragma unroll
            for (int mw = 0; mw < 3; mw++) {
                b0 = b0 + fmin(sb[mw], 0.0) * wb[mw][m];
                b1 = b1 + fmax(sb[mw], 0.0) * wb[mw][m];
            }
            bm[m] = b0;
            bp[m] = b1;
        }
    }
};

// Wave limiter phi(r) (src/2d/inlinelimiter.f90:71-97): 1 minmod,
// 2 superbee, 3 van Leer, 4 monotonized centred, 5 Beam-Warming.
template <class Mth>
__device__ static inline double philim(double r, int lim)
{
    switch (lim) {
    case 1: return fmax(0.0, fmin(1.0, r));
    case 2: return fmax(0.0, fmax(fmin(1.0, 2.0 * r), fmin(2.0, r)));
    case 3: return Mth::div(r + fabs(r), 1.0 + fabs(r));
    case 4: {
        double c = (1.0 + r) / 2.0;
        return fmax(0.0, fmin(c, fmin(2.0, 2.0 * r)));
    }
    default: return r;
    }
}

}  // namespace clawgpu
