// Device Riemann solvers for the 2D wave-propagation step (gfx950, FP64).
//
// These replace the rpn2/rpt2 plug-ins of the reference contract
// (src/2d/claw2.f:312-402; call sites src/2d/flux2.f90:110-111, :204-205,
// :220-221).  A GPU cannot call host Fortran procedures, so each solver is a
// compile-time policy selected by a solver id.  One call solves ONE interface
// between a left state qa (cell i-1) and a right state qb (cell i); the
// normal momentum sits in slot MU and the transverse one in slot MV (0-based),
// which is how the reference's ixy argument is expressed here.
//
// The arithmetic is written from LeVeque, "Finite Volume Methods for
// Hyperbolic Problems": constant-coefficient acoustics (ch. 18/21) and the
// Roe solver for the Euler equations with the Harten-Hyman entropy fix
// (sec. 15.3).  Expression order is fixed so that results are reproducible
// bit for bit by the CPU restatement that the tests compare against.
#pragma once

#include <hip/hip_runtime.h>

namespace clawgpu {

struct SolverParams {
    double rho, bulk, cc, zz;  // acoustics: density, bulk modulus, c, impedance
    double gamma, gamma1;      // Euler: gamma and gamma - 1
    int efix;                  // Euler: Harten-Hyman entropy fix
};

// Constant-coefficient acoustics, q = (p, u, v).  Waves r1 = (-Z, 1, 0) at
// speed -c and r2 = (Z, 1, 0) at speed +c along the normal direction.
struct Acoustics2 {
    static constexpr int M = 3;
    static constexpr int MW = 2;
    struct Cache {};

    template <int MU, int MV>
    __device__ static inline void rpn(const SolverParams& P, const double* qa, const double* qb,
                                      double w[MW][M], double s[MW], double am[M], double ap[M],
                                      Cache&)
    {
        const double zz = P.zz, cc = P.cc;
        double dp = qb[0] - qa[0];
        double du = qb[MU] - qa[MU];
        double a1 = (-dp + zz * du) / (2.0 * zz);
        double a2 = (dp + zz * du) / (2.0 * zz);
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
        double a1 = (-asdq[0] + zz * asdq[MV]) / (2.0 * zz);
        double a2 = (asdq[0] + zz * asdq[MV]) / (2.0 * zz);
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
struct Euler2 {
    static constexpr int M = 4;
    static constexpr int MW = 3;
    struct Cache {
        double u, v, enth, a, g1a2, euv, u2v2;
    };

    __device__ static inline double pressure(const double* q, double g1)
    {
        return g1 * (q[3] - 0.5 * (q[1] * q[1] + q[2] * q[2]) / q[0]);
    }

    template <int MU, int MV>
    __device__ static inline void roe(const SolverParams& P, const double* qa, const double* qb,
                                      Cache& r)
    {
        const double g1 = P.gamma1;
        double rhsqrtl = sqrt(qa[0]);
        double rhsqrtr = sqrt(qb[0]);
        double pl = pressure(qa, g1);
        double pr = pressure(qb, g1);
        double rhsq2 = rhsqrtl + rhsqrtr;
        r.u = (qa[MU] / rhsqrtl + qb[MU] / rhsqrtr) / rhsq2;
        r.v = (qa[MV] / rhsqrtl + qb[MV] / rhsqrtr) / rhsq2;
        r.enth = ((qa[3] + pl) / rhsqrtl + (qb[3] + pr) / rhsqrtr) / rhsq2;
        r.u2v2 = r.u * r.u + r.v * r.v;
        double a2 = g1 * (r.enth - 0.5 * r.u2v2);
        r.a = sqrt(a2);
        r.g1a2 = g1 / a2;
        r.euv = r.enth - r.u2v2;
    }

    template <int MU, int MV>
    __device__ static inline void rpn(const SolverParams& P, const double* qa, const double* qb,
                                      double w[MW][M], double s[MW], double am[M], double ap[M],
                                      Cache& r)
    {
        roe<MU, MV>(P, qa, qb, r);
        double d1 = qb[0] - qa[0];
        double d2 = qb[MU] - qa[MU];
        double d3 = qb[MV] - qa[MV];
        double d4 = qb[3] - qa[3];
        double a3 = r.g1a2 * (r.euv * d1 + r.u * d2 + r.v * d3 - d4);
        double a2 = d3 - r.v * d1;
        double a4 = (d2 + (r.a - r.u) * d1 - r.a * a3) / (2.0 * r.a);
        double a1 = d1 - a3 - a4;
        w[0][0] = a1;
        w[0][MU] = a1 * (r.u - r.a);
        w[0][MV] = a1 * r.v;
        w[0][3] = a1 * (r.enth - r.u * r.a);
        s[0] = r.u - r.a;
        w[1][0] = a3;
        w[1][MU] = a3 * r.u;
        w[1][MV] = a3 * r.v + a2;
        w[1][3] = a3 * 0.5 * r.u2v2 + a2 * r.v;
        s[1] = r.u;
        w[2][0] = a4;
        w[2][MU] = a4 * (r.u + r.a);
        w[2][MV] = a4 * r.v;
        w[2][3] = a4 * (r.enth + r.u * r.a);
        s[2] = r.u + r.a;

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
        const double gamma = P.gamma, g1 = P.gamma1;
        double amv[M] = {0.0, 0.0, 0.0, 0.0};
        do {
            double rhoa = qa[0];
            double pa = pressure(qa, g1);
            double ca = sqrt(gamma * pa / rhoa);
            double s0 = qa[MU] / rhoa - ca;
            if (s0 >= 0.0 && s[0] > 0.0) break;
            double qm[M];
#pragma unroll
            for (int m = 0; m < M; m++) qm[m] = qa[m] + w[0][m];
            double pm = pressure(qm, g1);
            double cm = sqrt(gamma * pm / qm[0]);
            double s1 = qm[MU] / qm[0] - cm;
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
            double rhob = qb[0];
            double pb = pressure(qb, g1);
            double cb = sqrt(gamma * pb / rhob);
            double s3 = qb[MU] / rhob + cb;
            double qn[M];
#pragma unroll
            for (int m = 0; m < M; m++) qn[m] = qb[m] - w[2][m];
            double pn = pressure(qn, g1);
            double cn = sqrt(gamma * pn / qn[0]);
            double s2 = qn[MU] / qn[0] + cn;
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
        double d1 = asdq[0], d2 = asdq[MU], d3 = asdq[MV], d4 = asdq[3];
        double a3 = r.g1a2 * (r.euv * d1 + r.u * d2 + r.v * d3 - d4);
        double a2 = d2 - r.u * d1;
        double a4 = (d3 + (r.a - r.v) * d1 - r.a * a3) / (2.0 * r.a);
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
__device__ static inline double philim(double r, int lim)
{
    switch (lim) {
    case 1: return fmax(0.0, fmin(1.0, r));
    case 2: return fmax(0.0, fmax(fmin(1.0, 2.0 * r), fmin(2.0, r)));
    case 3: return (r + fabs(r)) / (1.0 + fabs(r));
    case 4: {
        double c = (1.0 + r) / 2.0;
        return fmax(0.0, fmin(c, fmin(2.0, 2.0 * r)));
    }
    default: return r;
    }
}

}  // namespace clawgpu
