/*
 * oracle/claw2d_oracle.c -- CPU ORACLE for the 2D wave-propagation step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the
 * reference Fortran (clawpack/classic) 2D single-grid path, used by the
 * tests, by the smoke check and by the CPU-baseline leg of bench.py as the
 * checker.  It is never linked into, loaded by, or called from the product
 * library (classic_amd/).  It keeps the reference's slice structure and its
 * floating-point expression order on purpose, so that a build with
 * -ffp-contract=off reproduces the Fortran bit for bit.
 *
 * Functions and the reference code each one restates:
 *   orc_limiter      src/2d/inlinelimiter.f90:25-105
 *   orc_flux2        src/2d/flux2.f90:5-235
 *   orc_step2        src/2d/step2.f90:2-251 (x-sweeps step2_xslices :92-165,
 *                    y-sweeps step2_yslices :166-248; orc_step2_par shares
 *                    them among OpenMP threads by bands, same order per cell)
 *   orc_step2ds      src/2d/step2ds.f90:2-275
 *   orc_dimsp2       src/2d/dimsp2.f:4-98
 *   orc_bc2          src/2d/bc2.f:4-215
 *   orc_claw2        src/2d/claw2.f:494-784 (time loop, CFL retake, info codes)
 *   orc_annulus_aux  examples/advection_2d_annulus/setaux.f (test data)
 * Riemann solvers (the reference only fixes their contract,
 * src/2d/claw2.f:312-402; the arithmetic lives in the absent clawpack/riemann
 * repository, so these are written from LeVeque, "Finite Volume Methods for
 * Hyperbolic Problems" (FVMHP)):
 *   orc_rpn2_acoustics / orc_rpt2_acoustics   FVMHP 18.x / 21.x, constant rho, K
 *   orc_rpn2_euler / orc_rpt2_euler           Roe solver FVMHP 15.3 with the
 *                                             Harten-Hyman entropy fix 15.3.5
 *
 * Pinning (see tests/):
 *   - limiter/flux2/step2/step2ds/dimsp2/bc2/claw2: against the reference
 *     Fortran compiled from /root/reference (oracle/_ref) linked with these
 *     same solvers, bit for bit, and the committed golden vectors.
 *   - solvers: closed forms and invariants (sum of waves = jump, Roe property
 *     amdq+apdq = f(qr)-f(ql), rotational symmetry, Sod tube L1 convergence).
 *
 * Array layout follows the reference: q(meqn, 1-mbc:mx+mbc, 1-mbc:my+mbc),
 * column-major, m fastest.  Macros below take Fortran (1-based m) indices.
 */
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#include "claw2d_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- solvers */

static orc_solver_params g_par = {1.0, 4.0, 2.0, 2.0, 1.4, 1};

void orc_set_solver_params(const orc_solver_params* p)
{
    g_par = *p;
}

/* slice accessors: a(m, i) with a(meqn, 1-mbc:maxm+mbc) */
#define S1(a, m, i) (a)[((i) + mbc - 1) * meqn + ((m) - 1)]
#define SW(a, m, mw, i) (a)[(((i) + mbc - 1) * mwaves + ((mw) - 1)) * meqn + ((m) - 1)]
#define SS(a, mw, i) (a)[((i) + mbc - 1) * mwaves + ((mw) - 1)]

/* Acoustics, q = (p, u, v), constant rho and bulk modulus K (c = sqrt(K/rho),
 * Z = rho c).  Eigenvectors of the normal Jacobian: r1 = (-Z, 1, 0) with
 * speed -c, r2 = (Z, 1, 0) with speed +c (normal velocity in slot mu). */
void orc_rpn2_acoustics(int ixy, int maxm, int meqn, int mwaves, int maux, int mbc,
                        int mx, const double* ql, const double* qr, const double* auxl,
                        const double* auxr, double* wave, double* s, double* amdq,
                        double* apdq)
{
    (void)maxm; (void)maux; (void)auxl; (void)auxr;
    const double cc = g_par.cc, zz = g_par.zz;
    const int mu = (ixy == 1) ? 2 : 3;
    const int mv = (ixy == 1) ? 3 : 2;
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        double dp = S1(ql, 1, i) - S1(qr, 1, i - 1);
        double du = S1(ql, mu, i) - S1(qr, mu, i - 1);
        double a1 = (-dp + zz * du) / (2.0 * zz);
        double a2 = (dp + zz * du) / (2.0 * zz);
        SW(wave, 1, 1, i) = -a1 * zz;
        SW(wave, mu, 1, i) = a1;
        SW(wave, mv, 1, i) = 0.0;
        SS(s, 1, i) = -cc;
        SW(wave, 1, 2, i) = a2 * zz;
        SW(wave, mu, 2, i) = a2;
        SW(wave, mv, 2, i) = 0.0;
        SS(s, 2, i) = cc;
        for (int m = 1; m <= meqn; m++) {
            S1(amdq, m, i) = SS(s, 1, i) * SW(wave, m, 1, i);
            S1(apdq, m, i) = SS(s, 2, i) * SW(wave, m, 2, i);
        }
    }
}

/* Transverse split of asdq with the eigenvectors of the transverse Jacobian:
 * r1 = (-Z, 0, 1) speed -c, r2 = (Z, 0, 1) speed +c (transverse velocity mv). */
void orc_rpt2_acoustics(int ixy, int imp, int maxm, int meqn, int mwaves, int maux,
                        int mbc, int mx, const double* ql, const double* qr,
                        const double* aux1, const double* aux2, const double* aux3,
                        const double* asdq, double* bmasdq, double* bpasdq)
{
    (void)imp; (void)maxm; (void)mwaves; (void)maux; (void)ql; (void)qr;
    (void)aux1; (void)aux2; (void)aux3;
    const double cc = g_par.cc, zz = g_par.zz;
    const int mu = (ixy == 1) ? 2 : 3;
    const int mv = (ixy == 1) ? 3 : 2;
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        double a1 = (-S1(asdq, 1, i) + zz * S1(asdq, mv, i)) / (2.0 * zz);
        double a2 = (S1(asdq, 1, i) + zz * S1(asdq, mv, i)) / (2.0 * zz);
        S1(bmasdq, 1, i) = cc * a1 * zz;
        S1(bmasdq, mu, i) = 0.0;
        S1(bmasdq, mv, i) = -cc * a1;
        S1(bpasdq, 1, i) = cc * a2 * zz;
        S1(bpasdq, mu, i) = 0.0;
        S1(bpasdq, mv, i) = cc * a2;
    }
}

/* Euler, q = (rho, rho u, rho v, E), ideal gas.  Pressure always sums the
 * squares of components 2 then 3, so it does not depend on the sweep. */
static double euler_pressure(const double* q, double gamma1)
{
    return gamma1 * (q[3] - 0.5 * (q[1] * q[1] + q[2] * q[2]) / q[0]);
}

typedef struct {
    double u, v, enth, a, g1a2, euv, u2v2;
} roe_avg;

/* Roe averages between left state qa (cell i-1) and right state qb (cell i);
 * u is the velocity in slot mu, v in slot mv (FVMHP 15.3.2). */
static roe_avg euler_roe(const double* qa, const double* qb, int mu, int mv, double gamma1)
{
    roe_avg r;
    double rhsqrtl = sqrt(qa[0]);
    double rhsqrtr = sqrt(qb[0]);
    double pl = euler_pressure(qa, gamma1);
    double pr = euler_pressure(qb, gamma1);
    double rhsq2 = rhsqrtl + rhsqrtr;
    r.u = (qa[mu - 1] / rhsqrtl + qb[mu - 1] / rhsqrtr) / rhsq2;
    r.v = (qa[mv - 1] / rhsqrtl + qb[mv - 1] / rhsqrtr) / rhsq2;
    r.enth = ((qa[3] + pl) / rhsqrtl + (qb[3] + pr) / rhsqrtr) / rhsq2;
    r.u2v2 = r.u * r.u + r.v * r.v;
    double a2 = gamma1 * (r.enth - 0.5 * r.u2v2);
    r.a = sqrt(a2);
    r.g1a2 = gamma1 / a2;
    r.euv = r.enth - r.u2v2;
    return r;
}

void orc_rpn2_euler(int ixy, int maxm, int meqn, int mwaves, int maux, int mbc, int mx,
                    const double* ql, const double* qr, const double* auxl,
                    const double* auxr, double* wave, double* s, double* amdq,
                    double* apdq)
{
    (void)maxm; (void)maux; (void)auxl; (void)auxr;
    const double gamma = g_par.gamma, gamma1 = g_par.gamma - 1.0;
    const int mu = (ixy == 1) ? 2 : 3;
    const int mv = (ixy == 1) ? 3 : 2;
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        const double* qa = &S1(qr, 1, i - 1); /* left state  */
        const double* qb = &S1(ql, 1, i);     /* right state */
        roe_avg r = euler_roe(qa, qb, mu, mv, gamma1);
        double d1 = qb[0] - qa[0];
        double d2 = qb[mu - 1] - qa[mu - 1];
        double d3 = qb[mv - 1] - qa[mv - 1];
        double d4 = qb[3] - qa[3];
        /* strengths: entropy (a3), shear (a2), acoustic (a1, a4) */
        double a3 = r.g1a2 * (r.euv * d1 + r.u * d2 + r.v * d3 - d4);
        double a2 = d3 - r.v * d1;
        double a4 = (d2 + (r.a - r.u) * d1 - r.a * a3) / (2.0 * r.a);
        double a1 = d1 - a3 - a4;
        /* 1-wave: u-a */
        SW(wave, 1, 1, i) = a1;
        SW(wave, mu, 1, i) = a1 * (r.u - r.a);
        SW(wave, mv, 1, i) = a1 * r.v;
        SW(wave, 4, 1, i) = a1 * (r.enth - r.u * r.a);
        SS(s, 1, i) = r.u - r.a;
        /* 2-wave: entropy + shear, both at speed u */
        SW(wave, 1, 2, i) = a3;
        SW(wave, mu, 2, i) = a3 * r.u;
        SW(wave, mv, 2, i) = a3 * r.v + a2;
        SW(wave, 4, 2, i) = a3 * 0.5 * r.u2v2 + a2 * r.v;
        SS(s, 2, i) = r.u;
        /* 3-wave: u+a */
        SW(wave, 1, 3, i) = a4;
        SW(wave, mu, 3, i) = a4 * (r.u + r.a);
        SW(wave, mv, 3, i) = a4 * r.v;
        SW(wave, 4, 3, i) = a4 * (r.enth + r.u * r.a);
        SS(s, 3, i) = r.u + r.a;
    }

    if (!g_par.efix) {
        for (int i = 2 - mbc; i <= mx + mbc; i++) {
            for (int m = 1; m <= meqn; m++) {
                double am = 0.0, ap = 0.0;
                for (int mw = 1; mw <= mwaves; mw++) {
                    if (SS(s, mw, i) < 0.0)
                        am = am + SS(s, mw, i) * SW(wave, m, mw, i);
                    else
                        ap = ap + SS(s, mw, i) * SW(wave, m, mw, i);
                }
                S1(amdq, m, i) = am;
                S1(apdq, m, i) = ap;
            }
        }
        return;
    }

    /* Harten-Hyman entropy fix: amdq from the left-going parts, with a
     * transonic rarefaction split in the 1- or 3-family; apdq = df - amdq. */
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        const double* qa = &S1(qr, 1, i - 1);
        const double* qb = &S1(ql, 1, i);
        double am[4] = {0.0, 0.0, 0.0, 0.0};
        do {
            double rhoa = qa[0];
            double pa = euler_pressure(qa, gamma1);
            double ca = sqrt(gamma * pa / rhoa);
            double s0 = qa[mu - 1] / rhoa - ca;
            if (s0 >= 0.0 && SS(s, 1, i) > 0.0)
                break; /* fully supersonic to the right: amdq = 0 */
            double qm[4];
            for (int m = 1; m <= 4; m++) qm[m - 1] = qa[m - 1] + SW(wave, m, 1, i);
            double pm = euler_pressure(qm, gamma1);
            double cm = sqrt(gamma * pm / qm[0]);
            double s1 = qm[mu - 1] / qm[0] - cm;
            double sfract;
            if (s0 < 0.0 && s1 > 0.0)
                sfract = s0 * (s1 - SS(s, 1, i)) / (s1 - s0);
            else if (SS(s, 1, i) < 0.0)
                sfract = SS(s, 1, i);
            else
                sfract = 0.0;
            for (int m = 1; m <= 4; m++) am[m - 1] = sfract * SW(wave, m, 1, i);
            if (SS(s, 2, i) >= 0.0)
                break;
            for (int m = 1; m <= 4; m++) am[m - 1] = am[m - 1] + SS(s, 2, i) * SW(wave, m, 2, i);
            double rhob = qb[0];
            double pb = euler_pressure(qb, gamma1);
            double cb = sqrt(gamma * pb / rhob);
            double s3 = qb[mu - 1] / rhob + cb;
            double qn[4];
            for (int m = 1; m <= 4; m++) qn[m - 1] = qb[m - 1] - SW(wave, m, 3, i);
            double pn = euler_pressure(qn, gamma1);
            double cn = sqrt(gamma * pn / qn[0]);
            double s2 = qn[mu - 1] / qn[0] + cn;
            if (s2 < 0.0 && s3 > 0.0)
                sfract = s2 * (s3 - SS(s, 3, i)) / (s3 - s2);
            else if (SS(s, 3, i) < 0.0)
                sfract = SS(s, 3, i);
            else
                break;
            for (int m = 1; m <= 4; m++) am[m - 1] = am[m - 1] + sfract * SW(wave, m, 3, i);
        } while (0);
        for (int m = 1; m <= meqn; m++) {
            double df = 0.0;
            for (int mw = 1; mw <= mwaves; mw++) df = df + SS(s, mw, i) * SW(wave, m, mw, i);
            S1(amdq, m, i) = am[m - 1];
            S1(apdq, m, i) = df - am[m - 1];
        }
    }
}

void orc_rpt2_euler(int ixy, int imp, int maxm, int meqn, int mwaves, int maux, int mbc,
                    int mx, const double* ql, const double* qr, const double* aux1,
                    const double* aux2, const double* aux3, const double* asdq,
                    double* bmasdq, double* bpasdq)
{
    (void)imp; (void)maxm; (void)mwaves; (void)maux; (void)aux1; (void)aux2; (void)aux3;
    const double gamma1 = g_par.gamma - 1.0;
    const int mu = (ixy == 1) ? 2 : 3;
    const int mv = (ixy == 1) ? 3 : 2;
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        const double* qa = &S1(qr, 1, i - 1);
        const double* qb = &S1(ql, 1, i);
        roe_avg r = euler_roe(qa, qb, mu, mv, gamma1);
        double d1 = S1(asdq, 1, i);
        double d2 = S1(asdq, mu, i);
        double d3 = S1(asdq, mv, i);
        double d4 = S1(asdq, 4, i);
        double a3 = r.g1a2 * (r.euv * d1 + r.u * d2 + r.v * d3 - d4);
        double a2 = d2 - r.u * d1;
        double a4 = (d3 + (r.a - r.v) * d1 - r.a * a3) / (2.0 * r.a);
        double a1 = d1 - a3 - a4;
        double wb[3][4], sb[3];
        wb[0][0] = a1;
        wb[0][mu - 1] = a1 * r.u;
        wb[0][mv - 1] = a1 * (r.v - r.a);
        wb[0][3] = a1 * (r.enth - r.v * r.a);
        sb[0] = r.v - r.a;
        wb[1][0] = a3;
        wb[1][mu - 1] = a3 * r.u + a2;
        wb[1][mv - 1] = a3 * r.v;
        wb[1][3] = a3 * 0.5 * r.u2v2 + a2 * r.u;
        sb[1] = r.v;
        wb[2][0] = a4;
        wb[2][mu - 1] = a4 * r.u;
        wb[2][mv - 1] = a4 * (r.v + r.a);
        wb[2][3] = a4 * (r.enth + r.v * r.a);
        sb[2] = r.v + r.a;
        for (int m = 1; m <= meqn; m++) {
            double bm = 0.0, bp = 0.0;
            for (int mw = 0; mw < 3; mw++) {
                bm = bm + fmin(sb[mw], 0.0) * wb[mw][m - 1];
                bp = bp + fmax(sb[mw], 0.0) * wb[mw][m - 1];
            }
            S1(bmasdq, m, i) = bm;
            S1(bpasdq, m, i) = bp;
        }
    }
}

/* Variable-coefficient advection q_t + (u q)_x + (v q)_y = 0 with edge
 * velocities in aux: aux(1,i,j) = u at the left edge of cell (i,j),
 * aux(2,i,j) = v at its bottom edge (rpn2/rpt2_vc_advection of the absent
 * clawpack/riemann, named by examples/advection_2d_annulus/Makefile; FVMHP
 * 9.5 and 20.5).  The interface between cells i-1 and i moves with the edge
 * velocity of cell i; a transverse fluctuation in cell i1 (i-1 for imp = 1,
 * i for imp = 2) goes down with the bottom-edge velocity of that cell (aux2)
 * and up with the bottom-edge velocity of the cell above it (aux3). */
void orc_rpn2_vc_advection(int ixy, int maxm, int meqn, int mwaves, int maux, int mbc, int mx,
                           const double* ql, const double* qr, const double* auxl,
                           const double* auxr, double* wave, double* s, double* amdq,
                           double* apdq)
{
    (void)maxm; (void)meqn; (void)mwaves; (void)auxr;
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        double w = ql[(size_t)(i + mbc - 1)] - qr[(size_t)(i - 1 + mbc - 1)];
        double sp = auxl[(size_t)(i + mbc - 1) * maux + (ixy - 1)];
        wave[(size_t)(i + mbc - 1)] = w;
        s[(size_t)(i + mbc - 1)] = sp;
        amdq[(size_t)(i + mbc - 1)] = fmin(sp, 0.0) * w;
        apdq[(size_t)(i + mbc - 1)] = fmax(sp, 0.0) * w;
    }
}

void orc_rpt2_vc_advection(int ixy, int imp, int maxm, int meqn, int mwaves, int maux, int mbc,
                           int mx, const double* ql, const double* qr, const double* aux1,
                           const double* aux2, const double* aux3, const double* asdq,
                           double* bmasdq, double* bpasdq)
{
    (void)maxm; (void)meqn; (void)mwaves; (void)ql; (void)qr; (void)aux1;
    const int kv = 3 - ixy;
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        int i1 = i - 2 + imp;
        double a = asdq[(size_t)(i + mbc - 1)];
        bmasdq[(size_t)(i + mbc - 1)] = fmin(aux2[(size_t)(i1 + mbc - 1) * maux + (kv - 1)], 0.0) * a;
        bpasdq[(size_t)(i + mbc - 1)] = fmax(aux3[(size_t)(i1 + mbc - 1) * maux + (kv - 1)], 0.0) * a;
    }
}

orc_rpn2_t orc_get_rpn2(int solver_id)
{
    switch (solver_id) {
    case ORC_SOLVER_ACOUSTICS: return orc_rpn2_acoustics;
    case ORC_SOLVER_EULER: return orc_rpn2_euler;
    case ORC_SOLVER_VC_ADVECTION: return orc_rpn2_vc_advection;
    default: return NULL;
    }
}

orc_rpt2_t orc_get_rpt2(int solver_id)
{
    switch (solver_id) {
    case ORC_SOLVER_ACOUSTICS: return orc_rpt2_acoustics;
    case ORC_SOLVER_EULER: return orc_rpt2_euler;
    case ORC_SOLVER_VC_ADVECTION: return orc_rpt2_vc_advection;
    default: return NULL;
    }
}

/* examples/advection_2d_annulus/setaux.f with mapc2p.f and stream.f:
 * polar map (r, theta) -> (r cos theta, r sin theta), stream function
 * psi = pi (x^2 + y^2); aux(1) = (psi(top-left) - psi(bottom-left)) / dyc,
 * aux(2) = -(psi(bottom-right) - psi(bottom-left)) / dxc, aux(3) = physical
 * area / (dxc dyc) by the shoelace formula.  Used by the tests to build the
 * example's aux array exactly as the reference's setaux does. */
static double annulus_stream(double x, double y) { return 3.1415926535897931 * (x * x + y * y); }

void orc_annulus_aux(int mbc, int mx, int my, double xlower, double ylower, double dxc,
                     double dyc, double* aux)
{
    for (int j = 1 - mbc; j <= my + mbc; j++)
        for (int i = 1 - mbc; i <= mx + mbc; i++) {
            double xc[5], yc[5], xp[5], yp[5];
            xc[0] = xlower + (i - 1) * dxc;
            yc[0] = ylower + (j - 1) * dyc;
            xc[1] = xc[0];
            yc[1] = yc[0] + dyc;
            xc[2] = xc[0] + dxc;
            yc[2] = yc[0] + dyc;
            xc[3] = xc[0] + dxc;
            yc[3] = yc[0];
            for (int k = 0; k < 4; k++) {
                xp[k] = xc[k] * cos(yc[k]);
                yp[k] = xc[k] * sin(yc[k]);
            }
            double* a = aux + ((size_t)(j + mbc - 1) * (mx + 2 * mbc) + (i + mbc - 1)) * 3;
            a[0] = (annulus_stream(xp[1], yp[1]) - annulus_stream(xp[0], yp[0])) / dyc;
            a[1] = -(annulus_stream(xp[3], yp[3]) - annulus_stream(xp[0], yp[0])) / dxc;
            xp[4] = xp[0];
            yp[4] = yp[0];
            double area = 0.0;
            for (int ic = 0; ic < 4; ic++)
                area = area + 0.5 * (yp[ic] + yp[ic + 1]) * (xp[ic + 1] - xp[ic]);
            a[2] = area / (dxc * dyc);
        }
}

/* ---------------------------------------------------------------- limiter */

/* inlinelimiter.f90:25-105.  Dot products use unlimited neighbours: dotr of
 * interface i is formed before wave i is scaled, and wave i+1 is scaled only
 * in the next iteration. */
int orc_limiter(int maxm, int meqn, int mwaves, int mbc, int mx, double* wave,
                const double* s, const int* mthlim)
{
    (void)maxm;
    double dotr[16];
    for (int mw = 0; mw < mwaves; mw++) dotr[mw] = 0.0;
    for (int i = 0; i <= mx + 1; i++) {
        for (int mw = 1; mw <= mwaves; mw++) {
            if (mthlim[mw - 1] == 0) continue;
            double wnorm2 = 0.0;
            double dotl = dotr[mw - 1];
            dotr[mw - 1] = 0.0;
            for (int m = 1; m <= meqn; m++) {
                wnorm2 = wnorm2 + SW(wave, m, mw, i) * SW(wave, m, mw, i);
                dotr[mw - 1] = dotr[mw - 1] + SW(wave, m, mw, i) * SW(wave, m, mw, i + 1);
            }
            if (i == 0) continue;
            if (wnorm2 == 0.0) continue;
            double r;
            if (SS(s, mw, i) > 0.0)
                r = dotl / wnorm2;
            else
                r = dotr[mw - 1] / wnorm2;
            double wl;
            switch (mthlim[mw - 1]) {
            case 1: wl = fmax(0.0, fmin(1.0, r)); break;
            case 2: wl = fmax(0.0, fmax(fmin(1.0, 2.0 * r), fmin(2.0, r))); break;
            case 3: wl = (r + fabs(r)) / (1.0 + fabs(r)); break;
            case 4: {
                double c = (1.0 + r) / 2.0;
                wl = fmax(0.0, fmin(c, fmin(2.0, 2.0 * r)));
                break;
            }
            case 5: wl = r; break;
            default: return -1; /* reference stops: "Invalid limiter method." */
            }
            for (int m = 1; m <= meqn; m++) SW(wave, m, mw, i) = wl * SW(wave, m, mw, i);
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ flux2 */

typedef struct {
    double *wave, *s, *amdq, *apdq, *cqxx, *bmasdq, *bpasdq;
} flux_work;

/* flux2.f90:5-235.  fadd/gadd are zeroed by the caller's work array on the
 * reference side (claw2ez allocates fresh memory); here the caller passes
 * zero-initialised scratch, so order-1 and no-transverse runs read zeros. */
static int orc_flux2(int ixy, int maxm, int meqn, int mwaves, int maux, int mbc, int mx,
                     const double* q1d, double* dtdx1d, const double* aux1,
                     const double* aux2, const double* aux3, const int* method,
                     const int* mthlim, double* qadd, double* fadd, double* gadd,
                     double* cfl1d, flux_work* w, orc_rpn2_t rpn2, orc_rpt2_t rpt2,
                     int use_fwave)
{
    double *wave = w->wave, *s = w->s, *amdq = w->amdq, *apdq = w->apdq;
    double *cqxx = w->cqxx, *bmasdq = w->bmasdq, *bpasdq = w->bpasdq;
    int limit = 0;
    for (int mw = 0; mw < mwaves; mw++)
        if (mthlim[mw] > 0) limit = 1;
#define DT(i) dtdx1d[(i) + mbc - 1]
#define G(a, m, k, i) (a)[(((i) + mbc - 1) * 2 + ((k) - 1)) * meqn + ((m) - 1)]

    for (int i = 1 - mbc; i <= mx + mbc; i++)
        for (int m = 1; m <= meqn; m++) S1(qadd, m, i) = 0.0;

    rpn2(ixy, maxm, meqn, mwaves, maux, mbc, mx, q1d, q1d, aux2, aux2, wave, s, amdq, apdq);

    for (int i = 1; i <= mx + 1; i++) {
        for (int m = 1; m <= meqn; m++)
            S1(qadd, m, i - 1) = S1(qadd, m, i - 1) - DT(i - 1) * S1(amdq, m, i);
        for (int m = 1; m <= meqn; m++)
            S1(qadd, m, i) = S1(qadd, m, i) - DT(i) * S1(apdq, m, i);
    }

    double cfl = 0.0;
    int nan_speed = 0;
    for (int i = 1; i <= mx + 1; i++)
        for (int mw = 1; mw <= mwaves; mw++) {
            cfl = fmax(cfl, fmax(DT(i) * SS(s, mw, i), -DT(i - 1) * SS(s, mw, i)));
            if (SS(s, mw, i) != SS(s, mw, i)) nan_speed = 1;
        }
    /* Not in the reference: a NaN speed is reported as a NaN Courant number
     * (fmax would drop it) so that the time loop can stop, see orc_claw2. */
    *cfl1d = nan_speed ? NAN : cfl;

    if (method[1] != 1) {
        if (limit)
            if (orc_limiter(maxm, meqn, mwaves, mbc, mx, wave, s, mthlim) != 0) return -1;
        for (int i = 2 - mbc; i <= mx + mbc; i++) {
            double dtdxave = 0.5 * (DT(i - 1) + DT(i));
            for (int m = 1; m <= meqn; m++) S1(cqxx, m, i) = 0.0;
            for (int mw = 1; mw <= mwaves; mw++)
                for (int m = 1; m <= meqn; m++) {
                    double sw = use_fwave ? copysign(1.0, SS(s, mw, i)) : fabs(SS(s, mw, i));
                    S1(cqxx, m, i) = S1(cqxx, m, i) +
                                     sw * (1.0 - fabs(SS(s, mw, i)) * dtdxave) * SW(wave, m, mw, i);
                }
            for (int m = 1; m <= meqn; m++) S1(fadd, m, i) = 0.5 * S1(cqxx, m, i);
        }
    }

    if (method[2] <= 0) return 0;

    if (method[1] > 1 && method[2] == 2) {
        for (int i = 1; i <= mx + 1; i++)
            for (int m = 1; m <= meqn; m++) {
                S1(amdq, m, i) = S1(amdq, m, i) + S1(cqxx, m, i);
                S1(apdq, m, i) = S1(apdq, m, i) - S1(cqxx, m, i);
            }
    }

    rpt2(ixy, 1, maxm, meqn, mwaves, maux, mbc, mx, q1d, q1d, aux1, aux2, aux3, amdq,
         bmasdq, bpasdq);
    for (int i = 1; i <= mx + 1; i++) {
        for (int m = 1; m <= meqn; m++) G(gadd, m, 1, i - 1) = -0.5 * DT(i - 1) * S1(bmasdq, m, i);
        for (int m = 1; m <= meqn; m++) G(gadd, m, 2, i - 1) = -0.5 * DT(i - 1) * S1(bpasdq, m, i);
    }
    rpt2(ixy, 2, maxm, meqn, mwaves, maux, mbc, mx, q1d, q1d, aux1, aux2, aux3, apdq,
         bmasdq, bpasdq);
    for (int i = 1; i <= mx + 1; i++) {
        for (int m = 1; m <= meqn; m++)
            G(gadd, m, 1, i) = G(gadd, m, 1, i) - 0.5 * DT(i) * S1(bmasdq, m, i);
        for (int m = 1; m <= meqn; m++)
            G(gadd, m, 2, i) = G(gadd, m, 2, i) - 0.5 * DT(i) * S1(bpasdq, m, i);
    }
    return 0;
#undef DT
#undef G
}

/* ---------------------------------------------------- step2 / step2ds */

typedef struct {
    double *qadd, *fadd, *gadd, *q1d, *dtdx1d, *dtdy1d, *aux1, *aux2, *aux3;
    flux_work fw;
    void* block;
} step_work;

static int step_work_alloc(step_work* sw, int maxm, int meqn, int mwaves, int maux, int mbc)
{
    size_t n = (size_t)(maxm + 2 * mbc);
    int ma = maux > 0 ? maux : 1;
    size_t tot = n * (size_t)(meqn * 3 + 2 * meqn + meqn + 2 + 3 * ma +
                              meqn * mwaves + mwaves + 5 * meqn);
    double* b = (double*)calloc(tot, sizeof(double));
    if (!b) return -1;
    sw->block = b;
    sw->qadd = b; b += n * meqn;
    sw->fadd = b; b += n * meqn;
    sw->gadd = b; b += 2 * n * meqn;
    sw->q1d = b; b += n * meqn;
    sw->dtdx1d = b; b += n;
    sw->dtdy1d = b; b += n;
    sw->aux1 = b; b += n * ma;
    sw->aux2 = b; b += n * ma;
    sw->aux3 = b; b += n * ma;
    sw->fw.wave = b; b += n * meqn * mwaves;
    sw->fw.s = b; b += n * mwaves;
    sw->fw.amdq = b; b += n * meqn;
    sw->fw.apdq = b; b += n * meqn;
    sw->fw.cqxx = b; b += n * meqn;
    sw->fw.bmasdq = b; b += n * meqn;
    sw->fw.bpasdq = b; b += n * meqn;
    return 0;
}

#define Q2(a, m, i, j) (a)[((size_t)((j) + mbc - 1) * (mx + 2 * mbc) + ((i) + mbc - 1)) * meqn + ((m) - 1)]
#define A2(a, m, i, j) (a)[((size_t)((j) + mbc - 1) * (mx + 2 * mbc) + ((i) + mbc - 1)) * maux + ((m) - 1)]

/* step2.f90:2-251, the x-sweeps (:92-165): slices j = jlo..jhi, with the
 * reference's per-cell update expressions and their parenthesisation.  Only
 * updates that land in rows rlo..rhi are applied (the serial step passes the
 * whole range); a cell therefore receives the terms of slices j-1, j, j+1 in
 * that order whatever the band, which keeps the threaded step below equal to
 * the serial one bit for bit. */
#define G(a, m, k, i) (a)[(((i) + mbc - 1) * 2 + ((k) - 1)) * meqn + ((m) - 1)]
#define D1(a, i) (a)[(i) + mbc - 1]
#define INB(r, lo, hi) ((r) >= (lo) && (r) <= (hi))
static int step2_xslices(int maxm, int meqn, int mwaves, int maux, int mbc, int mx, int my,
                         const double* qold, double* qnew, const double* aux, double dtdx,
                         double dtdy, const int* method, const int* mthlim, int use_fwave,
                         orc_rpn2_t rpn2, orc_rpt2_t rpt2, step_work* sw, int jlo, int jhi,
                         int rlo, int rhi, double* cfl)
{
    const int index_capa = method[5];
    double cfl1d;
    double *qadd = sw->qadd, *fadd = sw->fadd, *gadd = sw->gadd, *q1d = sw->q1d;
    double* dtdx1d = sw->dtdx1d;
    (void)my;
    for (int j = jlo; j <= jhi; j++) {
        for (int i = 1 - mbc; i <= mx + mbc; i++)
            for (int m = 1; m <= meqn; m++) S1(q1d, m, i) = Q2(qold, m, i, j);
        if (index_capa > 0)
            for (int i = 1 - mbc; i <= mx + mbc; i++)
                D1(dtdx1d, i) = dtdx / A2(aux, index_capa, i, j);
        if (maux > 0)
            for (int i = 1 - mbc; i <= mx + mbc; i++)
                for (int ma = 1; ma <= maux; ma++) {
                    sw->aux1[(i + mbc - 1) * maux + ma - 1] = A2(aux, ma, i, j - 1);
                    sw->aux2[(i + mbc - 1) * maux + ma - 1] = A2(aux, ma, i, j);
                    sw->aux3[(i + mbc - 1) * maux + ma - 1] = A2(aux, ma, i, j + 1);
                }
        if (orc_flux2(1, maxm, meqn, mwaves, maux, mbc, mx, q1d, dtdx1d, sw->aux1, sw->aux2,
                      sw->aux3, method, mthlim, qadd, fadd, gadd, &cfl1d, &sw->fw, rpn2, rpt2,
                      use_fwave) != 0)
            return -1;
        *cfl = (cfl1d != cfl1d || *cfl != *cfl) ? NAN : fmax(*cfl, cfl1d);
        const int own0 = INB(j, rlo, rhi), ownm = INB(j - 1, rlo, rhi), ownp = INB(j + 1, rlo, rhi);
        if (index_capa == 0) {
            for (int i = 1; i <= mx; i++)
                for (int m = 1; m <= meqn; m++) {
                    if (own0)
                        Q2(qnew, m, i, j) = Q2(qnew, m, i, j) + S1(qadd, m, i) -
                                            dtdx * (S1(fadd, m, i + 1) - S1(fadd, m, i)) -
                                            dtdy * (G(gadd, m, 2, i) - G(gadd, m, 1, i));
                    if (ownm) Q2(qnew, m, i, j - 1) = Q2(qnew, m, i, j - 1) - dtdy * G(gadd, m, 1, i);
                    if (ownp) Q2(qnew, m, i, j + 1) = Q2(qnew, m, i, j + 1) + dtdy * G(gadd, m, 2, i);
                }
        } else {
            for (int i = 1; i <= mx; i++)
                for (int m = 1; m <= meqn; m++) {
                    if (own0)
                        Q2(qnew, m, i, j) = Q2(qnew, m, i, j) + S1(qadd, m, i) -
                                            (dtdx * (S1(fadd, m, i + 1) - S1(fadd, m, i)) +
                                             dtdy * (G(gadd, m, 2, i) - G(gadd, m, 1, i))) /
                                                A2(aux, index_capa, i, j);
                    if (ownm)
                        Q2(qnew, m, i, j - 1) = Q2(qnew, m, i, j - 1) -
                                                dtdy * G(gadd, m, 1, i) / A2(aux, index_capa, i, j - 1);
                    if (ownp)
                        Q2(qnew, m, i, j + 1) = Q2(qnew, m, i, j + 1) +
                                                dtdy * G(gadd, m, 2, i) / A2(aux, index_capa, i, j + 1);
                }
        }
    }
    return 0;
}

/* step2.f90:166-248, the y-sweeps: slices i = ilo..ihi, updates applied to
 * columns clo..chi only (as step2_xslices). */
static int step2_yslices(int maxm, int meqn, int mwaves, int maux, int mbc, int mx, int my,
                         const double* qold, double* qnew, const double* aux, double dtdx,
                         double dtdy, const int* method, const int* mthlim, int use_fwave,
                         orc_rpn2_t rpn2, orc_rpt2_t rpt2, step_work* sw, int ilo, int ihi,
                         int clo, int chi, double* cfl)
{
    const int index_capa = method[5];
    double cfl1d;
    double *qadd = sw->qadd, *fadd = sw->fadd, *gadd = sw->gadd, *q1d = sw->q1d;
    double* dtdy1d = sw->dtdy1d;
    for (int i = ilo; i <= ihi; i++) {
        for (int j = 1 - mbc; j <= my + mbc; j++)
            for (int m = 1; m <= meqn; m++) S1(q1d, m, j) = Q2(qold, m, i, j);
        if (index_capa > 0)
            for (int j = 1 - mbc; j <= my + mbc; j++)
                D1(dtdy1d, j) = dtdy / A2(aux, index_capa, i, j);
        if (maux > 0)
            for (int j = 1 - mbc; j <= my + mbc; j++)
                for (int ma = 1; ma <= maux; ma++) {
                    sw->aux1[(j + mbc - 1) * maux + ma - 1] = A2(aux, ma, i - 1, j);
                    sw->aux2[(j + mbc - 1) * maux + ma - 1] = A2(aux, ma, i, j);
                    sw->aux3[(j + mbc - 1) * maux + ma - 1] = A2(aux, ma, i + 1, j);
                }
        if (orc_flux2(2, maxm, meqn, mwaves, maux, mbc, my, q1d, dtdy1d, sw->aux1, sw->aux2,
                      sw->aux3, method, mthlim, qadd, fadd, gadd, &cfl1d, &sw->fw, rpn2, rpt2,
                      use_fwave) != 0)
            return -1;
        *cfl = (cfl1d != cfl1d || *cfl != *cfl) ? NAN : fmax(*cfl, cfl1d);
        const int own0 = INB(i, clo, chi), ownm = INB(i - 1, clo, chi), ownp = INB(i + 1, clo, chi);
        if (index_capa == 0) {
            for (int j = 1; j <= my; j++)
                for (int m = 1; m <= meqn; m++) {
                    if (own0)
                        Q2(qnew, m, i, j) = Q2(qnew, m, i, j) +
                                            (S1(qadd, m, j) - dtdy * (S1(fadd, m, j + 1) - S1(fadd, m, j)) -
                                             dtdx * (G(gadd, m, 2, j) - G(gadd, m, 1, j)));
                    if (ownm) Q2(qnew, m, i - 1, j) = Q2(qnew, m, i - 1, j) - dtdx * G(gadd, m, 1, j);
                    if (ownp) Q2(qnew, m, i + 1, j) = Q2(qnew, m, i + 1, j) + dtdx * G(gadd, m, 2, j);
                }
        } else {
            for (int j = 1; j <= my; j++)
                for (int m = 1; m <= meqn; m++) {
                    if (own0)
                        Q2(qnew, m, i, j) = Q2(qnew, m, i, j) + S1(qadd, m, j) -
                                            (dtdy * (S1(fadd, m, j + 1) - S1(fadd, m, j)) +
                                             dtdx * (G(gadd, m, 2, j) - G(gadd, m, 1, j))) /
                                                A2(aux, index_capa, i, j);
                    if (ownm)
                        Q2(qnew, m, i - 1, j) = Q2(qnew, m, i - 1, j) -
                                                dtdx * G(gadd, m, 1, j) / A2(aux, index_capa, i - 1, j);
                    if (ownp)
                        Q2(qnew, m, i + 1, j) = Q2(qnew, m, i + 1, j) +
                                                dtdx * G(gadd, m, 2, j) / A2(aux, index_capa, i + 1, j);
                }
        }
    }
    return 0;
}

/* without a capacity function dtdx1d/dtdy1d are the constant ratios
 * (step2.f90:80-86) */
static void step_work_ratios(step_work* sw, int maxm, int mbc, int index_capa, double dtdx,
                             double dtdy)
{
    if (index_capa == 0)
        for (int i = 1 - mbc; i <= maxm + mbc; i++) {
            D1(sw->dtdx1d, i) = dtdx;
            D1(sw->dtdy1d, i) = dtdy;
        }
}

static int g_threads = 1;

int orc_set_threads(int n)
{
#ifdef _OPENMP
    g_threads = n < 1 ? 1 : n;
#else
    (void)n;
    g_threads = 1;
#endif
    return g_threads;
}

int orc_threads(void) { return g_threads; }

/* The same step on g_threads OpenMP threads: the x-sweeps are shared out by
 * bands of rows (rows -1..my+2 receive updates), the y-sweeps by bands of
 * columns; a thread runs the slices of its band plus one on each side and
 * applies only the updates that land in its band, so each cell is updated
 * exactly as in the serial order.  The Courant number is a max over slices
 * (slices on band edges are counted twice, which a max does not see). */
static int orc_step2_par(int maxm, int meqn, int mwaves, int maux, int mbc, int mx, int my,
                         const double* qold, double* qnew, const double* aux, double dx,
                         double dy, double dt, const int* method, const int* mthlim,
                         double* cfl_out, int use_fwave, orc_rpn2_t rpn2, orc_rpt2_t rpt2)
{
#ifdef _OPENMP
    const double dtdx = dt / dx, dtdy = dt / dy;
    double cfl = 0.0;
    int bad = 0;
#pragma omp parallel num_threads(g_threads)
    {
        const int t = omp_get_thread_num(), nt = omp_get_num_threads();
        step_work sw;
        int ok = step_work_alloc(&sw, maxm, meqn, mwaves, maux, mbc) == 0;
        double c = 0.0;
        if (ok) {
            step_work_ratios(&sw, maxm, mbc, method[5], dtdx, dtdy);
            const int nr = my + 4;
            const int r0 = -1 + (int)((long)nr * t / nt), r1 = -2 + (int)((long)nr * (t + 1) / nt);
            const int j0 = r0 - 1 > 0 ? r0 - 1 : 0, j1 = r1 + 1 < my + 1 ? r1 + 1 : my + 1;
            if (r0 <= r1 && step2_xslices(maxm, meqn, mwaves, maux, mbc, mx, my, qold, qnew, aux,
                                          dtdx, dtdy, method, mthlim, use_fwave, rpn2, rpt2, &sw,
                                          j0, j1, r0, r1, &c))
                ok = 0;
        }
#pragma omp barrier
        if (ok) {
            const int nc = mx + 4;
            const int c0 = -1 + (int)((long)nc * t / nt), c1 = -2 + (int)((long)nc * (t + 1) / nt);
            const int i0 = c0 - 1 > 0 ? c0 - 1 : 0, i1 = c1 + 1 < mx + 1 ? c1 + 1 : mx + 1;
            if (c0 <= c1 && step2_yslices(maxm, meqn, mwaves, maux, mbc, mx, my, qold, qnew, aux,
                                          dtdx, dtdy, method, mthlim, use_fwave, rpn2, rpt2, &sw,
                                          i0, i1, c0, c1, &c))
                ok = 0;
        }
#pragma omp critical
        {
            cfl = (c != c || cfl != cfl) ? NAN : fmax(cfl, c);
            if (!ok) bad = 1;
        }
        if (sw.block) free(sw.block);
    }
    *cfl_out = cfl;
    return bad ? -1 : 0;
#else
    (void)maxm; (void)meqn; (void)mwaves; (void)maux; (void)mbc; (void)mx; (void)my;
    (void)qold; (void)qnew; (void)aux; (void)dx; (void)dy; (void)dt; (void)method;
    (void)mthlim; (void)cfl_out; (void)use_fwave; (void)rpn2; (void)rpt2;
    return -1;
#endif
}

static int orc_step2_w(int maxm, int meqn, int mwaves, int maux, int mbc, int mx, int my,
                       const double* qold, double* qnew, const double* aux, double dx,
                       double dy, double dt, const int* method, const int* mthlim,
                       double* cfl_out, int use_fwave, orc_rpn2_t rpn2, orc_rpt2_t rpt2,
                       step_work* sw)
{
    if (g_threads > 1)
        return orc_step2_par(maxm, meqn, mwaves, maux, mbc, mx, my, qold, qnew, aux, dx, dy, dt,
                             method, mthlim, cfl_out, use_fwave, rpn2, rpt2);
    const double dtdx = dt / dx, dtdy = dt / dy;
    double cfl = 0.0;
    step_work_ratios(sw, maxm, mbc, method[5], dtdx, dtdy);
    if (step2_xslices(maxm, meqn, mwaves, maux, mbc, mx, my, qold, qnew, aux, dtdx, dtdy, method,
                      mthlim, use_fwave, rpn2, rpt2, sw, 0, my + 1, INT_MIN, INT_MAX, &cfl))
        return -1;
    if (step2_yslices(maxm, meqn, mwaves, maux, mbc, mx, my, qold, qnew, aux, dtdx, dtdy, method,
                      mthlim, use_fwave, rpn2, rpt2, sw, 0, mx + 1, INT_MIN, INT_MAX, &cfl))
        return -1;
    *cfl_out = cfl;
    return 0;
}
#undef G
#undef D1
#undef INB

/* step2ds.f90:2-275: one direction of a dimensionally split step.  The
 * x-sweep runs over the ghost rows too; the y-sweep over ghost columns. */
static int orc_step2ds_w(int maxm, int meqn, int mwaves, int maux, int mbc, int mx, int my,
                         const double* qold, double* qnew, const double* aux, double dx,
                         double dy, double dt, const int* method, const int* mthlim,
                         double* cfl_out, int ids, int use_fwave, orc_rpn2_t rpn2,
                         orc_rpt2_t rpt2, step_work* sw)
{
    const int index_capa = method[5];
    double cfl = 0.0, cfl1d;
    double dtdx = dt / dx, dtdy = dt / dy;
    double *qadd = sw->qadd, *fadd = sw->fadd, *gadd = sw->gadd, *q1d = sw->q1d;
    double *dtdx1d = sw->dtdx1d, *dtdy1d = sw->dtdy1d;
#define D1(a, i) (a)[(i) + mbc - 1]
    if (index_capa == 0)
        for (int i = 1 - mbc; i <= maxm + mbc; i++) {
            D1(dtdx1d, i) = dtdx;
            D1(dtdy1d, i) = dtdy;
        }
    if (ids == 1) {
        for (int j = 1 - mbc; j <= my + mbc; j++) {
            for (int i = 1 - mbc; i <= mx + mbc; i++)
                for (int m = 1; m <= meqn; m++) S1(q1d, m, i) = Q2(qold, m, i, j);
            if (index_capa > 0)
                for (int i = 1 - mbc; i <= mx + mbc; i++)
                    D1(dtdx1d, i) = dtdx / A2(aux, index_capa, i, j);
            if (maux > 0) {
                for (int i = 1 - mbc; i <= mx + mbc; i++)
                    for (int ma = 1; ma <= maux; ma++)
                        sw->aux2[(i + mbc - 1) * maux + ma - 1] = A2(aux, ma, i, j);
                if (j != 1 - mbc)
                    for (int i = 1 - mbc; i <= mx + mbc; i++)
                        for (int ma = 1; ma <= maux; ma++)
                            sw->aux1[(i + mbc - 1) * maux + ma - 1] = A2(aux, ma, i, j - 1);
                if (j != my + mbc)
                    for (int i = 1 - mbc; i <= mx + mbc; i++)
                        for (int ma = 1; ma <= maux; ma++)
                            sw->aux3[(i + mbc - 1) * maux + ma - 1] = A2(aux, ma, i, j + 1);
            }
            if (orc_flux2(1, maxm, meqn, mwaves, maux, mbc, mx, q1d, dtdx1d, sw->aux1,
                          sw->aux2, sw->aux3, method, mthlim, qadd, fadd, gadd, &cfl1d, &sw->fw,
                          rpn2, rpt2, use_fwave) != 0)
                return -1;
            cfl = (cfl1d != cfl1d || cfl != cfl) ? NAN : fmax(cfl, cfl1d);
            for (int i = 1; i <= mx; i++)
                for (int m = 1; m <= meqn; m++) {
                    if (index_capa == 0)
                        Q2(qnew, m, i, j) = Q2(qnew, m, i, j) + S1(qadd, m, i) -
                                            dtdx * (S1(fadd, m, i + 1) - S1(fadd, m, i));
                    else
                        Q2(qnew, m, i, j) = Q2(qnew, m, i, j) + S1(qadd, m, i) -
                                            dtdx * (S1(fadd, m, i + 1) - S1(fadd, m, i)) /
                                                A2(aux, index_capa, i, j);
                }
        }
    }
    if (ids == 2) {
        for (int i = 1 - mbc; i <= mx + mbc; i++) {
            for (int j = 1 - mbc; j <= my + mbc; j++)
                for (int m = 1; m <= meqn; m++) S1(q1d, m, j) = Q2(qold, m, i, j);
            if (index_capa > 0)
                for (int j = 1 - mbc; j <= my + mbc; j++)
                    D1(dtdy1d, j) = dtdy / A2(aux, index_capa, i, j);
            if (maux > 0) {
                for (int j = 1 - mbc; j <= my + mbc; j++)
                    for (int ma = 1; ma <= maux; ma++)
                        sw->aux2[(j + mbc - 1) * maux + ma - 1] = A2(aux, ma, i, j);
                if (i != 1 - mbc)
                    for (int j = 1 - mbc; j <= my + mbc; j++)
                        for (int ma = 1; ma <= maux; ma++)
                            sw->aux1[(j + mbc - 1) * maux + ma - 1] = A2(aux, ma, i - 1, j);
                if (i != mx + mbc)
                    for (int j = 1 - mbc; j <= my + mbc; j++)
                        for (int ma = 1; ma <= maux; ma++)
                            sw->aux3[(j + mbc - 1) * maux + ma - 1] = A2(aux, ma, i + 1, j);
            }
            if (orc_flux2(2, maxm, meqn, mwaves, maux, mbc, my, q1d, dtdy1d, sw->aux1,
                          sw->aux2, sw->aux3, method, mthlim, qadd, fadd, gadd, &cfl1d, &sw->fw,
                          rpn2, rpt2, use_fwave) != 0)
                return -1;
            cfl = (cfl1d != cfl1d || cfl != cfl) ? NAN : fmax(cfl, cfl1d);
            for (int j = 1; j <= my; j++)
                for (int m = 1; m <= meqn; m++) {
                    if (index_capa == 0)
                        Q2(qnew, m, i, j) = Q2(qnew, m, i, j) + S1(qadd, m, j) -
                                            dtdy * (S1(fadd, m, j + 1) - S1(fadd, m, j));
                    else
                        Q2(qnew, m, i, j) = Q2(qnew, m, i, j) + S1(qadd, m, j) -
                                            dtdy * (S1(fadd, m, j + 1) - S1(fadd, m, j)) /
                                                A2(aux, index_capa, i, j);
                }
        }
    }
    *cfl_out = cfl;
    return 0;
#undef D1
}

/* dimsp2.f:4-98 (Godunov method(3)=-1 or Strang method(3)=-2). */
static int orc_dimsp2_w(int maxm, int meqn, int mwaves, int maux, int mbc, int mx, int my,
                        const double* qold, double* qnew, const double* aux, double dx,
                        double dy, double dt, const int* method, const int* mthlim,
                        double* cfl, const double* cflv, int use_fwave, orc_rpn2_t rpn2,
                        orc_rpt2_t rpt2, step_work* sw)
{
    double dt2 = dt / 2.0, cflx, cfly;
    double dtx = (method[2] == -2) ? dt2 : dt;
    if (orc_step2ds_w(maxm, meqn, mwaves, maux, mbc, mx, my, qold, qnew, aux, dx, dy, dtx,
                      method, mthlim, &cflx, 1, use_fwave, rpn2, rpt2, sw) != 0)
        return -1;
    if (cflx > cflv[0] || cflx != cflx) {
        *cfl = cflx;
        return 0;
    }
    if (orc_step2ds_w(maxm, meqn, mwaves, maux, mbc, mx, my, qnew, qnew, aux, dx, dy, dt,
                      method, mthlim, &cfly, 2, use_fwave, rpn2, rpt2, sw) != 0)
        return -1;
    *cfl = (cflx != cflx || cfly != cfly) ? NAN : fmax(cflx, cfly);
    if (method[2] == -2) {
        if (cfly > cflv[0]) {
            *cfl = cfly;
            return 0;
        }
        if (orc_step2ds_w(maxm, meqn, mwaves, maux, mbc, mx, my, qnew, qnew, aux, dx, dy, dt2,
                          method, mthlim, &cflx, 1, use_fwave, rpn2, rpt2, sw) != 0)
            return -1;
        *cfl = (cflx != cflx || *cfl != *cfl) ? NAN : fmax(*cfl, cflx);
    }
    return 0;
}

int orc_step2(int meqn, int mwaves, int maux, int mbc, int mx, int my, const double* qold,
              double* qnew, const double* aux, double dx, double dy, double dt,
              const int* method, const int* mthlim, int use_fwave, int solver_id,
              double* cfl)
{
    orc_rpn2_t rpn2 = orc_get_rpn2(solver_id);
    orc_rpt2_t rpt2 = orc_get_rpt2(solver_id);
    if (!rpn2 || !rpt2) return -2;
    int maxm = mx > my ? mx : my;
    step_work sw;
    if (step_work_alloc(&sw, maxm, meqn, mwaves, maux, mbc) != 0) return -3;
    int rc = orc_step2_w(maxm, meqn, mwaves, maux, mbc, mx, my, qold, qnew, aux, dx, dy, dt,
                         method, mthlim, cfl, use_fwave, rpn2, rpt2, &sw);
    free(sw.block);
    return rc;
}

int orc_dimsp2(int meqn, int mwaves, int maux, int mbc, int mx, int my, const double* qold,
               double* qnew, const double* aux, double dx, double dy, double dt,
               const int* method, const int* mthlim, const double* cflv, int use_fwave,
               int solver_id, double* cfl)
{
    orc_rpn2_t rpn2 = orc_get_rpn2(solver_id);
    orc_rpt2_t rpt2 = orc_get_rpt2(solver_id);
    if (!rpn2 || !rpt2) return -2;
    int maxm = mx > my ? mx : my;
    step_work sw;
    if (step_work_alloc(&sw, maxm, meqn, mwaves, maux, mbc) != 0) return -3;
    int rc = orc_dimsp2_w(maxm, meqn, mwaves, maux, mbc, mx, my, qold, qnew, aux, dx, dy, dt,
                          method, mthlim, cfl, cflv, use_fwave, rpn2, rpt2, &sw);
    free(sw.block);
    return rc;
}

/* -------------------------------------------------------------------- bc2 */

/* bc2.f:4-215: left, right, bottom, top, in that order (corners come from
 * the bottom/top pass over the already-filled ghost columns).  A wall
 * negates component 2 (x) or 3 (y); the reference does so whatever meqn is,
 * writing outside q when meqn is smaller -- here the negation is skipped
 * then. */
int orc_bc2(int meqn, int mbc, int mx, int my, double* q, const int* mthbc)
{
    for (int side = 0; side < 4; side++)
        if (mthbc[side] < 1 || mthbc[side] > 3) return -1;
    /* left */
    for (int j = 1 - mbc; j <= my + mbc; j++)
        for (int ibc = 1; ibc <= mbc; ibc++) {
            for (int m = 1; m <= meqn; m++) {
                if (mthbc[0] == 1) Q2(q, m, 1 - ibc, j) = Q2(q, m, 1, j);
                else if (mthbc[0] == 2) Q2(q, m, 1 - ibc, j) = Q2(q, m, mx + 1 - ibc, j);
                else Q2(q, m, 1 - ibc, j) = Q2(q, m, ibc, j);
            }
        }
    if (mthbc[0] == 3 && meqn >= 2)
        for (int j = 1 - mbc; j <= my + mbc; j++)
            for (int ibc = 1; ibc <= mbc; ibc++) Q2(q, 2, 1 - ibc, j) = -Q2(q, 2, ibc, j);
    /* right */
    for (int j = 1 - mbc; j <= my + mbc; j++)
        for (int ibc = 1; ibc <= mbc; ibc++) {
            for (int m = 1; m <= meqn; m++) {
                if (mthbc[1] == 1) Q2(q, m, mx + ibc, j) = Q2(q, m, mx, j);
                else if (mthbc[1] == 2) Q2(q, m, mx + ibc, j) = Q2(q, m, ibc, j);
                else Q2(q, m, mx + ibc, j) = Q2(q, m, mx + 1 - ibc, j);
            }
        }
    if (mthbc[1] == 3 && meqn >= 2)
        for (int j = 1 - mbc; j <= my + mbc; j++)
            for (int ibc = 1; ibc <= mbc; ibc++) Q2(q, 2, mx + ibc, j) = -Q2(q, 2, mx + 1 - ibc, j);
    /* bottom */
    for (int jbc = 1; jbc <= mbc; jbc++)
        for (int i = 1 - mbc; i <= mx + mbc; i++)
            for (int m = 1; m <= meqn; m++) {
                if (mthbc[2] == 1) Q2(q, m, i, 1 - jbc) = Q2(q, m, i, 1);
                else if (mthbc[2] == 2) Q2(q, m, i, 1 - jbc) = Q2(q, m, i, my + 1 - jbc);
                else Q2(q, m, i, 1 - jbc) = Q2(q, m, i, jbc);
            }
    if (mthbc[2] == 3 && meqn >= 3)
        for (int jbc = 1; jbc <= mbc; jbc++)
            for (int i = 1 - mbc; i <= mx + mbc; i++) Q2(q, 3, i, 1 - jbc) = -Q2(q, 3, i, jbc);
    /* top */
    for (int jbc = 1; jbc <= mbc; jbc++)
        for (int i = 1 - mbc; i <= mx + mbc; i++)
            for (int m = 1; m <= meqn; m++) {
                if (mthbc[3] == 1) Q2(q, m, i, my + jbc) = Q2(q, m, i, my);
                else if (mthbc[3] == 2) Q2(q, m, i, my + jbc) = Q2(q, m, i, jbc);
                else Q2(q, m, i, my + jbc) = Q2(q, m, i, my + 1 - jbc);
            }
    if (mthbc[3] == 3 && meqn >= 3)
        for (int jbc = 1; jbc <= mbc; jbc++)
            for (int i = 1 - mbc; i <= mx + mbc; i++) Q2(q, 3, i, my + jbc) = -Q2(q, 3, i, my + 1 - jbc);
    return 0;
}

/* ------------------------------------------------------------------ claw2 */

/* claw2.f:494-784.  Sources and b4step2 are the reference's no-op defaults
 * (src2.f90, b4step2.f90), so method(5) only changes the work-array size.
 * Returns info; tend, dtv, cflv, nv are updated like the reference. */
int orc_claw2(int meqn, int mwaves, int maux, int mbc, int mx, int my, double* q,
              const double* aux, double dx, double dy, double tstart, double* tend,
              double* dtv, double* cflv, int* nv, const int* method, const int* mthlim,
              const int* mthbc, int use_fwave, int solver_id, double* dt_trace,
              double* cfl_trace, int trace_cap, int* ntrace)
{
    orc_rpn2_t rpn2 = orc_get_rpn2(solver_id);
    orc_rpt2_t rpt2 = orc_get_rpt2(solver_id);
    if (!rpn2 || !rpt2) return -2;
    int maxm = mx > my ? mx : my;
    int info = 0;
    double t = tstart;
    int maxn = nv[0];
    double dt = dtv[0];
    double cflmax = 0.0, dtmin = dt, dtmax = dt, cfl = 0.0;
    nv[1] = 0;
    if (ntrace) *ntrace = 0;

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

    size_t nq = (size_t)(mx + 2 * mbc) * (my + 2 * mbc) * meqn;
    double* qwork = (double*)malloc(nq * sizeof(double));
    step_work sw;
    if (!qwork || step_work_alloc(&sw, maxm, meqn, mwaves, maux, mbc) != 0) {
        free(qwork);
        return -3;
    }
    if (maxn == 0) goto freed;
    for (int n = 1; n <= maxn; n++) {
        double told = t;
        if (told + dt > *tend && tstart < *tend) dt = *tend - told;
    retake:
        if (orc_bc2(meqn, mbc, mx, my, q, mthbc) != 0) { info = -4; goto freed; }
        memcpy(qwork, q, nq * sizeof(double));
        int rc;
        if (method[2] >= 0)
            rc = orc_step2_w(maxm, meqn, mwaves, maux, mbc, mx, my, qwork, q, aux, dx, dy, dt,
                             method, mthlim, &cfl, use_fwave, rpn2, rpt2, &sw);
        else
            rc = orc_dimsp2_w(maxm, meqn, mwaves, maux, mbc, mx, my, qwork, q, aux, dx, dy, dt,
                              method, mthlim, &cfl, cflv, use_fwave, rpn2, rpt2, &sw);
        if (rc != 0) { info = -1; goto freed; }
        /* Not in the reference: a NaN Courant number would make its retake
         * loop spin forever (NaN fails the cfl <= cfl_max test and fmin
         * drops the NaN from the new dt), so stop with info = -5. */
        if (cfl != cfl) { info = -5; goto freed; }
        t = told + dt;
        if (ntrace && *ntrace < trace_cap) {
            dt_trace[*ntrace] = dt;
            cfl_trace[*ntrace] = cfl;
            (*ntrace)++;
        }
        if (cfl <= cflv[0]) {
            cflmax = fmax(cfl, cflmax);
        } else {
            t = told;
            memcpy(q, qwork, nq * sizeof(double));
            if (method[0] == 1) {
                dt = fmin(dtv[1], dt * cflv[1] / cfl);
                goto retake;
            } else {
                cflmax = fmax(cfl, cflmax);
                goto freed;
            }
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
freed:
    free(qwork);
    free(sw.block);
done:
    if (info >= 0) {
        if (method[0] == 1 && t < *tend && nv[1] == maxn) info = 11;
        if (method[0] == 0 && cflmax > cflv[0]) info = 12;
    }
    *tend = t;
    cflv[2] = cflmax;
    cflv[3] = cfl;
    dtv[2] = dtmin;
    dtv[3] = dtmax;
    dtv[4] = dt;
    return info;
}
