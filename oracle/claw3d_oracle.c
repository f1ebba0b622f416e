/*
 * oracle/claw3d_oracle.c -- CPU ORACLE for the 3D wave-propagation step.
 *
 * TEST INFRASTRUCTURE ONLY.  A plain-C restatement of the reference Fortran
 * (clawpack/classic) 3D single-grid path, used by the tests and the smoke
 * check as the checker.  It is never linked into, loaded by, or called from
 * the product library (classic_amd/).  It keeps the reference's slice
 * structure, its blocked slice order and its floating-point expression order,
 * so that a build with -ffp-contract=off reproduces the Fortran bit for bit.
 *
 * Functions and the reference code each one restates:
 *   orc3_flux3       src/3d/flux3.f90:2-612
 *   orc3_step3       src/3d/step3.f90:3-873 (x/y/z sweeps over 2x2 slice
 *                    blocks in four offset passes, :127-206; the serial
 *                    order of those loops fixes each cell's summation order)
 *   orc3_step3ds     src/3d/step3ds.f90:3-408
 *   orc3_dimsp3      src/3d/dimsp3.f:5-85 (Godunov splitting only)
 *   orc3_bc3         src/3d/bc3.f (codes 1-3) and, for mthbc(6)=0, the
 *                    piston of examples/acoustics_3d_heterogeneous/bc3.f
 *   orc3_claw3       src/3d/claw3.f:738-904 (time loop, CFL retake, info),
 *                    with chkmth src/3d/chkmth.f:3-72
 *   limiter          orc_limiter of claw2d_oracle.c: src/3d/inlinelimiter.f90
 *                    is identical to the 2D file, and src/3d/limiter.f90 with
 *                    philim.f90 forms r = dot/wnorm2 and phi(r) the same way
 * Riemann solvers: the reference fixes only their contract
 * (src/3d/rpn3.f, rpt3.f, rptt3.f; claw3.f:292-516); the variable-coefficient
 * acoustics solvers its example names (rpn3/rpt3/rptt3_vc_acoustics of the
 * absent clawpack/riemann repository, version unpinned) are restated from
 * LeVeque, FVMHP 9.9 (interface Riemann problem for variable-coefficient
 * acoustics) and 21.x (transverse splitting into the transmitted parts of the
 * down- and up-going waves):
 *   aux(1,.) = impedance Z, aux(2,.) = sound speed c.
 *
 * Pinning (see tests/): step3/flux3/dimsp3/bc3/claw3 against the reference
 * Fortran compiled from /root/reference (oracle/_ref/libclawref3d.so) with
 * these solvers linked in, bit for bit; the solvers through the example's
 * regression frame sums (examples/acoustics_3d_heterogeneous/regression_data)
 * and through closed forms (homogeneous medium: waves are eigenvectors,
 * amdq + apdq = A dq; interface transmission coefficients).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "claw2d_oracle.h"
#include "claw3d_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

/* 1D slice arrays, Fortran indices, first index 1-mbc */
#define S1(a, m, i) (a)[(size_t)((i) + mbc - 1) * meqn + ((m) - 1)]
#define SW(w, m, mw, i) (w)[((size_t)((i) + mbc - 1) * mwaves + ((mw) - 1)) * meqn + ((m) - 1)]
#define SS(s, mw, i) (s)[(size_t)((i) + mbc - 1) * mwaves + ((mw) - 1)]
/* aux slices aux1/2/3(maux, 1-mbc:maxm+mbc, 3) */
#define AX(a, ma, i, kk) (a)[((size_t)((kk) - 1) * (maxm + 2 * mbc) + ((i) + mbc - 1)) * maux + ((ma) - 1)]
#define AL(a, ma, i) (a)[(size_t)((i) + mbc - 1) * maux + ((ma) - 1)]

/* normal, y-like and z-like velocity components for ixyz = 1, 2, 3 */
static void vc_components(int ixyz, int* mu, int* mv, int* mw)
{
    if (ixyz == 1) { *mu = 2; *mv = 3; *mw = 4; }
    else if (ixyz == 2) { *mu = 3; *mv = 4; *mw = 2; }
    else { *mu = 4; *mv = 2; *mw = 3; }
}

/* ------------------------------------------------ vc acoustics solvers */

/* Interface between cells i-1 and i: the jump in (p, u_n) splits into a
 * left-going wave in cell i-1 (eigenvector (-Z_{i-1}, 1), speed -c_{i-1}) and
 * a right-going wave in cell i ((Z_i, 1), speed c_i). */
void orc_rpn3_vc_acoustics(int ixyz, int maxm, int meqn, int mwaves, int maux, int mbc, int mx,
                           const double* ql, const double* qr, const double* auxl,
                           const double* auxr, double* wave, double* s, double* amdq,
                           double* apdq)
{
    (void)maxm;
    int mu, mv, mw;
    vc_components(ixyz, &mu, &mv, &mw);
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        double delta1 = S1(ql, 1, i) - S1(qr, 1, i - 1);
        double delta2 = S1(ql, mu, i) - S1(qr, mu, i - 1);
        double zi = AL(auxl, 1, i);
        double zim = AL(auxr, 1, i - 1);
        double a1 = (-delta1 + zi * delta2) / (zim + zi);
        double a2 = (delta1 + zim * delta2) / (zim + zi);
        for (int m = 1; m <= meqn; m++) {
            SW(wave, m, 1, i) = 0.0;
            SW(wave, m, 2, i) = 0.0;
        }
        SW(wave, 1, 1, i) = -a1 * zim;
        SW(wave, mu, 1, i) = a1;
        SS(s, 1, i) = -AL(auxr, 2, i - 1);
        SW(wave, 1, 2, i) = a2 * zi;
        SW(wave, mu, 2, i) = a2;
        SS(s, 2, i) = AL(auxl, 2, i);
    }
    for (int i = 2 - mbc; i <= mx + mbc; i++)
        for (int m = 1; m <= meqn; m++) {
            S1(amdq, m, i) = SS(s, 1, i) * SW(wave, m, 1, i);
            S1(apdq, m, i) = SS(s, 2, i) * SW(wave, m, 2, i);
        }
}

/* Split a flux difference living in cell i1 (i-1 for imp=1, i otherwise)
 * into the transmitted down-going part (speed -c of the neighbour below) and
 * up-going part (speed +c of the neighbour above) in the transverse direction
 * of component mt, from the impedances (zm, zz, zp) and sound speeds (cm, cp)
 * of the cells below, at and above. */
static void vc_split(int meqn, int mu, int mt, double a_p, double a_t, double zm, double zz,
                     double zp, double cm, double cp, double* bm, double* bp)
{
    double a1 = (-a_p + a_t * zz) / (zm + zz);
    double a2 = (a_p + a_t * zz) / (zz + zp);
    for (int m = 0; m < meqn; m++) {
        bm[m] = 0.0;
        bp[m] = 0.0;
    }
    bm[0] = cm * a1 * zm;
    bm[mu - 1] = 0.0;
    bm[mt - 1] = -cm * a1;
    bp[0] = cp * a2 * zp;
    bp[mu - 1] = 0.0;
    bp[mt - 1] = cp * a2;
}

void orc_rpt3_vc_acoustics(int ixyz, int icoor, int imp, int maxm, int meqn, int mwaves,
                           int maux, int mbc, int mx, const double* ql, const double* qr,
                           const double* aux1, const double* aux2, const double* aux3,
                           const double* asdq, double* bmasdq, double* bpasdq)
{
    (void)mwaves; (void)ql; (void)qr;
    int mu, mv, mw;
    vc_components(ixyz, &mu, &mv, &mw);
    int mt = icoor == 2 ? mv : mw;
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        int i1 = imp == 1 ? i - 1 : i;
        double zm, zz, zp, cm, cp;
        if (icoor == 2) { /* neighbours in the y-like direction: rows j-1, j+1 */
            zm = AX(aux1, 1, i1, 2); zz = AX(aux2, 1, i1, 2); zp = AX(aux3, 1, i1, 2);
            cm = AX(aux1, 2, i1, 2); cp = AX(aux3, 2, i1, 2);
        } else {          /* z-like: planes k-1, k+1 of row j */
            zm = AX(aux2, 1, i1, 1); zz = AX(aux2, 1, i1, 2); zp = AX(aux2, 1, i1, 3);
            cm = AX(aux2, 2, i1, 1); cp = AX(aux2, 2, i1, 3);
        }
        vc_split(meqn, mu, mt, S1(asdq, 1, i), S1(asdq, mt, i), zm, zz, zp, cm, cp,
                 &S1(bmasdq, 1, i), &S1(bpasdq, 1, i));
    }
}

/* Double-transverse split: bsasdq already moved into the neighbour plane
 * (icoor=2: z-like neighbour k-1 for impt=1, k+1 for impt=2) or row (icoor=3:
 * y-like neighbour j-1 / j+1), and is split there in the remaining
 * transverse direction. */
void orc_rptt3_vc_acoustics(int ixyz, int icoor, int imp, int impt, int maxm, int meqn,
                            int mwaves, int maux, int mbc, int mx, const double* ql,
                            const double* qr, const double* aux1, const double* aux2,
                            const double* aux3, const double* bsasdq, double* cmbsasdq,
                            double* cpbsasdq)
{
    (void)mwaves; (void)ql; (void)qr;
    int mu, mv, mw;
    vc_components(ixyz, &mu, &mv, &mw);
    int mt = icoor == 2 ? mv : mw;
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        int i1 = imp == 1 ? i - 1 : i;
        double zm, zz, zp, cm, cp;
        if (icoor == 2) {
            int kk = impt == 1 ? 1 : 3;
            zm = AX(aux1, 1, i1, kk); zz = AX(aux2, 1, i1, kk); zp = AX(aux3, 1, i1, kk);
            cm = AX(aux1, 2, i1, kk); cp = AX(aux3, 2, i1, kk);
        } else {
            const double* row = impt == 1 ? aux1 : aux3;
            zm = AX(row, 1, i1, 1); zz = AX(row, 1, i1, 2); zp = AX(row, 1, i1, 3);
            cm = AX(row, 2, i1, 1); cp = AX(row, 2, i1, 3);
        }
        vc_split(meqn, mu, mt, S1(bsasdq, 1, i), S1(bsasdq, mt, i), zm, zz, zp, cm, cp,
                 &S1(cmbsasdq, 1, i), &S1(cpbsasdq, 1, i));
    }
}

/* ----------------------------------------------------------------- flux3 */

enum {
    W_AMDQ, W_APDQ, W_CQXX,
    W_BMAMDQ, W_BMAPDQ, W_BPAMDQ, W_BPAPDQ, W_CMAMDQ, W_CMAPDQ, W_CPAMDQ, W_CPAPDQ,
    W_CMAMDQ2, W_CMAPDQ2, W_CPAMDQ2, W_CPAPDQ2,
    W_BMCQXXP, W_BPCQXXP, W_BMCQXXM, W_BPCQXXM, W_CMCQXXP, W_CPCQXXP, W_CMCQXXM, W_CPCQXXM,
    W_BMCMAMDQ, W_BMCMAPDQ, W_BPCMAMDQ, W_BPCMAPDQ, W_BMCPAMDQ, W_BMCPAPDQ, W_BPCPAMDQ,
    W_BPCPAPDQ,
    W_NARR
};

typedef struct {
    double* wave;
    double* s;
    double* a[W_NARR];
} flux3_work;

typedef struct {
    double *qadd, *fadd, *gadd, *hadd, *q1d, *dtdx1d, *aux1, *aux2, *aux3;
    flux3_work fw;
    void* block;
} step3_work;

static int step3_work_alloc(step3_work* sw, int maxm, int meqn, int mwaves, int maux, int mbc)
{
    size_t n = (size_t)(maxm + 2 * mbc);
    int ma = maux > 0 ? maux : 1;
    size_t tot = n * (size_t)(meqn * (2 + 12 + 1) + 1 + 9 * ma + meqn * mwaves + mwaves +
                              W_NARR * meqn);
    double* b = (double*)calloc(tot, sizeof(double));
    if (!b) return -1;
    sw->block = b;
    sw->qadd = b; b += n * meqn;
    sw->fadd = b; b += n * meqn;
    sw->gadd = b; b += 6 * n * meqn;
    sw->hadd = b; b += 6 * n * meqn;
    sw->q1d = b; b += n * meqn;
    sw->dtdx1d = b; b += n;
    sw->aux1 = b; b += 3 * n * ma;
    sw->aux2 = b; b += 3 * n * ma;
    sw->aux3 = b; b += 3 * n * ma;
    sw->fw.wave = b; b += n * meqn * mwaves;
    sw->fw.s = b; b += n * mwaves;
    for (int k = 0; k < W_NARR; k++) {
        sw->fw.a[k] = b;
        b += n * meqn;
    }
    return 0;
}

/* gadd/hadd(meqn, 2, -1:1, 1-mbc:maxm+mbc) */
#define GH(a, m, side, sl, i) (a)[((((size_t)((i) + mbc - 1) * 3 + ((sl) + 1)) * 2 + ((side) - 1)) * meqn) + ((m) - 1)]

/* flux3.f90:2-612.  dtdy/dtdz are the y-like and z-like dt/d ratios of the
 * slice direction (step3 passes them rotated). */
static int orc3_flux3(int ixyz, int maxm, int meqn, int mwaves, int mbc, int mx,
                      const double* q1d, const double* dtdx1d, double dtdy, double dtdz,
                      const double* aux1, const double* aux2, const double* aux3, int maux,
                      const int* method, const int* mthlim, double* qadd, double* fadd,
                      double* gadd, double* hadd, double* cfl1d, flux3_work* w, orc_rpn3_t rpn3,
                      orc_rpt3_t rpt3, orc_rptt3_t rptt3, int use_fwave)
{
    double *wave = w->wave, *s = w->s;
    double *amdq = w->a[W_AMDQ], *apdq = w->a[W_APDQ], *cqxx = w->a[W_CQXX];
    double *bmamdq = w->a[W_BMAMDQ], *bmapdq = w->a[W_BMAPDQ];
    double *bpamdq = w->a[W_BPAMDQ], *bpapdq = w->a[W_BPAPDQ];
    double *cmamdq = w->a[W_CMAMDQ], *cmapdq = w->a[W_CMAPDQ];
    double *cpamdq = w->a[W_CPAMDQ], *cpapdq = w->a[W_CPAPDQ];
    double *cmamdq2 = w->a[W_CMAMDQ2], *cmapdq2 = w->a[W_CMAPDQ2];
    double *cpamdq2 = w->a[W_CPAMDQ2], *cpapdq2 = w->a[W_CPAPDQ2];
    double *bmcqxxp = w->a[W_BMCQXXP], *bpcqxxp = w->a[W_BPCQXXP];
    double *bmcqxxm = w->a[W_BMCQXXM], *bpcqxxm = w->a[W_BPCQXXM];
    double *cmcqxxp = w->a[W_CMCQXXP], *cpcqxxp = w->a[W_CPCQXXP];
    double *cmcqxxm = w->a[W_CMCQXXM], *cpcqxxm = w->a[W_CPCQXXM];
    double *bmcmamdq = w->a[W_BMCMAMDQ], *bmcmapdq = w->a[W_BMCMAPDQ];
    double *bpcmamdq = w->a[W_BPCMAMDQ], *bpcmapdq = w->a[W_BPCMAPDQ];
    double *bmcpamdq = w->a[W_BMCPAMDQ], *bmcpapdq = w->a[W_BMCPAPDQ];
    double *bpcpamdq = w->a[W_BPCPAMDQ], *bpcpapdq = w->a[W_BPCPAPDQ];
#define DT(i) dtdx1d[(i) + mbc - 1]
    const double sixth = 1.0 / 6.0;
    int limit = 0;
    for (int mw = 0; mw < mwaves; mw++)
        if (mthlim[mw] > 0) limit = 1;

    for (int i = 1 - mbc; i <= mx + mbc; i++)
        for (int m = 1; m <= meqn; m++) {
            S1(qadd, m, i) = 0.0;
            S1(fadd, m, i) = 0.0;
            for (int side = 1; side <= 2; side++)
                for (int sl = -1; sl <= 1; sl++) {
                    GH(gadd, m, side, sl, i) = 0.0;
                    GH(hadd, m, side, sl, i) = 0.0;
                }
        }

    int m3, m4;
    if (method[2] < 0) { m3 = -1; m4 = 0; }
    else { m3 = method[2] / 10; m4 = method[2] - 10 * m3; }

    /* aux2(1,1-mbc,2): the middle plane of the slice's aux */
    const double* aux2mid = aux2 + (size_t)(maxm + 2 * mbc) * (maux > 0 ? maux : 1);
    rpn3(ixyz, maxm, meqn, mwaves, maux, mbc, mx, q1d, q1d, aux2mid, aux2mid, wave, s, amdq,
         apdq);

    /* forall: the first statement sees the zeroed qadd, the second the
     * result of the first */
    for (int i = 1; i <= mx + 1; i++)
        for (int m = 1; m <= meqn; m++) S1(qadd, m, i) = S1(qadd, m, i) - DT(i) * S1(apdq, m, i);
    for (int i = 1; i <= mx + 1; i++)
        for (int m = 1; m <= meqn; m++)
            S1(qadd, m, i - 1) = S1(qadd, m, i - 1) - DT(i - 1) * S1(amdq, m, i);

    double cfl = 0.0;
    int nan_speed = 0;
    for (int i = 1; i <= mx + 1; i++)
        for (int mw = 1; mw <= mwaves; mw++) {
            cfl = fmax(cfl, fmax(DT(i) * SS(s, mw, i), -DT(i - 1) * SS(s, mw, i)));
            if (SS(s, mw, i) != SS(s, mw, i)) nan_speed = 1;
        }
    /* Not in the reference: a NaN speed is reported as a NaN Courant number
     * so that the time loop can stop (see orc3_claw3). */
    *cfl1d = nan_speed ? NAN : cfl;

    if (method[1] != 1) {
        if (limit)
            if (orc_limiter(maxm, meqn, mwaves, mbc, mx, wave, s, mthlim) != 0) return -1;
        for (int i = 2 - mbc; i <= mx + mbc; i++) {
            double dtdxave = 0.5 * (DT(i - 1) + DT(i));
            for (int m = 1; m <= meqn; m++) S1(cqxx, m, i) = 0.0;
            for (int mw = 1; mw <= mwaves; mw++) {
                double sa = fabs(SS(s, mw, i));
                double lead = use_fwave ? 0.5 * copysign(1.0, SS(s, mw, i)) : 0.5 * sa;
                for (int m = 1; m <= meqn; m++)
                    S1(cqxx, m, i) = S1(cqxx, m, i) + lead * (1.0 - sa * dtdxave) * SW(wave, m, mw, i);
            }
            for (int m = 1; m <= meqn; m++) S1(fadd, m, i) = S1(fadd, m, i) + S1(cqxx, m, i);
        }
    }

    if (m3 <= 0) return 0;

#define RPT(icoor, imp, in, outm, outp) \
    rpt3(ixyz, icoor, imp, maxm, meqn, mwaves, maux, mbc, mx, q1d, q1d, aux1, aux2, aux3, in, outm, outp)
#define RPTT(icoor, imp, impt, in, outm, outp) \
    rptt3(ixyz, icoor, imp, impt, maxm, meqn, mwaves, maux, mbc, mx, q1d, q1d, aux1, aux2, aux3, in, outm, outp)

    RPT(2, 1, amdq, bmamdq, bpamdq);
    RPT(2, 2, apdq, bmapdq, bpapdq);
    RPT(3, 1, amdq, cmamdq, cpamdq);
    RPT(3, 2, apdq, cmapdq, cpapdq);

    if (m3 == 2) {
        if (maux > 0) {
            RPT(2, 1, cqxx, bmcqxxm, bpcqxxm);
            RPT(2, 2, cqxx, bmcqxxp, bpcqxxp);
            /* flux3.f90:331 passes an undefined scalar as aux3 here; the
             * z-like transverse split reads only aux2, so the correct slice
             * is passed. */
            RPT(3, 1, cqxx, cmcqxxm, cpcqxxm);
            RPT(3, 2, cqxx, cmcqxxp, cpcqxxp);
        } else {
            RPT(2, 0, cqxx, bmcqxxm, bpcqxxm);
            RPT(3, 0, cqxx, cmcqxxm, cpcqxxm);
            for (int i = 0; i <= mx + 2; i++)
                for (int m = 1; m <= meqn; m++) {
                    S1(bmcqxxp, m, i) = S1(bmcqxxm, m, i);
                    S1(bpcqxxp, m, i) = S1(bpcqxxm, m, i);
                    S1(cmcqxxp, m, i) = S1(cmcqxxm, m, i);
                    S1(cpcqxxp, m, i) = S1(cpcqxxm, m, i);
                }
        }
    }

    if (m4 == 1) {
        for (int i = 0; i <= mx + 2; i++)
            for (int m = 1; m <= meqn; m++) {
                S1(cpapdq2, m, i) = S1(cpapdq, m, i);
                S1(cpamdq2, m, i) = S1(cpamdq, m, i);
                S1(cmapdq2, m, i) = S1(cmapdq, m, i);
                S1(cmamdq2, m, i) = S1(cmamdq, m, i);
            }
    } else if (m4 == 2) {
        for (int i = 0; i <= mx + 2; i++)
            for (int m = 1; m <= meqn; m++) {
                S1(cpapdq2, m, i) = S1(cpapdq, m, i) - 3.0 * S1(cpcqxxp, m, i);
                S1(cpamdq2, m, i) = S1(cpamdq, m, i) + 3.0 * S1(cpcqxxm, m, i);
                S1(cmapdq2, m, i) = S1(cmapdq, m, i) - 3.0 * S1(cmcqxxp, m, i);
                S1(cmamdq2, m, i) = S1(cmamdq, m, i) + 3.0 * S1(cmcqxxm, m, i);
            }
    }

    if (m4 > 0) {
        RPTT(2, 2, 2, cpapdq2, bmcpapdq, bpcpapdq);
        RPTT(2, 1, 2, cpamdq2, bmcpamdq, bpcpamdq);
        RPTT(2, 2, 1, cmapdq2, bmcmapdq, bpcmapdq);
        RPTT(2, 1, 1, cmamdq2, bmcmamdq, bpcmamdq);
    }

    /* G fluxes (flux3.f90:378-467): per interface i, the A^- side updates
     * position i-1 and the A^+ side position i, in the reference's order */
    for (int i = 1; i <= mx + 1; i++)
        for (int m = 1; m <= meqn; m++) {
            GH(gadd, m, 1, 0, i - 1) = GH(gadd, m, 1, 0, i - 1) - 0.5 * DT(i - 1) * S1(bmamdq, m, i);
            GH(gadd, m, 2, 0, i - 1) = GH(gadd, m, 2, 0, i - 1) - 0.5 * DT(i - 1) * S1(bpamdq, m, i);
            GH(gadd, m, 1, 0, i) = GH(gadd, m, 1, 0, i) - 0.5 * DT(i) * S1(bmapdq, m, i);
            GH(gadd, m, 2, 0, i) = GH(gadd, m, 2, 0, i) - 0.5 * DT(i) * S1(bpapdq, m, i);
            if (m4 > 0) {
                double cp_ = sixth * DT(i) * dtdz, cm_ = sixth * DT(i - 1) * dtdz;
                GH(gadd, m, 2, 0, i) = GH(gadd, m, 2, 0, i) + cp_ * (S1(bpcpapdq, m, i) - S1(bpcmapdq, m, i));
                GH(gadd, m, 1, 0, i) = GH(gadd, m, 1, 0, i) + cp_ * (S1(bmcpapdq, m, i) - S1(bmcmapdq, m, i));
                GH(gadd, m, 2, 1, i) = GH(gadd, m, 2, 1, i) - cp_ * S1(bpcpapdq, m, i);
                GH(gadd, m, 1, 1, i) = GH(gadd, m, 1, 1, i) - cp_ * S1(bmcpapdq, m, i);
                GH(gadd, m, 2, -1, i) = GH(gadd, m, 2, -1, i) + cp_ * S1(bpcmapdq, m, i);
                GH(gadd, m, 1, -1, i) = GH(gadd, m, 1, -1, i) + cp_ * S1(bmcmapdq, m, i);
                GH(gadd, m, 2, 0, i - 1) = GH(gadd, m, 2, 0, i - 1) + cm_ * (S1(bpcpamdq, m, i) - S1(bpcmamdq, m, i));
                GH(gadd, m, 1, 0, i - 1) = GH(gadd, m, 1, 0, i - 1) + cm_ * (S1(bmcpamdq, m, i) - S1(bmcmamdq, m, i));
                GH(gadd, m, 2, 1, i - 1) = GH(gadd, m, 2, 1, i - 1) - cm_ * S1(bpcpamdq, m, i);
                GH(gadd, m, 1, 1, i - 1) = GH(gadd, m, 1, 1, i - 1) - cm_ * S1(bmcpamdq, m, i);
                GH(gadd, m, 2, -1, i - 1) = GH(gadd, m, 2, -1, i - 1) + cm_ * S1(bpcmamdq, m, i);
                GH(gadd, m, 1, -1, i - 1) = GH(gadd, m, 1, -1, i - 1) + cm_ * S1(bmcmamdq, m, i);
            }
            if (m3 < 2) continue;
            GH(gadd, m, 2, 0, i) = GH(gadd, m, 2, 0, i) + DT(i) * S1(bpcqxxp, m, i);
            GH(gadd, m, 1, 0, i) = GH(gadd, m, 1, 0, i) + DT(i) * S1(bmcqxxp, m, i);
            GH(gadd, m, 2, 0, i - 1) = GH(gadd, m, 2, 0, i - 1) - DT(i - 1) * S1(bpcqxxm, m, i);
            GH(gadd, m, 1, 0, i - 1) = GH(gadd, m, 1, 0, i - 1) - DT(i - 1) * S1(bmcqxxm, m, i);
        }

    if (m4 == 2) {
        for (int i = 0; i <= mx + 2; i++)
            for (int m = 1; m <= meqn; m++) {
                S1(bpapdq, m, i) = S1(bpapdq, m, i) - 3.0 * S1(bpcqxxp, m, i);
                S1(bpamdq, m, i) = S1(bpamdq, m, i) + 3.0 * S1(bpcqxxm, m, i);
                S1(bmapdq, m, i) = S1(bmapdq, m, i) - 3.0 * S1(bmcqxxp, m, i);
                S1(bmamdq, m, i) = S1(bmamdq, m, i) + 3.0 * S1(bmcqxxm, m, i);
            }
    }

    /* the z-like splits of the y-like parts reuse the storage of the
     * previous double-transverse results, as the reference does */
    if (m4 > 0) {
        RPTT(3, 2, 2, bpapdq, bmcpapdq, bpcpapdq);
        RPTT(3, 1, 2, bpamdq, bmcpamdq, bpcpamdq);
        RPTT(3, 2, 1, bmapdq, bmcmapdq, bpcmapdq);
        RPTT(3, 1, 1, bmamdq, bmcmamdq, bpcmamdq);
    }

    /* H fluxes (flux3.f90:489-608) */
    for (int i = 1; i <= mx + 1; i++)
        for (int m = 1; m <= meqn; m++) {
            GH(hadd, m, 1, 0, i - 1) = GH(hadd, m, 1, 0, i - 1) - 0.5 * DT(i - 1) * S1(cmamdq, m, i);
            GH(hadd, m, 2, 0, i - 1) = GH(hadd, m, 2, 0, i - 1) - 0.5 * DT(i - 1) * S1(cpamdq, m, i);
            GH(hadd, m, 1, 0, i) = GH(hadd, m, 1, 0, i) - 0.5 * DT(i) * S1(cmapdq, m, i);
            GH(hadd, m, 2, 0, i) = GH(hadd, m, 2, 0, i) - 0.5 * DT(i) * S1(cpapdq, m, i);
            if (m4 > 0) {
                double cp_ = sixth * DT(i) * dtdy, cm_ = sixth * DT(i - 1) * dtdy;
                GH(hadd, m, 2, 0, i) = GH(hadd, m, 2, 0, i) + cp_ * (S1(bpcpapdq, m, i) - S1(bpcmapdq, m, i));
                GH(hadd, m, 1, 0, i) = GH(hadd, m, 1, 0, i) + cp_ * (S1(bmcpapdq, m, i) - S1(bmcmapdq, m, i));
                GH(hadd, m, 2, 1, i) = GH(hadd, m, 2, 1, i) - cp_ * S1(bpcpapdq, m, i);
                GH(hadd, m, 1, 1, i) = GH(hadd, m, 1, 1, i) - cp_ * S1(bmcpapdq, m, i);
                GH(hadd, m, 2, -1, i) = GH(hadd, m, 2, -1, i) + cp_ * S1(bpcmapdq, m, i);
                GH(hadd, m, 1, -1, i) = GH(hadd, m, 1, -1, i) + cp_ * S1(bmcmapdq, m, i);
                GH(hadd, m, 2, 0, i - 1) = GH(hadd, m, 2, 0, i - 1) + cm_ * (S1(bpcpamdq, m, i) - S1(bpcmamdq, m, i));
                GH(hadd, m, 1, 0, i - 1) = GH(hadd, m, 1, 0, i - 1) + cm_ * (S1(bmcpamdq, m, i) - S1(bmcmamdq, m, i));
                GH(hadd, m, 2, 1, i - 1) = GH(hadd, m, 2, 1, i - 1) - cm_ * S1(bpcpamdq, m, i);
                GH(hadd, m, 1, 1, i - 1) = GH(hadd, m, 1, 1, i - 1) - cm_ * S1(bmcpamdq, m, i);
                GH(hadd, m, 2, -1, i - 1) = GH(hadd, m, 2, -1, i - 1) + cm_ * S1(bpcmamdq, m, i);
                GH(hadd, m, 1, -1, i - 1) = GH(hadd, m, 1, -1, i - 1) + cm_ * S1(bmcmamdq, m, i);
            }
            if (m3 < 2) continue;
            GH(hadd, m, 2, 0, i) = GH(hadd, m, 2, 0, i) + DT(i) * S1(cpcqxxp, m, i);
            GH(hadd, m, 1, 0, i) = GH(hadd, m, 1, 0, i) + DT(i) * S1(cmcqxxp, m, i);
            GH(hadd, m, 2, 0, i - 1) = GH(hadd, m, 2, 0, i - 1) - DT(i - 1) * S1(cpcqxxm, m, i);
            GH(hadd, m, 1, 0, i - 1) = GH(hadd, m, 1, 0, i - 1) - DT(i - 1) * S1(cmcqxxm, m, i);
        }
    return 0;
#undef RPT
#undef RPTT
#undef DT
}

/* ----------------------------------------------------------------- step3 */

typedef struct {
    int meqn, mbc, mx, my, mz, maux;
} dims3;

static inline size_t cell3(const dims3* d, int i, int j, int k)
{
    return ((size_t)(k + d->mbc - 1) * (d->my + 2 * d->mbc) + (size_t)(j + d->mbc - 1)) *
               (d->mx + 2 * d->mbc) +
           (size_t)(i + d->mbc - 1);
}

/* The nine cells a slice position updates, with the step3.f90 expressions.
 * a, b = offsets of the target cell in the y-like (t1) and z-like (t2)
 * directions of the slice; dn, d1, d2 = dt/dx ratios along the slice, t1 and
 * t2.  Without capa the updates are two sequential operations on qnew
 * (step3.f90:264-335, 522-579, 763-823); with capa the flux terms are summed
 * and divided by the capacity (:337-424, 580-672, 824-866). */
static void apply9(double* qn, const double* gadd, const double* hadd, const double* qadd,
                   const double* fadd, int meqn, int mbc, int i, int a, int b, double dn,
                   double d1, double d2, double kappa, int capa)
{
    for (int m = 1; m <= meqn; m++) {
        double q = qn[m - 1];
#define G_(side, sl) GH(gadd, m, side, sl, i)
#define H_(side, sl) GH(hadd, m, side, sl, i)
        if (!capa) {
            if (b == -1) {
                if (a == -1) q = q - d1 * G_(1, -1) - d2 * H_(1, -1);
                else if (a == 0) q = q - d1 * (G_(2, -1) - G_(1, -1)) - d2 * H_(1, 0);
                else q = q + d1 * G_(2, -1) - d2 * H_(1, 1);
            } else if (b == 0) {
                if (a == -1) q = q - d1 * G_(1, 0) - d2 * (H_(2, -1) - H_(1, -1));
                else if (a == 0)
                    q = q + S1(qadd, m, i) - dn * (S1(fadd, m, i + 1) - S1(fadd, m, i)) -
                        d1 * (G_(2, 0) - G_(1, 0)) - d2 * (H_(2, 0) - H_(1, 0));
                else q = q + d1 * G_(2, 0) - d2 * (H_(2, 1) - H_(1, 1));
            } else {
                if (a == -1) q = q - d1 * G_(1, 1) + d2 * H_(2, -1);
                else if (a == 0) q = q - d1 * (G_(2, 1) - G_(1, 1)) + d2 * H_(2, 0);
                else q = q + d1 * G_(2, 1) + d2 * H_(2, 1);
            }
        } else {
            if (b == -1) {
                if (a == -1) q = q - (d1 * G_(1, -1) + d2 * H_(1, -1)) / kappa;
                else if (a == 0) q = q - (d1 * (G_(2, -1) - G_(1, -1)) + d2 * H_(1, 0)) / kappa;
                else q = q + (d1 * G_(2, -1) - d2 * H_(1, 1)) / kappa;
            } else if (b == 0) {
                if (a == -1) q = q - (d1 * G_(1, 0) + d2 * (H_(2, -1) - H_(1, -1))) / kappa;
                else if (a == 0)
                    q = q + S1(qadd, m, i) -
                        (dn * (S1(fadd, m, i + 1) - S1(fadd, m, i)) + d1 * (G_(2, 0) - G_(1, 0)) +
                         d2 * (H_(2, 0) - H_(1, 0))) / kappa;
                else q = q + (d1 * G_(2, 0) - d2 * (H_(2, 1) - H_(1, 1))) / kappa;
            } else {
                if (a == -1) q = q - (d1 * G_(1, 1) - d2 * H_(2, -1)) / kappa;
                else if (a == 0) q = q - (d1 * (G_(2, 1) - G_(1, 1)) - d2 * H_(2, 0)) / kappa;
                else q = q + (d1 * G_(2, 1) + d2 * H_(2, 1)) / kappa;
            }
        }
#undef G_
#undef H_
        qn[m - 1] = q;
    }
}

/* One slice: gather q and aux, call flux3, scatter into qnew.
 * dir = 1, 2, 3; (p1, p2) = the slice's fixed indices: (j, k) for x, (i, k)
 * for y, (i, j) for z. */
static int slice3(int dir, int p1, int p2, const dims3* d, int maxm, int mwaves,
                  const double* qold, double* qnew, const double* aux, double dtdx, double dtdy,
                  double dtdz, const int* method, const int* mthlim, int use_fwave, int split,
                  double* cfl, orc_rpn3_t rpn3, orc_rpt3_t rpt3, orc_rptt3_t rptt3,
                  step3_work* sw)
{
    const int meqn = d->meqn, mbc = d->mbc, maux = d->maux;
    const int index_capa = method[5];
    int n = dir == 1 ? d->mx : dir == 2 ? d->my : d->mz;
    double dn = dir == 1 ? dtdx : dir == 2 ? dtdy : dtdz;
    double d1 = dir == 1 ? dtdy : dir == 2 ? dtdz : dtdx; /* y-like */
    double d2 = dir == 1 ? dtdz : dir == 2 ? dtdx : dtdy; /* z-like */
    double* q1d = sw->q1d;
    double* dt1d = sw->dtdx1d;
#define AT(l, u, v) (dir == 1 ? cell3(d, l, u, v) : dir == 2 ? cell3(d, u, l, v) : cell3(d, u, v, l))
    /* AT(l, p1, p2) is position l of this slice; for the neighbour slices
     * the transverse offsets (a in t1, b in t2) map to (p1, p2) offsets: x:
     * (j+a, k+b); y: t1 = z, t2 = x -> (i+b, k+a); z: t1 = x, t2 = y ->
     * (i+a, j+b). */
    for (int l = 1 - mbc; l <= n + mbc; l++)
        for (int m = 1; m <= meqn; m++) S1(q1d, m, l) = qold[AT(l, p1, p2) * meqn + m - 1];
    for (int l = 1 - mbc; l <= maxm + mbc; l++) dt1d[l + mbc - 1] = dn;
    if (index_capa > 0)
        for (int l = 1 - mbc; l <= n + mbc; l++)
            dt1d[l + mbc - 1] = dn / aux[AT(l, p1, p2) * maux + index_capa - 1];
    if (maux > 0) {
        double* A[3] = {sw->aux1, sw->aux2, sw->aux3};
        for (int r = 0; r < 3; r++)
            for (int kk = 1; kk <= 3; kk++) {
                if (split && (r != 1)) continue;
                for (int l = 1 - mbc; l <= n + mbc; l++) {
                    size_t c;
                    /* step3: x: aux1/2/3 = rows j-1/j/j+1, kk = plane k-1..k+1;
                     * y: aux1/2/3 = planes k-1/k/k+1, kk = column i-1..i+1;
                     * z: aux1/2/3 = columns i-1/i/i+1, kk = row j-1..j+1.
                     * step3ds only fills aux2 with the kk neighbours. */
                    if (dir == 1) c = cell3(d, l, p1 + r - 1, p2 + kk - 2);
                    else if (dir == 2) c = cell3(d, p1 + kk - 2, l, p2 + r - 1);
                    else c = cell3(d, p1 + r - 1, p2 + kk - 2, l);
                    for (int ma = 1; ma <= maux; ma++) AX(A[r], ma, l, kk) = aux[c * maux + ma - 1];
                }
            }
    }
    double cfl1d;
    if (orc3_flux3(dir, maxm, meqn, mwaves, mbc, n, q1d, dt1d, d1, d2, sw->aux1, sw->aux2,
                   sw->aux3, maux, method, mthlim, sw->qadd, sw->fadd, sw->gadd, sw->hadd,
                   &cfl1d, &sw->fw, rpn3, rpt3, rptt3, use_fwave) != 0)
        return -1;
    *cfl = (cfl1d != cfl1d || *cfl != *cfl) ? NAN : fmax(*cfl, cfl1d);

    if (split) { /* step3ds.f90: only the slice itself changes */
        for (int l = 1; l <= n; l++) {
            size_t c = AT(l, p1, p2);
            for (int m = 1; m <= meqn; m++) {
                double* qp = &qnew[c * meqn + m - 1];
                if (index_capa == 0)
                    *qp = *qp + S1(sw->qadd, m, l) - dn * (S1(sw->fadd, m, l + 1) - S1(sw->fadd, m, l));
                else
                    *qp = *qp + S1(sw->qadd, m, l) -
                          dn * (S1(sw->fadd, m, l + 1) - S1(sw->fadd, m, l)) / aux[c * maux + index_capa - 1];
            }
        }
        return 0;
    }
    for (int b = -1; b <= 1; b++)
        for (int a = -1; a <= 1; a++)
            for (int l = 1; l <= n; l++) {
                size_t c;
                if (dir == 1) c = cell3(d, l, p1 + a, p2 + b);
                else if (dir == 2) c = cell3(d, p1 + b, l, p2 + a);
                else c = cell3(d, p1 + a, p2 + b, l);
                double kappa = index_capa > 0 ? aux[c * maux + index_capa - 1] : 1.0;
                apply9(&qnew[c * meqn], sw->gadd, sw->hadd, sw->qadd, sw->fadd, meqn, mbc, l, a,
                       b, dn, d1, d2, kappa, index_capa > 0);
            }
    return 0;
#undef AT
}

/* One pass of a sweep direction of step3.f90:127-866: the 2x2 blocks of
 * slices whose block indices (outer ob, inner ib) are oo, oo+2, ... and io,
 * io+2, ..., the slices inside a block with the outer index slowest; slice3
 * takes (inner, outer) = (j, k) for x, (i, k) for y, (i, j) for z.  The
 * increments of a slice reach the slices next to it only, so the blocks of
 * one pass touch disjoint cells: with orc_threads() > 1 they run on OpenMP
 * threads (as the reference's own OpenMP loop does, step3.f90:186) and each
 * cell still receives its terms in the serial order, bit for bit. */
static int pass3(int dir, int oo, int io, int nbo, int nbin, int pomax, int pimax,
                 const dims3* d, int maxm, int mwaves, const double* qold, double* qnew,
                 const double* aux, double dtdx, double dtdy, double dtdz, const int* method,
                 const int* mthlim, int use_fwave, double* cfl, orc_rpn3_t rpn3,
                 orc_rpt3_t rpt3, orc_rptt3_t rptt3, step3_work* sw)
{
#define BLOCK_SLICES(W, CFL, FAIL)                                                             \
    for (int po = 2 * ob; po <= (2 * ob + 1 < pomax ? 2 * ob + 1 : pomax); po++)               \
        for (int pi = 2 * ib; pi <= (2 * ib + 1 < pimax ? 2 * ib + 1 : pimax); pi++)           \
            if (slice3(dir, pi, po, d, maxm, mwaves, qold, qnew, aux, dtdx, dtdy, dtdz, method, \
                       mthlim, use_fwave, 0, CFL, rpn3, rpt3, rptt3, W))                        \
                FAIL;
#ifdef _OPENMP
    if (orc_threads() > 1) {
        int bad = 0;
#pragma omp parallel num_threads(orc_threads())
        {
            step3_work w;
            int ok = step3_work_alloc(&w, maxm, d->meqn, mwaves, d->maux, d->mbc) == 0;
            double c = 0.0;
#pragma omp for collapse(2) schedule(dynamic)
            for (int ob = oo; ob <= nbo - 1; ob += 2)
                for (int ib = io; ib <= nbin - 1; ib += 2) {
                    if (!ok) continue;
                    BLOCK_SLICES(&w, &c, ok = 0)
                }
#pragma omp critical
            {
                *cfl = (c != c || *cfl != *cfl) ? NAN : fmax(*cfl, c);
                if (!ok) bad = 1;
            }
            if (w.block) free(w.block);
        }
        return bad ? -1 : 0;
    }
#endif
    for (int ob = oo; ob <= nbo - 1; ob += 2)
        for (int ib = io; ib <= nbin - 1; ib += 2) {
            BLOCK_SLICES(sw, cfl, return -1)
        }
    return 0;
#undef BLOCK_SLICES
}

/* step3.f90:127-866: x sweeps (slices (j, k)), y sweeps ((i, k)), z sweeps
 * ((i, j)), each as four passes over 2x2 blocks of slices in the reference's
 * serial order. */
static int orc3_step3_w(const dims3* d, int maxm, int mwaves, const double* qold, double* qnew,
                        const double* aux, double dx, double dy, double dz, double dt,
                        const int* method, const int* mthlim, double* cfl_out, int use_fwave,
                        orc_rpn3_t rpn3, orc_rpt3_t rpt3, orc_rptt3_t rptt3, step3_work* sw)
{
    double cfl = 0.0;
    double dtdx = dt / dx, dtdy = dt / dy, dtdz = dt / dz;
    const int nbi = (d->mx + 3) / 2, nbj = (d->my + 3) / 2, nbk = (d->mz + 3) / 2;
    /* (direction, outer blocks, inner blocks, last outer slice, last inner slice) */
    const int sweeps[3][5] = {{1, nbk, nbj, d->mz + 1, d->my + 1},
                              {2, nbk, nbi, d->mz + 1, d->mx + 1},
                              {3, nbj, nbi, d->my + 1, d->mx + 1}};
    for (int sdir = 0; sdir < 3; sdir++)
        for (int oo = 0; oo <= 1; oo++)
            for (int io = 0; io <= 1; io++)
                if (pass3(sweeps[sdir][0], oo, io, sweeps[sdir][1], sweeps[sdir][2],
                          sweeps[sdir][3], sweeps[sdir][4], d, maxm, mwaves, qold, qnew, aux,
                          dtdx, dtdy, dtdz, method, mthlim, use_fwave, &cfl, rpn3, rpt3, rptt3,
                          sw))
                    return -1;
    *cfl_out = cfl;
    return 0;
}

/* step3ds.f90:3-408: one direction, every slice including the ghost ones,
 * qnew may alias qold (y and z sweeps of dimsp3 run in place). */
static int orc3_step3ds_w(const dims3* d, int maxm, int mwaves, const double* qold,
                          double* qnew, const double* aux, double dx, double dy, double dz,
                          double dt, const int* method, const int* mthlim, double* cfl_out,
                          int idir, int use_fwave, orc_rpn3_t rpn3, orc_rpt3_t rpt3,
                          orc_rptt3_t rptt3, step3_work* sw)
{
    double cfl = 0.0;
    double dtdx = dt / dx, dtdy = dt / dy, dtdz = dt / dz;
    if (idir == 1) {
        for (int k = 0; k <= d->mz + 1; k++)
            for (int j = 0; j <= d->my + 1; j++)
                if (slice3(1, j, k, d, maxm, mwaves, qold, qnew, aux, dtdx, dtdy, dtdz, method,
                           mthlim, use_fwave, 1, &cfl, rpn3, rpt3, rptt3, sw))
                    return -1;
    } else if (idir == 2) {
        for (int k = 0; k <= d->mz + 1; k++)
            for (int i = 0; i <= d->mx + 1; i++)
                if (slice3(2, i, k, d, maxm, mwaves, qold, qnew, aux, dtdx, dtdy, dtdz, method,
                           mthlim, use_fwave, 1, &cfl, rpn3, rpt3, rptt3, sw))
                    return -1;
    } else {
        for (int j = 0; j <= d->my + 1; j++)
            for (int i = 0; i <= d->mx + 1; i++)
                if (slice3(3, i, j, d, maxm, mwaves, qold, qnew, aux, dtdx, dtdy, dtdz, method,
                           mthlim, use_fwave, 1, &cfl, rpn3, rpt3, rptt3, sw))
                    return -1;
    }
    *cfl_out = cfl;
    return 0;
}

/* dimsp3.f:5-85 */
static int orc3_dimsp3_w(const dims3* d, int maxm, int mwaves, const double* qold, double* qnew,
                         const double* aux, double dx, double dy, double dz, double dt,
                         const int* method, const int* mthlim, double* cfl, const double* cflv,
                         int use_fwave, orc_rpn3_t rpn3, orc_rpt3_t rpt3, orc_rptt3_t rptt3,
                         step3_work* sw)
{
    double cflx, cfly, cflz;
    if (orc3_step3ds_w(d, maxm, mwaves, qold, qnew, aux, dx, dy, dz, dt, method, mthlim, &cflx,
                       1, use_fwave, rpn3, rpt3, rptt3, sw))
        return -1;
    if (cflx > cflv[0] || cflx != cflx) { *cfl = cflx; return 0; }
    if (orc3_step3ds_w(d, maxm, mwaves, qnew, qnew, aux, dx, dy, dz, dt, method, mthlim, &cfly,
                       2, use_fwave, rpn3, rpt3, rptt3, sw))
        return -1;
    if (cfly > cflv[0] || cfly != cfly) { *cfl = cfly; return 0; }
    if (orc3_step3ds_w(d, maxm, mwaves, qnew, qnew, aux, dx, dy, dz, dt, method, mthlim, &cflz,
                       3, use_fwave, rpn3, rpt3, rptt3, sw))
        return -1;
    *cfl = cflz != cflz ? NAN : fmax(cflx, fmax(cfly, cflz));
    return 0;
}

static int get_solver3(int id, orc_rpn3_t* n, orc_rpt3_t* t, orc_rptt3_t* tt)
{
    if (id != ORC3_SOLVER_VC_ACOUSTICS) return -2;
    *n = orc_rpn3_vc_acoustics;
    *t = orc_rpt3_vc_acoustics;
    *tt = orc_rptt3_vc_acoustics;
    return 0;
}

int orc3_step3(int meqn, int mwaves, int maux, int mbc, int mx, int my, int mz,
               const double* qold, double* qnew, const double* aux, double dx, double dy,
               double dz, double dt, const int* method, const int* mthlim, int use_fwave,
               int solver_id, double* cfl)
{
    orc_rpn3_t rpn3; orc_rpt3_t rpt3; orc_rptt3_t rptt3;
    if (get_solver3(solver_id, &rpn3, &rpt3, &rptt3)) return -2;
    dims3 d = {meqn, mbc, mx, my, mz, maux};
    int maxm = mx > my ? (mx > mz ? mx : mz) : (my > mz ? my : mz);
    step3_work sw;
    if (step3_work_alloc(&sw, maxm, meqn, mwaves, maux, mbc)) return -3;
    int rc = orc3_step3_w(&d, maxm, mwaves, qold, qnew, aux, dx, dy, dz, dt, method, mthlim, cfl,
                          use_fwave, rpn3, rpt3, rptt3, &sw);
    free(sw.block);
    return rc;
}

int orc3_dimsp3(int meqn, int mwaves, int maux, int mbc, int mx, int my, int mz,
                const double* qold, double* qnew, const double* aux, double dx, double dy,
                double dz, double dt, const int* method, const int* mthlim, const double* cflv,
                int use_fwave, int solver_id, double* cfl)
{
    orc_rpn3_t rpn3; orc_rpt3_t rpt3; orc_rptt3_t rptt3;
    if (get_solver3(solver_id, &rpn3, &rpt3, &rptt3)) return -2;
    dims3 d = {meqn, mbc, mx, my, mz, maux};
    int maxm = mx > my ? (mx > mz ? mx : mz) : (my > mz ? my : mz);
    step3_work sw;
    if (step3_work_alloc(&sw, maxm, meqn, mwaves, maux, mbc)) return -3;
    int rc = orc3_dimsp3_w(&d, maxm, mwaves, qold, qnew, aux, dx, dy, dz, dt, method, mthlim, cfl,
                           cflv, use_fwave, rpn3, rpt3, rptt3, &sw);
    free(sw.block);
    return rc;
}

/* ------------------------------------------------------------------- bc3 */

/* src/3d/bc3.f: xlower, xupper, ylower, yupper, zlower, zupper in that
 * order, each over the full extent of the other two indices (ghosts
 * included), so later faces overwrite edges and corners.  mthbc(6) = 0 is
 * the solid wall with the moving piston of the heterogeneous-acoustics
 * example (examples/acoustics_3d_heterogeneous/bc3.f, label 600-606). */
int orc3_bc3(int meqn, int mbc, int mx, int my, int mz, const orc3_geom* g, double* q,
             double t, const int* mthbc)
{
    dims3 d = {meqn, mbc, mx, my, mz, 0};
    for (int f = 0; f < 6; f++)
        if (mthbc[f] < (f == 5 ? 0 : 1) || mthbc[f] > 3) return -1;
    const int lo[3] = {1 - mbc, 1 - mbc, 1 - mbc};
    const int hi[3] = {mx + mbc, my + mbc, mz + mbc};
    const int nn[3] = {mx, my, mz};
    for (int f = 0; f < 6; f++) {
        int ax = f / 2, upper = f % 2, code = mthbc[f];
        int b1 = ax == 0 ? 1 : 0, b2 = ax == 2 ? 1 : 2; /* the other two axes */
        for (int u = lo[b1]; u <= hi[b1]; u++)
            for (int v = lo[b2]; v <= hi[b2]; v++)
                for (int ib = 1; ib <= mbc; ib++) {
                    int ghost = upper ? nn[ax] + ib : 1 - ib;
                    int src;
                    if (code == 1) src = upper ? nn[ax] : 1;
                    else if (code == 2) src = upper ? ib : nn[ax] + 1 - ib;
                    else src = upper ? nn[ax] + 1 - ib : ib; /* 3 and the piston wall */
                    int ig[3], is[3];
                    ig[ax] = ghost; is[ax] = src;
                    ig[b1] = is[b1] = u;
                    ig[b2] = is[b2] = v;
                    size_t cg = cell3(&d, ig[0], ig[1], ig[2]), cs = cell3(&d, is[0], is[1], is[2]);
                    for (int m = 0; m < meqn; m++) q[cg * meqn + m] = q[cs * meqn + m];
                    if (code == 3) q[cg * meqn + ax + 1] = -q[cs * meqn + ax + 1];
                    if (code == 0) {
                        double s = 0.0;
                        if (t < 0.1) {
                            double xi = g->xlower + (ig[0] - 0.5) * g->dx;
                            double yj = g->ylower + (ig[1] - 0.5) * g->dy;
                            double r2 = (xi * xi + yj * yj);
                            s = -exp(-500.0 * r2) * exp(-(500.0 * (t * t))); /* -500*t**2 */
                            if (fabs(s) < 1e-6) s = 0.0;
                        }
                        q[cg * meqn + 3] = 2.0 * s - q[cs * meqn + 3];
                    }
                }
    }
    return 0;
}

/* ------------------------------------------------------------------ claw3 */

/* chkmth.f:3-72 */
static int chkmth3(const int* method)
{
    int info = 0;
    if (method[0] != 0 && method[0] != 1) info = 6;
    if (method[1] != 1 && method[1] != 2) info = 6;
    if (method[1] == 1 && method[2] >= 0)
        if (method[2] != 0 && method[2] != 10 && method[2] != 11) info = 6;
    if (method[1] == 2 && method[2] >= 0)
        if (method[2] != 0 && method[2] != 10 && method[2] != 11 && method[2] != 20 &&
            method[2] != 21 && method[2] != 22)
            info = 6;
    return info;
}

/* claw3.f:596-904.  Sources and b4step3 are the reference's no-op defaults,
 * so method(5) only changes the work-array size there. */
int orc3_claw3(int meqn, int mwaves, int maux, int mbc, int mx, int my, int mz, double* q,
               const double* aux, const orc3_geom* g, double tstart, double* tend, double* dtv,
               double* cflv, int* nv, const int* method, const int* mthlim, const int* mthbc,
               int use_fwave, int solver_id, double* dt_trace, double* cfl_trace, int trace_cap,
               int* ntrace)
{
    orc_rpn3_t rpn3; orc_rpt3_t rpt3; orc_rptt3_t rptt3;
    if (get_solver3(solver_id, &rpn3, &rpt3, &rptt3)) return -2;
    dims3 d = {meqn, mbc, mx, my, mz, maux};
    int maxm = mx > my ? (mx > mz ? mx : mz) : (my > mz ? my : mz);
    int info = 0;
    double t = tstart;
    int maxn = nv[0];
    double dt = dtv[0];
    double cflmax = 0.0, dtmin = dt, dtmax = dt, cfl = 0.0;
    double* qwork = NULL;
    step3_work sw = {0};
    nv[1] = 0;
    if (ntrace) *ntrace = 0;

    info = chkmth3(method);
    if (info == 6) goto done;
    if (mbc < 2) { info = 1; goto done; }
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

    size_t nq = (size_t)(mx + 2 * mbc) * (my + 2 * mbc) * (mz + 2 * mbc) * meqn;
    qwork = (double*)malloc(nq * sizeof(double));
    if (!qwork || step3_work_alloc(&sw, maxm, meqn, mwaves, maux, mbc) != 0) {
        free(qwork);
        return -3;
    }
    if (maxn == 0) goto freed;
    for (int n = 1; n <= maxn; n++) {
        double told = t;
        if (told + dt > *tend && tstart < *tend) dt = *tend - told;
    retake:
        if (orc3_bc3(meqn, mbc, mx, my, mz, g, q, told, mthbc) != 0) { info = -4; goto freed; }
        memcpy(qwork, q, nq * sizeof(double));
        int rc;
        if (method[2] >= 0)
            rc = orc3_step3_w(&d, maxm, mwaves, qwork, q, aux, g->dx, g->dy, g->dz, dt, method,
                              mthlim, &cfl, use_fwave, rpn3, rpt3, rptt3, &sw);
        else
            rc = orc3_dimsp3_w(&d, maxm, mwaves, qwork, q, aux, g->dx, g->dy, g->dz, dt, method,
                               mthlim, &cfl, cflv, use_fwave, rpn3, rpt3, rptt3, &sw);
        if (rc != 0) { info = -1; goto freed; }
        /* Not in the reference: a NaN Courant number would retake forever */
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
