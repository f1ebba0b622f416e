/*
 * oracle/claw1d_oracle.c -- CPU ORACLE for the 1D wave-propagation step.
 *
 * TEST INFRASTRUCTURE ONLY.  A plain-C restatement of the reference Fortran
 * (clawpack/classic) 1D path, used by the tests as the checker; never linked
 * into or called from the product library (classic_amd/).  It keeps the
 * reference's loop order and expression order so that a build with
 * -ffp-contract=off reproduces the Fortran bit for bit.
 *
 * Functions and the reference code each one restates:
 *   orc1_step1   src/1d/step1.f90 (Godunov update in place, Courant number,
 *                limiter, second-order fluxes f, capacity dtdx(i))
 *   orc1_bc1     src/1d/bc1.f (codes 1 extrap, 2 periodic, 3 wall)
 *   orc1_claw1   src/1d/claw1.f (time loop: the new dt is chosen before the
 *                Courant check, a rejected step restores q and retakes
 *                without clamping dt to tend again; info codes 2, 3, 11, 12)
 *   limiter      orc_limiter of claw2d_oracle.c (src/1d/inlinelimiter.f90 is
 *                identical to the 2D file)
 * Riemann solvers: the reference fixes only the rp1 contract (src/1d/claw1.f
 * header; call site step1.f90); the solvers its examples name live in the
 * absent clawpack/riemann repository (version unpinned) and are restated from
 * LeVeque, FVMHP: rp1_acoustics (2.8 / 3.9, constant rho, K), rp1_acoustics
 * _variable (9.9, aux = (Z, c)), rp1_advection (constant u).  Pinning: the
 * examples' regression frame sums (acoustics_1d_example1,
 * acoustics_1d_heterogeneous), the reference Fortran compiled from
 * /root/reference (oracle/_ref/libclawref1d.so), closed forms.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "claw2d_oracle.h"
#include "claw1d_oracle.h"

#define S1(a, m, i) (a)[(size_t)((i) + mbc - 1) * meqn + ((m) - 1)]
#define SW(w, m, mw, i) (w)[((size_t)((i) + mbc - 1) * mwaves + ((mw) - 1)) * meqn + ((m) - 1)]
#define SS(s, mw, i) (s)[(size_t)((i) + mbc - 1) * mwaves + ((mw) - 1)]
#define AU(a, ma, i) (a)[(size_t)((i) + mbc - 1) * maux + ((ma) - 1)]

static orc1_params g1 = {1.0, 4.0, 2.0, 2.0, 1.0};

void orc1_set_params(const orc1_params* p) { g1 = *p; }

/* constant-coefficient acoustics, q = (p, u): waves (-Z, 1) at -c, (Z, 1) at c */
void orc_rp1_acoustics(int maxm, int meqn, int mwaves, int maux, int mbc, int mx,
                       const double* ql, const double* qr, const double* auxl,
                       const double* auxr, double* wave, double* s, double* amdq, double* apdq)
{
    (void)maxm; (void)maux; (void)auxl; (void)auxr;
    const double zz = g1.zz, cc = g1.cc;
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        double delta1 = S1(ql, 1, i) - S1(qr, 1, i - 1);
        double delta2 = S1(ql, 2, i) - S1(qr, 2, i - 1);
        double a1 = (-delta1 + zz * delta2) / (2.0 * zz);
        double a2 = (delta1 + zz * delta2) / (2.0 * zz);
        SW(wave, 1, 1, i) = -a1 * zz;
        SW(wave, 2, 1, i) = a1;
        SS(s, 1, i) = -cc;
        SW(wave, 1, 2, i) = a2 * zz;
        SW(wave, 2, 2, i) = a2;
        SS(s, 2, i) = cc;
    }
    for (int i = 2 - mbc; i <= mx + mbc; i++)
        for (int m = 1; m <= meqn; m++) {
            S1(amdq, m, i) = SS(s, 1, i) * SW(wave, m, 1, i);
            S1(apdq, m, i) = SS(s, 2, i) * SW(wave, m, 2, i);
        }
}

/* variable-coefficient acoustics, aux = (Z, c): left-going wave in cell i-1,
 * right-going wave in cell i */
void orc_rp1_acoustics_variable(int maxm, int meqn, int mwaves, int maux, int mbc, int mx,
                                const double* ql, const double* qr, const double* auxl,
                                const double* auxr, double* wave, double* s, double* amdq,
                                double* apdq)
{
    (void)maxm;
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        double delta1 = S1(ql, 1, i) - S1(qr, 1, i - 1);
        double delta2 = S1(ql, 2, i) - S1(qr, 2, i - 1);
        double zi = AU(auxl, 1, i), zim = AU(auxr, 1, i - 1);
        double a1 = (-delta1 + zi * delta2) / (zim + zi);
        double a2 = (delta1 + zim * delta2) / (zim + zi);
        SW(wave, 1, 1, i) = -a1 * zim;
        SW(wave, 2, 1, i) = a1;
        SS(s, 1, i) = -AU(auxr, 2, i - 1);
        SW(wave, 1, 2, i) = a2 * zi;
        SW(wave, 2, 2, i) = a2;
        SS(s, 2, i) = AU(auxl, 2, i);
    }
    for (int i = 2 - mbc; i <= mx + mbc; i++)
        for (int m = 1; m <= meqn; m++) {
            S1(amdq, m, i) = SS(s, 1, i) * SW(wave, m, 1, i);
            S1(apdq, m, i) = SS(s, 2, i) * SW(wave, m, 2, i);
        }
}

/* scalar advection q_t + u q_x = 0 */
void orc_rp1_advection(int maxm, int meqn, int mwaves, int maux, int mbc, int mx,
                       const double* ql, const double* qr, const double* auxl,
                       const double* auxr, double* wave, double* s, double* amdq, double* apdq)
{
    (void)maxm; (void)maux; (void)auxl; (void)auxr; (void)mwaves;
    const double u = g1.u;
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        SW(wave, 1, 1, i) = S1(ql, 1, i) - S1(qr, 1, i - 1);
        SS(s, 1, i) = u;
        S1(amdq, 1, i) = fmin(u, 0.0) * SW(wave, 1, 1, i);
        S1(apdq, 1, i) = fmax(u, 0.0) * SW(wave, 1, 1, i);
    }
}

orc_rp1_t orc1_get_rp1(int solver_id)
{
    switch (solver_id) {
    case ORC1_SOLVER_ACOUSTICS: return orc_rp1_acoustics;
    case ORC1_SOLVER_VC_ACOUSTICS: return orc_rp1_acoustics_variable;
    case ORC1_SOLVER_ADVECTION: return orc_rp1_advection;
    default: return 0;
    }
}

typedef struct {
    double *f, *wave, *s, *amdq, *apdq, *dtdx;
    void* block;
} step1_work;

static int step1_work_alloc(step1_work* w, int meqn, int mwaves, int mbc, int mx)
{
    size_t n = (size_t)(mx + 2 * mbc);
    double* b = (double*)calloc(n * (size_t)(meqn * (3 + mwaves) + mwaves + 1), sizeof(double));
    if (!b) return -1;
    w->block = b;
    w->f = b; b += n * meqn;
    w->wave = b; b += n * meqn * mwaves;
    w->s = b; b += n * mwaves;
    w->amdq = b; b += n * meqn;
    w->apdq = b; b += n * meqn;
    w->dtdx = b;
    return 0;
}

/* step1.f90: q is updated in place; f is the caller's scratch (the
 * reference reuses claw1's work array, so f(mx+2) -- read for the ghost cell
 * mx+1 only -- keeps whatever an earlier step left there). */
static int orc1_step1_w(int meqn, int mwaves, int maux, int mbc, int mx, double* q,
                        const double* aux, double dx, double dt, const int* method,
                        const int* mthlim, double* cfl_out, int use_fwave, orc_rp1_t rp1,
                        step1_work* w)
{
    double *f = w->f, *wave = w->wave, *s = w->s, *amdq = w->amdq, *apdq = w->apdq;
#define DT(i) w->dtdx[(i) + mbc - 1]
    int limit = 0;
    for (int mw = 0; mw < mwaves; mw++)
        if (mthlim[mw] > 0) limit = 1;
    const int index_capa = method[5];
    for (int i = 1 - mbc; i <= mx + mbc; i++) {
        if (index_capa > 0) {
            if (AU(aux, index_capa, i) <= 0.0) return -2; /* reference stops */
            DT(i) = dt / (dx * AU(aux, index_capa, i));
        } else {
            DT(i) = dt / dx;
        }
    }
    rp1(mx, meqn, mwaves, maux, mbc, mx, q, q, aux, aux, wave, s, amdq, apdq);
    for (int i = 1; i <= mx + 1; i++) {
        for (int m = 1; m <= meqn; m++) S1(q, m, i - 1) = S1(q, m, i - 1) - DT(i - 1) * S1(amdq, m, i);
        for (int m = 1; m <= meqn; m++) S1(q, m, i) = S1(q, m, i) - DT(i) * S1(apdq, m, i);
    }
    double cfl = 0.0;
    int nan_speed = 0;
    for (int i = 1; i <= mx + 1; i++)
        for (int mw = 1; mw <= mwaves; mw++) {
            cfl = fmax(cfl, fmax(DT(i) * SS(s, mw, i), -DT(i - 1) * SS(s, mw, i)));
            if (SS(s, mw, i) != SS(s, mw, i)) nan_speed = 1;
        }
    /* Not in the reference: a NaN speed gives a NaN Courant number. */
    *cfl_out = nan_speed ? NAN : cfl;
    if (method[1] == 1) return 0;
    if (limit)
        if (orc_limiter(mx, meqn, mwaves, mbc, mx, wave, s, mthlim) != 0) return -1;
    for (int i = 1; i <= mx + 1; i++) {
        for (int m = 1; m <= meqn; m++) S1(f, m, i) = 0.0;
        double dtdxave = 0.5 * (DT(i - 1) + DT(i));
        for (int mw = 1; mw <= mwaves; mw++) {
            double sa = fabs(SS(s, mw, i));
            double lead = use_fwave ? 0.5 * copysign(1.0, SS(s, mw, i)) : 0.5 * sa;
            for (int m = 1; m <= meqn; m++)
                S1(f, m, i) = S1(f, m, i) + lead * (1.0 - sa * dtdxave) * SW(wave, m, mw, i);
        }
    }
    for (int i = 1; i <= mx + 1; i++)
        for (int m = 1; m <= meqn; m++)
            S1(q, m, i) = S1(q, m, i) - DT(i) * (S1(f, m, i + 1) - S1(f, m, i));
    return 0;
#undef DT
}

int orc1_step1(int meqn, int mwaves, int maux, int mbc, int mx, double* q, const double* aux,
               double dx, double dt, const int* method, const int* mthlim, int use_fwave,
               int solver_id, double* cfl)
{
    orc_rp1_t rp1 = orc1_get_rp1(solver_id);
    if (!rp1) return -2;
    step1_work w;
    if (step1_work_alloc(&w, meqn, mwaves, mbc, mx)) return -3;
    int rc = orc1_step1_w(meqn, mwaves, maux, mbc, mx, q, aux, dx, dt, method, mthlim, cfl,
                          use_fwave, rp1, &w);
    free(w.block);
    return rc;
}

/* bc1.f: left then right; the wall negates q(2) */
int orc1_bc1(int meqn, int mbc, int mx, double* q, const int* mthbc)
{
    for (int side = 0; side < 2; side++)
        if (mthbc[side] < 1 || mthbc[side] > 3) return -1;
    for (int ibc = 1; ibc <= mbc; ibc++) {
        for (int m = 1; m <= meqn; m++) {
            if (mthbc[0] == 1) S1(q, m, 1 - ibc) = S1(q, m, 1);
            else if (mthbc[0] == 2) S1(q, m, 1 - ibc) = S1(q, m, mx + 1 - ibc);
            else S1(q, m, 1 - ibc) = S1(q, m, ibc);
        }
        if (mthbc[0] == 3) S1(q, 2, 1 - ibc) = -S1(q, 2, ibc);
    }
    for (int ibc = 1; ibc <= mbc; ibc++) {
        for (int m = 1; m <= meqn; m++) {
            if (mthbc[1] == 1) S1(q, m, mx + ibc) = S1(q, m, mx);
            else if (mthbc[1] == 2) S1(q, m, mx + ibc) = S1(q, m, ibc);
            else S1(q, m, mx + ibc) = S1(q, m, mx + 1 - ibc);
        }
        if (mthbc[1] == 3) S1(q, 2, mx + ibc) = -S1(q, 2, mx + 1 - ibc);
    }
    return 0;
}

/* claw1.f time loop.  The reference's work array starts zeroed (claw1ez
 * allocates it), which is what f(mx+2) reads on the first step. */
int orc1_claw1(int meqn, int mwaves, int maux, int mbc, int mx, double* q, const double* aux,
               double dx, double tstart, double* tend, double* dtv, double* cflv, int* nv,
               const int* method, const int* mthlim, const int* mthbc, int use_fwave,
               int solver_id, double* dt_trace, double* cfl_trace, int trace_cap, int* ntrace)
{
    orc_rp1_t rp1 = orc1_get_rp1(solver_id);
    if (!rp1) return -2;
    int info = 0;
    double t = tstart;
    int maxn = nv[0];
    double dt = dtv[0];
    double cflmax = 0.0, dtmin = dt, dtmax = dt, cfl = 0.0;
    double* qwork = NULL;
    step1_work w = {0};
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
    size_t nq = (size_t)(mx + 2 * mbc) * meqn;
    qwork = (double*)calloc(nq, sizeof(double));
    if (!qwork || step1_work_alloc(&w, meqn, mwaves, mbc, mx)) {
        free(qwork);
        return -3;
    }
    if (maxn == 0) goto freed;
    for (int n = 1; n <= maxn; n++) {
        double told = t;
        if (told + dt > *tend && tstart < *tend) dt = *tend - told;
        if (method[0] == 1) memcpy(qwork, q, nq * sizeof(double));
    retake:
        t = told + dt;
        if (orc1_bc1(meqn, mbc, mx, q, mthbc)) { info = -4; goto freed; }
        int rc = orc1_step1_w(meqn, mwaves, maux, mbc, mx, q, aux, dx, dt, method, mthlim, &cfl,
                              use_fwave, rp1, &w);
        if (rc) { info = -1; goto freed; }
        if (cfl != cfl) { info = -5; goto freed; } /* not in the reference: NaN speed */
        if (ntrace && *ntrace < trace_cap) {
            dt_trace[*ntrace] = dt;
            cfl_trace[*ntrace] = cfl;
            (*ntrace)++;
        }
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
        } else {
            t = told;
            memcpy(q, qwork, nq * sizeof(double));
            if (method[0] == 1) goto retake;
            cflmax = fmax(cfl, cflmax);
            goto freed;
        }
        nv[1] = nv[1] + 1;
        if (t >= *tend) break;
    }
freed:
    free(qwork);
    free(w.block);
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
