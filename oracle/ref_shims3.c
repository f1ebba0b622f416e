/*
 * oracle/ref_shims3.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Glue that lets the reference 3D Fortran (compiled from /root/reference into
 * oracle/_ref by oracle/Makefile) run with this project's variable-coefficient
 * acoustics solvers:
 *   - Fortran-ABI wrappers around the oracle's rpn3/rpt3/rptt3, passed to the
 *     reference claw3/step3/dimsp3 as the external procedures of the plug-in
 *     contract (src/3d/claw3.f:292-516);
 *   - C entry points that size and zero the work array as src/3d/claw3ez.f
 *     does and call the reference subroutines.  The piston boundary condition
 *     is the example's own bc3 (examples/acoustics_3d_heterogeneous/bc3.f),
 *     compiled with its symbol renamed to bc3piston_.
 */
#include <stdlib.h>

#include "claw3d_oracle.h"

extern void claw3_(int* meqn, int* mwaves, int* mbc, int* mx, int* my, int* mz, int* maux,
                   double* q, double* aux, double* xlower, double* ylower, double* zlower,
                   double* dx, double* dy, double* dz, double* tstart, double* tend, double* dtv,
                   double* cflv, int* nv, int* method, int* mthlim, int* mthbc, double* work,
                   int* mwork, int* use_fwave, int* info, void* bc3, void* rpn3, void* rpt3,
                   void* rptt3, void* src3, void* b4step3);
extern void step3_(int* maxm, int* meqn, int* mwaves, int* mbc, int* mx, int* my, int* mz,
                   double* qold, double* qnew, double* aux, double* dx, double* dy, double* dz,
                   double* dt, int* method, int* mthlim, double* cfl, double* qadd, double* fadd,
                   double* gadd, double* hadd, double* q1d, double* dtdx1d, double* dtdy1d,
                   double* dtdz1d, double* aux1, double* aux2, double* aux3, int* maux,
                   double* work, int* mwork, int* use_fwave, void* rpn3, void* rpt3,
                   void* rptt3);
extern void dimsp3_(int* maxm, int* meqn, int* mwaves, int* mbc, int* mx, int* my, int* mz,
                    double* qold, double* qnew, double* aux, double* dx, double* dy, double* dz,
                    double* dt, int* method, int* mthlim, double* cfl, double* cflv, double* qadd,
                    double* fadd, double* gadd, double* hadd, double* q1d, double* dtdx1d,
                    double* dtdy1d, double* dtdz1d, double* aux1, double* aux2, double* aux3,
                    int* maux, double* work, int* mwork, int* use_fwave, void* rpn3, void* rpt3,
                    void* rptt3);
extern void bc3_(int* meqn, int* mbc, int* mx, int* my, int* mz, double* xlower, double* ylower,
                 double* zlower, double* dx, double* dy, double* dz, double* q, int* maux,
                 double* aux, double* t, double* dt, int* mthbc);
extern void bc3piston_(int* meqn, int* mbc, int* mx, int* my, int* mz, double* xlower,
                       double* ylower, double* zlower, double* dx, double* dy, double* dz,
                       double* q, int* maux, double* aux, double* t, double* dt, int* mthbc);
extern void src3_(void);
extern void b4step3_(void);

static void f_rpn3(int* ixyz, int* maxm, int* meqn, int* mwaves, int* maux, int* mbc, int* mx,
                   double* ql, double* qr, double* auxl, double* auxr, double* wave, double* s,
                   double* amdq, double* apdq)
{
    orc_rpn3_vc_acoustics(*ixyz, *maxm, *meqn, *mwaves, *maux, *mbc, *mx, ql, qr, auxl, auxr,
                          wave, s, amdq, apdq);
}

static void f_rpt3(int* ixyz, int* icoor, int* imp, int* maxm, int* meqn, int* mwaves,
                   int* maux, int* mbc, int* mx, double* ql, double* qr, double* aux1,
                   double* aux2, double* aux3, double* asdq, double* bmasdq, double* bpasdq)
{
    orc_rpt3_vc_acoustics(*ixyz, *icoor, *imp, *maxm, *meqn, *mwaves, *maux, *mbc, *mx, ql, qr,
                          aux1, aux2, aux3, asdq, bmasdq, bpasdq);
}

static void f_rptt3(int* ixyz, int* icoor, int* imp, int* impt, int* maxm, int* meqn,
                    int* mwaves, int* maux, int* mbc, int* mx, double* ql, double* qr,
                    double* aux1, double* aux2, double* aux3, double* bsasdq, double* cmbsasdq,
                    double* cpbsasdq)
{
    orc_rptt3_vc_acoustics(*ixyz, *icoor, *imp, *impt, *maxm, *meqn, *mwaves, *maux, *mbc, *mx,
                           ql, qr, aux1, aux2, aux3, bsasdq, cmbsasdq, cpbsasdq);
}

/* the same wrappers under the external names claw3ez passes to claw3 */
void rpn3_(int* ixyz, int* maxm, int* meqn, int* mwaves, int* maux, int* mbc, int* mx, double* ql,
           double* qr, double* auxl, double* auxr, double* wave, double* s, double* amdq,
           double* apdq)
{
    f_rpn3(ixyz, maxm, meqn, mwaves, maux, mbc, mx, ql, qr, auxl, auxr, wave, s, amdq, apdq);
}

void rpt3_(int* ixyz, int* icoor, int* imp, int* maxm, int* meqn, int* mwaves, int* maux,
           int* mbc, int* mx, double* ql, double* qr, double* aux1, double* aux2, double* aux3,
           double* asdq, double* bmasdq, double* bpasdq)
{
    f_rpt3(ixyz, icoor, imp, maxm, meqn, mwaves, maux, mbc, mx, ql, qr, aux1, aux2, aux3, asdq,
           bmasdq, bpasdq);
}

void rptt3_(int* ixyz, int* icoor, int* imp, int* impt, int* maxm, int* meqn, int* mwaves,
            int* maux, int* mbc, int* mx, double* ql, double* qr, double* aux1, double* aux2,
            double* aux3, double* bsasdq, double* cmbsasdq, double* cpbsasdq)
{
    f_rptt3(ixyz, icoor, imp, impt, maxm, meqn, mwaves, maux, mbc, mx, ql, qr, aux1, aux2, aux3,
            bsasdq, cmbsasdq, cpbsasdq);
}

static int maxm3(int mx, int my, int mz)
{
    int m = mx > my ? mx : my;
    return m > mz ? m : mz;
}

int ref3_claw3(int meqn, int mwaves, int maux, int mbc, int mx, int my, int mz, double* q,
               double* aux, const orc3_geom* g, double tstart, double* tend, double* dtv,
               double* cflv, int* nv, int* method, int* mthlim, int* mthbc, int use_fwave,
               int piston)
{
    int narray = method[4] < 2 ? 1 : 2;
    int maxm = maxm3(mx, my, mz);
    /* claw3ez.f:193-196 with one thread */
    int mwork = (maxm + 2 * mbc) * (46 * meqn + mwaves + meqn * mwaves + 9 * maux + 3) +
                narray * (mx + 2 * mbc) * (my + 2 * mbc) * (mz + 2 * mbc) * meqn;
    double* work = (double*)calloc((size_t)mwork, sizeof(double));
    double xl = g->xlower, yl = g->ylower, zl = g->zlower, dx = g->dx, dy = g->dy, dz = g->dz;
    int info = 0, fw = use_fwave ? 1 : 0;
    if (!work) return -3;
    claw3_(&meqn, &mwaves, &mbc, &mx, &my, &mz, &maux, q, aux, &xl, &yl, &zl, &dx, &dy, &dz,
           &tstart, tend, dtv, cflv, nv, method, mthlim, mthbc, work, &mwork, &fw, &info,
           piston ? (void*)bc3piston_ : (void*)bc3_, (void*)f_rpn3, (void*)f_rpt3,
           (void*)f_rptt3, (void*)src3_, (void*)b4step3_);
    free(work);
    return info;
}

/* Work arrays of step3/dimsp3 for `nt` threads: the reference indexes every
 * slice array by thread in its last dimension (step3.f90:26-38) and the
 * flux3 scratch in `work` by thread (:71-113). */
static int ref3_step_common(int split, int nt, int meqn, int mwaves, int maux, int mbc, int mx,
                            int my, int mz, double* qold, double* qnew, double* aux, double dx,
                            double dy, double dz, double dt, int* method, int* mthlim,
                            double* cflv, int use_fwave, double* cfl)
{
    int maxm = maxm3(mx, my, mz);
    size_t n = (size_t)(maxm + 2 * mbc);
    int ma = maux > 0 ? maux : 1;
    if (nt < 1) nt = 1;
    int mwork = (int)(n * (size_t)(meqn * mwaves + mwaves + 40 * meqn) * (size_t)nt) + 16;
    size_t tot = n * (size_t)(15 * meqn + 3 + 9 * ma) * (size_t)nt + (size_t)mwork;
    double* b = (double*)calloc(tot, sizeof(double));
    int fw = use_fwave ? 1 : 0;
    if (!b) return -3;
    double* qadd = b;
    double* fadd = qadd + n * meqn * nt;
    double* gadd = fadd + n * meqn * nt;
    double* hadd = gadd + 6 * n * meqn * nt;
    double* q1d = hadd + 6 * n * meqn * nt;
    double* dtdx1d = q1d + n * meqn * nt;
    double* dtdy1d = dtdx1d + n * nt;
    double* dtdz1d = dtdy1d + n * nt;
    double* aux1 = dtdz1d + n * nt;
    double* aux2 = aux1 + 3 * n * ma * nt;
    double* aux3 = aux2 + 3 * n * ma * nt;
    double* work = aux3 + 3 * n * ma * nt;
    if (!split)
        step3_(&maxm, &meqn, &mwaves, &mbc, &mx, &my, &mz, qold, qnew, aux, &dx, &dy, &dz, &dt,
               method, mthlim, cfl, qadd, fadd, gadd, hadd, q1d, dtdx1d, dtdy1d, dtdz1d, aux1,
               aux2, aux3, &maux, work, &mwork, &fw, (void*)f_rpn3, (void*)f_rpt3,
               (void*)f_rptt3);
    else
        dimsp3_(&maxm, &meqn, &mwaves, &mbc, &mx, &my, &mz, qold, qnew, aux, &dx, &dy, &dz, &dt,
                method, mthlim, cfl, cflv, qadd, fadd, gadd, hadd, q1d, dtdx1d, dtdy1d, dtdz1d,
                aux1, aux2, aux3, &maux, work, &mwork, &fw, (void*)f_rpn3, (void*)f_rpt3,
                (void*)f_rptt3);
    free(b);
    return 0;
}

/* step3 with work arrays for nt OpenMP threads (the reference's own OpenMP
 * loop, step3.f90:132-869, in a build compiled with -fopenmp; the caller
 * sets the thread count through OMP_NUM_THREADS) */
int ref3_step3_nt(int nt, int meqn, int mwaves, int maux, int mbc, int mx, int my, int mz,
                  double* qold, double* qnew, double* aux, double dx, double dy, double dz,
                  double dt, int* method, int* mthlim, int use_fwave, double* cfl)
{
    return ref3_step_common(0, nt, meqn, mwaves, maux, mbc, mx, my, mz, qold, qnew, aux, dx, dy,
                            dz, dt, method, mthlim, NULL, use_fwave, cfl);
}

int ref3_step3(int meqn, int mwaves, int maux, int mbc, int mx, int my, int mz, double* qold,
               double* qnew, double* aux, double dx, double dy, double dz, double dt,
               int* method, int* mthlim, int use_fwave, double* cfl)
{
    return ref3_step_common(0, 1, meqn, mwaves, maux, mbc, mx, my, mz, qold, qnew, aux, dx, dy, dz,
                            dt, method, mthlim, NULL, use_fwave, cfl);
}

int ref3_dimsp3(int meqn, int mwaves, int maux, int mbc, int mx, int my, int mz, double* qold,
                double* qnew, double* aux, double dx, double dy, double dz, double dt,
                int* method, int* mthlim, double* cflv, int use_fwave, double* cfl)
{
    return ref3_step_common(1, 1, meqn, mwaves, maux, mbc, mx, my, mz, qold, qnew, aux, dx, dy, dz,
                            dt, method, mthlim, cflv, use_fwave, cfl);
}

int ref3_bc3(int meqn, int mbc, int mx, int my, int mz, const orc3_geom* g, double* q, double t,
             int* mthbc, int piston)
{
    double xl = g->xlower, yl = g->ylower, zl = g->zlower, dx = g->dx, dy = g->dy, dz = g->dz;
    double dt = 0.0, dummy_aux[1] = {0.0};
    int maux = 0;
    if (piston)
        bc3piston_(&meqn, &mbc, &mx, &my, &mz, &xl, &yl, &zl, &dx, &dy, &dz, q, &maux, dummy_aux,
                   &t, &dt, mthbc);
    else
        bc3_(&meqn, &mbc, &mx, &my, &mz, &xl, &yl, &zl, &dx, &dy, &dz, q, &maux, dummy_aux, &t,
             &dt, mthbc);
    return 0;
}

/* Test user routines for the plug-in contract: b4step3 and src3 append
 * (kind, t, dt) to a log (kind 1 = b4step3, 2 = src3); src3 applies the decay
 * q(1) <- q(1) - (k dt) q(1) on the interior cells.  ref3_claw3_hooks runs
 * the reference claw3 with them. */
static double g_hook3_k = 0.0;
static double* g_hook3_log = 0;
static int g_hook3_n = 0, g_hook3_cap = 0;

void ref3_hooks_set(double k, double* log, int cap)
{
    g_hook3_k = k;
    g_hook3_log = log;
    g_hook3_cap = cap;
    g_hook3_n = 0;
}

int ref3_hooks_count(void) { return g_hook3_n; }

static void hook3_log(double kind, double t, double dt)
{
    if (g_hook3_n < g_hook3_cap) {
        g_hook3_log[3 * g_hook3_n] = kind;
        g_hook3_log[3 * g_hook3_n + 1] = t;
        g_hook3_log[3 * g_hook3_n + 2] = dt;
    }
    g_hook3_n++;
}

static void test_b4step3(int* mbc, int* mx, int* my, int* mz, int* meqn, double* q,
                         double* xlower, double* ylower, double* zlower, double* dx, double* dy,
                         double* dz, double* t, double* dt, int* maux, double* aux)
{
    (void)mbc; (void)mx; (void)my; (void)mz; (void)meqn; (void)q; (void)xlower; (void)ylower;
    (void)zlower; (void)dx; (void)dy; (void)dz; (void)maux; (void)aux;
    hook3_log(1.0, *t, *dt);
}

static void test_src3(int* meqn, int* mbc, int* mx, int* my, int* mz, double* xlower,
                      double* ylower, double* zlower, double* dx, double* dy, double* dz,
                      double* q, int* maux, double* aux, double* t, double* dt)
{
    (void)xlower; (void)ylower; (void)zlower; (void)dx; (void)dy; (void)dz; (void)maux;
    (void)aux;
    hook3_log(2.0, *t, *dt);
    const double kd = g_hook3_k * *dt;
    const long nx = *mx + 2 * *mbc, ny = *my + 2 * *mbc;
    for (int k = 1; k <= *mz; k++)
        for (int j = 1; j <= *my; j++)
            for (int i = 1; i <= *mx; i++) {
                const long cell = ((long)(k + *mbc - 1) * ny + (j + *mbc - 1)) * nx + (i + *mbc - 1);
                double* v = &q[cell * *meqn];
                v[0] = v[0] - kd * v[0];
            }
}

int ref3_claw3_hooks(int meqn, int mwaves, int maux, int mbc, int mx, int my, int mz, double* q,
                     double* aux, const orc3_geom* g, double tstart, double* tend, double* dtv,
                     double* cflv, int* nv, int* method, int* mthlim, int* mthbc, int use_fwave)
{
    int narray = method[4] < 2 ? 1 : 2;
    int maxm = maxm3(mx, my, mz);
    int mwork = (maxm + 2 * mbc) * (46 * meqn + mwaves + meqn * mwaves + 9 * maux + 3) +
                narray * (mx + 2 * mbc) * (my + 2 * mbc) * (mz + 2 * mbc) * meqn;
    double* work = (double*)calloc((size_t)mwork, sizeof(double));
    double xl = g->xlower, yl = g->ylower, zl = g->zlower, dx = g->dx, dy = g->dy, dz = g->dz;
    int info = 0, fw = use_fwave ? 1 : 0;
    if (!work) return -3;
    claw3_(&meqn, &mwaves, &mbc, &mx, &my, &mz, &maux, q, aux, &xl, &yl, &zl, &dx, &dy, &dz,
           &tstart, tend, dtv, cflv, nv, method, mthlim, mthbc, work, &mwork, &fw, &info,
           (void*)bc3_, (void*)f_rpn3, (void*)f_rpt3, (void*)f_rptt3, (void*)test_src3,
           (void*)test_b4step3);
    free(work);
    return info;
}
