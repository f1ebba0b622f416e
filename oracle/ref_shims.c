/*
 * oracle/ref_shims.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Glue that lets the reference Fortran (compiled from /root/reference into
 * oracle/_ref by oracle/Makefile) run with this project's Riemann solvers:
 *   - Fortran-ABI wrappers (every argument by reference) around the oracle's
 *     rpn2/rpt2, passed to the reference claw2/step2 as the external
 *     procedures of the plug-in contract (src/2d/claw2.f:312-402);
 *   - C entry points ref_claw2 / ref_step2 / ref_dimsp2 / ref_bc2 that size
 *     and zero the work array exactly as src/2d/claw2ez.f:175-184 does and
 *     call the reference subroutines.
 * The reference's claw2/step2/flux2/limiter/bc2 arithmetic runs unmodified.
 */
#include <stdlib.h>
#include <string.h>

#include "claw2d_oracle.h"

/* reference subroutines (flang appends one underscore) */
extern void claw2_(int* meqn, int* mwaves, int* maux, int* mbc, int* mx, int* my, double* q,
                   double* aux, double* xlower, double* ylower, double* dx, double* dy,
                   double* tstart, double* tend, double* dtv, double* cflv, int* nv,
                   int* method, int* mthlim, int* mthbc, double* work, int* mwork,
                   int* use_fwaves, int* info, void* bc2, void* rpn2, void* rpt2, void* src2,
                   void* b4step2);
extern void step2_(int* maxm, int* meqn, int* mwaves, int* maux, int* mbc, int* mx, int* my,
                   double* qold, double* qnew, double* aux, double* dx, double* dy, double* dt,
                   int* method, int* mthlim, double* cfl, double* qadd, double* fadd,
                   double* gadd, double* q1d, double* dtdx1d, double* dtdy1d, double* aux1,
                   double* aux2, double* aux3, double* work, int* mwork, int* use_fwave,
                   void* rpn2, void* rpt2);
extern void dimsp2_(int* maxm, int* meqn, int* mwaves, int* maux, int* mbc, int* mx, int* my,
                    double* qold, double* qnew, double* aux, double* dx, double* dy, double* dt,
                    int* method, int* mthlim, double* cfl, double* cflv, double* qadd,
                    double* fadd, double* gadd, double* q1d, double* dtdx1d, double* dtdy1d,
                    double* aux1, double* aux2, double* aux3, double* work, int* mwork,
                    int* use_fwave, void* rpn2, void* rpt2);
extern void bc2_(int* meqn, int* mbc, int* mx, int* my, double* xlower, double* ylower,
                 double* dx, double* dy, double* q, int* maux, double* aux, double* t,
                 double* dt, int* mthbc);
extern void src2_(void);
extern void b4step2_(void);

static int g_solver = ORC_SOLVER_ACOUSTICS;

static void f_rpn2(int* ixy, int* maxm, int* meqn, int* mwaves, int* maux, int* mbc, int* mx,
                   double* ql, double* qr, double* auxl, double* auxr, double* wave, double* s,
                   double* amdq, double* apdq)
{
    orc_get_rpn2(g_solver)(*ixy, *maxm, *meqn, *mwaves, *maux, *mbc, *mx, ql, qr, auxl, auxr,
                           wave, s, amdq, apdq);
}

static void f_rpt2(int* ixy, int* imp, int* maxm, int* meqn, int* mwaves, int* maux, int* mbc,
                   int* mx, double* ql, double* qr, double* aux1, double* aux2, double* aux3,
                   double* asdq, double* bmasdq, double* bpasdq)
{
    orc_get_rpt2(g_solver)(*ixy, *imp, *maxm, *meqn, *mwaves, *maux, *mbc, *mx, ql, qr, aux1,
                           aux2, aux3, asdq, bmasdq, bpasdq);
}

/* the same wrappers under the external names the reference drivers pass
 * to claw2 (claw2ez.f:15, :270-273) */
void rpn2_(int* ixy, int* maxm, int* meqn, int* mwaves, int* maux, int* mbc, int* mx, double* ql,
           double* qr, double* auxl, double* auxr, double* wave, double* s, double* amdq,
           double* apdq)
{
    f_rpn2(ixy, maxm, meqn, mwaves, maux, mbc, mx, ql, qr, auxl, auxr, wave, s, amdq, apdq);
}

void rpt2_(int* ixy, int* imp, int* maxm, int* meqn, int* mwaves, int* maux, int* mbc, int* mx,
           double* ql, double* qr, double* aux1, double* aux2, double* aux3, double* asdq,
           double* bmasdq, double* bpasdq)
{
    f_rpt2(ixy, imp, maxm, meqn, mwaves, maux, mbc, mx, ql, qr, aux1, aux2, aux3, asdq, bmasdq,
           bpasdq);
}

void ref_set_solver(int solver_id, const orc_solver_params* p)
{
    g_solver = solver_id;
    orc_set_solver_params(p);
}

static int mwork_of(int meqn, int mwaves, int maux, int mbc, int mx, int my, int narray)
{
    int maxm = mx > my ? mx : my;
    return (maxm + 2 * mbc) * (10 * meqn + mwaves + meqn * mwaves + 3 * maux + 2) +
           narray * (mx + 2 * mbc) * (my + 2 * mbc) * meqn;
}

int ref_claw2(int meqn, int mwaves, int maux, int mbc, int mx, int my, double* q, double* aux,
              double xlower, double ylower, double dx, double dy, double tstart, double* tend,
              double* dtv, double* cflv, int* nv, int* method, int* mthlim, int* mthbc,
              int use_fwave)
{
    int narray = method[4] < 2 ? 1 : 2;
    int mwork = mwork_of(meqn, mwaves, maux, mbc, mx, my, narray);
    double* work = (double*)calloc((size_t)mwork, sizeof(double));
    double dummy_aux[1] = {0.0};
    int info = 0, fw = use_fwave ? 1 : 0;
    if (!work) return -3;
    claw2_(&meqn, &mwaves, &maux, &mbc, &mx, &my, q, maux > 0 ? aux : dummy_aux, &xlower,
           &ylower, &dx, &dy, &tstart, tend, dtv, cflv, nv, method, mthlim, mthbc, work, &mwork,
           &fw, &info, (void*)bc2_, (void*)f_rpn2, (void*)f_rpt2, (void*)src2_,
           (void*)b4step2_);
    free(work);
    return info;
}

/* Test user routines for the plug-in contract: b4step2 and src2 append
 * (kind, t, dt) to a log (kind 1 = b4step2, 2 = src2); src2 applies the decay
 * q(1) <- q(1) - (k dt) q(1) on the interior cells.  ref_claw2_hooks runs the
 * reference claw2 with them. */
static double g_hook_k = 0.0;
static double* g_hook_log = 0;
static int g_hook_n = 0, g_hook_cap = 0;

void ref_hooks_set(double k, double* log, int cap)
{
    g_hook_k = k;
    g_hook_log = log;
    g_hook_cap = cap;
    g_hook_n = 0;
}

int ref_hooks_count(void) { return g_hook_n; }

static void hook_log(double kind, double t, double dt)
{
    if (g_hook_n < g_hook_cap) {
        g_hook_log[3 * g_hook_n] = kind;
        g_hook_log[3 * g_hook_n + 1] = t;
        g_hook_log[3 * g_hook_n + 2] = dt;
    }
    g_hook_n++;
}

static void test_b4step2(int* mbc, int* mx, int* my, int* meqn, double* q, double* xlower,
                         double* ylower, double* dx, double* dy, double* t, double* dt, int* maux,
                         double* aux)
{
    (void)mbc; (void)mx; (void)my; (void)meqn; (void)q; (void)xlower; (void)ylower; (void)dx;
    (void)dy; (void)maux; (void)aux;
    hook_log(1.0, *t, *dt);
}

static void test_src2(int* meqn, int* mbc, int* mx, int* my, double* xlower, double* ylower,
                      double* dx, double* dy, double* q, int* maux, double* aux, double* t,
                      double* dt)
{
    (void)xlower; (void)ylower; (void)dx; (void)dy; (void)maux; (void)aux;
    hook_log(2.0, *t, *dt);
    const double kd = g_hook_k * *dt;
    const long nx = *mx + 2 * *mbc;
    for (int j = 1; j <= *my; j++)
        for (int i = 1; i <= *mx; i++) {
            double* v = &q[((long)(j + *mbc - 1) * nx + (i + *mbc - 1)) * *meqn];
            v[0] = v[0] - kd * v[0];
        }
}

int ref_claw2_hooks(int meqn, int mwaves, int maux, int mbc, int mx, int my, double* q,
                    double* aux, double xlower, double ylower, double dx, double dy, double tstart,
                    double* tend, double* dtv, double* cflv, int* nv, int* method, int* mthlim,
                    int* mthbc, int use_fwave)
{
    int narray = method[4] < 2 ? 1 : 2;
    int mwork = mwork_of(meqn, mwaves, maux, mbc, mx, my, narray);
    double* work = (double*)calloc((size_t)mwork, sizeof(double));
    double dummy_aux[1] = {0.0};
    int info = 0, fw = use_fwave ? 1 : 0;
    if (!work) return -3;
    claw2_(&meqn, &mwaves, &maux, &mbc, &mx, &my, q, maux > 0 ? aux : dummy_aux, &xlower,
           &ylower, &dx, &dy, &tstart, tend, dtv, cflv, nv, method, mthlim, mthbc, work, &mwork,
           &fw, &info, (void*)bc2_, (void*)f_rpn2, (void*)f_rpt2, (void*)test_src2,
           (void*)test_b4step2);
    free(work);
    return info;
}

/* one step2 or dimsp2 call with claw2's work partition (claw2.f:565-587) */
static int ref_step_common(int split, int meqn, int mwaves, int maux, int mbc, int mx, int my,
                           double* qold, double* qnew, double* aux, double dx, double dy,
                           double dt, int* method, int* mthlim, double* cflv, int use_fwave,
                           double* cfl)
{
    int maxm = mx > my ? mx : my;
    int n = maxm + 2 * mbc;
    int ma = maux > 0 ? maux : 1;
    int mwork = n * (meqn * mwaves + mwaves + 5 * meqn) + 16;
    size_t tot = (size_t)n * (4 * meqn + meqn + 2 + 3 * ma) + (size_t)mwork;
    double* b = (double*)calloc(tot, sizeof(double));
    double dummy_aux[1] = {0.0};
    int fw = use_fwave ? 1 : 0;
    if (!b) return -3;
    double* qadd = b;
    double* fadd = qadd + (size_t)n * meqn;
    double* gadd = fadd + (size_t)n * meqn;
    double* q1d = gadd + (size_t)2 * n * meqn;
    double* dtdx1d = q1d + (size_t)n * meqn;
    double* dtdy1d = dtdx1d + n;
    double* aux1 = dtdy1d + n;
    double* aux2 = aux1 + (size_t)n * ma;
    double* aux3 = aux2 + (size_t)n * ma;
    double* work = aux3 + (size_t)n * ma;
    double* a = maux > 0 ? aux : dummy_aux;
    if (!split)
        step2_(&maxm, &meqn, &mwaves, &maux, &mbc, &mx, &my, qold, qnew, a, &dx, &dy, &dt,
               method, mthlim, cfl, qadd, fadd, gadd, q1d, dtdx1d, dtdy1d, aux1, aux2, aux3, work,
               &mwork, &fw, (void*)f_rpn2, (void*)f_rpt2);
    else
        dimsp2_(&maxm, &meqn, &mwaves, &maux, &mbc, &mx, &my, qold, qnew, a, &dx, &dy, &dt,
                method, mthlim, cfl, cflv, qadd, fadd, gadd, q1d, dtdx1d, dtdy1d, aux1, aux2,
                aux3, work, &mwork, &fw, (void*)f_rpn2, (void*)f_rpt2);
    free(b);
    return 0;
}

int ref_step2(int meqn, int mwaves, int maux, int mbc, int mx, int my, double* qold,
              double* qnew, double* aux, double dx, double dy, double dt, int* method,
              int* mthlim, int use_fwave, double* cfl)
{
    return ref_step_common(0, meqn, mwaves, maux, mbc, mx, my, qold, qnew, aux, dx, dy, dt,
                           method, mthlim, NULL, use_fwave, cfl);
}

int ref_dimsp2(int meqn, int mwaves, int maux, int mbc, int mx, int my, double* qold,
               double* qnew, double* aux, double dx, double dy, double dt, int* method,
               int* mthlim, double* cflv, int use_fwave, double* cfl)
{
    return ref_step_common(1, meqn, mwaves, maux, mbc, mx, my, qold, qnew, aux, dx, dy, dt,
                           method, mthlim, cflv, use_fwave, cfl);
}

int ref_bc2(int meqn, int mbc, int mx, int my, double* q, int* mthbc)
{
    double z = 0.0, dummy_aux[1] = {0.0};
    int maux = 0;
    bc2_(&meqn, &mbc, &mx, &my, &z, &z, &z, &z, q, &maux, dummy_aux, &z, &z, mthbc);
    return 0;
}
