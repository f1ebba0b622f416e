/*
 * oracle/ref_shims1.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Glue that lets the reference 1D Fortran (src/1d, compiled into oracle/_ref
 * by oracle/Makefile) run with this project's rp1 solvers: a Fortran-ABI
 * wrapper around the oracle's rp1 passed to claw1/step1 as the external
 * procedure of the plug-in contract, and C entry points sizing the work
 * array like src/1d/claw1ez.f.
 */
#include <stdlib.h>

#include "claw1d_oracle.h"

extern void claw1_(int* meqn, int* mwaves, int* maux, int* mbc, int* mx, double* q, double* aux,
                   double* xlower, double* dx, double* tstart, double* tend, double* dtv,
                   double* cflv, int* nv, int* method, int* mthlim, int* mthbc, double* work,
                   int* mwork, int* use_fwave, int* info, void* bc1, void* rp1, void* src1,
                   void* b4step1);
extern void step1_(int* meqn, int* mwaves, int* mbc, int* maux, int* mx, double* q, double* aux,
                   double* dx, double* dt, int* method, int* mthlim, double* cfl, double* f,
                   double* wave, double* s, double* amdq, double* apdq, double* dtdx,
                   int* use_fwave, void* rp1);
extern void bc1_(int* meqn, int* mbc, int* mx, double* xlower, double* dx, double* q, int* maux,
                 double* aux, double* t, double* dt, int* mthbc);
extern void src1_(void);
extern void b4step1_(void);

static int g_solver1 = ORC1_SOLVER_ACOUSTICS;

static void f_rp1(int* maxm, int* meqn, int* mwaves, int* maux, int* mbc, int* mx, double* ql,
                  double* qr, double* auxl, double* auxr, double* wave, double* s, double* amdq,
                  double* apdq)
{
    orc1_get_rp1(g_solver1)(*maxm, *meqn, *mwaves, *maux, *mbc, *mx, ql, qr, auxl, auxr, wave,
                            s, amdq, apdq);
}

/* the same wrapper under the external name claw1ez passes to claw1 */
void rp1_(int* maxm, int* meqn, int* mwaves, int* maux, int* mbc, int* mx, double* ql, double* qr,
          double* auxl, double* auxr, double* wave, double* s, double* amdq, double* apdq)
{
    f_rp1(maxm, meqn, mwaves, maux, mbc, mx, ql, qr, auxl, auxr, wave, s, amdq, apdq);
}

void ref1_set_solver(int solver_id, const orc1_params* p)
{
    g_solver1 = solver_id;
    orc1_set_params(p);
}

int ref1_claw1(int meqn, int mwaves, int maux, int mbc, int mx, double* q, double* aux,
               double xlower, double dx, double tstart, double* tend, double* dtv, double* cflv,
               int* nv, int* method, int* mthlim, int* mthbc, int use_fwave)
{
    /* claw1ez.f: mwork = (mx+2mbc)*(2+4*meqn+mwaves+meqn*mwaves) */
    int mwork = (mx + 2 * mbc) * (2 + 4 * meqn + mwaves + meqn * mwaves);
    double* work = (double*)calloc((size_t)mwork, sizeof(double));
    double dummy_aux[1] = {0.0};
    int info = 0, fw = use_fwave ? 1 : 0;
    if (!work) return -3;
    claw1_(&meqn, &mwaves, &maux, &mbc, &mx, q, maux > 0 ? aux : dummy_aux, &xlower, &dx,
           &tstart, tend, dtv, cflv, nv, method, mthlim, mthbc, work, &mwork, &fw, &info,
           (void*)bc1_, (void*)f_rp1, (void*)src1_, (void*)b4step1_);
    free(work);
    return info;
}

int ref1_step1(int meqn, int mwaves, int maux, int mbc, int mx, double* q, double* aux,
               double dx, double dt, int* method, int* mthlim, int use_fwave, double* cfl)
{
    size_t n = (size_t)(mx + 2 * mbc);
    double* b = (double*)calloc(n * (size_t)(meqn * (3 + mwaves) + mwaves + 1), sizeof(double));
    double dummy_aux[1] = {0.0};
    int fw = use_fwave ? 1 : 0;
    if (!b) return -3;
    double* f = b;
    double* wave = f + n * meqn;
    double* s = wave + n * meqn * mwaves;
    double* amdq = s + n * mwaves;
    double* apdq = amdq + n * meqn;
    double* dtdx = apdq + n * meqn;
    step1_(&meqn, &mwaves, &mbc, &maux, &mx, q, maux > 0 ? aux : dummy_aux, &dx, &dt, method,
           mthlim, cfl, f, wave, s, amdq, apdq, dtdx, &fw, (void*)f_rp1);
    free(b);
    return 0;
}

int ref1_bc1(int meqn, int mbc, int mx, double* q, int* mthbc)
{
    double z = 0.0, dummy_aux[1] = {0.0};
    int maux = 0;
    bc1_(&meqn, &mbc, &mx, &z, &z, q, &maux, dummy_aux, &z, &z, mthbc);
    return 0;
}

/* Test user routines for the plug-in contract: b4step1 and src1 append
 * (kind, t, dt) to a log (kind 1 = b4step1, 2 = src1); src1 applies the decay
 * q(1) <- q(1) - (k dt) q(1) on the interior cells.  ref1_claw1_hooks runs
 * the reference claw1 with them. */
static double g_hook1_k = 0.0;
static double* g_hook1_log = 0;
static int g_hook1_n = 0, g_hook1_cap = 0;

void ref1_hooks_set(double k, double* log, int cap)
{
    g_hook1_k = k;
    g_hook1_log = log;
    g_hook1_cap = cap;
    g_hook1_n = 0;
}

int ref1_hooks_count(void) { return g_hook1_n; }

static void hook1_log(double kind, double t, double dt)
{
    if (g_hook1_n < g_hook1_cap) {
        g_hook1_log[3 * g_hook1_n] = kind;
        g_hook1_log[3 * g_hook1_n + 1] = t;
        g_hook1_log[3 * g_hook1_n + 2] = dt;
    }
    g_hook1_n++;
}

static void test_b4step1(int* mbc, int* mx, int* meqn, double* q, double* xlower, double* dx,
                         double* t, double* dt, int* maux, double* aux)
{
    (void)mbc; (void)mx; (void)meqn; (void)q; (void)xlower; (void)dx; (void)maux; (void)aux;
    hook1_log(1.0, *t, *dt);
}

static void test_src1(int* meqn, int* mbc, int* mx, double* xlower, double* dx, double* q,
                      int* maux, double* aux, double* t, double* dt)
{
    (void)xlower; (void)dx; (void)maux; (void)aux;
    hook1_log(2.0, *t, *dt);
    const double kd = g_hook1_k * *dt;
    for (int i = 1; i <= *mx; i++) {
        double* v = &q[(long)(i + *mbc - 1) * *meqn];
        v[0] = v[0] - kd * v[0];
    }
}

int ref1_claw1_hooks(int meqn, int mwaves, int maux, int mbc, int mx, double* q, double* aux,
                     double xlower, double dx, double tstart, double* tend, double* dtv,
                     double* cflv, int* nv, int* method, int* mthlim, int* mthbc, int use_fwave)
{
    int mwork = (mx + 2 * mbc) * (2 + 4 * meqn + mwaves + meqn * mwaves);
    double* work = (double*)calloc((size_t)mwork, sizeof(double));
    double dummy_aux[1] = {0.0};
    int info = 0, fw = use_fwave ? 1 : 0;
    if (!work) return -3;
    claw1_(&meqn, &mwaves, &maux, &mbc, &mx, q, maux > 0 ? aux : dummy_aux, &xlower, &dx,
           &tstart, tend, dtv, cflv, nv, method, mthlim, mthbc, work, &mwork, &fw, &info,
           (void*)bc1_, (void*)f_rp1, (void*)test_src1, (void*)test_b4step1);
    free(work);
    return info;
}
