/*
 * oracle/ref_shims3_ez.c -- TEST INFRASTRUCTURE ONLY.
 *
 * qinit / setaux / setprob for running the reference driver claw3ez
 * (src/3d/claw3ez.f) unmodified from the tests: qinit and setaux copy arrays
 * registered with ref3_ez_set_init (reference layout, ghost cells included).
 * The Riemann solvers come from ref_shims3.c (rpn3_/rpt3_/rptt3_).
 */
#include <string.h>

static const double* g_q0 = 0;
static long g_nq = 0;
static const double* g_aux0 = 0;
static long g_naux = 0;

void ref3_ez_set_init(const double* q, long nq, const double* aux, long naux)
{
    g_q0 = q;
    g_nq = nq;
    g_aux0 = aux;
    g_naux = naux;
}

void qinit_(int* meqn, int* mbc, int* mx, int* my, int* mz, double* xlower, double* ylower,
            double* zlower, double* dx, double* dy, double* dz, double* q, int* maux, double* aux)
{
    (void)meqn; (void)mbc; (void)mx; (void)my; (void)mz; (void)xlower; (void)ylower;
    (void)zlower; (void)dx; (void)dy; (void)dz; (void)maux; (void)aux;
    if (g_q0) memcpy(q, g_q0, (size_t)g_nq * sizeof(double));
}

void setaux_(int* mbc, int* mx, int* my, int* mz, double* xlower, double* ylower, double* zlower,
             double* dx, double* dy, double* dz, int* maux, double* aux)
{
    (void)mbc; (void)mx; (void)my; (void)mz; (void)xlower; (void)ylower; (void)zlower; (void)dx;
    (void)dy; (void)dz; (void)maux;
    if (g_aux0) memcpy(aux, g_aux0, (size_t)g_naux * sizeof(double));
}

void setprob_(void) {}

extern void claw3ez_(void);
extern void ref_close_unit(int iu);

void ref_claw3ez(void)
{
    claw3ez_();
    ref_close_unit(10);
}
