/*
 * oracle/ref_shims1_ez.c -- TEST INFRASTRUCTURE ONLY.
 *
 * qinit / setaux / setprob for running the reference driver claw1ez
 * (src/1d/claw1ez.f) unmodified from the tests: qinit and setaux copy arrays
 * registered with ref1_ez_set_init (reference layout, ghost cells included).
 * The Riemann solver comes from ref_shims1.c (rp1_).
 */
#include <string.h>

static const double* g_q0 = 0;
static long g_nq = 0;
static const double* g_aux0 = 0;
static long g_naux = 0;

void ref1_ez_set_init(const double* q, long nq, const double* aux, long naux)
{
    g_q0 = q;
    g_nq = nq;
    g_aux0 = aux;
    g_naux = naux;
}

void qinit_(int* meqn, int* mbc, int* mx, double* xlower, double* dx, double* q, int* maux,
            double* aux)
{
    (void)meqn; (void)mbc; (void)mx; (void)xlower; (void)dx; (void)maux; (void)aux;
    if (g_q0) memcpy(q, g_q0, (size_t)g_nq * sizeof(double));
}

void setaux_(int* mbc, int* mx, double* xlower, double* dx, int* maux, double* aux)
{
    (void)mbc; (void)mx; (void)xlower; (void)dx; (void)maux;
    if (g_aux0) memcpy(aux, g_aux0, (size_t)g_naux * sizeof(double));
}

void setprob_(void) {}

extern void claw1ez_(void);
extern void ref_close_unit(int iu);

void ref_claw1ez(void)
{
    claw1ez_();
    ref_close_unit(10);
}
