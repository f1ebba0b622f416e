/*
 * oracle/ref_shims_ez.c -- TEST INFRASTRUCTURE ONLY.
 *
 * The user routines that the reference driver claw2ez (src/2d/claw2ez.f)
 * calls -- qinit, setaux, setprob -- for running that driver unmodified
 * from the tests: qinit and setaux copy arrays registered beforehand with
 * ref_ez_set_init (reference layout, ghost cells included), setprob does
 * nothing.  The Riemann solvers come from ref_shims.c (rpn2_/rpt2_).
 */
#include <string.h>

static const double* g_q0 = 0;
static long g_nq = 0;
static const double* g_aux0 = 0;
static long g_naux = 0;

void ref_ez_set_init(const double* q, long nq, const double* aux, long naux)
{
    g_q0 = q;
    g_nq = nq;
    g_aux0 = aux;
    g_naux = naux;
}

void qinit_(int* meqn, int* mbc, int* mx, int* my, double* xlower, double* ylower, double* dx,
            double* dy, double* q, int* maux, double* aux)
{
    (void)meqn; (void)mbc; (void)mx; (void)my; (void)xlower; (void)ylower; (void)dx; (void)dy;
    (void)maux; (void)aux;
    if (g_q0) memcpy(q, g_q0, (size_t)g_nq * sizeof(double));
}

void setaux_(int* mbc, int* mx, int* my, double* xlower, double* ylower, double* dx, double* dy,
             int* maux, double* aux)
{
    (void)mbc; (void)mx; (void)my; (void)xlower; (void)ylower; (void)dx; (void)dy; (void)maux;
    if (g_aux0) memcpy(aux, g_aux0, (size_t)g_naux * sizeof(double));
}

void setprob_(void) {}

extern void claw2ez_(void);
extern void ref_close_unit(int iu);

/* Run the reference driver in the current directory (claw.data must exist:
 * the reference stops the process otherwise) and close its fort.info unit so
 * the file is complete and the unit free for the next run. */
void ref_claw2ez(void)
{
    claw2ez_();
    ref_close_unit(10);
}
