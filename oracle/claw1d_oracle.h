/*
 * oracle/claw1d_oracle.h -- CPU ORACLE (test infrastructure only; never part
 * of the product library).  Plain-C restatement of the reference 1D path;
 * see claw1d_oracle.c for the reference file:line each function follows.
 */
#ifndef CLAW1D_ORACLE_H
#define CLAW1D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC1_SOLVER_ACOUSTICS = 1, ORC1_SOLVER_VC_ACOUSTICS = 3, ORC1_SOLVER_ADVECTION = 4 };

typedef struct {
    double rho, bulk, cc, zz; /* constant acoustics */
    double u;                 /* advection speed */
} orc1_params;

typedef void (*orc_rp1_t)(int maxm, int meqn, int mwaves, int maux, int mbc, int mx,
                          const double* ql, const double* qr, const double* auxl,
                          const double* auxr, double* wave, double* s, double* amdq,
                          double* apdq);

void orc1_set_params(const orc1_params* p);
orc_rp1_t orc1_get_rp1(int solver_id);
void orc_rp1_acoustics(int maxm, int meqn, int mwaves, int maux, int mbc, int mx,
                       const double* ql, const double* qr, const double* auxl,
                       const double* auxr, double* wave, double* s, double* amdq, double* apdq);
void orc_rp1_acoustics_variable(int maxm, int meqn, int mwaves, int maux, int mbc, int mx,
                                const double* ql, const double* qr, const double* auxl,
                                const double* auxr, double* wave, double* s, double* amdq,
                                double* apdq);
void orc_rp1_advection(int maxm, int meqn, int mwaves, int maux, int mbc, int mx,
                       const double* ql, const double* qr, const double* auxl,
                       const double* auxr, double* wave, double* s, double* amdq, double* apdq);

int orc1_step1(int meqn, int mwaves, int maux, int mbc, int mx, double* q, const double* aux,
               double dx, double dt, const int* method, const int* mthlim, int use_fwave,
               int solver_id, double* cfl);
int orc1_bc1(int meqn, int mbc, int mx, double* q, const int* mthbc);
int orc1_claw1(int meqn, int mwaves, int maux, int mbc, int mx, double* q, const double* aux,
               double dx, double tstart, double* tend, double* dtv, double* cflv, int* nv,
               const int* method, const int* mthlim, const int* mthbc, int use_fwave,
               int solver_id, double* dt_trace, double* cfl_trace, int trace_cap, int* ntrace);

#ifdef __cplusplus
}
#endif
#endif
