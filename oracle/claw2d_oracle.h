/*
 * oracle/claw2d_oracle.h -- CPU ORACLE (test infrastructure only; never part
 * of the product library).  Plain-C restatement of the reference 2D path;
 * see claw2d_oracle.c for the reference file:line each function follows.
 */
#ifndef CLAW2D_ORACLE_H
#define CLAW2D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_SOLVER_ACOUSTICS = 1, ORC_SOLVER_EULER = 2, ORC_SOLVER_VC_ADVECTION = 5 };

typedef struct {
    double rho, bulk, cc, zz; /* acoustics: density, bulk modulus, sound speed, impedance */
    double gamma;             /* Euler: ratio of specific heats */
    int efix;                 /* Euler: Harten-Hyman entropy fix on/off */
} orc_solver_params;

typedef void (*orc_rpn2_t)(int ixy, int maxm, int meqn, int mwaves, int maux, int mbc, int mx,
                           const double* ql, const double* qr, const double* auxl,
                           const double* auxr, double* wave, double* s, double* amdq,
                           double* apdq);
typedef void (*orc_rpt2_t)(int ixy, int imp, int maxm, int meqn, int mwaves, int maux, int mbc,
                           int mx, const double* ql, const double* qr, const double* aux1,
                           const double* aux2, const double* aux3, const double* asdq,
                           double* bmasdq, double* bpasdq);

void orc_set_solver_params(const orc_solver_params* p);
orc_rpn2_t orc_get_rpn2(int solver_id);
orc_rpt2_t orc_get_rpt2(int solver_id);

void orc_rpn2_acoustics(int ixy, int maxm, int meqn, int mwaves, int maux, int mbc, int mx,
                        const double* ql, const double* qr, const double* auxl,
                        const double* auxr, double* wave, double* s, double* amdq,
                        double* apdq);
void orc_rpt2_acoustics(int ixy, int imp, int maxm, int meqn, int mwaves, int maux, int mbc,
                        int mx, const double* ql, const double* qr, const double* aux1,
                        const double* aux2, const double* aux3, const double* asdq,
                        double* bmasdq, double* bpasdq);
void orc_rpn2_euler(int ixy, int maxm, int meqn, int mwaves, int maux, int mbc, int mx,
                    const double* ql, const double* qr, const double* auxl, const double* auxr,
                    double* wave, double* s, double* amdq, double* apdq);
void orc_rpt2_euler(int ixy, int imp, int maxm, int meqn, int mwaves, int maux, int mbc, int mx,
                    const double* ql, const double* qr, const double* aux1, const double* aux2,
                    const double* aux3, const double* asdq, double* bmasdq, double* bpasdq);

void orc_rpn2_vc_advection(int ixy, int maxm, int meqn, int mwaves, int maux, int mbc, int mx,
                           const double* ql, const double* qr, const double* auxl,
                           const double* auxr, double* wave, double* s, double* amdq,
                           double* apdq);
void orc_rpt2_vc_advection(int ixy, int imp, int maxm, int meqn, int mwaves, int maux, int mbc,
                           int mx, const double* ql, const double* qr, const double* aux1,
                           const double* aux2, const double* aux3, const double* asdq,
                           double* bmasdq, double* bpasdq);
void orc_annulus_aux(int mbc, int mx, int my, double xlower, double ylower, double dxc,
                     double dyc, double* aux);

int orc_limiter(int maxm, int meqn, int mwaves, int mbc, int mx, double* wave, const double* s,
                const int* mthlim);

int orc_step2(int meqn, int mwaves, int maux, int mbc, int mx, int my, const double* qold,
              double* qnew, const double* aux, double dx, double dy, double dt,
              const int* method, const int* mthlim, int use_fwave, int solver_id, double* cfl);
int orc_dimsp2(int meqn, int mwaves, int maux, int mbc, int mx, int my, const double* qold,
               double* qnew, const double* aux, double dx, double dy, double dt,
               const int* method, const int* mthlim, const double* cflv, int use_fwave,
               int solver_id, double* cfl);
int orc_bc2(int meqn, int mbc, int mx, int my, double* q, const int* mthbc);
/* OpenMP threads of the unsplit steps (the OpenMP build only; 1 otherwise):
 * the threaded steps update every cell in the serial order, bit for bit. */
int orc_set_threads(int n);
int orc_threads(void);
int orc_claw2(int meqn, int mwaves, int maux, int mbc, int mx, int my, double* q,
              const double* aux, double dx, double dy, double tstart, double* tend, double* dtv,
              double* cflv, int* nv, const int* method, const int* mthlim, const int* mthbc,
              int use_fwave, int solver_id, double* dt_trace, double* cfl_trace, int trace_cap,
              int* ntrace);

#ifdef __cplusplus
}
#endif
#endif
