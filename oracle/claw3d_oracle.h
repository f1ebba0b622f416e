/*
 * oracle/claw3d_oracle.h -- CPU ORACLE (test infrastructure only; never part
 * of the product library).  Plain-C restatement of the reference 3D path;
 * see claw3d_oracle.c for the reference file:line each function follows.
 */
#ifndef CLAW3D_ORACLE_H
#define CLAW3D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC3_SOLVER_VC_ACOUSTICS = 3 };

/* Riemann solvers with the 1D-slice contract of src/3d/rpn3.f, rpt3.f and
 * rptt3.f (arrays indexed from 1-mbc, aux slices dimensioned with maxm). */
typedef void (*orc_rpn3_t)(int ixyz, int maxm, int meqn, int mwaves, int maux, int mbc, int mx,
                           const double* ql, const double* qr, const double* auxl,
                           const double* auxr, double* wave, double* s, double* amdq,
                           double* apdq);
typedef void (*orc_rpt3_t)(int ixyz, int icoor, int imp, int maxm, int meqn, int mwaves,
                           int maux, int mbc, int mx, const double* ql, const double* qr,
                           const double* aux1, const double* aux2, const double* aux3,
                           const double* asdq, double* bmasdq, double* bpasdq);
typedef void (*orc_rptt3_t)(int ixyz, int icoor, int imp, int impt, int maxm, int meqn,
                            int mwaves, int maux, int mbc, int mx, const double* ql,
                            const double* qr, const double* aux1, const double* aux2,
                            const double* aux3, const double* bsasdq, double* cmbsasdq,
                            double* cpbsasdq);

void orc_rpn3_vc_acoustics(int ixyz, int maxm, int meqn, int mwaves, int maux, int mbc, int mx,
                           const double* ql, const double* qr, const double* auxl,
                           const double* auxr, double* wave, double* s, double* amdq,
                           double* apdq);
void orc_rpt3_vc_acoustics(int ixyz, int icoor, int imp, int maxm, int meqn, int mwaves,
                           int maux, int mbc, int mx, const double* ql, const double* qr,
                           const double* aux1, const double* aux2, const double* aux3,
                           const double* asdq, double* bmasdq, double* bpasdq);
void orc_rptt3_vc_acoustics(int ixyz, int icoor, int imp, int impt, int maxm, int meqn,
                            int mwaves, int maux, int mbc, int mx, const double* ql,
                            const double* qr, const double* aux1, const double* aux2,
                            const double* aux3, const double* bsasdq, double* cmbsasdq,
                            double* cpbsasdq);

/* Geometry needed by the boundary conditions (the piston of the
 * heterogeneous-acoustics example reads the lower corner and the spacing). */
typedef struct {
    double xlower, ylower, zlower, dx, dy, dz;
} orc3_geom;

int orc3_step3(int meqn, int mwaves, int maux, int mbc, int mx, int my, int mz,
               const double* qold, double* qnew, const double* aux, double dx, double dy,
               double dz, double dt, const int* method, const int* mthlim, int use_fwave,
               int solver_id, double* cfl);
int orc3_dimsp3(int meqn, int mwaves, int maux, int mbc, int mx, int my, int mz,
                const double* qold, double* qnew, const double* aux, double dx, double dy,
                double dz, double dt, const int* method, const int* mthlim, const double* cflv,
                int use_fwave, int solver_id, double* cfl);
int orc3_bc3(int meqn, int mbc, int mx, int my, int mz, const orc3_geom* g, double* q,
             double t, const int* mthbc);
int orc3_claw3(int meqn, int mwaves, int maux, int mbc, int mx, int my, int mz, double* q,
               const double* aux, const orc3_geom* g, double tstart, double* tend, double* dtv,
               double* cflv, int* nv, const int* method, const int* mthlim, const int* mthbc,
               int use_fwave, int solver_id, double* dt_trace, double* cfl_trace, int trace_cap,
               int* ntrace);

#ifdef __cplusplus
}
#endif
#endif
