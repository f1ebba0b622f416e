/*
 * clawgpu.h -- C-ABI of libclawgpu.so, the MI355X (gfx950) implementation of
 * Clawpack Classic's single-grid wave-propagation step.
 *
 * Plain C types only: pointers, sizes, doubles and ints.  Host arrays use the
 * reference layout q(meqn, 1-mbc:mx+mbc, 1-mbc:my+mbc) (Fortran order, m
 * fastest).  The library keeps q resident on the GPU between calls in a
 * context; only the Courant number and frame copies cross to the host.
 *
 * Every entry point returns CLAWGPU_OK (0) or a negative error code and never
 * aborts the process; clawgpu_last_error() describes the last failure of the
 * calling thread.  Calls on one context must come from one host thread at a
 * time; they are ordered on the context's own HIP stream.
 *
 * Riemann solvers are selected by id (the GPU cannot call the host rpn2/rpt2
 * procedure pointers of the reference contract, src/2d/claw2.f:312-402):
 *   CLAWGPU_SOLVER_ACOUSTICS  meqn 3, mwaves 2, params (rho, bulk)
 *   CLAWGPU_SOLVER_EULER      meqn 4, mwaves 3, params (gamma, efix 0/1)
 */
#ifndef CLAWGPU_H
#define CLAWGPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define CLAWGPU_ABI_VERSION 1

enum {
    CLAWGPU_OK = 0,
    CLAWGPU_ERR_ARG = -1,         /* invalid argument */
    CLAWGPU_ERR_UNSUPPORTED = -2, /* valid for the reference, not implemented here */
    CLAWGPU_ERR_HIP = -3,         /* HIP runtime error (no device, out of memory, ...) */
    CLAWGPU_ERR_NONFINITE = -4    /* a wave speed was NaN: the step cannot be judged */
};

enum { CLAWGPU_SOLVER_ACOUSTICS = 1, CLAWGPU_SOLVER_EULER = 2 };

/* Arithmetic of the step kernel.  EXACT (the default) performs the
 * reference's floating-point operations in the reference's order with no
 * fused multiply-add, so results equal the Fortran step2/flux2 bit for bit.
 * FAST fuses multiply-adds and reuses reciprocals; it agrees with EXACT to
 * within 1e-12 relative per field (tested over whole runs). */
enum { CLAWGPU_ARITH_EXACT = 0, CLAWGPU_ARITH_FAST = 1 };

typedef struct clawgpu_ctx clawgpu_ctx;

/* Library identification and the last error message of this thread. */
int clawgpu_abi_version(void);
const char* clawgpu_last_error(void);

/* Create / destroy a device context for one grid.  Replaces the allocation
 * of q, aux and the work array in src/2d/claw2ez.f:173-263 (the library owns
 * two device copies of q used as qold/qnew in turn, so copyq2,
 * src/2d/copyq2.f:4-20, becomes a buffer swap).  maux must be 0 for the
 * built-in solvers. */
int clawgpu_create(clawgpu_ctx** ctx, int meqn, int mwaves, int maux, int mbc, int mx, int my,
                   int solver_id, const double* params, int nparams);
int clawgpu_destroy(clawgpu_ctx* ctx);

/* Copy q (reference layout, ghost cells included) to / from the device. */
int clawgpu_set_q(clawgpu_ctx* ctx, const double* q);
int clawgpu_get_q(clawgpu_ctx* ctx, double* q);

/* Fill the ghost cells of the device q: src/2d/bc2.f:4-215 with
 * mthbc(1:4) = (left, right, bottom, top) in {1 extrap, 2 periodic, 3 wall}.
 * mthbc = 0 (user BCs) is not supported on the device. */
int clawgpu_bc2(clawgpu_ctx* ctx, const int* mthbc);

/* One unsplit step, the contract of step2 (src/2d/step2.f90:2-251): the
 * device q (ghost cells already set) is qold, the result replaces it, and
 * *cfl receives the Courant number of the step.  method(1:7) and
 * mthlim(1:mwaves) as in src/2d/claw2.f:164-270; method(3) in {0,1,2},
 * method(6) = 0.  The scratch arguments of step2 (qadd, fadd, gadd, q1d,
 * dtdx1d, dtdy1d, aux1-3, work) and the rpn2/rpt2 procedure arguments have
 * no counterpart: scratch lives in registers, the solver is the context's. */
int clawgpu_step2(clawgpu_ctx* ctx, double dx, double dy, double dt, const int* method,
                  const int* mthlim, int use_fwave, double* cfl);

/* One dimensionally split step, the contract of dimsp2
 * (src/2d/dimsp2.f:4-98) with step2ds (src/2d/step2ds.f90:2-275):
 * method(3) = -1 Godunov (x then y) or -2 Strang (x dt/2, y dt, x dt/2); the
 * x-sweeps cover the ghost rows and the y-sweep the ghost columns, as in the
 * reference.  *cfl follows dimsp2's early-abort rule with cflv(1). */
int clawgpu_dimsp2(clawgpu_ctx* ctx, double dx, double dy, double dt, const int* method,
                   const int* mthlim, const double* cflv, int use_fwave, double* cfl);

/* The claw2 time loop (src/2d/claw2.f:5-8 signature, :494-784 semantics) on
 * the device q: bc2, step2 or dimsp2 (method(3) < 0), Courant check with retake on rejection, variable
 * or fixed dt, and the outputs tend, dtv(3:5), cflv(3:4), nv(2) and info
 * (0, 2, 3, 5, 6, 11, 12 as in :288-300).  info = -5 is returned (with
 * CLAWGPU_ERR_NONFINITE) when a wave speed is NaN, where the reference would
 * loop forever. */
int clawgpu_claw2(clawgpu_ctx* ctx, double dx, double dy, double tstart, double* tend,
                  double* dtv, double* cflv, int* nv, const int* method, const int* mthlim,
                  const int* mthbc, int use_fwave, int* info);

/* Tuning and measurement helpers (no reference counterpart): rows per
 * wavefront segment of the step kernel, the arithmetic mode, and nsteps fixed-dt steps timed with
 * HIP events on the context's stream without per-step host synchronisation:
 * *total_ms spans every launch, *step_kernel_ms sums the step-kernel launches
 * alone, *cfl_max receives the largest Courant number of those steps. */
int clawgpu_set_segment_rows(clawgpu_ctx* ctx, int rows);
int clawgpu_set_arith(clawgpu_ctx* ctx, int mode);
int clawgpu_time_steps(clawgpu_ctx* ctx, double dx, double dy, double dt, const int* method,
                       const int* mthlim, const int* mthbc, int use_fwave, int nsteps,
                       float* total_ms, float* step_kernel_ms, double* cfl_max);

/* Solver-level evaluation for tests (no reference counterpart): the device
 * rpn2 on n interfaces between ql[k] and qr[k] (each meqn doubles) and the
 * device rpt2 split of asdq[k] with that interface's eigenstructure, in the
 * context's solver and arithmetic mode.  Outputs follow the reference's
 * per-interface layout: wave(meqn, mwaves), s(mwaves), amdq, apdq, bmasdq,
 * bpasdq (meqn each).  ixy = 1 for an x-slice, 2 for a y-slice. */
int clawgpu_solver_probe(clawgpu_ctx* ctx, int ixy, int n, const double* ql, const double* qr,
                         const double* asdq, double* wave, double* s, double* amdq, double* apdq,
                         double* bmasdq, double* bpasdq);

/* Learned flux limiter (no reference counterpart: it replaces phi(theta) of
 * src/2d/inlinelimiter.f90:71-97 as limiter method 6).  A 5-32-32-1 GELU
 * network, parameters as 1281 floats in the order W1[5][32], b1[32],
 * W2[32][32], b2[32], W3[32], b3, evaluated per interface wave on the matrix
 * cores in BF16 with FP32 accumulation, clipped to the Sweby TVD region with
 * MC where the clip triggers.  With the parameters set, mthlim(mw) = 6 in
 * clawgpu_step2 / clawgpu_claw2 (unsplit) uses it.  The remaining calls
 * expose the pre-pass and the training gradients:
 *   features: 5 features, theta and gradient coefficients of every interface
 *             wave of the device q (ghost cells set) at this dt, MC defaults
 *             in the phi tables;
 *   phi:      mode 0 BF16-MFMA network, 1 FP32 network, 2 MC;
 *   read/write: tables 0 features (float, 8 per wave), 1 phi x, 2 phi y
 *             (double), 3 theta (double), 4 gradient coefficients (float,
 *             meqn per wave), 5 dL/dphi (float); waves per direction from
 *             table_size (slot (i, J) for i in [-1, mx+2], J in [-1, my+2]);
 *   grad:     parameter gradient (host, 1281 floats) and loss of mode 0
 *             (sum over waves of (out - phi_MC)^2) or mode 1 (one unsplit
 *             step against the target q, mean squared error per cell). */
int clawgpu_learned_set_params(clawgpu_ctx* ctx, const float* params, int n);
int clawgpu_learned_features(clawgpu_ctx* ctx, double dx, double dy, double dt);
int clawgpu_learned_phi(clawgpu_ctx* ctx, int mode);
int clawgpu_learned_table_size(clawgpu_ctx* ctx, long* waves_per_direction);
int clawgpu_learned_read(clawgpu_ctx* ctx, int what, void* out, long nbytes);
int clawgpu_learned_write(clawgpu_ctx* ctx, int what, const void* in, long nbytes);
int clawgpu_learned_set_target(clawgpu_ctx* ctx, const double* q);
int clawgpu_learned_grad(clawgpu_ctx* ctx, double dx, double dy, double dt, const int* method,
                         const int* mthbc, int mode, float* grad, double* loss);

#ifdef __cplusplus
}
#endif
#endif
