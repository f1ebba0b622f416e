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
 *   CLAWGPU_SOLVER_ACOUSTICS     meqn 3, mwaves 2, params (rho, bulk)
 *   CLAWGPU_SOLVER_EULER         meqn 4, mwaves 3, params (gamma, efix 0/1)
 *   CLAWGPU_SOLVER_VC_ADVECTION  meqn 1, mwaves 1, maux >= 2 with aux(1) = u
 *                                at the left edge, aux(2) = v at the bottom
 *                                edge of each cell (the annulus example);
 *                                the capacity function (method(6) > 0) is
 *                                supported with this solver
 * for the 3D path (clawgpu3_*, src/3d/claw3.f:292-516):
 *   CLAWGPU_SOLVER_VC_ACOUSTICS  meqn 4, mwaves 2, maux >= 2 with
 *                                aux(1) = impedance Z, aux(2) = sound speed c
 * and for the 1D path (clawgpu1_*, rp1 of src/1d/claw1.f):
 *   CLAWGPU_SOLVER_ACOUSTICS     meqn 2, mwaves 2, params (rho, bulk)
 *   CLAWGPU_SOLVER_VC_ACOUSTICS  meqn 2, mwaves 2, aux(1) = Z, aux(2) = c
 *   CLAWGPU_SOLVER_ADVECTION     meqn 1, mwaves 1, params (u)
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

enum {
    CLAWGPU_SOLVER_ACOUSTICS = 1,
    CLAWGPU_SOLVER_EULER = 2,
    CLAWGPU_SOLVER_VC_ACOUSTICS = 3,
    CLAWGPU_SOLVER_ADVECTION = 4,
    CLAWGPU_SOLVER_VC_ADVECTION = 5,
    CLAWGPU_SOLVER_USER = 100     /* a context of clawgpu_create_user */
};

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
 * acoustics and Euler solvers and >= 2 for variable-coefficient advection. */
int clawgpu_create(clawgpu_ctx** ctx, int meqn, int mwaves, int maux, int mbc, int mx, int my,
                   int solver_id, const double* params, int nparams);
int clawgpu_destroy(clawgpu_ctx* ctx);

/* Copy q (reference layout, ghost cells included) to / from the device. */
int clawgpu_set_q(clawgpu_ctx* ctx, const double* q);
int clawgpu_get_q(clawgpu_ctx* ctx, double* q);

/* Copy aux(maux, 1-mbc:mx+mbc, 1-mbc:my+mbc) to the device: what setaux
 * fills in src/2d/claw2ez.f:229-232 (contexts created with maux > 0). */
int clawgpu_set_aux(clawgpu_ctx* ctx, const double* aux);

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

/* No counterpart in the reference: clawgpu_step2 for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No context: the caller passes device
 * buffers and the call allocates nothing, its only scratch (the Courant
 * accumulator) being a 256-byte array the library holds per GPU until the
 * process ends.  One unsplit step of solver_id (ACOUSTICS or EULER, params
 * as clawgpu_create, arith EXACT or FAST) from qold (ghost cells set) into
 * the interior cells of qnew, both in field-plane layout: element (m, i, j),
 * 1-based cell indices, at m*plane + (j+mbc-1)*pitch + (i+mbc-1), with
 * pitch >= mx+2*mbc, plane >= pitch*(my+2*mbc), qold != qnew.  The Courant
 * number is written to the device double *cfl.  Everything is enqueued on
 * `stream` (a hipStream_t of the current device; NULL = the null stream) and
 * nothing waits: read *cfl after the stream is done.  mx, my <= 1048576,
 * 2 <= mbc <= 64; method and mthlim as clawgpu_step2 with limiters 0..5.
 * Calls that overlap in time on one GPU (unordered streams or host threads)
 * share the scratch and race; calls ordered on one stream do not. */
int clawgpu_step2_noalloc(int solver_id, const double* params, int nparams, int arith, int mx,
                          int my, int mbc, long pitch, long plane, const double* qold,
                          double* qnew, double dx, double dy, double dt, const int* method,
                          const int* mthlim, int use_fwave, double* cfl, void* stream);

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

/* User routines of the reference's plug-in contract (src/2d/claw2.f:404-447:
 * bc2 with user boundary conditions, b4step2, src2), as C callbacks with a
 * user pointer first and the reference's arguments by value (arrays by
 * pointer, reference layout, ghost cells included; aux is NULL when
 * maux = 0).  The library copies q to the host for each call and back after
 * it (and aux after b4step2), so hooks cost two transfers of q per call.
 * Any member may be NULL: bc2 = NULL uses the built-in codes 1-3, a NULL
 * b4step2 / src2 does nothing (the reference's stubs). */
typedef struct clawgpu_hooks {
    void* user;
    void (*bc2)(void* user, int meqn, int mbc, int mx, int my, double xlower, double ylower,
                double dx, double dy, double* q, int maux, double* aux, double t, double dt,
                const int* mthbc);
    void (*b4step2)(void* user, int mbc, int mx, int my, int meqn, double* q, double xlower,
                    double ylower, double dx, double dy, double t, double dt, int maux,
                    double* aux);
    void (*src2)(void* user, int meqn, int mbc, int mx, int my, double xlower, double ylower,
                 double dx, double dy, double* q, int maux, double* aux, double t, double dt);
} clawgpu_hooks;

/* clawgpu_claw2 with user routines called where claw2 calls them
 * (claw2.f:620-737): bc2 then b4step2 before every step attempt (again on a
 * retake); with method(5) = 2 src2 over dt/2 before the step (its input
 * restored on a retake) and over dt/2 at t + dt after it, with method(5) = 1
 * src2 over dt after it.  xlower, ylower are passed to the hooks. */
int clawgpu_claw2_hooks(clawgpu_ctx* ctx, double dx, double dy, double xlower, double ylower,
                        double tstart, double* tend, double* dtv, double* cflv, int* nv,
                        const int* method, const int* mthlim, const int* mthbc, int use_fwave,
                        const clawgpu_hooks* hooks, int* info);

/* Device-side user routines: the same plug-in contract (bc2, b4step2, src2
 * of claw2.f:620-737) as HIP kernels in a gfx950 code object, launched by
 * the library on the device q at the points clawgpu_claw2_hooks calls the
 * host routines, so q never leaves the GPU.  Each kernel takes one
 * clawgpu_dev_grid by value and is launched with one thread per cell,
 * ghost cells included: thread t of the 1D grid (blocks of 256 threads)
 * handles cell (i, j) with t = (j + mbc - 1) * (mx + 2 mbc) + (i + mbc - 1),
 * threads with t beyond the grid must return (include/clawgpu_device.h has
 * the index helpers).  Build a code object with
 *   hipcc --genco --offload-arch=gfx950 -O3 -I include hooks.hip -o hooks.co
 * (add -ffp-contract=off for the reference's rounding). */
typedef struct clawgpu_dev_grid {
    double* q;       /* q(m, i, j) at q[m * fs + (j + mbc - 1) * pitch + (i + mbc - 1)] */
    double* aux;     /* aux(m, i, j) likewise; NULL when maux = 0 */
    double* user;    /* the hooks' user array (clawgpu_dev_hooks_user), on the device */
    long fs;         /* doubles per field plane */
    int pitch;       /* doubles per row */
    int meqn, maux, mbc, mx, my;
    int mthbc[4];
    double xlower, ylower, dx, dy, t, dt;
} clawgpu_dev_grid;

typedef struct clawgpu_dev_hooks clawgpu_dev_hooks;

/* Load the code object at path (no reference counterpart: device code
 * cannot be linked into the Fortran loop) and look up the named kernels;
 * a NULL name leaves that routine out (built-in bc2 for NULL bc2).  nuser
 * doubles of device memory, zeroed, are passed to the kernels as
 * clawgpu_dev_grid.user.  Load on the device the contexts run on. */
int clawgpu_dev_hooks_load(clawgpu_dev_hooks** hooks, const char* path, const char* bc2,
                           const char* b4step2, const char* src2, int nuser);
int clawgpu_dev_hooks_destroy(clawgpu_dev_hooks* hooks);
/* Copy n doubles of the user array from the host (in, if not NULL), then
 * back to the host (out, if not NULL), after the context's pending work. */
int clawgpu_dev_hooks_user(clawgpu_dev_hooks* hooks, const double* in, double* out, int n);

/* clawgpu_claw2_hooks with device routines: bc2 then b4step2 on the device
 * q before every step attempt (again on a retake), src2 over dt/2 before
 * and after the step (method(5) = 2; the state before the first half step
 * is kept on the device for a retake) or over dt after it (method(5) = 1). */
int clawgpu_claw2_dev_hooks(clawgpu_ctx* ctx, double dx, double dy, double xlower,
                            double ylower, double tstart, double* tend, double* dtv,
                            double* cflv, int* nv, const int* method, const int* mthlim,
                            const int* mthbc, int use_fwave, const clawgpu_dev_hooks* hooks,
                            int* info);

/* User Riemann solvers: the rpn2/rpt2 a caller links into claw2
 * (claw2.f:312-402, called from flux2.f90:121-235 and step2ds).  Here a
 * solver is a policy struct compiled into the library's kernel templates by
 * include/clawgpu_solver.h, built into a gfx950 code object that exports
 * solver `name`.  Loading checks the kernel-argument layout the object was
 * built against.  A context of clawgpu_create_user runs every 2D path of a
 * built-in solver (clawgpu_step2, clawgpu_dimsp2, clawgpu_claw2 with its
 * device-side loop, the hook variants, clawgpu_solver_probe) with the
 * user's kernels, in the arithmetic the object was built with
 * (clawgpu_solver2_shape); it has no aux arrays, capacity function or
 * learned limiter.  Keep the solver loaded while its contexts exist. */
typedef struct clawgpu_solver2 clawgpu_solver2;
int clawgpu_solver2_load(clawgpu_solver2** solver, const char* path, const char* name);
int clawgpu_solver2_destroy(clawgpu_solver2* solver);
/* meqn, mwaves and arith (CLAWGPU_ARITH_EXACT or _FAST) of a loaded solver */
int clawgpu_solver2_shape(const clawgpu_solver2* solver, int* meqn, int* mwaves, int* arith);
/* clawgpu_create for a user solver: params[0..nparams-1] (nparams <= 16)
 * reach the solver as SolverParams.user. */
int clawgpu_create_user(clawgpu_ctx** ctx, const clawgpu_solver2* solver, int mbc, int mx,
                        int my, const double* params, int nparams);

/* Tuning and measurement helpers (no reference counterpart): rows per
 * wavefront segment of the step kernel (0, the default: chosen per grid so
 * that the kernel's waves fill whole rounds of the chip; get_segment_rows
 * reports the rows the last unsplit step used, 0 before the first), the
 * arithmetic mode, and nsteps fixed-dt steps timed with
 * HIP events on the context's stream without per-step host synchronisation:
 * *total_ms spans every launch, *step_kernel_ms sums the step-kernel launches
 * alone, *cfl_max receives the largest Courant number of those steps. */
int clawgpu_set_segment_rows(clawgpu_ctx* ctx, int rows);
int clawgpu_get_segment_rows(clawgpu_ctx* ctx, int* rows);
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
 *             meqn per wave; read only), 5 dL/dphi (float); waves per direction from
 *             table_size (slot (i, J) for i in [-1, mx+2], J in [-1, my+2]);
 *   grad:     parameter gradient (host, 1281 floats) and loss of mode 0
 *             (sum over waves of (out - phi_MC)^2), mode 2 (the same with
 *             superbee's phi in place of MC's), mode 1 (one unsplit
 *             step against the target q, mean squared error per cell) or
 *             mode 3 (as mode 1, mean absolute error summed over fields). */
int clawgpu_learned_set_params(clawgpu_ctx* ctx, const float* params, int n);
int clawgpu_learned_features(clawgpu_ctx* ctx, double dx, double dy, double dt);
int clawgpu_learned_phi(clawgpu_ctx* ctx, int mode);
int clawgpu_learned_table_size(clawgpu_ctx* ctx, long* waves_per_direction);
int clawgpu_learned_read(clawgpu_ctx* ctx, int what, void* out, long nbytes);
int clawgpu_learned_write(clawgpu_ctx* ctx, int what, const void* in, long nbytes);
int clawgpu_learned_set_target(clawgpu_ctx* ctx, const double* q);
int clawgpu_learned_grad(clawgpu_ctx* ctx, double dx, double dy, double dt, const int* method,
                         const int* mthbc, int mode, float* grad, double* loss);
/* Measurement helper (no reference counterpart): reps fused pre-passes (the
 * one kernel a step with limiter 6 runs: features and network, phi tables
 * only), then one feature pass and reps stand-alone BF16-MFMA inference
 * passes (phi mode 0) on the current q, timed with HIP events on the
 * context's stream; *fused_ms and *infer_ms receive the averages per launch.
 * Leaves the tables of one feature + inference pass. */
int clawgpu_learned_time(clawgpu_ctx* ctx, double dx, double dy, double dt, int reps,
                         float* fused_ms, float* infer_ms);


/* Benchmark helper (no reference counterpart): reps plain device copies of
 * ndoubles doubles; *ms_avg = average time per copy.  A kernel of known
 * traffic for calibrating the FETCH_SIZE / WRITE_SIZE counters. */
int clawgpu_copy_calibration(long ndoubles, int reps, float* ms_avg);

/* Measurement helpers (no reference counterpart): the ceilings the step
 * kernels are measured against on the device in use.
 * clawgpu_copy_bandwidth: reps device copies of ndoubles doubles with copy
 *   kernel `variant` (0: the calibration copy above, one 16-byte load per
 *   lane per iteration; 1: four 16-byte loads in flight per lane; 2: as 1
 *   with nontemporal loads and stores; 3: eight in flight, nontemporal);
 *   *ms_avg = average time per copy (2 x 8 x ndoubles bytes move).
 * clawgpu_fp64_fma_peak: FP64 fused multiply-add throughput, eight
 *   independent chains per lane over every SIMD of the device, iters
 *   steps per chain; *tflops = 2 flops per FMA / time. */
int clawgpu_copy_bandwidth(long ndoubles, int variant, int reps, float* ms_avg);
int clawgpu_fp64_fma_peak(int iters, float* ms, double* tflops);

/* --------------------------------------------------------- batched 2D */

/* Many independent 2D grids advanced together (no reference counterpart as
 * one call: each problem follows its own claw2 loop, src/2d/claw2.f:494-784,
 * exactly as clawgpu_claw2 would run it, and results equal that call's bit
 * for bit in EXACT arithmetic).  One batched iteration performs, for every
 * unfinished problem, one step attempt: dt clamp to tend, bc2, the unsplit
 * step of clawgpu_step2, the Courant check with retake and the next dt, all
 * decided on the device; iterations are replayed from a captured HIP graph.
 *
 * clawgpu_batch_create: nprob problems; per problem k: solver[k] (ACOUSTICS
 *   or EULER), mx[k], my[k], params[2k..2k+1] (as clawgpu_create),
 *   mthbc[4k..4k+3], dx[k], dy[k]; mbc is shared.
 * clawgpu_batch_set_q / get_q: q of problem k in the reference layout.
 * clawgpu_batch_start: begin claw2 runs of every problem from tstart to
 *   tend[k] with the shared dtv(1:2), cflv(1:2), nv(1) = maxn, method(1:7)
 *   (method(3) in {0,1,2}, method(6) = 0), mthlim(1:3) (waves 1..mwaves).
 * clawgpu_batch_advance: at most max_iters batched iterations, stopping
 *   early when every problem has finished; *nleft = unfinished problems.
 * clawgpu_batch_result: claw2's outputs of problem k so far: tend (*t),
 *   dtv(3:5), cflv(3:4), nv(1:2), info (0, 2, 11, 12, or -5 for a NaN
 *   wave speed).
 * clawgpu_batch_set_schedule / get_schedule: sched(1:4) = rows per tile,
 *   blocks of each persistent step kernel (0: two per CU), tile order (1:
 *   problems by decreasing size) and the acoustics kernel's register budget
 *   (1: three waves per SIMD); clawgpu_batch_tune measures every candidate
 *   schedule on state-preserving iterations and keeps the fastest
 *   (cand_out: the four values per candidate; ms_out: ms per iteration). */
typedef struct clawgpu_batch clawgpu_batch;

int clawgpu_batch_create(clawgpu_batch** batch, int nprob, int mbc, const int* solver, const int* mx,
                         const int* my, const double* params, const int* mthbc, const double* dx,
                         const double* dy);
int clawgpu_batch_destroy(clawgpu_batch* batch);
int clawgpu_batch_set_q(clawgpu_batch* batch, int k, const double* q);
int clawgpu_batch_get_q(clawgpu_batch* batch, int k, double* q);
int clawgpu_batch_set_arith(clawgpu_batch* batch, int mode);
int clawgpu_batch_start(clawgpu_batch* batch, double tstart, const double* tend, const double* dtv,
                        const double* cflv, int maxn, const int* method, const int* mthlim,
                        int use_fwave);
int clawgpu_batch_advance(clawgpu_batch* batch, int max_iters, int use_graph, int* nleft);
int clawgpu_batch_result(clawgpu_batch* batch, int k, double* t, double* dtv, double* cflv, int* nv,
                         int* info);
int clawgpu_batch_set_schedule(clawgpu_batch* batch, const int* sched);
int clawgpu_batch_get_schedule(clawgpu_batch* batch, int* sched);
int clawgpu_batch_tune(clawgpu_batch* batch, int reps, int* ncand, int* cand_out, float* ms_out,
                       int max_cand);
/* Measurement helper: iters batched iterations launched without the graph,
 * HIP events around each step-kernel launch; *total_ms spans every launch,
 * *step_kernel_ms sums the step kernels. */
int clawgpu_batch_time(clawgpu_batch* batch, int iters, float* total_ms, float* step_kernel_ms);

/* ------------------------------------------------------------------ 3D */

typedef struct clawgpu3_ctx clawgpu3_ctx;

/* Create / destroy a device context for one 3D grid: q(4, 1-mbc:mx+mbc,
 * 1-mbc:my+mbc, 1-mbc:mz+mbc) and aux(maux, ...) in the reference layout on
 * the host.  Replaces the allocations of src/3d/claw3ez.f:189-255 (three
 * device copies of q replace copyq3's work copy, src/3d/copyq3.f).  The
 * solver must be CLAWGPU_SOLVER_VC_ACOUSTICS; mbc must be 2. */
int clawgpu3_create(clawgpu3_ctx** ctx, int meqn, int mwaves, int maux, int mbc, int mx, int my,
                    int mz, int solver_id);
int clawgpu3_destroy(clawgpu3_ctx* ctx);

/* Copy q / aux (reference layout, ghost cells included) to and from the
 * device; set_aux is what src/3d/claw3ez.f:231-234 (setaux) fills. */
int clawgpu3_set_q(clawgpu3_ctx* ctx, const double* q);
int clawgpu3_get_q(clawgpu3_ctx* ctx, double* q);
int clawgpu3_set_aux(clawgpu3_ctx* ctx, const double* aux);

/* Fill the ghost cells of the device q at time t: src/3d/bc3.f with
 * mthbc(1:6) = (x-lo, x-hi, y-lo, y-hi, z-lo, z-hi) in {1 extrap, 2 periodic,
 * 3 wall}; mthbc(6) = 0 selects the moving-piston wall of the
 * heterogeneous-acoustics example (examples/acoustics_3d_heterogeneous/bc3.f),
 * whose velocity needs geom = (xlower, ylower, zlower, dx, dy, dz). */
int clawgpu3_bc3(clawgpu3_ctx* ctx, const double* geom, double t, const int* mthbc);

/* One unsplit step, the contract of step3 (src/3d/step3.f90:3-873): the
 * device q with ghost cells set is qold and the interior of the result
 * replaces it (ghost cells of the result are unspecified; the reference
 * leaves partial updates there); *cfl is the Courant number.  method(2) in
 * {1, 2}, method(3) in {0, 10, 11, 20, 21, 22} (src/3d/chkmth.f:3-72),
 * mthlim(1:2) in 0..5.  method(6) > 0 names the aux component holding the
 * capacity function (step3.f90:119-202, 337-424; dtdx1d = dtdx / kappa).
 * Fast arithmetic (the default, method(6) = 0): summation order differs
 * from the reference's slice order, results agree to round-off (tested to
 * 1e-12).  With a capacity function, or in exact arithmetic
 * (clawgpu3_set_arith), the step follows the reference's rounding and the
 * order of its passes over blocks of slices: bit for bit. */
int clawgpu3_step3(clawgpu3_ctx* ctx, double dx, double dy, double dz, double dt,
                   const int* method, const int* mthlim, int use_fwave, double* cfl);

/* One Godunov-split step, the contract of dimsp3 (src/3d/dimsp3.f:5-85,
 * step3ds src/3d/step3ds.f90:3-408) with method(3) = -1, including its early
 * return after a sweep whose Courant number exceeds cflv(1); method(6) > 0:
 * capacity function (step3ds.f90:125-173, 217-228).  Bit for bit with the
 * reference, ghost cells included. */
int clawgpu3_dimsp3(clawgpu3_ctx* ctx, double dx, double dy, double dz, double dt,
                    const int* method, const int* mthlim, const double* cflv, int use_fwave,
                    double* cfl);

/* The claw3 time loop (src/3d/claw3.f:596-904) on the device: bc3, step3 or
 * dimsp3, CFL check with retake, new dt, with the reference's dtv(1:5),
 * cflv(1:4), nv(1:2) and info codes (0, 2, 3, 5, 6, 11, 12; -5 for a NaN
 * sound speed).  geom = (xlower, ylower, zlower, dx, dy, dz). */
int clawgpu3_claw3(clawgpu3_ctx* ctx, const double* geom, double tstart, double* tend,
                   double* dtv, double* cflv, int* nv, const int* method, const int* mthlim,
                   const int* mthbc, int use_fwave, int* info);

/* User routines of the 3D plug-in contract (src/3d/claw3.f:766-885: bc3,
 * b4step3, src3), with the conventions of clawgpu_hooks; a NULL bc3 takes the
 * built-in bc3 (codes 1-3, the piston with mthbc(6) = 0). */
typedef struct clawgpu3_hooks {
    void* user;
    void (*bc3)(void* user, int meqn, int mbc, int mx, int my, int mz, double xlower,
                double ylower, double zlower, double dx, double dy, double dz, double* q, int maux,
                double* aux, double t, double dt, const int* mthbc);
    void (*b4step3)(void* user, int mbc, int mx, int my, int mz, int meqn, double* q,
                    double xlower, double ylower, double zlower, double dx, double dy, double dz,
                    double t, double dt, int maux, double* aux);
    void (*src3)(void* user, int meqn, int mbc, int mx, int my, int mz, double xlower,
                 double ylower, double zlower, double dx, double dy, double dz, double* q,
                 int maux, double* aux, double t, double dt);
} clawgpu3_hooks;

/* clawgpu3_claw3 with user routines called where claw3 calls them
 * (claw3.f:766-885): bc3 then b4step3 before every attempt; with
 * method(5) = 2 src3 over dt/2 before the step (its input restored on a
 * rejection) and over dt/2 at t + dt after an accepted step, with
 * method(5) = 1 src3 over dt after it.  aux changed by b4step3 is copied back
 * (and its sound speeds and material table recomputed). */
int clawgpu3_claw3_hooks(clawgpu3_ctx* ctx, const double* geom, double tstart, double* tend,
                         double* dtv, double* cflv, int* nv, const int* method, const int* mthlim,
                         const int* mthbc, int use_fwave, const clawgpu3_hooks* hooks, int* info);

/* Device-side 3D user routines (bc3, b4step3, src3 of claw3.f:766-885) as
 * HIP kernels, loaded with clawgpu_dev_hooks_load (their names in the bc,
 * b4step and src positions).  Each takes one clawgpu_dev_grid3 by value and
 * is launched with one thread per cell, ghost cells included: thread t
 * handles cell (i, j, k) with t = ((k + mbc - 1) * (my + 2 mbc) + (j + mbc -
 * 1)) * (mx + 2 mbc) + (i + mbc - 1) (include/clawgpu_device.h). */
typedef struct clawgpu_dev_grid3 {
    double* q;       /* q(m, i, j, k) at q[m * fs + t(i, j, k)] */
    double* aux;     /* aux(m, i, j, k) likewise */
    double* user;    /* the hooks' user array, on the device */
    long fs;         /* doubles per field plane: (mx + 2 mbc)(my + 2 mbc)(mz + 2 mbc) */
    int meqn, maux, mbc, mx, my, mz;
    int mthbc[6];
    double xlower, ylower, zlower, dx, dy, dz, t, dt;
} clawgpu_dev_grid3;

/* clawgpu3_claw3_hooks with device routines, called at the same points on
 * the device q (the state before a Strang source half step is kept on the
 * device for a retake).  After a device b4step3 the sound-speed maximum
 * (the Courant number) is recomputed on the device and the material table
 * is checked against aux; if b4step3 changed a material, later steps read
 * aux directly. */
int clawgpu3_claw3_dev_hooks(clawgpu3_ctx* ctx, const double* geom, double tstart, double* tend,
                             double* dtv, double* cflv, int* nv, const int* method,
                             const int* mthlim, const int* mthbc, int use_fwave,
                             const clawgpu_dev_hooks* hooks, int* info);

/* User Riemann solvers in 3D: the rpn3 a caller links into claw3
 * (claw3.f:292-330, called from flux3.f90:205 in the sweeps of step3ds.f90
 * that dimsp3.f drives).  A solver is a policy struct compiled by
 * CLAWGPU_EXPORT_SOLVER3 (include/clawgpu_solver.h) into a gfx950 code
 * object that exports solver `name`; loading checks the kernel-argument
 * layout the object was built against.  A context of clawgpu3_create_user
 * runs the dimensionally split step (method(3) = -1: clawgpu3_dimsp3,
 * clawgpu3_claw3 and its hook variants) with the user's sweeps over meqn
 * fields, mwaves waves and the solver's aux planes, in the arithmetic the
 * object was built with; it has no transverse terms (rpt3, rptt3), no
 * capacity function and no piston boundary.  claw3 takes the step before
 * its Courant number is known (the sweeps form it) and uses the result
 * only if the step is accepted.  Keep the solver loaded while its
 * contexts exist. */
typedef struct clawgpu_solver3 clawgpu_solver3;
int clawgpu_solver3_load(clawgpu_solver3** solver, const char* path, const char* name);
int clawgpu_solver3_destroy(clawgpu_solver3* solver);
/* meqn, mwaves, maux (aux planes it reads) and arith of a loaded solver */
int clawgpu_solver3_shape(const clawgpu_solver3* solver, int* meqn, int* mwaves, int* maux,
                          int* arith);
/* clawgpu3_create for a user solver (mbc = 2; maux >= the solver's):
 * params[0..nparams-1] (nparams <= 16) reach it as SolverParams.user. */
int clawgpu3_create_user(clawgpu3_ctx** ctx, const clawgpu_solver3* solver, int maux, int mx,
                         int my, int mz, const double* params, int nparams);

/* Cells per segment of the y/z split sweeps (tuning; default 64). */
int clawgpu3_set_segment(clawgpu3_ctx* ctx, int cells);

/* Arithmetic of the unsplit 3D step, as clawgpu_set_arith: mode 1 (the
 * default) fast (fused kernels, FMA contraction, within 1e-12 of the
 * reference), mode 0 exact (one kernel per slice position and direction
 * plus one per cell: the reference's rounding and order, bit for bit with
 * step3; slower).  Steps with a capacity function always run exact. */
int clawgpu3_set_arith(clawgpu3_ctx* ctx, int mode);

/* Tile shape of the unsplit kernel (tuning): 0 = 60x4x4, 1 = 60x4x2,
 * 2 = 60x2x2, 3 = 60x8x2, 4 = 60x2x4 cells per wavefront;
 * 5, 6 (default) = z-marching blocks of 60x10 or 60x6 columns (these need
 * the material table and otherwise run shape 4);
 * xcd_order (tile order, default 3) bit 0 hands each XCD a contiguous
 * range of tiles, bit 1 numbers tiles with z fastest instead of x (tile
 * kernels only). */
int clawgpu3_set_tile(clawgpu3_ctx* ctx, int tile, int xcd_order);
/* The tile shape and order set (tuning). */
int clawgpu3_get_tile(const clawgpu3_ctx* ctx, int* tile, int* xcd_order);

/* Material table (no reference counterpart; same values, fewer bytes):
 * clawgpu3_set_aux finds the distinct (Z, c) pairs of aux; when there are at
 * most 64, the unsplit kernel reads a one-byte material index per cell and a
 * table of the pairs instead of the two aux doubles, with identical results.
 * enable = 0 makes it read aux regardless; *nmat (if not null) receives the
 * number of materials found (0: more than 64, aux is read). */
int clawgpu3_set_aux_table(clawgpu3_ctx* ctx, int enable, int* nmat);

/* Benchmark helper (no reference counterpart): nsteps fixed-dt unsplit
 * steps (bc3 + step) timed with HIP events, as clawgpu_time_steps. */
int clawgpu3_time_steps(clawgpu3_ctx* ctx, const double* geom, double dt, const int* method,
                        const int* mthlim, const int* mthbc, int use_fwave, int nsteps,
                        float* total_ms, float* step_kernel_ms, double* cfl);


/* ------------------------------------------------------------------ 1D */

typedef struct clawgpu1_ctx clawgpu1_ctx;

/* Create / destroy a device context for one 1D grid, q(meqn, 1-mbc:mx+mbc)
 * and aux(maux, 1-mbc:mx+mbc) in the reference layout; replaces the
 * allocations of src/1d/claw1ez.f:150-220.  mbc must be 2. */
int clawgpu1_create(clawgpu1_ctx** ctx, int meqn, int mwaves, int maux, int mbc, int mx,
                    int solver_id, const double* params, int nparams);
int clawgpu1_destroy(clawgpu1_ctx* ctx);
int clawgpu1_set_q(clawgpu1_ctx* ctx, const double* q);
int clawgpu1_get_q(clawgpu1_ctx* ctx, double* q);
int clawgpu1_set_aux(clawgpu1_ctx* ctx, const double* aux);

/* Ghost cells of the device q: src/1d/bc1.f with mthbc(1:2) in {1, 2, 3}. */
int clawgpu1_bc1(clawgpu1_ctx* ctx, const int* mthbc);

/* One step, the contract of step1 (src/1d/step1.f90): the device q with
 * ghost cells set is updated (interior cells; the reference also changes
 * ghost cells 0 and mx+1, which the next bc1 overwrites), *cfl is the Courant
 * number.  method(2) in {1, 2}, method(6) = capacity index (0 = none),
 * mthlim(1:mwaves) in 0..5.  Bit for bit with the reference. */
int clawgpu1_step1(clawgpu1_ctx* ctx, double dx, double dt, const int* method, const int* mthlim,
                   int use_fwave, double* cfl);

/* The claw1 time loop (src/1d/claw1.f) on the device, with its order of
 * operations (the new dt is chosen before the Courant check; a retake does
 * not clamp dt to tend again) and info codes (0, 2, 3, 11, 12; -5 for a NaN
 * wave speed). */
int clawgpu1_claw1(clawgpu1_ctx* ctx, double dx, double tstart, double* tend, double* dtv,
                   double* cflv, int* nv, const int* method, const int* mthlim, const int* mthbc,
                   int use_fwave, int* info);

/* User routines of the 1D plug-in contract (src/1d/claw1.f:386, called at
 * claw1.f:478-510: bc1, b4step1, src1), with the conventions of
 * clawgpu_hooks: user pointer first, the reference's arguments by value,
 * arrays in the reference layout with ghost cells; NULL members take the
 * built-in bc1 (codes 1-3) or do nothing. */
typedef struct clawgpu1_hooks {
    void* user;
    void (*bc1)(void* user, int meqn, int mbc, int mx, double xlower, double dx, double* q,
                int maux, double* aux, double t, double dt, const int* mthbc);
    void (*b4step1)(void* user, int mbc, int mx, int meqn, double* q, double xlower, double dx,
                    double t, double dt, int maux, double* aux);
    void (*src1)(void* user, int meqn, int mbc, int mx, double xlower, double dx, double* q,
                 int maux, double* aux, double t, double dt);
} clawgpu1_hooks;

/* clawgpu1_claw1 with user routines called where claw1 calls them
 * (claw1.f:478-510): bc1 then b4step1 before every attempt; with
 * method(5) = 2 src1 over dt/2 at t before the step and over dt/2 at
 * t + dt/2 after it, with method(5) = 1 src1 over dt at t + dt after it.
 * The post-step source precedes the Courant check, so a rejected attempt
 * calls it too, and a rejection restores q as it was when the step began
 * (claw1.f:459 and 539).  aux changed by b4step1 is copied back. */
int clawgpu1_claw1_hooks(clawgpu1_ctx* ctx, double dx, double xlower, double tstart, double* tend,
                         double* dtv, double* cflv, int* nv, const int* method, const int* mthlim,
                         const int* mthbc, int use_fwave, const clawgpu1_hooks* hooks, int* info);

#ifdef __cplusplus
}
#endif
#endif
