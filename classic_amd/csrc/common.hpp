// Types shared by the host API and the step kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace clawgpu {

constexpr int kMaxWaves = 8;
constexpr int kWavesPerBlock = 4;
constexpr int kStripOut = 60;  // output columns per 64-lane strip (lanes 2..61)

constexpr int kUserParams = 16;  // parameters of a user solver

struct SolverParams {
    double rho, bulk, cc, zz;  // acoustics: density, bulk modulus, c, impedance
    double gamma, gamma1;      // Euler: gamma and gamma - 1
    int efix;                  // Euler: Harten-Hyman entropy fix
    int nuser;                 // user solver: parameters given
    double user[kUserParams];  // user solver: its parameters, zero beyond nuser
};

// Field-plane layout of a grid with mbc ghost cells: element (m, i, j) at
// m*fs + (j+mbc-1)*pitch + (i+mbc-1), 1-based Fortran cell indices.
struct Layout2 {
    int mx, my, mbc, pitch;
    long fs;
    __host__ __device__ long off(int i, int j) const
    {
        return (long)(j + mbc - 1) * pitch + (i + mbc - 1);
    }
};

struct StepArgs {
    const double* qold;
    double* qnew;
    Layout2 L;
    double dtdx, dtdy;
    int order, trans, use_fwave;
    int lim[kMaxWaves];
    int nstrips, seglen, nsegs;
    unsigned long long* cfl_bits;  // max Courant number of the step, as raw bits
    SolverParams P;
    // limiter 6: phi read from per-interface tables (learned limiter);
    // entry (i, J, mw) at ((J + 1) * (mx + 4) + (i + 1)) * mwaves + mw for
    // i in [-1, mx+2], J in [-1, my+2]; x: interface between cells (i-1, J)
    // and (i, J); y: between (i, J-1) and (i, J)
    const double* phix;
    const double* phiy;
    // bcmap != 0: the ghost cells of qold are not filled; a load outside
    // the interior reads the interior cell the bc2 rules copy (codes
    // mthbc: 1 extrapolation, 2 periodic, 3 wall negating the normal
    // momentum), which is the value bc2 would have stored there
    int bcmap;
    int mthbc[4];
};

// State of the device-side claw2 loop (src/2d/claw2.f:596-753) before one
// attempt of a step: the buffer holding q, the attempt's dt and start time,
// and the loop's running outputs.
struct LoopState {
    double t, told, dt;             // time reached, start of the attempt, its dt
    double cflmax, dtmin, dtmax, cfl;
    int n;     // accepted steps so far
    int cur;   // buffer holding q (the attempt reads it, writes (cur+1)%3)
    int done;  // nothing left to do: every later attempt is a no-op
    int info;  // 0, or -5 for a non-finite Courant number
};

// One launch of the loop: attempt k.  st[2] holds the states before
// attempts of alternating parity, cflb[3] the Courant maxima (as raw bits)
// of the attempts k % 3; attempt k decides attempt k-1 from st[(k-1)&1] and
// cflb[(k-1)%3], stores the result in st[k&1] and clears cflb[(k+1)%3].
struct LoopArgs {
    double* q[3];
    double dx, dy, tstart, tend;
    double dtv2, cflv1, cflv2;  // dt_max, cfl_max, cfl_desired
    int maxn, vardt;
    LoopState* st;
    unsigned long long* cflb;
    int k;
};

// limiter id of the table-driven (learned) limiter
constexpr int kLimTable = 6;

__host__ __device__ inline long phi_index(int i, int J, int mx, int my, int mw, int mwaves)
{
    i = i < -1 ? -1 : (i > mx + 2 ? mx + 2 : i);
    J = J < -1 ? -1 : (J > my + 2 ? my + 2 : J);
    return ((long)(J + 1) * (mx + 4) + (i + 1)) * mwaves + mw;
}

enum { ARITH_EXACT = 0, ARITH_FAST = 1 };

struct SweepArgs;
int launch_sweep_exact(int solver, int dir, const SweepArgs& a, int nblocks, hipStream_t s);
int launch_sweep_fast(int solver, int dir, const SweepArgs& a, int nblocks, hipStream_t s);

// Launch the unsplit step kernel for solver id `solver`; returns a
// hipError_t value.  exact: reference rounding (FMA contraction off,
// correctly rounded division); fast: FMA contraction and reciprocal reuse.
int launch_step2_exact(int solver, const StepArgs& a, int nblocks, hipStream_t s);
int launch_step2_fast(int solver, const StepArgs& a, int nblocks, hipStream_t s);
// One attempt of the device-side claw2 loop (a: everything but the buffers,
// dt ratios and Courant slot, which come from the loop state) and the
// loop's last decision (one wave).
int launch_loop2_exact(int solver, const StepArgs& a, const LoopArgs& l, int nblocks, hipStream_t s);
int launch_loop2_fast(int solver, const StepArgs& a, const LoopArgs& l, int nblocks, hipStream_t s);
int launch_loop2_finish(const LoopArgs& l, hipStream_t s);
// Wavefronts of the step kernel launch_step2_* would pick for a that one CU
// holds at once (0 if the occupancy query fails).
int step2_waves_per_cu_exact(int solver, const StepArgs& a);
int step2_waves_per_cu_fast(int solver, const StepArgs& a);

int launch_probe_exact(int solver, int n, int ixy, const SolverParams& P, const double* ql,
                       const double* qr, const double* asdq, double* wave, double* s,
                       double* amdq, double* apdq, double* bm, double* bp, hipStream_t st);
int launch_probe_fast(int solver, int n, int ixy, const SolverParams& P, const double* ql,
                      const double* qr, const double* asdq, double* wave, double* s,
                      double* amdq, double* apdq, double* bm, double* bp, hipStream_t st);

}  // namespace clawgpu
