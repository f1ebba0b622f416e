// Types shared by the host API and the step kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace clawgpu {

constexpr int kMaxWaves = 8;
constexpr int kWavesPerBlock = 4;
constexpr int kStripOut = 60;  // output columns per 64-lane strip (lanes 2..61)

struct SolverParams {
    double rho, bulk, cc, zz;  // acoustics: density, bulk modulus, c, impedance
    double gamma, gamma1;      // Euler: gamma and gamma - 1
    int efix;                  // Euler: Harten-Hyman entropy fix
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
