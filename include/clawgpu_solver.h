/* User Riemann solvers for the 2D kernels (the rpn2/rpt2 plug-in contract of
 * src/2d/claw2.f:312-402, src/2d/flux2.f90:121-235).
 *
 * A solver is a policy struct that the library's kernel templates are
 * instantiated with; CLAWGPU_EXPORT_SOLVER2 turns it into the kernels of a
 * gfx950 code object, which the host loads with clawgpu_solver2_load and
 * uses through clawgpu_create_user (include/clawgpu.h).  Build with
 *   hipcc --genco --offload-arch=gfx950 -O3 -std=c++17 -ffp-contract=off \
 *         -I include -I classic_amd/csrc solver.hip -o solver.co
 * (a translation unit that puts "#pragma clang fp contract(fast)" before the
 * includes and instantiates with clawgpu::FastMath gets the library's FAST
 * arithmetic instead).
 *
 * The policy (the built-in ones are Acoustics2 and Euler2 in riemann2d.hpp):
 *
 *   template <class Mth> struct MySolver {
 *     using Math = Mth;                 // clawgpu::ExactMath or FastMath
 *     static constexpr int M = ...;     // meqn, 1..8
 *     static constexpr int MW = ...;    // mwaves, 1..8
 *     struct Cell { double q[...]; ... };   // what rpn needs of one cell
 *     static constexpr int XSHIFT = ...;    // leading doubles of Cell that
 *                                           // rpn reads of the left cell
 *     static constexpr bool kOuterSpeedsBound = ...;  // |s| of the first or
 *                                           // last wave bounds every |s|
 *     struct Cache { ... };             // rpn -> rpt values of an interface
 *     __device__ static Cell cell(const clawgpu::SolverParams& P, const double* q);
 *     template <int MU, int MV>         // normal / transverse momentum index
 *     __device__ static void rpn(const clawgpu::SolverParams& P, const Cell& L,
 *                                const Cell& R, double w[MW][M], double s[MW],
 *                                double amdq[M], double apdq[M], Cache& c);
 *     template <int MU, int MV>
 *     __device__ static void rpt(const clawgpu::SolverParams& P, const Cache& c,
 *                                const double* asdq, double bmasdq[M],
 *                                double bpasdq[M]);
 *   };
 *
 * rpn solves the interface between L (cell i-1) and R (cell i) in the
 * direction whose momentum is component MU (MU = 1: x, MU = 2: y; 0-based
 * component indices), returning the waves, speeds and fluctuations as the
 * reference rpn2 does for one interface; rpt splits asdq in the transverse
 * direction (component MV) as the reference rpt2 does.  The solver's
 * parameters are P.user[0 .. P.nuser-1] (clawgpu_create_user's params).
 * Wall boundaries (mthbc = 3) negate components 2 and 3 (1-based), as bc2
 * does, and need M >= 3.
 */
#ifndef CLAWGPU_SOLVER_H
#define CLAWGPU_SOLVER_H

#include "clawgpu.h"
#include "step2d_kernel.hpp"
#include "split2d_kernel.hpp"
#include "step3d_user.hpp"

/* version of the kernel-argument layout a code object was built against */
#define CLAWGPU_SOLVER2_ABI 1

/* The kernels of solver NAME (a policy type without commas: use an alias
 * for a template instance): NAME_info (the solver's shape and the argument
 * layout, checked at load), NAME_step (unsplit step), NAME_loop1 / NAME_loop2
 * (one attempt of the device-side claw2 loop without / with walls),
 * NAME_sweep_x / NAME_sweep_y (dimensional splitting), NAME_probe (rpn and
 * rpt on given states, for tests). */
#define CLAWGPU_EXPORT_SOLVER2(NAME, S)                                                          \
    static_assert(S::M >= 1 && S::M <= clawgpu::kMaxWaves, "meqn out of range");                \
    static_assert(S::MW >= 1 && S::MW <= clawgpu::kMaxWaves, "mwaves out of range");           \
    extern "C" __device__ __attribute__((used)) int NAME##_info[8] = {                          \
        CLAWGPU_SOLVER2_ABI, S::M, S::MW, S::Math::fast ? 1 : 0, (int)sizeof(clawgpu::StepArgs), \
        (int)sizeof(clawgpu::LoopArgs), (int)sizeof(clawgpu::SweepArgs),                        \
        (int)sizeof(clawgpu::SolverParams)};                                                    \
    extern "C" __global__ __launch_bounds__(256, CLAWGPU_MIN_WAVES) void NAME##_step(           \
        clawgpu::StepArgs a)                                                                    \
    {                                                                                           \
        clawgpu::step2_unsplit_body<S, clawgpu::CfgAny>(a);                                     \
    }                                                                                           \
    extern "C" __global__ __launch_bounds__(256, S::Math::fast ? 2 : CLAWGPU_MIN_WAVES) void    \
        NAME##_loop1(clawgpu::StepArgs a, clawgpu::LoopArgs l)                                  \
    {                                                                                           \
        clawgpu::step2_loop_body<S, clawgpu::CfgAny, 1>(a, l);                                  \
    }                                                                                           \
    extern "C" __global__ __launch_bounds__(256, S::Math::fast ? 2 : CLAWGPU_MIN_WAVES) void    \
        NAME##_loop2(clawgpu::StepArgs a, clawgpu::LoopArgs l)                                  \
    {                                                                                           \
        clawgpu::step2_loop_body<S, clawgpu::CfgAny, 2>(a, l);                                  \
    }                                                                                           \
    extern "C" __global__ __launch_bounds__(256) void NAME##_sweep_x(clawgpu::SweepArgs a)      \
    {                                                                                           \
        clawgpu::sweep_x_body<S>(a);                                                            \
    }                                                                                           \
    extern "C" __global__ __launch_bounds__(256) void NAME##_sweep_y(clawgpu::SweepArgs a)      \
    {                                                                                           \
        clawgpu::sweep_y_body<S>(a);                                                            \
    }                                                                                           \
    extern "C" __global__ void NAME##_probe(int n, int ixy, clawgpu::SolverParams P,            \
                                            const double* ql, const double* qr,                 \
                                            const double* asdq, double* wave, double* s,        \
                                            double* amdq, double* apdq, double* bm, double* bp) \
    {                                                                                           \
        clawgpu::solver_probe_body<S>(n, ixy, P, ql, qr, asdq, wave, s, amdq, apdq, bm, bp);    \
    }

/* 3D solvers (rpn3 of src/3d/claw3.f:292-335) for the dimensionally split
 * step (method(3) = -1: dimsp3/step3ds, which has no transverse terms, so
 * no rpt3/rptt3):
 *
 *   struct MySolver3 {
 *     using Math = clawgpu::ExactMath;  // or FastMath
 *     static constexpr int M = ..., MW = ...;  // meqn, mwaves (1..8)
 *     static constexpr int MAUX = ...;         // aux components read (0..8)
 *     struct Cell { double ...; };             // what rpn needs of one cell
 *     __device__ static Cell cell(const clawgpu::SolverParams& P,
 *                                 const double* q, const double* aux);
 *     template <int MU>                        // normal momentum: 1 x, 2 y, 3 z
 *     __device__ static void rpn(const clawgpu::SolverParams& P, const Cell& L,
 *                                const Cell& R, double w[MW][M], double s[MW],
 *                                double amdq[M], double apdq[M]);
 *   };
 *
 * cell() gets q(1..M) and aux(1..MAUX) of one cell; rpn solves the interface
 * between L (cell i-1) and R (cell i) as the reference rpn3 does. */
#define CLAWGPU_SOLVER3_ABI clawgpu::s3u::kAbi

/* Kernels of 3D solver NAME: NAME_info3, NAME_sweep3_x, NAME_sweep3_yz. */
#define CLAWGPU_EXPORT_SOLVER3(NAME, S)                                                          \
    static_assert(S::M >= 1 && S::M <= clawgpu::kMaxWaves, "meqn out of range");                \
    static_assert(S::MW >= 1 && S::MW <= clawgpu::kMaxWaves, "mwaves out of range");           \
    static_assert(S::MAUX >= 0 && S::MAUX <= 8, "MAUX out of range");                          \
    extern "C" __device__ __attribute__((used)) int NAME##_info3[8] = {                         \
        CLAWGPU_SOLVER3_ABI, S::M, S::MW, S::MAUX, S::Math::fast ? 1 : 0,                        \
        (int)sizeof(clawgpu::s3u::Sweep3U), (int)sizeof(clawgpu::SolverParams), 0};            \
    extern "C" __global__ __launch_bounds__(256) void NAME##_sweep3_x(clawgpu::s3u::Sweep3U a)  \
    {                                                                                           \
        clawgpu::s3u::sweep3u_x_body<S>(a);                                                     \
    }                                                                                           \
    extern "C" __global__ __launch_bounds__(256) void NAME##_sweep3_yz(clawgpu::s3u::Sweep3U a, \
                                                                       int dir)                 \
    {                                                                                           \
        clawgpu::s3u::sweep3u_yz_body<S>(a, dir);                                               \
    }

#endif
