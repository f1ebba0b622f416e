// User Riemann solvers for the plug-in tests, built by __graft_entry__.build()
// into tests/user_solvers.co (EXACT arithmetic) and, with
// -DUSER_SOLVERS_FAST, tests/user_solvers_fast.co (FAST).
//   user_acoustics  constant-coefficient acoustics, params (Z, c): the
//                   algorithm of the built-in acoustics solver, so every 2D
//                   path must give the built-in results.
//   user_advection  q_t + u q_x + v q_y = 0, params (u, v): the
//                   variable-coefficient advection solver of the reference's
//                   annulus example with constant edge velocities.
//   user_vc_acoustics3  3D variable-coefficient acoustics, aux (Z, c): the
//                   algorithm of the built-in 3D solver (rpn3 of the
//                   reference's heterogeneous acoustics example), so dimsp3
//                   and claw3 must give the built-in split results.
//   user_advection3 q_t + u q_x + v q_y + w q_z = 0, params (u, v, w).
#ifdef USER_SOLVERS_FAST
#pragma clang fp contract(fast)
#endif
#include "clawgpu_solver.h"

#ifdef USER_SOLVERS_FAST
using Mth = clawgpu::FastMath;
#else
using Mth = clawgpu::ExactMath;
#endif

using clawgpu::SolverParams;

struct UserAcoustics {
    using Math = Mth;
    static constexpr int M = 3, MW = 2;
    struct Cell {
        double q[3];
    };
    static constexpr int XSHIFT = 3;
    static constexpr bool kOuterSpeedsBound = true;
    struct Cache {};

    __device__ static inline Cell cell(const SolverParams&, const double* q)
    {
        return Cell{{q[0], q[1], q[2]}};
    }

    // waves r1 = (-Z, 1) at -c and r2 = (Z, 1) at +c in (p, normal velocity)
    template <int MU, int MV>
    __device__ static inline void rpn(const SolverParams& P, const Cell& L, const Cell& R,
                                      double w[MW][M], double s[MW], double am[M], double ap[M],
                                      Cache&)
    {
        const double Z = P.user[0], c = P.user[1];
        const double dp = R.q[0] - L.q[0], du = R.q[MU] - L.q[MU];
        double a1, a2;
        if constexpr (Mth::fast) {
            const double h = 0.5 / Z;
            a1 = (-dp + Z * du) * h;
            a2 = (dp + Z * du) * h;
        } else {
            a1 = (-dp + Z * du) / (2.0 * Z);
            a2 = (dp + Z * du) / (2.0 * Z);
        }
        w[0][0] = -a1 * Z;
        w[0][MU] = a1;
        w[0][MV] = 0.0;
        w[1][0] = a2 * Z;
        w[1][MU] = a2;
        w[1][MV] = 0.0;
        s[0] = -c;
        s[1] = c;
        for (int m = 0; m < M; m++) {
            am[m] = s[0] * w[0][m];
            ap[m] = s[1] * w[1][m];
        }
    }

    template <int MU, int MV>
    __device__ static inline void rpt(const SolverParams& P, const Cache&, const double* asdq,
                                      double bm[M], double bp[M])
    {
        const double Z = P.user[0], c = P.user[1];
        double a1, a2;
        if constexpr (Mth::fast) {
            const double h = 0.5 / Z;
            a1 = (-asdq[0] + Z * asdq[MV]) * h;
            a2 = (asdq[0] + Z * asdq[MV]) * h;
        } else {
            a1 = (-asdq[0] + Z * asdq[MV]) / (2.0 * Z);
            a2 = (asdq[0] + Z * asdq[MV]) / (2.0 * Z);
        }
        bm[0] = c * a1 * Z;
        bm[MU] = 0.0;
        bm[MV] = -c * a1;
        bp[0] = c * a2 * Z;
        bp[MU] = 0.0;
        bp[MV] = c * a2;
    }
};

struct UserAdvection {
    using Math = Mth;
    static constexpr int M = 1, MW = 1;
    struct Cell {
        double q[1];
    };
    static constexpr int XSHIFT = 1;
    static constexpr bool kOuterSpeedsBound = true;
    struct Cache {};

    __device__ static inline Cell cell(const SolverParams&, const double* q) { return Cell{{q[0]}}; }

    // one wave, the jump, at the normal velocity (u for x, v for y)
    template <int MU, int MV>
    __device__ static inline void rpn(const SolverParams& P, const Cell& L, const Cell& R,
                                      double w[MW][M], double s[MW], double am[M], double ap[M],
                                      Cache&)
    {
        const double u = MU == 1 ? P.user[0] : P.user[1];
        w[0][0] = R.q[0] - L.q[0];
        s[0] = u;
        am[0] = fmin(u, 0.0) * w[0][0];
        ap[0] = fmax(u, 0.0) * w[0][0];
    }

    // the transverse velocity carries asdq down (negative part) and up
    template <int MU, int MV>
    __device__ static inline void rpt(const SolverParams& P, const Cache&, const double* asdq,
                                      double bm[M], double bp[M])
    {
        const double v = MU == 1 ? P.user[1] : P.user[0];
        bm[0] = fmin(v, 0.0) * asdq[0];
        bp[0] = fmax(v, 0.0) * asdq[0];
    }
};

CLAWGPU_EXPORT_SOLVER2(user_acoustics, UserAcoustics)
CLAWGPU_EXPORT_SOLVER2(user_advection, UserAdvection)

// 3D split-step solvers (CLAWGPU_EXPORT_SOLVER3): rpn<MU> solves the
// Riemann problem normal to direction MU (1 x, 2 y, 3 z) between cells L and R.
struct UserVcAcoustics3 {
    using Math = Mth;
    static constexpr int M = 4, MW = 2, MAUX = 2;
    struct Cell {
        double q[4];
        double Z, c;
    };

    __device__ static inline Cell cell(const SolverParams&, const double* q, const double* aux)
    {
        return Cell{{q[0], q[1], q[2], q[3]}, aux[0], aux[1]};
    }

    // the jump in (p, normal velocity) splits into (-Z_L, 1) a1 at -c_L and
    // (Z_R, 1) a2 at +c_R
    template <int MU>
    __device__ static inline void rpn(const SolverParams&, const Cell& L, const Cell& R,
                                      double w[MW][M], double s[MW], double am[M], double ap[M])
    {
        const double d1 = R.q[0] - L.q[0], d2 = R.q[MU] - L.q[MU];
        const double den = L.Z + R.Z;
        const double a1 = Mth::div(-d1 + R.Z * d2, den);
        const double a2 = Mth::div(d1 + L.Z * d2, den);
        for (int m = 0; m < M; m++) {
            w[0][m] = 0.0;
            w[1][m] = 0.0;
        }
        w[0][0] = -a1 * L.Z;
        w[0][MU] = a1;
        w[1][0] = a2 * R.Z;
        w[1][MU] = a2;
        s[0] = -L.c;
        s[1] = R.c;
        for (int m = 0; m < M; m++) {
            am[m] = s[0] * w[0][m];
            ap[m] = s[1] * w[1][m];
        }
    }
};

struct UserAdvection3 {
    using Math = Mth;
    static constexpr int M = 1, MW = 1, MAUX = 0;
    struct Cell {
        double q[1];
    };

    __device__ static inline Cell cell(const SolverParams&, const double* q, const double*)
    {
        return Cell{{q[0]}};
    }

    template <int MU>
    __device__ static inline void rpn(const SolverParams& P, const Cell& L, const Cell& R,
                                      double w[MW][M], double s[MW], double am[M], double ap[M])
    {
        const double u = P.user[MU - 1];
        w[0][0] = R.q[0] - L.q[0];
        s[0] = u;
        am[0] = fmin(u, 0.0) * w[0][0];
        ap[0] = fmax(u, 0.0) * w[0][0];
    }
};

CLAWGPU_EXPORT_SOLVER3(user_vc_acoustics3, UserVcAcoustics3)
CLAWGPU_EXPORT_SOLVER3(user_advection3, UserAdvection3)
