// Device-side user routines for the tests (clawgpu_claw2_dev_hooks,
// clawgpu3_claw3_dev_hooks): the test routines of tests/hooks_cases.py
// written as HIP kernels, built into tests/device_hooks.co by the package
// build.
//   user array: [0] k, [1] number of logged calls, [2] room for calls, then
//   (kind, t, dt) per call (kind 1 = b4step2, 2 = src2)
//   src2: q(1) <- q(1) - (k dt) q(1) on the interior cells
//   bc2: bc2's extrapolation (mthbc = 1 on every side)
#include "clawgpu_device.h"

template <class G>
__device__ inline void log_call(const G& g, double kind)
{
    if (!clawgpu_dev_first()) return;
    const int n = (int)g.user[1];
    if (n < (int)g.user[2]) {
        g.user[3 + 3 * n] = kind;
        g.user[4 + 3 * n] = g.t;
        g.user[5 + 3 * n] = g.dt;
    }
    g.user[1] = n + 1;
}

extern "C" __global__ void test_b4step2(clawgpu_dev_grid g) { log_call(g, 1.0); }

extern "C" __global__ void test_src2(clawgpu_dev_grid g)
{
    log_call(g, 2.0);
    int i, j;
    if (!clawgpu_dev_cell(g, &i, &j)) return;
    if (i < 1 || i > g.mx || j < 1 || j > g.my) return;
    const double kd = g.user[0] * g.dt;
    double* q = g.q + clawgpu_dev_off(g, i, j);
    q[0] = q[0] - kd * q[0];
}

extern "C" __global__ void test_bc2_extrap(clawgpu_dev_grid g)
{
    int i, j;
    if (!clawgpu_dev_cell(g, &i, &j)) return;
    if (i >= 1 && i <= g.mx && j >= 1 && j <= g.my) return;
    const int si = i < 1 ? 1 : (i > g.mx ? g.mx : i);
    const int sj = j < 1 ? 1 : (j > g.my ? g.my : j);
    const long d = clawgpu_dev_off(g, i, j), s = clawgpu_dev_off(g, si, sj);
    for (int m = 0; m < g.meqn; m++) g.q[m * g.fs + d] = g.q[m * g.fs + s];
}

// ---- 3D (the same user array)
extern "C" __global__ void test_b4step3(clawgpu_dev_grid3 g) { log_call(g, 1.0); }

extern "C" __global__ void test_src3(clawgpu_dev_grid3 g)
{
    log_call(g, 2.0);
    int i, j, k;
    if (!clawgpu_dev_cell3(g, &i, &j, &k)) return;
    if (i < 1 || i > g.mx || j < 1 || j > g.my || k < 1 || k > g.mz) return;
    const double kd = g.user[0] * g.dt;
    double* q = g.q + clawgpu_dev_off3(g, i, j, k);
    q[0] = q[0] - kd * q[0];
}

// bc3's extrapolation (mthbc = 1 on every face)
extern "C" __global__ void test_bc3_extrap(clawgpu_dev_grid3 g)
{
    int i, j, k;
    if (!clawgpu_dev_cell3(g, &i, &j, &k)) return;
    if (i >= 1 && i <= g.mx && j >= 1 && j <= g.my && k >= 1 && k <= g.mz) return;
    const int si = i < 1 ? 1 : (i > g.mx ? g.mx : i);
    const int sj = j < 1 ? 1 : (j > g.my ? g.my : j);
    const int sk = k < 1 ? 1 : (k > g.mz ? g.mz : k);
    const long d = clawgpu_dev_off3(g, i, j, k), s = clawgpu_dev_off3(g, si, sj, sk);
    for (int m = 0; m < g.meqn; m++) g.q[m * g.fs + d] = g.q[m * g.fs + s];
}

// b4step3 that changes the sound speed of every cell (aux(2)) as
// c <- c * (1 + user[0] * dt), logging the call
extern "C" __global__ void test_b4step3_speed(clawgpu_dev_grid3 g)
{
    log_call(g, 1.0);
    int i, j, k;
    if (!clawgpu_dev_cell3(g, &i, &j, &k)) return;
    const long o = clawgpu_dev_off3(g, i, j, k);
    g.aux[g.fs + o] = g.aux[g.fs + o] * (1.0 + g.user[0] * g.dt);
}
