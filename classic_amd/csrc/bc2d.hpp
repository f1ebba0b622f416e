// Ghost-cell fill of a 2D grid (src/2d/bc2.f:4-215) as a per-cell device
// function, shared by the single-grid and the batched paths.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace clawgpu {

struct BcArgs {
    double* q;
    Layout2 L;
    int meqn;
    int mthbc[4];
    unsigned long long* cfl_bits;
};

__device__ inline int bc_map(int k, int n, int mbc, int lo, int hi, bool& flip)
{
    (void)mbc;
    flip = false;
    if (k < 1) {
        int g = 1 - k;  // ghost layer 1..mbc
        if (lo == 1) return 1;
        if (lo == 2) return n + 1 - g;
        flip = true;
        return g;
    }
    if (k > n) {
        int g = k - n;
        if (hi == 1) return n;
        if (hi == 2) return g;
        flip = true;
        return n + 1 - g;
    }
    return k;
}

// Ghost cell number t of the grid (rows of ghost cells first, then the ghost
// columns of the interior rows); 2*mbc*(mx+2mbc) + 2*mbc*my cells in all.
// The value is the reference's: a cell outside in both directions ends with
// the bottom/top pass over the filled ghost columns, i.e. the source cell
// has both indices mapped, and each wall negates its normal momentum.
__device__ inline long bc2_count(const Layout2& L)
{
    return 2L * L.mbc * (L.mx + 2 * L.mbc) + 2L * L.mbc * L.my;
}

__device__ inline void bc2_cell(double* q, const Layout2& L, int meqn, const int* mthbc, long t)
{
    const int nx = L.mx + 2 * L.mbc;
    if (t < 0 || t >= bc2_count(L)) return;
    int i, j;
    if (t < 2L * L.mbc * nx) {
        int r = (int)(t / nx);  // 0..2mbc-1
        i = 1 - L.mbc + (int)(t % nx);
        j = r < L.mbc ? 1 - L.mbc + r : L.my + 1 + (r - L.mbc);
    } else {
        long u = t - 2L * L.mbc * nx;
        int c = (int)(u / L.my);  // 0..2mbc-1
        j = 1 + (int)(u % L.my);
        i = c < L.mbc ? 1 - L.mbc + c : L.mx + 1 + (c - L.mbc);
    }
    bool fy, fx;
    int js = bc_map(j, L.my, L.mbc, mthbc[2], mthbc[3], fy);
    int is = bc_map(i, L.mx, L.mbc, mthbc[0], mthbc[1], fx);
    long src = L.off(is, js), dst = L.off(i, j);
    for (int m = 0; m < meqn; m++) {
        double v = q[m * L.fs + src];
        if ((m == 1 && fx) || (m == 2 && fy)) v = -v;
        q[m * L.fs + dst] = v;
    }
}

}  // namespace clawgpu
