/* Helpers for device-side user routines (clawgpu_claw2_dev_hooks,
 * clawgpu3_claw3_dev_hooks): the cell a thread of the library's launch
 * handles and the address of q(m, i, j) / q(m, i, j, k).
 * Include from HIP code built into a code object with hipcc --genco. */
#ifndef CLAWGPU_DEVICE_H
#define CLAWGPU_DEVICE_H

#include <hip/hip_runtime.h>

#include "clawgpu.h"

/* The cell (i, j) of this thread, 1-based with ghost cells (i in
 * 1-mbc..mx+mbc); false for threads beyond the grid. */
__device__ inline bool clawgpu_dev_cell(const clawgpu_dev_grid& g, int* i, int* j)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int nx = g.mx + 2 * g.mbc, ny = g.my + 2 * g.mbc;
    if (t >= (long)nx * ny) return false;
    *i = (int)(t % nx) + 1 - g.mbc;
    *j = (int)(t / nx) + 1 - g.mbc;
    return true;
}

/* Offset of cell (i, j) within a field plane of q or aux. */
__device__ inline long clawgpu_dev_off(const clawgpu_dev_grid& g, int i, int j)
{
    return (long)(j + g.mbc - 1) * g.pitch + (i + g.mbc - 1);
}

/* 3D (clawgpu3_claw3_dev_hooks): the cell (i, j, k) of this thread, 1-based
 * with ghost cells; false for threads beyond the grid. */
__device__ inline bool clawgpu_dev_cell3(const clawgpu_dev_grid3& g, int* i, int* j, int* k)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nx = g.mx + 2 * g.mbc, ny = g.my + 2 * g.mbc;
    if (t >= g.fs) return false;
    *i = (int)(t % nx) + 1 - g.mbc;
    *j = (int)((t / nx) % ny) + 1 - g.mbc;
    *k = (int)(t / (nx * ny)) + 1 - g.mbc;
    return true;
}

/* Offset of cell (i, j, k) within a field plane of q or aux (3D). */
__device__ inline long clawgpu_dev_off3(const clawgpu_dev_grid3& g, int i, int j, int k)
{
    const long nx = g.mx + 2 * g.mbc, ny = g.my + 2 * g.mbc;
    return ((long)(k + g.mbc - 1) * ny + (j + g.mbc - 1)) * nx + (i + g.mbc - 1);
}

/* Whether this thread is the first of the launch (for work done once). */
__device__ inline bool clawgpu_dev_first(void)
{
    return blockIdx.x == 0 && threadIdx.x == 0;
}

#endif
