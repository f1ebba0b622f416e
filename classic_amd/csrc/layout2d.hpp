// Conversion between the reference layout q(meqn, 1-mbc:mx+mbc, 1-mbc:my+mbc)
// (Fortran order, m fastest) and the device layout (one padded plane per
// field).  Each translation unit that includes this gets its own copy of the
// kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace clawgpu {
namespace {

__global__ void aos_to_planes(const double* __restrict__ src, double* __restrict__ dst, Layout2 L,
                              int meqn)
{
    const int nx = L.mx + 2 * L.mbc, ny = L.my + 2 * L.mbc;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nx * ny) return;
    int ii = (int)(t % nx), jj = (int)(t / nx);
    for (int m = 0; m < meqn; m++) dst[m * L.fs + (long)jj * L.pitch + ii] = src[t * meqn + m];
}

__global__ void planes_to_aos(const double* __restrict__ src, double* __restrict__ dst, Layout2 L,
                              int meqn)
{
    const int nx = L.mx + 2 * L.mbc, ny = L.my + 2 * L.mbc;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nx * ny) return;
    int ii = (int)(t % nx), jj = (int)(t / nx);
    for (int m = 0; m < meqn; m++) dst[t * meqn + m] = src[m * L.fs + (long)jj * L.pitch + ii];
}

}  // namespace
}  // namespace clawgpu
