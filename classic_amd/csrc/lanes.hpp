// Cross-lane shifts of doubles within one 64-lane wavefront (gfx950 DPP).
#pragma once

#include <hip/hip_runtime.h>

#ifndef CLAWGPU_DPP
#define CLAWGPU_DPP 1
#endif

namespace clawgpu {

// Whole-wave shift by one lane: up(v) gives lane l the value of lane l-1,
// dn(v) the value of lane l+1 (lanes with no source get 0 or their own).
// The DPP moves zero-fill through bound_ctrl instead of merging into an
// "old" operand of 0, which would cost a v_mov of the destination per
// 32-bit half (one instruction in seven of the 2D Euler step).
__device__ inline double lane_up(double v)
{
#if CLAWGPU_DPP
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_update_dpp(0, lo, 0x138, 0xf, 0xf, true);  // wave_shr:1
    hi = __builtin_amdgcn_update_dpp(0, hi, 0x138, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
#else
    return __shfl_up(v, 1);
#endif
}

__device__ inline double lane_dn(double v)
{
#if CLAWGPU_DPP
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_update_dpp(0, lo, 0x130, 0xf, 0xf, true);  // wave_shl:1
    hi = __builtin_amdgcn_update_dpp(0, hi, 0x130, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
#else
    return __shfl_down(v, 1);
#endif
}

// the same shifts for one 32-bit integer
__device__ inline int lane_up_i(int v)
{
#if CLAWGPU_DPP
    return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, true);
#else
    return __shfl_up(v, 1);
#endif
}

__device__ inline int lane_dn_i(int v)
{
#if CLAWGPU_DPP
    return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, true);
#else
    return __shfl_down(v, 1);
#endif
}

template <class T>
__device__ inline T lane_up_t(const T& v)
{
    T r;
    const double* s = reinterpret_cast<const double*>(&v);
    double* d = reinterpret_cast<double*>(&r);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / sizeof(double)); k++) d[k] = lane_up(s[k]);
    return r;
}

// The wave's maximum of a non-negative value, stored by lane 0 in part[w]
// (every lane of the wave must call it).  Kernels with a wave per few cells
// collect their Courant numbers this way and reduce them afterwards with
// wave_max_reduce: one atomic per wave on one address serialises at about
// 11 ns each.
__device__ inline void store_wave_max(double v, double* part, long w)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fmax(v, __shfl_xor(v, d));
    if ((threadIdx.x & 63) == 0) part[w] = v;
}

// Body of a reduction kernel: the max of part[0..n-1] (non-negative) into
// *out as raw bits, one atomic per wave of that kernel.
__device__ inline void wave_max_reduce(const double* part, long n, unsigned long long* out)
{
    double m = 0.0;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (long)gridDim.x * blockDim.x)
        m = fmax(m, part[e]);
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) m = fmax(m, __shfl_xor(m, d));
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

// Max of a non-negative double over the wave, then one atomic per wave.
__device__ inline void publish_cfl(double cfl, unsigned long long* bits)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) cfl = fmax(cfl, __shfl_xor(cfl, d));
    if ((threadIdx.x & 63) == 0) atomicMax(bits, (unsigned long long)__double_as_longlong(cfl));
}

// Courant contribution max(dtdx*s, -dtdx*s) (flux2.f90:129-130); a NaN
// speed becomes +inf so that the host refuses the step instead of
// silently dropping it from the max.
__device__ inline double cfl_term(double dtdx, double s)
{
    double t = fmax(dtdx * s, -(dtdx * s));
    return (t == t) ? t : __builtin_huge_val();
}

}  // namespace clawgpu
