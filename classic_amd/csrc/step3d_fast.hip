// Unsplit 3D step, FMA-contracted arithmetic (held to 1e-12 of the reference).
#pragma clang fp contract(fast)
#include "step3d.hpp"
#include "step3d_zmarch.hpp"

namespace clawgpu {
namespace s3 {

template <int TY, int TZ, class Cfg>
static int launch_cfg(Args3 a, hipStream_t s)
{
    a.nty = (a.L.my + TY - 1) / TY;
    a.ntz = (a.L.mz + TZ - 1) / TZ;
    long nblocks = (long)a.ntx * a.nty * a.ntz;
    if (nblocks <= 0) return 0;
    if (a.xcd) nblocks = 8 * ((nblocks + 7) / 8);
    // the material-table variants are compiled for the two fastest tiles
    constexpr bool tables = (TY == 4 && TZ == 2) || (TY == 2 && TZ == 4);
    const dim3 g((unsigned)nblocks);
    if (tables && a.nmat > 0 && a.nmat <= kMaxKcMat)
        hipLaunchKernelGGL((step3_tile_kernel<TY, TZ, Cfg, tables, tables>), g, dim3(64), 0, s, a);
    else if (tables && a.nmat > 0 && a.nmat <= kMaxMat)
        hipLaunchKernelGGL((step3_tile_kernel<TY, TZ, Cfg, tables, false>), g, dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL((step3_tile_kernel<TY, TZ, Cfg, false, false>), g, dim3(64), 0, s, a);
    return (int)hipGetLastError();
}

// the (2, 22) method with the same limiter on both waves is compiled per
// limiter; everything else takes the flags from the arguments
template <int TY, int TZ>
static int launch_tile(const Args3& a, hipStream_t s)
{
    if (a.m2 == 2 && a.m3 == 2 && a.m4 == 2 && a.lim0 == a.lim1) {
        switch (a.lim0) {
        case 1: return launch_cfg<TY, TZ, Cfg3Fix<2, 2, 2, 1, 1>>(a, s);
        case 2: return launch_cfg<TY, TZ, Cfg3Fix<2, 2, 2, 2, 2>>(a, s);
        case 3: return launch_cfg<TY, TZ, Cfg3Fix<2, 2, 2, 3, 3>>(a, s);
        case 4: return launch_cfg<TY, TZ, Cfg3Fix<2, 2, 2, 4, 4>>(a, s);
        default: break;
        }
    }
    return launch_cfg<TY, TZ, Cfg3Any>(a, s);
}

// z-marching kernel: rows per block TY, planes per block kZSeg
constexpr int kZSeg = 64;

template <int TY, class Cfg>
static int launch_zm_cfg(Args3 a, hipStream_t s)
{
    a.nty = (a.L.my + TY - 1) / TY;
    const int nzs = (a.L.mz + kZSeg - 1) / kZSeg;
    long nblocks = (long)a.ntx * a.nty * nzs;
    if (nblocks <= 0) return 0;
    if (a.xcd & 1) nblocks = 8 * ((nblocks + 7) / 8);
    const dim3 g((unsigned)nblocks), b(64 * (TY + 2));
    if (a.nmat <= kMaxKcMat)
        hipLaunchKernelGGL((step3_zmarch_kernel<TY, Cfg, true>), g, b, 0, s, a, kZSeg);
    else
        hipLaunchKernelGGL((step3_zmarch_kernel<TY, Cfg, false>), g, b, 0, s, a, kZSeg);
    return (int)hipGetLastError();
}

template <int TY>
static int launch_zmarch(const Args3& a, hipStream_t s)
{
    if (a.m2 == 2 && a.m3 == 2 && a.m4 == 2 && a.lim0 == a.lim1) {
        switch (a.lim0) {
        case 1: return launch_zm_cfg<TY, Cfg3Fix<2, 2, 2, 1, 1>>(a, s);
        case 2: return launch_zm_cfg<TY, Cfg3Fix<2, 2, 2, 2, 2>>(a, s);
        case 3: return launch_zm_cfg<TY, Cfg3Fix<2, 2, 2, 3, 3>>(a, s);
        case 4: return launch_zm_cfg<TY, Cfg3Fix<2, 2, 2, 4, 4>>(a, s);
        default: break;
        }
    }
    return launch_zm_cfg<TY, Cfg3Any>(a, s);
}

int launch_step3_tile(const Args3& a, int tile, hipStream_t s)
{
    // the z-marching kernel reads materials through the material table
    if (tile >= 5 && (a.nmat <= 0 || a.nmat > kMaxMat)) tile = 4;
    switch (tile) {
    case 5: return launch_zmarch<10>(a, s);
    case 6: return launch_zmarch<6>(a, s);
    case 0: return launch_tile<4, 4>(a, s);
    case 2: return launch_tile<2, 2>(a, s);
    case 3: return launch_tile<8, 2>(a, s);
    case 4: return launch_tile<2, 4>(a, s);
    default: return launch_tile<4, 2>(a, s);
    }
}

}  // namespace s3
}  // namespace clawgpu
