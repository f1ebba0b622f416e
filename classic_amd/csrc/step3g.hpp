// Host-side interface of the 3D records path (step3g.hip): the unsplit step
// with the reference's rounding and order, for the capacity function and the
// exact arithmetic mode, and the Courant numbers of a step with capacity.
#pragma once

#include <hip/hip_runtime.h>

#include "step3d_kernel.hpp"

namespace clawgpu {
namespace s3 {

struct Args3g {
    const double* q;     // qold planes
    double* qn;          // qnew planes
    const double* aux;   // aux planes: Z, c, ...
    const double* capa;  // capacity plane, or nullptr
    double* rec;         // records, rec3_doubles(L) doubles
    Layout3 L;
    double dt3[3];       // dt/dx, dt/dy, dt/dz
    int m2, m3, m4, fwave, lim0, lim1;
    int dir;             // set by the launcher
};

// doubles of records the largest direction needs
long rec3_doubles(const Layout3& L);
// qnew = step3(qold): copy, then records + combine for x, y, z
int launch_step3_records(Args3g A, hipStream_t s);
// Courant numbers (dt/d / kappa) * c of the three directions: raw bits in
// bits[0..2], NaN flag in bits[3] (cleared by the caller)
int launch_capa_cfl(const double* c, const double* kappa, const Layout3& L, const double* d3,
                    unsigned long long* bits, hipStream_t s);

}  // namespace s3
}  // namespace clawgpu
