// Unsplit 3D step with the reference's rounding and summation order
// ("records path"): the capacity function (method(6) > 0) and the exact
// arithmetic mode of the 3D unsplit step.
//
// step3 (src/3d/step3.f90:127-866) calls flux3 on every x-, y- and z-slice
// and adds each slice's qadd/fadd/gadd/hadd to the nine cells around every
// position of the slice right away, in the order of its block passes.  The
// increments depend on qold only, so the work splits into two kernels per
// direction:
//   * rec3_kernel, one thread per slice position (l, p1, p2): flux3's values
//     at position l (src/3d/flux3.f90:205-608) -- qadd(l), fadd(l+1) -
//     fadd(l), and gadd/hadd(:, side, -1:1, l) summed in the order flux3's
//     loop over the interfaces forms them (interface l's A^+ side, then
//     interface l+1's A^- side) -- written to a record;
//   * comb3_kernel, one thread per interior cell: the nine records of the
//     slices that touch the cell, applied with step3's expressions
//     (:264-424 and the rotated forms) in the order of step3's passes over
//     2x2 blocks of slices (both block offsets, then blocks, then slices).
// All arithmetic is IEEE without contraction, so qnew equals the reference
// bit for bit.  Solver: variable-coefficient acoustics, q = (p, u, v, w),
// aux = (Z, c, ...); every flux difference the reference splits
// transversely has a zero transverse velocity, which the splits below keep
// as an explicit zero operand.
#include <hip/hip_runtime.h>

#include "ctx.hpp"
#include "step3d.hpp"
#include "step3g.hpp"

namespace clawgpu {
namespace s3 {
namespace {

using EM = ExactMath;

// doubles per slice position: qadd[4], dF[4], G[side 2][slice 3][4], H[2][3][4]
constexpr int kRec = 56;
constexpr int kG = 8, kH = 32;

// Geometry of a sweep direction: the slice's length n, the extents n1, n2 of
// the slice indices (p1, p2) -- step3's (inner, outer) loop indices: (j, k)
// for x, (i, k) for y, (i, j) for z -- the normal and transverse components
// and the dt/d ratios along the normal, y-like and z-like directions (step3
// passes flux3 its arguments rotated: y-like / z-like are y / z for x, z / x
// for y, x / y for z).
struct Geo {
    int n, n1, n2;
    int mu, mv, mw;  // 0-based components: normal, y-like, z-like
    double dn, d1, d2;
};

__device__ inline Geo geo(const Args3g& A)
{
    const Layout3& L = A.L;
    if (A.dir == 0) return {L.mx, L.my, L.mz, 1, 2, 3, A.dt3[0], A.dt3[1], A.dt3[2]};
    if (A.dir == 1) return {L.my, L.mx, L.mz, 2, 3, 1, A.dt3[1], A.dt3[2], A.dt3[0]};
    return {L.mz, L.mx, L.my, 3, 1, 2, A.dt3[2], A.dt3[0], A.dt3[1]};
}

// Cell at position l of slice (p1, p2), offset a in the y-like and b in the
// z-like direction (x: (l, p1+a, p2+b); y: (p1+b, l, p2+a); z: (p1+a, p2+b, l)).
__device__ inline long cell_at(const Args3g& A, int l, int p1, int p2, int a, int b)
{
    if (A.dir == 0) return A.L.off(l, p1 + a, p2 + b);
    if (A.dir == 1) return A.L.off(p1 + b, l, p2 + a);
    return A.L.off(p1 + a, p2 + b, l);
}

// (p, t) parts of a transverse split result: pressure and the transverse
// velocity of the split's direction
struct Pt {
    double p, t;
};

// The transverse split of a flux difference with pressure ap and a zero
// transverse velocity in the cell with impedance zz, between the cells lo
// (below) and hi (above): flux3's rpt3/rptt3 for this solver
// (down-going at -c_lo, up-going at +c_hi).
__device__ inline void vsplit(double ap, const Mat& lo, double zz, const Mat& hi, Pt& bm, Pt& bp)
{
    const double at = 0.0;
    const double a1 = (-ap + at * zz) / (lo.Z + zz);
    const double a2 = (ap + at * zz) / (zz + hi.Z);
    bm = {lo.c * a1 * lo.Z, -lo.c * a1};
    bp = {hi.c * a2 * hi.Z, hi.c * a2};
}

__global__ __launch_bounds__(256) void rec3_kernel(Args3g A)
{
    const Geo g = geo(A);
    const long np = (long)g.n * (g.n1 + 2) * (g.n2 + 2);
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= np) return;
    const int l = (int)(e % g.n) + 1;
    const long r = e / g.n;
    const int p1 = (int)(r % (g.n1 + 2)), p2 = (int)(r / (g.n1 + 2));
    const long fs = A.L.fs, afs = A.L.fs;
    double* out = A.rec + e * kRec;

    // slice cells l-2 .. l+2 and the dt/dx ratios of l-1 .. l+1
    double pq[5], nq[5], DT[3];
    Mat mt[5];
    for (int o = 0; o < 5; o++) {
        const long c = cell_at(A, l - 2 + o, p1, p2, 0, 0);
        pq[o] = A.q[c];
        nq[o] = A.q[g.mu * fs + c];
        mt[o] = {A.aux[c], A.aux[afs + c], 0};
        if (o >= 1 && o <= 3) DT[o - 1] = A.capa ? g.dn / A.capa[c] : g.dn;
    }
    // interfaces l-1, l, l+1, l+2 (between cells k-1 and k)
    Iface f[4];
    for (int k = 0; k < 4; k++) f[k] = solve<EM>(pq[k], nq[k], mt[k], pq[k + 1], nq[k + 1], mt[k + 1]);
    Wv cql = {0.0, 0.0}, cqr = {0.0, 0.0};  // cqxx of interfaces l and l+1
    if (A.m2 == 2) {
        const Wv a1 = limit<EM>(f[1].w1, f[0].w1, f[2].w1, f[1].s1, A.lim0);
        const Wv a2 = limit<EM>(f[1].w2, f[0].w2, f[2].w2, f[1].s2, A.lim1);
        cql = cqxx(a1, f[1].s1, a2, f[1].s2, 0.5 * (DT[0] + DT[1]), A.fwave);
        const Wv b1 = limit<EM>(f[2].w1, f[1].w1, f[3].w1, f[2].s1, A.lim0);
        const Wv b2 = limit<EM>(f[2].w2, f[1].w2, f[3].w2, f[2].s2, A.lim1);
        cqr = cqxx(b1, f[2].s1, b2, f[2].s2, 0.5 * (DT[1] + DT[2]), A.fwave);
    }
    for (int k = 0; k < kRec; k++) out[k] = 0.0;
    // qadd(l) = (0 - DT(l) apdq(l)) - DT(l) amdq(l+1); fadd(l+1) - fadd(l)
    out[0] = (0.0 - DT[1] * f[1].ap.p) - DT[1] * f[2].am.p;
    out[g.mu] = (0.0 - DT[1] * f[1].ap.n) - DT[1] * f[2].am.n;
    out[4] = (0.0 + cqr.p) - (0.0 + cql.p);
    out[4 + g.mu] = (0.0 + cqr.n) - (0.0 + cql.n);
    if (A.m3 <= 0) return;

    // materials of the 3x3 transverse neighbourhood of position l: N[a+1][b+1]
    Mat N[3][3];
    for (int a = -1; a <= 1; a++)
        for (int b = -1; b <= 1; b++) {
            const long c = cell_at(A, l, p1, p2, a, b);
            N[a + 1][b + 1] = {A.aux[c], A.aux[afs + c], 0};
        }
    const double zc = N[1][1].Z;
    // y-like split (rpt3 icoor 2) and z-like split (icoor 3) at position l
    auto ysplit = [&](double ap, Pt& bm, Pt& bp) { vsplit(ap, N[0][1], zc, N[2][1], bm, bp); };
    auto zsplit = [&](double ap, Pt& bm, Pt& bp) { vsplit(ap, N[1][0], zc, N[1][2], bm, bp); };
    // double-transverse splits (rptt3): y-like in the plane b, z-like in the row a
    auto yplane = [&](double ap, int b, Pt& bm, Pt& bp) {
        vsplit(ap, N[0][b + 1], N[1][b + 1].Z, N[2][b + 1], bm, bp);
    };
    auto zrow = [&](double ap, int a, Pt& bm, Pt& bp) {
        vsplit(ap, N[a + 1][0], N[a + 1][1].Z, N[a + 1][2], bm, bp);
    };

    // A^+ side of interface l ("p") and A^- side of interface l+1 ("m")
    Pt bmapdq, bpapdq, cmapdq, cpapdq, bmamdq, bpamdq, cmamdq, cpamdq;
    ysplit(f[1].ap.p, bmapdq, bpapdq);
    ysplit(f[2].am.p, bmamdq, bpamdq);
    zsplit(f[1].ap.p, cmapdq, cpapdq);
    zsplit(f[2].am.p, cmamdq, cpamdq);
    Pt bmcqxxp{0, 0}, bpcqxxp{0, 0}, cmcqxxp{0, 0}, cpcqxxp{0, 0};
    Pt bmcqxxm{0, 0}, bpcqxxm{0, 0}, cmcqxxm{0, 0}, cpcqxxm{0, 0};
    if (A.m3 == 2) {
        ysplit(cqr.p, bmcqxxm, bpcqxxm);
        ysplit(cql.p, bmcqxxp, bpcqxxp);
        zsplit(cqr.p, cmcqxxm, cpcqxxm);
        zsplit(cql.p, cmcqxxp, cpcqxxp);
    }
    Pt cpapdq2{0, 0}, cpamdq2{0, 0}, cmapdq2{0, 0}, cmamdq2{0, 0};
    if (A.m4 == 1) {
        cpapdq2 = cpapdq;
        cpamdq2 = cpamdq;
        cmapdq2 = cmapdq;
        cmamdq2 = cmamdq;
    } else if (A.m4 == 2) {
        cpapdq2 = {cpapdq.p - 3.0 * cpcqxxp.p, cpapdq.t - 3.0 * cpcqxxp.t};
        cpamdq2 = {cpamdq.p + 3.0 * cpcqxxm.p, cpamdq.t + 3.0 * cpcqxxm.t};
        cmapdq2 = {cmapdq.p - 3.0 * cmcqxxp.p, cmapdq.t - 3.0 * cmcqxxp.t};
        cmamdq2 = {cmamdq.p + 3.0 * cmcqxxm.p, cmamdq.t + 3.0 * cmcqxxm.t};
    }
    Pt bmcpapdq{0, 0}, bpcpapdq{0, 0}, bmcpamdq{0, 0}, bpcpamdq{0, 0};
    Pt bmcmapdq{0, 0}, bpcmapdq{0, 0}, bmcmamdq{0, 0}, bpcmamdq{0, 0};
    if (A.m4 > 0) {
        yplane(cpapdq2.p, 1, bmcpapdq, bpcpapdq);
        yplane(cpamdq2.p, 1, bmcpamdq, bpcpamdq);
        yplane(cmapdq2.p, -1, bmcmapdq, bpcmapdq);
        yplane(cmamdq2.p, -1, bmcmamdq, bpcmamdq);
    }
    const double sixth = 1.0 / 6.0;
    const double dl = DT[1];
    // records entry (side 1..2, slice -1..1, component m)
    auto G = [&](int side, int sl, int m) -> double& { return out[kG + ((side - 1) * 3 + sl + 1) * 4 + m]; };
    auto H = [&](int side, int sl, int m) -> double& { return out[kH + ((side - 1) * 3 + sl + 1) * 4 + m]; };

    // G (flux3.f90:378-467): in the y-like split the transverse component is mv
    for (int pass = 0; pass < 2; pass++) {
        const int comp = pass == 0 ? 0 : g.mv;
        auto pick = [&](const Pt& v) { return pass == 0 ? v.p : v.t; };
        // interface l, A^+ side
        G(1, 0, comp) = G(1, 0, comp) - 0.5 * dl * pick(bmapdq);
        G(2, 0, comp) = G(2, 0, comp) - 0.5 * dl * pick(bpapdq);
        if (A.m4 > 0) {
            const double cp_ = sixth * dl * g.d2;
            G(2, 0, comp) = G(2, 0, comp) + cp_ * (pick(bpcpapdq) - pick(bpcmapdq));
            G(1, 0, comp) = G(1, 0, comp) + cp_ * (pick(bmcpapdq) - pick(bmcmapdq));
            G(2, 1, comp) = G(2, 1, comp) - cp_ * pick(bpcpapdq);
            G(1, 1, comp) = G(1, 1, comp) - cp_ * pick(bmcpapdq);
            G(2, -1, comp) = G(2, -1, comp) + cp_ * pick(bpcmapdq);
            G(1, -1, comp) = G(1, -1, comp) + cp_ * pick(bmcmapdq);
        }
        if (A.m3 >= 2) {
            G(2, 0, comp) = G(2, 0, comp) + dl * pick(bpcqxxp);
            G(1, 0, comp) = G(1, 0, comp) + dl * pick(bmcqxxp);
        }
        // interface l+1, A^- side
        G(1, 0, comp) = G(1, 0, comp) - 0.5 * dl * pick(bmamdq);
        G(2, 0, comp) = G(2, 0, comp) - 0.5 * dl * pick(bpamdq);
        if (A.m4 > 0) {
            const double cm_ = sixth * dl * g.d2;
            G(2, 0, comp) = G(2, 0, comp) + cm_ * (pick(bpcpamdq) - pick(bpcmamdq));
            G(1, 0, comp) = G(1, 0, comp) + cm_ * (pick(bmcpamdq) - pick(bmcmamdq));
            G(2, 1, comp) = G(2, 1, comp) - cm_ * pick(bpcpamdq);
            G(1, 1, comp) = G(1, 1, comp) - cm_ * pick(bmcpamdq);
            G(2, -1, comp) = G(2, -1, comp) + cm_ * pick(bpcmamdq);
            G(1, -1, comp) = G(1, -1, comp) + cm_ * pick(bmcmamdq);
        }
        if (A.m3 >= 2) {
            G(2, 0, comp) = G(2, 0, comp) - dl * pick(bpcqxxm);
            G(1, 0, comp) = G(1, 0, comp) - dl * pick(bmcqxxm);
        }
    }

    // y-like parts with the correction wave (m4 = 2), then their z-like
    // splits in the rows a = +-1 (reusing the double-transverse storage)
    if (A.m4 == 2) {
        bpapdq = {bpapdq.p - 3.0 * bpcqxxp.p, bpapdq.t - 3.0 * bpcqxxp.t};
        bpamdq = {bpamdq.p + 3.0 * bpcqxxm.p, bpamdq.t + 3.0 * bpcqxxm.t};
        bmapdq = {bmapdq.p - 3.0 * bmcqxxp.p, bmapdq.t - 3.0 * bmcqxxp.t};
        bmamdq = {bmamdq.p + 3.0 * bmcqxxm.p, bmamdq.t + 3.0 * bmcqxxm.t};
    }
    if (A.m4 > 0) {
        zrow(bpapdq.p, 1, bmcpapdq, bpcpapdq);
        zrow(bpamdq.p, 1, bmcpamdq, bpcpamdq);
        zrow(bmapdq.p, -1, bmcmapdq, bpcmapdq);
        zrow(bmamdq.p, -1, bmcmamdq, bpcmamdq);
    }

    // H (flux3.f90:489-608): in the z-like split the transverse component is mw
    for (int pass = 0; pass < 2; pass++) {
        const int comp = pass == 0 ? 0 : g.mw;
        auto pick = [&](const Pt& v) { return pass == 0 ? v.p : v.t; };
        H(1, 0, comp) = H(1, 0, comp) - 0.5 * dl * pick(cmapdq);
        H(2, 0, comp) = H(2, 0, comp) - 0.5 * dl * pick(cpapdq);
        if (A.m4 > 0) {
            const double cp_ = sixth * dl * g.d1;
            H(2, 0, comp) = H(2, 0, comp) + cp_ * (pick(bpcpapdq) - pick(bpcmapdq));
            H(1, 0, comp) = H(1, 0, comp) + cp_ * (pick(bmcpapdq) - pick(bmcmapdq));
            H(2, 1, comp) = H(2, 1, comp) - cp_ * pick(bpcpapdq);
            H(1, 1, comp) = H(1, 1, comp) - cp_ * pick(bmcpapdq);
            H(2, -1, comp) = H(2, -1, comp) + cp_ * pick(bpcmapdq);
            H(1, -1, comp) = H(1, -1, comp) + cp_ * pick(bmcmapdq);
        }
        if (A.m3 >= 2) {
            H(2, 0, comp) = H(2, 0, comp) + dl * pick(cpcqxxp);
            H(1, 0, comp) = H(1, 0, comp) + dl * pick(cmcqxxp);
        }
        H(1, 0, comp) = H(1, 0, comp) - 0.5 * dl * pick(cmamdq);
        H(2, 0, comp) = H(2, 0, comp) - 0.5 * dl * pick(cpamdq);
        if (A.m4 > 0) {
            const double cm_ = sixth * dl * g.d1;
            H(2, 0, comp) = H(2, 0, comp) + cm_ * (pick(bpcpamdq) - pick(bpcmamdq));
            H(1, 0, comp) = H(1, 0, comp) + cm_ * (pick(bmcpamdq) - pick(bmcmamdq));
            H(2, 1, comp) = H(2, 1, comp) - cm_ * pick(bpcpamdq);
            H(1, 1, comp) = H(1, 1, comp) - cm_ * pick(bmcpamdq);
            H(2, -1, comp) = H(2, -1, comp) + cm_ * pick(bpcmamdq);
            H(1, -1, comp) = H(1, -1, comp) + cm_ * pick(bmcmamdq);
        }
        if (A.m3 >= 2) {
            H(2, 0, comp) = H(2, 0, comp) - dl * pick(cpcqxxm);
            H(1, 0, comp) = H(1, 0, comp) - dl * pick(cmcqxxm);
        }
    }
}

// One interior cell: the nine slices of this direction that touch it, in
// the order of step3's passes, each applied with step3's expression for the
// cell's offset (a, b) from the slice.  qin may equal qn (y and z sweeps).
__global__ __launch_bounds__(256) void comb3_kernel(Args3g A, const double* qin)
{
    const Layout3& L = A.L;
    const long ncell = (long)L.mx * L.my * L.mz;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ncell) return;
    const int i = (int)(e % L.mx) + 1, j = (int)((e / L.mx) % L.my) + 1,
              k = (int)(e / ((long)L.mx * L.my)) + 1;
    const Geo g = geo(A);
    const long c = L.off(i, j, k);
    // the slice (p1, p2) that holds this cell at offset (a, b) (cell_at
    // inverted) and the position l along it
    const int l = A.dir == 0 ? i : A.dir == 1 ? j : k;
    auto slice_of = [&](int a, int b, int& p1, int& p2) {
        if (A.dir == 0) { p1 = j - a; p2 = k - b; }
        else if (A.dir == 1) { p1 = i - b; p2 = k - a; }
        else { p1 = i - a; p2 = j - b; }
    };
    // pass keys: step3's order over the slices
    int key[9], ord[9];
    for (int s = 0; s < 9; s++) {
        const int a = s % 3 - 1, b = s / 3 - 1;
        int p1, p2;
        slice_of(a, b, p1, p2);
        const int pass = ((p2 >> 1) & 1) * 2 + ((p1 >> 1) & 1);
        key[s] = (((pass * 4096 + (p2 >> 1)) * 4096 + (p1 >> 1)) * 2 + (p2 & 1)) * 2 + (p1 & 1);
        ord[s] = s;
    }
    for (int s = 1; s < 9; s++)
        for (int u = s; u > 0 && key[ord[u]] < key[ord[u - 1]]; u--) {
            const int tmp = ord[u];
            ord[u] = ord[u - 1];
            ord[u - 1] = tmp;
        }
    const double kappa = A.capa ? A.capa[c] : 1.0;
    const bool capa = A.capa != nullptr;
    const double dn = g.dn, d1 = g.d1, d2 = g.d2;
    double q[4];
    for (int m = 0; m < 4; m++) q[m] = qin[m * L.fs + c];
    for (int s = 0; s < 9; s++) {
        const int a = ord[s] % 3 - 1, b = ord[s] / 3 - 1;
        int p1, p2;
        slice_of(a, b, p1, p2);
        const double* R = A.rec + (((long)p2 * (g.n1 + 2) + p1) * g.n + (l - 1)) * kRec;
        auto G = [&](int side, int sl, int m) { return R[kG + ((side - 1) * 3 + sl + 1) * 4 + m]; };
        auto H = [&](int side, int sl, int m) { return R[kH + ((side - 1) * 3 + sl + 1) * 4 + m]; };
        for (int m = 0; m < 4; m++) {
            double v = q[m];
            if (!capa) {
                if (b == -1) {
                    if (a == -1) v = v - d1 * G(1, -1, m) - d2 * H(1, -1, m);
                    else if (a == 0) v = v - d1 * (G(2, -1, m) - G(1, -1, m)) - d2 * H(1, 0, m);
                    else v = v + d1 * G(2, -1, m) - d2 * H(1, 1, m);
                } else if (b == 0) {
                    if (a == -1) v = v - d1 * G(1, 0, m) - d2 * (H(2, -1, m) - H(1, -1, m));
                    else if (a == 0)
                        v = v + R[m] - dn * R[4 + m] - d1 * (G(2, 0, m) - G(1, 0, m)) -
                            d2 * (H(2, 0, m) - H(1, 0, m));
                    else v = v + d1 * G(2, 0, m) - d2 * (H(2, 1, m) - H(1, 1, m));
                } else {
                    if (a == -1) v = v - d1 * G(1, 1, m) + d2 * H(2, -1, m);
                    else if (a == 0) v = v - d1 * (G(2, 1, m) - G(1, 1, m)) + d2 * H(2, 0, m);
                    else v = v + d1 * G(2, 1, m) + d2 * H(2, 1, m);
                }
            } else {
                if (b == -1) {
                    if (a == -1) v = v - (d1 * G(1, -1, m) + d2 * H(1, -1, m)) / kappa;
                    else if (a == 0)
                        v = v - (d1 * (G(2, -1, m) - G(1, -1, m)) + d2 * H(1, 0, m)) / kappa;
                    else v = v + (d1 * G(2, -1, m) - d2 * H(1, 1, m)) / kappa;
                } else if (b == 0) {
                    if (a == -1)
                        v = v - (d1 * G(1, 0, m) + d2 * (H(2, -1, m) - H(1, -1, m))) / kappa;
                    else if (a == 0)
                        v = v + R[m] -
                            (dn * R[4 + m] + d1 * (G(2, 0, m) - G(1, 0, m)) +
                             d2 * (H(2, 0, m) - H(1, 0, m))) / kappa;
                    else v = v + (d1 * G(2, 0, m) - d2 * (H(2, 1, m) - H(1, 1, m))) / kappa;
                } else {
                    if (a == -1) v = v - (d1 * G(1, 1, m) - d2 * H(2, -1, m)) / kappa;
                    else if (a == 0)
                        v = v - (d1 * (G(2, 1, m) - G(1, 1, m)) - d2 * H(2, 0, m)) / kappa;
                    else v = v + (d1 * G(2, 1, m) + d2 * H(2, 1, m)) / kappa;
                }
            }
            q[m] = v;
        }
    }
    for (int m = 0; m < 4; m++) A.qn[m * L.fs + c] = q[m];
}

__global__ void copy3_kernel(const double* __restrict__ s, double* __restrict__ d, long n)
{
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
        d[e] = s[e];
}

// max over the cells with every index in [0, n+1] of (dt/d / kappa) * c for
// each direction d: flux3's Courant terms dtdx1d(i) s(mw, i) and
// -dtdx1d(i-1) s(mw, i) with s = -c(i-1), c(i) (flux3.f90:217-224); raw bits
// in out[0..2] (values >= 0), a NaN flag in out[3]
__global__ void capa_cfl_kernel(const double* __restrict__ cc, const double* __restrict__ kap,
                                Layout3 L, double d0, double d1, double d2, unsigned long long* out)
{
    const long n = (long)(L.mx + 2) * (L.my + 2) * (L.mz + 2);
    double m[3] = {0.0, 0.0, 0.0};
    bool nan = false;
    const double d[3] = {d0, d1, d2};
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const int i = (int)(e % (L.mx + 2)), j = (int)((e / (L.mx + 2)) % (L.my + 2)),
                  k = (int)(e / ((long)(L.mx + 2) * (L.my + 2)));
        const long c = L.off(i, j, k);
        for (int r = 0; r < 3; r++) {
            const double v = (d[r] / kap[c]) * cc[c];
            if (v != v) nan = true;
            else m[r] = fmax(m[r], v);
        }
    }
    for (int r = 0; r < 3; r++)
        for (int s = 32; s >= 1; s >>= 1) m[r] = fmax(m[r], __shfl_xor(m[r], s));
    for (int s = 32; s >= 1; s >>= 1) nan = nan || __shfl_xor((int)nan, s);
    if ((threadIdx.x & 63) == 0) {
        for (int r = 0; r < 3; r++) atomicMax(&out[r], (unsigned long long)__double_as_longlong(m[r]));
        if (nan) atomicOr(&out[3], 1ull);
    }
}

}  // namespace

long rec3_doubles(const Layout3& L)
{
    const long nx = (long)L.mx * (L.my + 2) * (L.mz + 2);
    const long ny = (long)L.my * (L.mz + 2) * (L.mx + 2);
    const long nz = (long)L.mz * (L.mx + 2) * (L.my + 2);
    long m = nx > ny ? nx : ny;
    m = m > nz ? m : nz;
    return m * kRec;
}

int launch_step3_records(Args3g A, hipStream_t s)
{
    const long nq = 4 * A.L.fs;
    hipLaunchKernelGGL(copy3_kernel, dim3(1024), dim3(256), 0, s, A.q, A.qn, nq);
    const long ncell = (long)A.L.mx * A.L.my * A.L.mz;
    for (int dir = 0; dir < 3; dir++) {
        A.dir = dir;
        const Layout3& L = A.L;
        const long np = dir == 0 ? (long)L.mx * (L.my + 2) * (L.mz + 2)
                      : dir == 1 ? (long)L.my * (L.mz + 2) * (L.mx + 2)
                                 : (long)L.mz * (L.mx + 2) * (L.my + 2);
        hipLaunchKernelGGL(rec3_kernel, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, A);
        hipLaunchKernelGGL(comb3_kernel, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, s, A,
                           dir == 0 ? A.q : (const double*)A.qn);
    }
    return (int)hipGetLastError();
}

int launch_capa_cfl(const double* c, const double* kappa, const Layout3& L, const double* d3,
                    unsigned long long* bits, hipStream_t s)
{
    hipLaunchKernelGGL(capa_cfl_kernel, dim3(1024), dim3(256), 0, s, c, kappa, L, d3[0], d3[1],
                       d3[2], bits);
    return (int)hipGetLastError();
}

}  // namespace s3
}  // namespace clawgpu
