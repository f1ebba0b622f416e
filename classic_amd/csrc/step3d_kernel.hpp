// 3D wave-propagation step for gfx950 (FP64): variable-coefficient acoustics.
//
// Replaces the reference's step3 -> flux3 -> rpn3/limiter/rpt3/rptt3 path
// (src/3d/step3.f90:3-873, src/3d/flux3.f90:2-612) and the split
// dimsp3 -> step3ds path (src/3d/dimsp3.f:5-85, src/3d/step3ds.f90:3-408) for
// q = (p, u, v, w), aux = (Z, c) (impedance, sound speed).
//
// Solver structure used throughout.  The normal solve at an interface only
// involves p and the normal velocity: both waves, amdq and apdq have exactly
// those two non-zero components.  A transverse split of a fluctuation with
// zero transverse velocity (every input the reference passes to rpt3/rptt3
// for this solver has that form) is linear in its pressure alone:
//     B^- a = a * (-c_m Z_m/(Z_m+Z), c_m/(Z_m+Z)),
//     B^+ a = a * ( c_p Z_p/(Z+Z_p), c_p/(Z+Z_p))      in (p, v_t),
// with (Z_m, c_m), (Z, c), (Z_p, c_p) the cells below, at and above in the
// transverse direction.  So every split costs four multiplications by
// coefficients that depend on the materials only (struct Kc), and every G/H
// flux increment has two non-zero components (p and one velocity).
//
// Unsplit step ("tile kernel").  One 64-lane wavefront owns a tile of
// 60 x TY x TZ cells (lanes 2..61 along x; lanes 0, 1, 62, 63 are halo) and
// computes, for that tile, every slice position whose flux increments reach
// it: x-slices (j', k') for j' in [j0-1, j0+TY], k' in [k0-1, k0+TZ] with one
// interface per lane, and y- and z-slices marched along j and k with the
// limiter pipeline in registers.  Each slice position turns its two
// interfaces into qadd, the fadd difference and the gadd/hadd increments of
// flux3 (G/H of 2 sides x 3 neighbouring slices), and adds the resulting
// change of the nine cells it touches (step3.f90:264-335 and the rotated
// forms) to an LDS accumulator of the tile.  qnew = qold + accumulator.  The
// sums are therefore formed per cell in a different order than the
// reference's slice-by-slice updates: results agree to round-off, not bit
// for bit.  Cross-lane data (left cell, neighbouring waves, the A^- side of
// the right interface, x-neighbour materials) moves with DPP shifts.
//
// Split step (sweep kernels).  One direction at a time, out of place, every
// slice including the ghost ones, with the reference's expression order
// (qadd = (0 - d*apdq) - d*amdq, q + qadd - d*(F(i+1) - F(i))) and, in the
// exact translation unit, no contraction: bit for bit with step3ds.
//
// Device layout: field planes, element (m, i, j, k) at
// m*fs + ((k+mbc-1)*ny + (j+mbc-1))*nx + (i+mbc-1), 1-based Fortran indices,
// nx = mx+2mbc, ny = my+2mbc; aux planes likewise (Z, then c).
#pragma once

#include <hip/hip_runtime.h>

#include "lanes.hpp"
#include "riemann2d.hpp"

namespace clawgpu {
namespace s3 {

constexpr int kOut = 60;

struct Layout3 {
    int mx, my, mz, mbc, nx, ny, nz;
    long fs;
    __host__ __device__ long off(int i, int j, int k) const
    {
        return ((long)(k + mbc - 1) * ny + (j + mbc - 1)) * nx + (i + mbc - 1);
    }
    // clamped to the allocated range [1-mbc, n+mbc]
    __device__ long offc(int i, int j, int k) const
    {
        i = i < 1 - mbc ? 1 - mbc : (i > mx + mbc ? mx + mbc : i);
        j = j < 1 - mbc ? 1 - mbc : (j > my + mbc ? my + mbc : j);
        k = k < 1 - mbc ? 1 - mbc : (k > mz + mbc ? mz + mbc : k);
        return off(i, j, k);
    }
};

// Material table: when the media holds at most kMaxMat distinct (Z, c)
// pairs, mat[c] indexes lut (Z, c interleaved) and the unsplit kernel reads
// one byte per cell instead of two doubles; the values are the same doubles.
constexpr int kMaxMat = 64;
// up to this many materials the transverse coefficients of every material
// triple are tabulated too (kMaxKcMat^3 entries of 32 bytes in LDS)
constexpr int kMaxKcMat = 4;

struct Args3 {
    const double* __restrict__ q;    // 4 planes
    double* __restrict__ qn;         // 4 planes
    const double* __restrict__ aux;  // planes Z, c
    const unsigned char* __restrict__ mat;  // material index per cell (nmat > 0)
    const double* __restrict__ lut;         // 2 * nmat doubles
    int nmat;                               // 0: read aux
    Layout3 L;
    long afs;                        // aux plane stride
    double dtdx, dtdy, dtdz;
    int m2, m3, m4, fwave, lim0, lim1;
    int ntx, nty, ntz;
    int xcd;  // bit 0: each of the 8 XCDs gets a contiguous range of tiles;
              // bit 1: tile index with z fastest (else x fastest)
};

struct Mat {
    double Z, c;
    int id;  // material index (material-table mode; 0 otherwise)
};
struct Wv {
    double p, n;  // pressure and normal velocity components
};
struct Iface {
    Wv w1, w2, am, ap;
    double s1, s2;
};
struct Kc {
    double mp, mt, pp, pt;
};
struct Rec {
    double gp[2][3], gt[2][3], hp[2][3], ht[2][3];  // [side][slice -1, 0, +1]
};

// Method flags and limiters: read from the arguments (Cfg3Any) or fixed at
// compile time (Cfg3Fix) so that the common configurations lose their
// uniform branches and the slice code becomes one basic block.
struct Cfg3Any {
    __device__ static int m2(const struct Args3& a);
    __device__ static int m3(const struct Args3& a);
    __device__ static int m4(const struct Args3& a);
    __device__ static int lim0(const struct Args3& a);
    __device__ static int lim1(const struct Args3& a);
};
template <int M2, int M3, int M4, int L0, int L1>
struct Cfg3Fix {
    __device__ static constexpr int m2(const struct Args3&) { return M2; }
    __device__ static constexpr int m3(const struct Args3&) { return M3; }
    __device__ static constexpr int m4(const struct Args3&) { return M4; }
    __device__ static constexpr int lim0(const struct Args3&) { return L0; }
    __device__ static constexpr int lim1(const struct Args3&) { return L1; }
};
__device__ inline int Cfg3Any::m2(const Args3& a) { return a.m2; }
__device__ inline int Cfg3Any::m3(const Args3& a) { return a.m3; }
__device__ inline int Cfg3Any::m4(const Args3& a) { return a.m4; }
__device__ inline int Cfg3Any::lim0(const Args3& a) { return a.lim0; }
__device__ inline int Cfg3Any::lim1(const Args3& a) { return a.lim1; }

__device__ inline Mat mat_up(Mat m) { return {lane_up(m.Z), lane_up(m.c), lane_up_i(m.id)}; }
__device__ inline Mat mat_dn(Mat m) { return {lane_dn(m.Z), lane_dn(m.c), lane_dn_i(m.id)}; }
__device__ inline Wv wv_up(Wv w) { return {lane_up(w.p), lane_up(w.n)}; }
__device__ inline Wv wv_dn(Wv w) { return {lane_dn(w.p), lane_dn(w.n)}; }

__device__ inline Mat load_mat(const Args3& a, long c) { return {a.aux[c], a.aux[a.afs + c], 0}; }

// Interface between cell L (i-1) and R (i): the jump in (p, u_n) splits into
// the left-going wave (-Z_L, 1) a1 at speed -c_L and the right-going wave
// (Z_R, 1) a2 at speed c_R.
template <class M>
__device__ inline Iface solve(double pL, double nL, Mat mL, double pR, double nR, Mat mR)
{
    double d1 = pR - pL, d2 = nR - nL;
    double den = mL.Z + mR.Z;
    double a1 = M::div(-d1 + mR.Z * d2, den);
    double a2 = M::div(d1 + mL.Z * d2, den);
    Iface f;
    f.w1 = {-a1 * mL.Z, a1};
    f.s1 = -mL.c;
    f.w2 = {a2 * mR.Z, a2};
    f.s2 = mR.c;
    f.am = {f.s1 * f.w1.p, f.s1 * f.w1.n};
    f.ap = {f.s2 * f.w2.p, f.s2 * f.w2.n};
    return f;
}

// Limiter of one wave from the raw waves of the neighbouring interfaces of
// the same family (inlinelimiter.f90:51-100; the squares and dot products
// over the four components reduce to p and u_n).
template <class M>
__device__ inline Wv limit(Wv w, Wv wl, Wv wr, double s, int lim)
{
    if (lim == 0) return w;
    double wn2 = w.p * w.p + w.n * w.n;
    if (wn2 == 0.0) return w;
    double dot = s > 0.0 ? wl.p * w.p + wl.n * w.n : w.p * wr.p + w.n * wr.n;
    double r = M::div(dot, wn2);
    double phi = philim<M>(r, lim);
    return {phi * w.p, phi * w.n};
}

// cqxx of flux3.f90:233-262 (the factor 0.5 is part of it in 3D)
__device__ inline Wv cqxx(Wv w1, double s1, Wv w2, double s2, double d, int fwave)
{
    double a1 = fabs(s1), a2 = fabs(s2);
    double l1 = fwave ? 0.5 * copysign(1.0, s1) : 0.5 * a1;
    double l2 = fwave ? 0.5 * copysign(1.0, s2) : 0.5 * a2;
    double f1 = l1 * (1.0 - a1 * d), f2 = l2 * (1.0 - a2 * d);
    return {f1 * w1.p + f2 * w2.p, f1 * w1.n + f2 * w2.n};
}

template <class M>
__device__ inline Kc kcoef(Mat lo, Mat mid, Mat hi)
{
    double rm = M::rcp(lo.Z + mid.Z), rp = M::rcp(mid.Z + hi.Z);
    return {-(lo.c * lo.Z) * rm, lo.c * rm, (hi.c * hi.Z) * rp, hi.c * rp};
}

// The flux3 G and H increments (flux3.f90:378-467 and :489-608) that one
// side of a slice position contributes: sgn = +1 for the A^+ fluctuation of
// its low interface, -1 for the A^- fluctuation of its high interface.  a and
// cq are the pressure components of that fluctuation and of cqxx.
__device__ inline void rec_side(Rec& R, double a, double cq, double sgn, const Kc& Y0,
                                const Kc& Z0, const Kc& Ym, const Kc& Yp, const Kc& Zm,
                                const Kc& Zp, double d, double d1, double d2, int m3, int m4)
{
    const double hd = 0.5 * d;
    double bmp = a * Y0.mp, bmt = a * Y0.mt, bpp = a * Y0.pp, bpt = a * Y0.pt;
    double cmp = a * Z0.mp, cmt = a * Z0.mt, cpp = a * Z0.pp, cpt = a * Z0.pt;
    R.gp[0][1] -= hd * bmp; R.gt[0][1] -= hd * bmt;
    R.gp[1][1] -= hd * bpp; R.gt[1][1] -= hd * bpt;
    R.hp[0][1] -= hd * cmp; R.ht[0][1] -= hd * cmt;
    R.hp[1][1] -= hd * cpp; R.ht[1][1] -= hd * cpt;
    double qbm = 0.0, qbp = 0.0, qcm = 0.0, qcp = 0.0;  // splits of cqxx (pressure parts)
    if (m3 == 2) {
        qbm = cq * Y0.mp; qbp = cq * Y0.pp; qcm = cq * Z0.mp; qcp = cq * Z0.pp;
    }
    if (m4 > 0) {
        const double t3 = m4 == 2 ? 3.0 * sgn : 0.0;
        // C^+/C^- parts (with the split correction wave when m4 = 2) moved to
        // the t2-neighbour plane, split there in t1
        double cp2 = cpp - t3 * qcp, cm2 = cmp - t3 * qcm;
        double c = (1.0 / 6.0) * d * d2;
        double bmcp_p = cp2 * Yp.mp, bmcp_t = cp2 * Yp.mt, bpcp_p = cp2 * Yp.pp, bpcp_t = cp2 * Yp.pt;
        double bmcm_p = cm2 * Ym.mp, bmcm_t = cm2 * Ym.mt, bpcm_p = cm2 * Ym.pp, bpcm_t = cm2 * Ym.pt;
        R.gp[1][1] += c * (bpcp_p - bpcm_p); R.gt[1][1] += c * (bpcp_t - bpcm_t);
        R.gp[0][1] += c * (bmcp_p - bmcm_p); R.gt[0][1] += c * (bmcp_t - bmcm_t);
        R.gp[1][2] -= c * bpcp_p; R.gt[1][2] -= c * bpcp_t;
        R.gp[0][2] -= c * bmcp_p; R.gt[0][2] -= c * bmcp_t;
        R.gp[1][0] += c * bpcm_p; R.gt[1][0] += c * bpcm_t;
        R.gp[0][0] += c * bmcm_p; R.gt[0][0] += c * bmcm_t;
        // B^+/B^- parts moved to the t1-neighbour row, split there in t2
        double bp2 = bpp - t3 * qbp, bm2 = bmp - t3 * qbm;
        double e = (1.0 / 6.0) * d * d1;
        double cmbp_p = bp2 * Zp.mp, cmbp_t = bp2 * Zp.mt, cpbp_p = bp2 * Zp.pp, cpbp_t = bp2 * Zp.pt;
        double cmbm_p = bm2 * Zm.mp, cmbm_t = bm2 * Zm.mt, cpbm_p = bm2 * Zm.pp, cpbm_t = bm2 * Zm.pt;
        R.hp[1][1] += e * (cpbp_p - cpbm_p); R.ht[1][1] += e * (cpbp_t - cpbm_t);
        R.hp[0][1] += e * (cmbp_p - cmbm_p); R.ht[0][1] += e * (cmbp_t - cmbm_t);
        R.hp[1][2] -= e * cpbp_p; R.ht[1][2] -= e * cpbp_t;
        R.hp[0][2] -= e * cmbp_p; R.ht[0][2] -= e * cmbp_t;
        R.hp[1][0] += e * cpbm_p; R.ht[1][0] += e * cpbm_t;
        R.hp[0][0] += e * cmbm_p; R.ht[0][0] += e * cmbm_t;
    }
    if (m3 == 2) {
        const double sd = sgn * d;
        R.gp[1][1] += sd * qbp; R.gt[1][1] += sd * (cq * Y0.pt);
        R.gp[0][1] += sd * qbm; R.gt[0][1] += sd * (cq * Y0.mt);
        R.hp[1][1] += sd * qcp; R.ht[1][1] += sd * (cq * Z0.pt);
        R.hp[0][1] += sd * qcm; R.ht[0][1] += sd * (cq * Z0.mt);
    }
}

// N[a+1][b+1] = materials of the cell at offset a along t1 and b along t2.
// With KL the transverse coefficients of a (lo, mid, hi) material triple
// come from the table klut (built by kcoef itself, so the values are the
// same bits) instead of two reciprocals each.
template <class M, bool KL>
__device__ inline void make_record(Rec& R, double ap, double cqp, double am, double cqm,
                                   const Mat (&N)[3][3], double d, double d1, double d2, int m3,
                                   int m4, const Kc* klut, int nm)
{
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
        for (int t = 0; t < 3; t++) R.gp[s][t] = R.gt[s][t] = R.hp[s][t] = R.ht[s][t] = 0.0;
    if (m3 <= 0) return;
    auto kc = [&](const Mat& lo, const Mat& mid, const Mat& hi) -> Kc {
        if constexpr (KL) return klut[(lo.id * nm + mid.id) * nm + hi.id];
        else return kcoef<M>(lo, mid, hi);
    };
    Kc Y0 = kc(N[0][1], N[1][1], N[2][1]);
    Kc Z0 = kc(N[1][0], N[1][1], N[1][2]);
    Kc Ym = {0, 0, 0, 0}, Yp = Ym, Zm = Ym, Zp = Ym;
    if (m4 > 0) {
        Ym = kc(N[0][0], N[1][0], N[2][0]);
        Yp = kc(N[0][2], N[1][2], N[2][2]);
        Zm = kc(N[0][0], N[0][1], N[0][2]);
        Zp = kc(N[2][0], N[2][1], N[2][2]);
    }
    rec_side(R, ap, cqp, 1.0, Y0, Z0, Ym, Yp, Zm, Zp, d, d1, d2, m3, m4);
    rec_side(R, am, cqm, -1.0, Y0, Z0, Ym, Yp, Zm, Zp, d, d1, d2, m3, m4);
}

// The change of the nine cells a slice position touches (step3.f90:264-335
// written as increments).  Target t = 3*(b+1) + (a+1) is the cell at t1
// offset a and t2 offset b; v[t] = (p, u_n, v_t1, v_t2), u_n non-zero only
// for the slice's own cell (t = 4).
struct T9 {
    double p[9], t1[9], t2[9], n;
};

__device__ inline void apply_targets(const Rec& R, double qp, double qn, double dFp, double dFn,
                                     double d, double d1, double d2, bool trans, T9& T)
{
#pragma unroll
    for (int t = 0; t < 9; t++) T.p[t] = T.t1[t] = T.t2[t] = 0.0;
    T.n = qn - d * dFn;
    T.p[4] = qp - d * dFp;
    if (!trans) return;
    // G slice index = t2 offset b, H slice index = t1 offset a; side 0 = 1, 1 = 2
    T.p[0] = -d1 * R.gp[0][0] - d2 * R.hp[0][0];
    T.t1[0] = -d1 * R.gt[0][0];
    T.t2[0] = -d2 * R.ht[0][0];
    T.p[1] = -d1 * (R.gp[1][0] - R.gp[0][0]) - d2 * R.hp[0][1];
    T.t1[1] = -d1 * (R.gt[1][0] - R.gt[0][0]);
    T.t2[1] = -d2 * R.ht[0][1];
    T.p[2] = d1 * R.gp[1][0] - d2 * R.hp[0][2];
    T.t1[2] = d1 * R.gt[1][0];
    T.t2[2] = -d2 * R.ht[0][2];
    T.p[3] = -d1 * R.gp[0][1] - d2 * (R.hp[1][0] - R.hp[0][0]);
    T.t1[3] = -d1 * R.gt[0][1];
    T.t2[3] = -d2 * (R.ht[1][0] - R.ht[0][0]);
    T.p[4] = T.p[4] - d1 * (R.gp[1][1] - R.gp[0][1]) - d2 * (R.hp[1][1] - R.hp[0][1]);
    T.t1[4] = -d1 * (R.gt[1][1] - R.gt[0][1]);
    T.t2[4] = -d2 * (R.ht[1][1] - R.ht[0][1]);
    T.p[5] = d1 * R.gp[1][1] - d2 * (R.hp[1][2] - R.hp[0][2]);
    T.t1[5] = d1 * R.gt[1][1];
    T.t2[5] = -d2 * (R.ht[1][2] - R.ht[0][2]);
    T.p[6] = -d1 * R.gp[0][2] + d2 * R.hp[1][0];
    T.t1[6] = -d1 * R.gt[0][2];
    T.t2[6] = d2 * R.ht[1][0];
    T.p[7] = -d1 * (R.gp[1][2] - R.gp[0][2]) + d2 * R.hp[1][1];
    T.t1[7] = -d1 * (R.gt[1][2] - R.gt[0][2]);
    T.t2[7] = d2 * R.ht[1][1];
    T.p[8] = d1 * R.gp[1][2] + d2 * R.hp[1][2];
    T.t1[8] = d1 * R.gt[1][2];
    T.t2[8] = d2 * R.ht[1][2];
}

// Accumulator layout: word 1 + ((kz*TY + jy)*4 + m)*64 + lane for cell
// (kz, jy) of the tile, component m, lane -1..64.  The slots of lanes 0, 1,
// 62 and 63 belong to halo cells whose result is never stored, so the
// increments a lane owes lane -1 or 64 may land in a neighbouring row's slot
// of lane 63 or 0, and increments for cells outside the tile all go to word
// 0 (+ 64 m): a don't-care slot for every component.  Component m of a cell
// is 64 m words after component 0.
constexpr int kAccRow = 64;

// Add T to the accumulator: idx(t) gives the LDS index of component 0 of
// target t (the don't-care word 0 for targets outside the tile; that test is
// uniform across the wave).  LANE
// names the target offset that moves to another lane (0: none, 1: the t2
// offset b = t/3 - 1, 2: the t1 offset a = t%3 - 1).  Targets are processed
// in rounds of one lane offset: within a round every lane touches distinct
// words (its own cells, or all its neighbours on the same side), so the
// round's reads can all be issued before its writes; across rounds the
// read-modify-writes are ordered, which keeps the increments that two lanes
// add to one cell.  A target whose slot folds to the constant word 0 once
// the sweep loops are unrolled (outside the tile for every wave) is skipped
// altogether, and with it the arithmetic that only feeds it.
__device__ inline bool known_outside(int slot) { return __builtin_constant_p(slot) && slot == 0; }

template <int LANE, class Idx>
__device__ inline void scatter(double* acc, const T9& T, bool trans, int cn, int c1, int c2,
                               Idx&& idx)
{
    if (!trans) {
        const int ip = idx(4), in = ip + kAccRow * cn;
        if (known_outside(ip)) return;
        const double op = acc[ip], on = acc[in];
        acc[ip] = op + T.p[4];
        acc[in] = on + T.n;
        return;
    }
    constexpr int rounds = LANE == 0 ? 1 : 3;
#pragma unroll
    for (int g = 0; g < rounds; g++) {
        int ts[9];
        int nt = 0;
#pragma unroll
        for (int t = 0; t < 9; t++) {
            const int off = LANE == 1 ? t / 3 : t % 3;
            if (LANE == 0 || off == g) ts[nt++] = t;
        }
        int ip[9], i1[9], i2[9];
        double op[9], o1[9], o2[9];
#pragma unroll
        for (int k = 0; k < nt; k++) {
            ip[k] = idx(ts[k]);
            i1[k] = ip[k] + kAccRow * c1;
            i2[k] = ip[k] + kAccRow * c2;
        }
        const bool centre = LANE == 0 || g == 1;
        const int in = centre ? idx(4) + kAccRow * cn : 0;
        const bool cdead = centre && known_outside(idx(4));
#pragma unroll
        for (int k = 0; k < nt; k++) {
            if (known_outside(ip[k])) continue;
            op[k] = acc[ip[k]];
            o1[k] = acc[i1[k]];
            o2[k] = acc[i2[k]];
        }
        const double on = centre && !cdead ? acc[in] : 0.0;
#pragma unroll
        for (int k = 0; k < nt; k++) {
            if (known_outside(ip[k])) continue;
            acc[ip[k]] = op[k] + T.p[ts[k]];
            acc[i1[k]] = o1[k] + T.t1[ts[k]];
            acc[i2[k]] = o2[k] + T.t2[ts[k]];
        }
        if (centre && !cdead) acc[in] = on + T.n;
    }
}

// One slice position from its two limited interfaces: low interface lo (A^+
// side, cqxx cqlo) and high interface hi (A^- side, cqxx cqhi).
template <class M, class Cfg, bool KL>
__device__ inline void slice_position(const Args3& a, const Iface& lo, Wv cqlo, const Iface& hi,
                                      Wv cqhi, const Mat (&N)[3][3], double d, double d1,
                                      double d2, T9& T, const Kc* klut, int nm)
{
    Rec R;
    make_record<M, KL>(R, lo.ap.p, cqlo.p, hi.am.p, cqhi.p, N, d, d1, d2, Cfg::m3(a), Cfg::m4(a),
                       klut, nm);
    double qp = (-d * lo.ap.p) - d * hi.am.p;
    double qn = (-d * lo.ap.n) - d * hi.am.n;
    apply_targets(R, qp, qn, cqhi.p - cqlo.p, cqhi.n - cqlo.n, d, d1, d2, Cfg::m3(a) > 0, T);
}

template <class M, class Cfg>
__device__ inline Wv limited_cq(const Args3& a, const Iface& f, const Iface& fl, const Iface& fr,
                                double d)
{
    if (Cfg::m2(a) != 2) return {0.0, 0.0};
    Wv w1 = limit<M>(f.w1, fl.w1, fr.w1, f.s1, Cfg::lim0(a));
    Wv w2 = limit<M>(f.w2, fl.w2, fr.w2, f.s2, Cfg::lim1(a));
    return cqxx(w1, f.s1, w2, f.s2, d, a.fwave);
}

template <int TY, int TZ, class Cfg, bool MI, bool KL>
__global__ __launch_bounds__(64) void step3_tile_kernel(Args3 a)
{
    using M = FastMath;
    static_assert(MI || !KL, "the coefficient table needs the material table");
    // accumulator of the tile (layout above kAccRow), so that the slice code
    // has no branches
    constexpr int kAccWords = TZ * TY * 4 * kAccRow + 2;
    __shared__ double acc[kAccWords];
    __shared__ double lut[MI ? 2 * kMaxMat : 1];
    __shared__ Kc klut_s[KL ? kMaxKcMat * kMaxKcMat * kMaxKcMat : 1];
    const Kc* klut = klut_s;
    const int l = threadIdx.x;
    if constexpr (MI) {
        for (int e = l; e < 2 * a.nmat; e += 64) lut[e] = a.lut[e];
        if constexpr (KL) {
            // transverse coefficients of every (lo, mid, hi) material triple
            const int n = a.nmat;
            for (int e = l; e < n * n * n; e += 64) {
                const int lo = e / (n * n), mid = (e / n) % n, hi = e % n;
                klut_s[e] = kcoef<M>(Mat{a.lut[2 * lo], a.lut[2 * lo + 1], lo},
                                     Mat{a.lut[2 * mid], a.lut[2 * mid + 1], mid},
                                     Mat{a.lut[2 * hi], a.lut[2 * hi + 1], hi});
            }
        }
        __syncthreads();
    }
    auto load_mat = [&](const Args3& aa, long c) -> Mat {
        if constexpr (MI) {
            const int m = aa.mat[c];
            return {lut[2 * m], lut[2 * m + 1], m};
        } else {
            return {aa.aux[c], aa.aux[aa.afs + c], 0};
        }
    };
    int t = blockIdx.x;
    if (a.xcd & 1) {
        // blocks are dealt to the XCDs round-robin; renumber so that XCD x
        // works on tiles [x*per, (x+1)*per), neighbours sharing its L2
        const int ntiles = a.ntx * a.nty * a.ntz;
        const int per = (ntiles + 7) / 8;
        t = (t & 7) * per + (t >> 3);
        if (t >= ntiles) return;
    }
    int tx, ty, tz;
    if (a.xcd & 2) {
        // z fastest: the tiles that share a tile's halo planes and rows run
        // close together in time
        tz = t % a.ntz;
        t /= a.ntz;
        tx = t % a.ntx;
        ty = t / a.ntx;
    } else {
        tx = t % a.ntx;
        t /= a.ntx;
        ty = t % a.nty;
        tz = t / a.nty;
    }
    const int i = kOut * tx + l - 1, j0 = TY * ty + 1, k0 = TZ * tz + 1;
    const Layout3& L = a.L;
    const long fs = L.fs;
#pragma unroll
    for (int e = 0; e < TZ * TY * 4; e++) acc[1 + e * kAccRow + l] = 0.0;
#define ACC(kz, jy, m, lane) acc[1 + (((kz) * TY + (jy)) * 4 + (m)) * kAccRow + (lane)]
    // component-0 slot of cell (kz, jy) at lane ln in [-1, 64]; ok is uniform
#define SLOTI(ok, kz, jy, ln) ((ok) ? 1 + ((kz) * TY + (jy)) * 4 * kAccRow + (ln) : 0)
    const bool trans = Cfg::m3(a) > 0;

    // ---- x sweeps: slice (j', k'), one interface per lane (its low face)
    {
        const double d = a.dtdx, d1 = a.dtdy, d2 = a.dtdz;
        for (int kk = 0; kk < TZ + 2; kk++) {
            const int kp = k0 - 1 + kk;
            Mat W[3][3];  // rows j'-1..j'+1 (t1) x planes k'-1..k'+1 (t2) of cell i
            Mat Wn[3];    // row j'+2, loaded one slice ahead
            double pn, un;  // p, u of row j'+1, loaded one slice ahead
#pragma unroll
            for (int b = 0; b < 3; b++) {
                W[1][b] = load_mat(a, L.offc(i, j0 - 2, kp + b - 1));
                W[2][b] = load_mat(a, L.offc(i, j0 - 1, kp + b - 1));
                Wn[b] = load_mat(a, L.offc(i, j0, kp + b - 1));
            }
            {
                const long c = L.offc(i, j0 - 1, kp);
                pn = a.q[c];
                un = a.q[fs + c];
            }
#pragma unroll
            for (int jj = 0; jj < TY + 2; jj++) {
                const int jp = j0 - 1 + jj;
#pragma unroll
                for (int b = 0; b < 3; b++) {
                    W[0][b] = W[1][b];
                    W[1][b] = W[2][b];
                    W[2][b] = Wn[b];
                }
                const double p = pn, u = un;
                if (jj + 1 < TY + 2) {  // next slice's data, in flight during this one
#pragma unroll
                    for (int b = 0; b < 3; b++) Wn[b] = load_mat(a, L.offc(i, jp + 2, kp + b - 1));
                    const long cn = L.offc(i, jp + 1, kp);
                    pn = a.q[cn];
                    un = a.q[fs + cn];
                }
                const Mat me = W[1][1];
                Iface f = solve<M>(lane_up(p), lane_up(u), mat_up(me), p, u, me);
                Iface fl, fr;
                fl.w1 = wv_up(f.w1); fl.w2 = wv_up(f.w2);
                fr.w1 = wv_dn(f.w1); fr.w2 = wv_dn(f.w2);
                Wv cq = limited_cq<M, Cfg>(a, f, fl, fr, d);
                Iface hi;
                hi.am = wv_dn(f.am);
                Wv cqhi = wv_dn(cq);
                T9 T;
                slice_position<M, Cfg, KL>(a, f, cq, hi, cqhi, W, d, d1, d2, T, klut, a.nmat);
                // x: n = u, t1 = v (rows, offset a), t2 = w (planes, offset b)
                scatter<0>(acc, T, trans, 1, 2, 3, [&](int t) {
                    const int jy = jp + (t % 3 - 1) - j0, kz = kp + (t / 3 - 1) - k0;
                    const bool ok = jy >= 0 && jy < TY && kz >= 0 && kz < TZ;
                    return SLOTI(ok, kz, jy, l);
                });
            }
        }
    }

    // ---- y sweeps: slice (i, k') marched along j; t1 = z, t2 = x
    {
        const double d = a.dtdy, d1 = a.dtdz, d2 = a.dtdx;
        for (int kk = 0; kk < TZ + 2; kk++) {
            const int kp = k0 - 1 + kk;
            double P[TY + 4], V[TY + 4];
            Mat C3[TY + 4][3];
            Iface F[TY + 4];
            Wv CQ[TY + 4];
            // every load of the column first, so that they are all in flight
            // before the first solve needs one
#pragma unroll
            for (int cc = 0; cc < TY + 4; cc++) {
                const int j = j0 - 2 + cc;
                const long c = L.offc(i, j, kp);
                P[cc] = a.q[c];
                V[cc] = a.q[2 * fs + c];
#pragma unroll
                for (int s = 0; s < 3; s++) C3[cc][s] = load_mat(a, L.offc(i, j, kp + s - 1));
            }
#pragma unroll
            for (int cc = 0; cc < TY + 4; cc++) {
                if (cc >= 1) F[cc] = solve<M>(P[cc - 1], V[cc - 1], C3[cc - 1][1], P[cc], V[cc], C3[cc][1]);
                if (cc >= 3) CQ[cc - 1] = limited_cq<M, Cfg>(a, F[cc - 1], F[cc - 2], F[cc], d);
                if (cc >= 4) {
                    const int r = cc - 2, jy = cc - 4;
                    Mat N[3][3];
#pragma unroll
                    for (int s = 0; s < 3; s++) {
                        N[s][1] = C3[r][s];
                        N[s][0] = mat_up(C3[r][s]);
                        N[s][2] = mat_dn(C3[r][s]);
                    }
                    T9 T;
                    slice_position<M, Cfg, KL>(a, F[r], CQ[r], F[r + 1], CQ[r + 1], N, d, d1, d2, T, klut, a.nmat);
                    // y: n = v, t1 = w (planes, offset a), t2 = u (lanes, offset b)
                    scatter<1>(acc, T, trans, 2, 3, 1, [&](int t) {
                        const int kz = kp + (t % 3 - 1) - k0, ln = l + (t / 3 - 1);
                        const bool ok = kz >= 0 && kz < TZ;
                        return SLOTI(ok, kz, jy, ln);
                    });
                }
            }
        }
    }

    // ---- z sweeps: slice (i, j') marched along k; t1 = x, t2 = y
    {
        const double d = a.dtdz, d1 = a.dtdx, d2 = a.dtdy;
        for (int jj = 0; jj < TY + 2; jj++) {
            const int jp = j0 - 1 + jj;
            double P[TZ + 4], W_[TZ + 4];
            Mat C3[TZ + 4][3];
            Iface F[TZ + 4];
            Wv CQ[TZ + 4];
#pragma unroll
            for (int cc = 0; cc < TZ + 4; cc++) {
                const int k = k0 - 2 + cc;
                const long c = L.offc(i, jp, k);
                P[cc] = a.q[c];
                W_[cc] = a.q[3 * fs + c];
#pragma unroll
                for (int s = 0; s < 3; s++) C3[cc][s] = load_mat(a, L.offc(i, jp + s - 1, k));
            }
#pragma unroll
            for (int cc = 0; cc < TZ + 4; cc++) {
                if (cc >= 1) F[cc] = solve<M>(P[cc - 1], W_[cc - 1], C3[cc - 1][1], P[cc], W_[cc], C3[cc][1]);
                if (cc >= 3) CQ[cc - 1] = limited_cq<M, Cfg>(a, F[cc - 1], F[cc - 2], F[cc], d);
                if (cc >= 4) {
                    const int r = cc - 2, kz = cc - 4;
                    Mat N[3][3];
#pragma unroll
                    for (int s = 0; s < 3; s++) {
                        N[1][s] = C3[r][s];
                        N[0][s] = mat_up(C3[r][s]);
                        N[2][s] = mat_dn(C3[r][s]);
                    }
                    T9 T;
                    slice_position<M, Cfg, KL>(a, F[r], CQ[r], F[r + 1], CQ[r + 1], N, d, d1, d2, T, klut, a.nmat);
                    // z: n = w, t1 = u (lanes, offset a), t2 = v (rows, offset b)
                    scatter<2>(acc, T, trans, 3, 1, 2, [&](int t) {
                        const int jy = jp + (t / 3 - 1) - j0, ln = l + (t % 3 - 1);
                        const bool ok = jy >= 0 && jy < TY;
                        return SLOTI(ok, kz, jy, ln);
                    });
                }
            }
        }
    }

    // ---- qnew = qold + accumulated change, output lanes only
    if (l >= 2 && l < 62 && i <= L.mx) {
#pragma unroll
        for (int kz = 0; kz < TZ; kz++)
#pragma unroll
            for (int jy = 0; jy < TY; jy++) {
                const int j = j0 + jy, k = k0 + kz;
                if (j > L.my || k > L.mz) continue;
                const long c = L.off(i, j, k);
#pragma unroll
                for (int m = 0; m < 4; m++) a.qn[m * fs + c] = a.q[m * fs + c] + ACC(kz, jy, m, l);
            }
    }
#undef ACC
#undef SLOTI
}

// ------------------------------------------------------------ split sweeps

// A sweep reads field m from plane qf[m] and writes it to plane qo[m]: the
// pressure and its own velocity component are updated, another field is
// copied when qo[m] is set and otherwise not touched (a split step then
// passes that field's plane on to the next sweep).  Materials come from the
// byte index mat and the table lut (the same doubles as aux) when mat is
// set.
struct SweepArgs3 {
    const double* qf[4];
    double* qo[4];
    const double* __restrict__ aux;
    const unsigned char* __restrict__ mat;
    const double* __restrict__ lut;
    const double* __restrict__ capa;  // capacity plane (method(6) > 0), or nullptr
    Layout3 L;
    long afs;
    double d;  // dt / (spacing along the sweep)
    int m2, fwave, lim0, lim1;
    int nstrips;  // x: 60-cell strips; y/z: 64-cell lane blocks
    int seg;      // y/z: cells per segment along the sweep
    int nsegs;
};

__device__ inline Mat sweep_mat(const SweepArgs3& a, long c)
{
    if (a.mat) {
        const int id = a.mat[c];
        return {a.lut[2 * id], a.lut[2 * id + 1], id};
    }
    return {a.aux[c], a.aux[a.afs + c], 0};
}

// update of one slice cell: (q + qadd) - d*(F(hi) - F(lo)), step3ds.f90;
// with capacity kappa the flux difference term is divided by it
__device__ inline double ds_update(double q, double qadd, double Fhi, double Flo, double d)
{
    return q + qadd - d * (Fhi - Flo);
}
__device__ inline double ds_update_capa(double q, double qadd, double Fhi, double Flo, double d,
                                        double kappa)
{
    return q + qadd - d * (Fhi - Flo) / kappa;
}

// x sweep of step3ds (idir = 1): one wave per 60-cell strip of one x-slice
// (j, k), over all slices including the ghost ones; lanes hold cells and get
// their neighbours through DPP.  Cells outside the updated range are copied.
// With capacity, the cell's dt/dx is d / kappa (step3ds.f90: dtdx1d), the
// interface's correction uses the mean of its two cells' ratios and the flux
// difference is divided by kappa.
template <class M>
__global__ __launch_bounds__(256) void sweep3_x_kernel(SweepArgs3 a)
{
    const int l = threadIdx.x & 63;
    const long w = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const Layout3& L = a.L;
    const int nyt = L.my + 2 * L.mbc;
    const long nsl = (long)nyt * (L.mz + 2 * L.mbc);
    if (w >= nsl * a.nstrips) return;  // uniform per wave
    const int strip = (int)(w % a.nstrips);
    const long sl = w / a.nstrips;
    const int j = (int)(sl % nyt) + 1 - L.mbc;
    const int k = (int)(sl / nyt) + 1 - L.mbc;
    const int i = kOut * strip + l - 1;
    const long c = L.offc(i, j, k);
    const double p = a.qf[0][c], u = a.qf[1][c];
    const Mat mm = sweep_mat(a, c);
    Iface f = solve<M>(lane_up(p), lane_up(u), mat_up(mm), p, u, mm);
    const double kap = a.capa ? a.capa[c] : 1.0;
    const double dc = a.capa ? a.d / kap : a.d;  // dtdx1d of this cell
    Wv cq = {0.0, 0.0};
    if (a.m2 == 2) {
        Wv w1 = limit<M>(f.w1, wv_up(f.w1), wv_dn(f.w1), f.s1, a.lim0);
        Wv w2 = limit<M>(f.w2, wv_up(f.w2), wv_dn(f.w2), f.s2, a.lim1);
        const double dave = a.capa ? 0.5 * (lane_up(dc) + dc) : a.d;
        cq = cqxx(w1, f.s1, w2, f.s2, dave, a.fwave);
    }
    const Wv amhi = wv_dn(f.am), cqhi = wv_dn(cq);
    const bool own = (l >= 2 && l < 62) || (strip == 0 && l < 2);
    if (!own || i > L.mx + L.mbc) return;
    const bool upd = i >= 1 && i <= L.mx && j >= 0 && j <= L.my + 1 && k >= 0 && k <= L.mz + 1;
    const long o = L.off(i, j, k);
    if (upd && a.capa) {
        const double qp = (0.0 - dc * f.ap.p) - dc * amhi.p;
        const double qu = (0.0 - dc * f.ap.n) - dc * amhi.n;
        a.qo[0][o] = ds_update_capa(p, qp, cqhi.p, cq.p, a.d, kap);
        a.qo[1][o] = ds_update_capa(u, qu, cqhi.n, cq.n, a.d, kap);
    } else if (upd) {
        const double qp = (0.0 - a.d * f.ap.p) - a.d * amhi.p;
        const double qu = (0.0 - a.d * f.ap.n) - a.d * amhi.n;
        a.qo[0][o] = ds_update(p, qp, cqhi.p, cq.p, a.d);
        a.qo[1][o] = ds_update(u, qu, cqhi.n, cq.n, a.d);
    } else {
        a.qo[0][o] = p;
        a.qo[1][o] = u;
    }
    if (a.qo[2]) a.qo[2][o] = a.qf[2][o];
    if (a.qo[3]) a.qo[3][o] = a.qf[3][o];
}

// y (dir = 1) or z (dir = 2) sweep of step3ds: lanes hold 64 consecutive
// x-cells; each lane marches along the sweep direction over one segment of
// its slice, keeping the raw interfaces s, s+1, s+2 and the limited cqxx of
// s, s+1 in registers.
template <class M>
__global__ __launch_bounds__(256) void sweep3_yz_kernel(SweepArgs3 a, int dir)
{
    const int l = threadIdx.x & 63;
    const long w = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const Layout3& L = a.L;
    const int nt = dir == 1 ? L.mz + 2 * L.mbc : L.my + 2 * L.mbc;  // other transverse extent
    if (w >= (long)a.nstrips * nt * a.nsegs) return;               // uniform per wave
    const int blk = (int)(w % a.nstrips);
    const long r = w / a.nstrips;
    const int seg = (int)(r % a.nsegs);
    const int o2 = (int)(r / a.nsegs) + 1 - L.mbc;  // k for dir 1, j for dir 2
    const int i = 64 * blk + l + 1 - L.mbc;
    if (i > L.mx + L.mbc) return;
    const int n = dir == 1 ? L.my : L.mz;
    const int n2 = dir == 1 ? L.mz : L.my;
    const int comp = dir == 1 ? 2 : 3;
    const bool slice_ok = i >= 0 && i <= L.mx + 1 && o2 >= 0 && o2 <= n2 + 1;
    const int s0 = seg == 0 ? 1 - L.mbc : 1 + seg * a.seg;
    const int s1 = seg == a.nsegs - 1 ? n + L.mbc : (seg + 1) * a.seg;
    auto idx = [&](int s) -> long {
        s = s < 1 - L.mbc ? 1 - L.mbc : (s > n + L.mbc ? n + L.mbc : s);
        return dir == 1 ? L.off(i, s, o2) : L.off(i, o2, s);
    };
    // cell s: p, normal velocity, material and (with capacity) dt/d / kappa
    auto cell = [&](int s, double& pv, double& nv, Mat& mv, double& dv) {
        const long c = idx(s);
        pv = a.qf[0][c];
        nv = a.qf[comp][c];
        mv = sweep_mat(a, c);
        dv = a.capa ? a.d / a.capa[c] : a.d;
    };
    // limited cqxx of interface f between cells with ratios dl, dr
    auto lim_cq = [&](const Iface& f, const Iface& fl, const Iface& fr, double dl, double dr) -> Wv {
        if (a.m2 != 2) return {0.0, 0.0};
        Wv w1 = limit<M>(f.w1, fl.w1, fr.w1, f.s1, a.lim0);
        Wv w2 = limit<M>(f.w2, fl.w2, fr.w2, f.s2, a.lim1);
        return cqxx(w1, f.s1, w2, f.s2, a.capa ? 0.5 * (dl + dr) : a.d, a.fwave);
    };
    double pm2, nm2, pm1, nm1, p0, n0, p1, n1, dm2, dm1, d0, d1;
    Mat mm2, mm1, m0, m1;
    cell(s0 - 2, pm2, nm2, mm2, dm2);
    cell(s0 - 1, pm1, nm1, mm1, dm1);
    cell(s0, p0, n0, m0, d0);
    cell(s0 + 1, p1, n1, m1, d1);
    Iface fprev = solve<M>(pm2, nm2, mm2, pm1, nm1, mm1);  // interface s0-1
    Iface fcur = solve<M>(pm1, nm1, mm1, p0, n0, m0);      // interface s0
    Iface fnext = solve<M>(p0, n0, m0, p1, n1, m1);        // interface s0+1
    Wv cqcur = lim_cq(fcur, fprev, fnext, dm1, d0);
    for (int s = s0; s <= s1; s++) {
        double p2, n2v, d2;
        Mat m2;
        cell(s + 2, p2, n2v, m2, d2);
        const Iface fnn = solve<M>(p1, n1, m1, p2, n2v, m2);  // interface s+2
        const Wv cqnext = lim_cq(fnext, fcur, fnn, d0, d1);
        const long o = idx(s);
        if (slice_ok && s >= 1 && s <= n && a.capa) {
            const double kap = a.capa[o];
            const double qp = (0.0 - d0 * fcur.ap.p) - d0 * fnext.am.p;
            const double qv = (0.0 - d0 * fcur.ap.n) - d0 * fnext.am.n;
            a.qo[0][o] = ds_update_capa(p0, qp, cqnext.p, cqcur.p, a.d, kap);
            a.qo[comp][o] = ds_update_capa(n0, qv, cqnext.n, cqcur.n, a.d, kap);
        } else if (slice_ok && s >= 1 && s <= n) {
            const double qp = (0.0 - a.d * fcur.ap.p) - a.d * fnext.am.p;
            const double qv = (0.0 - a.d * fcur.ap.n) - a.d * fnext.am.n;
            a.qo[0][o] = ds_update(p0, qp, cqnext.p, cqcur.p, a.d);
            a.qo[comp][o] = ds_update(n0, qv, cqnext.n, cqcur.n, a.d);
        } else {
            a.qo[0][o] = p0;
            a.qo[comp][o] = n0;
        }
#pragma unroll
        for (int m = 1; m < 4; m++)
            if (m != comp && a.qo[m]) a.qo[m][o] = a.qf[m][o];
        fcur = fnext;
        fnext = fnn;
        cqcur = cqnext;
        p0 = p1; n0 = n1; m0 = m1; d0 = d1;
        p1 = p2; n1 = n2v; m1 = m2; d1 = d2;
    }
}

}  // namespace s3
}  // namespace clawgpu
