// 3D unsplit step, z-marching form (vc acoustics, material table): every
// x-, y- and z-slice position of step3/flux3 (src/3d/step3.f90:179-868,
// src/3d/flux3.f90:233-608) is evaluated once per block instead of once per
// tile that its increments reach.
//
// A block owns a column of 60 x TY cells (x lanes 2..61 of a 64-lane strip,
// rows j0..j0+TY-1) over a segment of planes kb..ke, and marches along z.
// It has TY + 2 waves; wave r owns row j = j0 - 1 + r (rows j0-1 and
// j0+TY are halo rows whose positions are needed only for the increments
// they send into the tile).  At plane k every wave evaluates
//   * the x-slice position (j, k): one interface per lane, neighbours by DPP;
//   * the z-slice position (j, k): the z interfaces k-1/2, k+1/2 and k+3/2 of
//     its own column and the limited cqxx of k-1/2 are carried in registers
//     from plane to plane, so each z interface is solved once;
//   * the y interface j+1/2 at plane k, published in LDS; after a barrier the
//     rows of the tile read the four y interfaces around them and evaluate
//     the y-slice position (j, k).
// A position changes the nine cells of its 3x3 transverse neighbourhood
// (step3.f90:264-335): those in the wave's own row (x positions: planes
// k-1..k+1; y positions: planes k-1..k+1 and lanes l-1..l+1; z positions:
// lanes l-1..l+1) are added to register accumulators of planes k-1, k, k+1;
// those in rows j-1 and j+1 are summed per cell and written to a per-wave
// LDS mailbox that the neighbouring wave adds after the barrier.  When plane
// k is done, plane k-1 of the wave's row is complete and is stored (each
// plane's accumulators start from its qold, added when the plane is the
// current one, so the store needs no second read of q).  Sums per cell run in another order than the
// reference's slice-by-slice updates (as in the tile kernel): agreement to
// round-off, not bit for bit.
#pragma once

#include "step3d_kernel.hpp"

namespace clawgpu {
namespace s3 {

// mailbox slots: plane k-1 (p, v, w), plane k (p, u, v, w), plane k+1 (p, v, w)
constexpr int kBox = 10;

struct Raw {
    Wv w1, w2;
};

// an interface from its raw waves and the materials on either side, with
// the operations of solve()
__device__ inline Iface from_raw(const Raw& w, const Mat& lo, const Mat& hi)
{
    Iface f;
    f.w1 = w.w1;
    f.w2 = w.w2;
    f.s1 = -lo.c;
    f.s2 = hi.c;
    f.am = {f.s1 * f.w1.p, f.s1 * f.w1.n};
    f.ap = {f.s2 * f.w2.p, f.s2 * f.w2.n};
    return f;
}

// An interface carried as its wave strengths (a1, a2): the waves are
// a1 (-Z_lo, 1) and a2 (Z_hi, 1), rebuilt with the products solve() forms,
// so they are the same doubles.
struct Alpha {
    double a1, a2;
};

__device__ inline Iface from_alpha(const Alpha& w, const Mat& lo, const Mat& hi)
{
    return from_raw(Raw{{-w.a1 * lo.Z, w.a1}, {w.a2 * hi.Z, w.a2}}, lo, hi);
}

__device__ inline Iface waves_only(const Raw& w)
{
    Iface f{};
    f.w1 = w.w1;
    f.w2 = w.w2;
    return f;
}

// The nine-cell change of one slice position whose transverse neighbourhood
// is one material (Z, c), from the same inputs as slice_position.  There
// every transverse coefficient set of kcoef is (mp, mt, pp, pt) =
// (-h, g, h, g) with h = cZ/(2Z), g = c/(2Z), and the G/H records of
// flux3.f90:378-608 collapse: summed over both sides (A^+ of the low
// interface, sgn +1; A^- of the high interface, sgn -1) with
//   S = a+ + a-,  D = cq+ - cq- (method(3) = 2x), B = S - 3 D (x = 2),
// the centre-slice records are
//   gp[1][1] = -gp[0][1] = G1 = -d/2 h S + 2 cc h^2 B + d h D,
//   gt[0][1] = gt[1][1] = Gt1 = -d/2 g S + 2 cc h g B + d g D,
// (hp, ht likewise with ee in place of cc), and the other slices hold
// +-cc h^2 B, -cc h g B (and the ee forms), cc = d d2/6, ee = d d1/6 (the
// double-transverse terms, method(3) = x1, x2).  The targets of
// step3.f90:264-335 then take the few distinct values written out below.
template <class M, class Cfg>
__device__ inline void homog_position(const Args3& a, double h, double g, const Iface& lo, Wv cqlo,
                                      const Iface& hi, Wv cqhi, double d, double d1, double d2,
                                      T9& T)
{
    const int m3 = Cfg::m3(a), m4 = Cfg::m4(a);
    const double qp = (-d * lo.ap.p) - d * hi.am.p;
    const double qn = (-d * lo.ap.n) - d * hi.am.n;
    T.n = qn - d * (cqhi.n - cqlo.n);
    const double base = qp - d * (cqhi.p - cqlo.p);
#pragma unroll
    for (int t = 0; t < 9; t++) T.p[t] = T.t1[t] = T.t2[t] = 0.0;
    T.p[4] = base;
    if (m3 <= 0) return;
    const double hd = 0.5 * d;
    const double S = lo.ap.p + hi.am.p;
    const double D = m3 == 2 ? cqlo.p - cqhi.p : 0.0;
    const double B = (m3 == 2 && m4 == 2) ? S - 3.0 * D : S;
    double G1 = -hd * h * S, Gt1 = -hd * g * S;
    double H1 = G1, Ht1 = Gt1;
    double Gx = 0.0, Gy = 0.0, Hx = 0.0, Hy = 0.0;
    if (m4 > 0) {
        const double cc = (1.0 / 6.0) * d * d2, ee = (1.0 / 6.0) * d * d1;
        const double hB = h * B;
        Gx = cc * h * hB;
        Gy = cc * g * hB;
        Hx = ee * h * hB;
        Hy = ee * g * hB;
        G1 += 2.0 * Gx;
        Gt1 += 2.0 * Gy;
        H1 += 2.0 * Hx;
        Ht1 += 2.0 * Hy;
    }
    if (m3 == 2) {
        const double dh = d * h * D, dg = d * g * D;
        G1 += dh;
        H1 += dh;
        Gt1 += dg;
        Ht1 += dg;
    }
    const double X = -(d1 * Gx + d2 * Hx);
    T.p[0] = T.p[2] = T.p[6] = T.p[8] = X;
    T.p[1] = T.p[7] = 2.0 * d1 * Gx + d2 * H1;
    T.p[3] = T.p[5] = d1 * G1 + 2.0 * d2 * Hx;
    T.p[4] = base - 2.0 * (d1 * G1 + d2 * H1);
    T.t1[0] = T.t1[6] = d1 * Gy;
    T.t1[2] = T.t1[8] = -(d1 * Gy);
    T.t1[3] = -(d1 * Gt1);
    T.t1[5] = d1 * Gt1;
    T.t2[0] = T.t2[2] = d2 * Hy;
    T.t2[6] = T.t2[8] = -(d2 * Hy);
    T.t2[1] = -(d2 * Ht1);
    T.t2[7] = d2 * Ht1;
}

template <int TY>
constexpr int zmarch_lds_bytes()
{
    return ((TY + 3) * 4 * 64 + 2 * (TY + 2) * kBox * 64 + 2 * kMaxMat) * 8 +
           kMaxKcMat * kMaxKcMat * kMaxKcMat * (int)sizeof(Kc);
}

template <int TY, class Cfg, bool KL>
__global__ __launch_bounds__(64 * (TY + 2)) void step3_zmarch_kernel(Args3 a, int kseg)
{
    using M = FastMath;
    constexpr int NW = TY + 2;
    enum { P = 0, U = 1, V = 2, W = 3 };
    __shared__ double lut[2 * kMaxMat];
    __shared__ Kc klut_s[KL ? kMaxKcMat * kMaxKcMat * kMaxKcMat : 1];
    __shared__ double ybuf[(NW + 1) * 4 * 64];  // raw waves of y interfaces j0-3/2 .. j0+TY+1/2
    __shared__ double ubox[NW * kBox * 64];     // wave r -> row above (wave r+1)
    __shared__ double dbox[NW * kBox * 64];     // wave r -> row below (wave r-1)
    const int tid = threadIdx.x, l = tid & 63;
    const int r = __builtin_amdgcn_readfirstlane(tid >> 6);
    for (int e = tid; e < 2 * a.nmat; e += 64 * NW) lut[e] = a.lut[e];
    if constexpr (KL) {
        const int n = a.nmat;
        for (int e = tid; e < n * n * n; e += 64 * NW) {
            const int lo = e / (n * n), mid = (e / n) % n, hi = e % n;
            klut_s[e] = kcoef<M>(Mat{a.lut[2 * lo], a.lut[2 * lo + 1], lo},
                                 Mat{a.lut[2 * mid], a.lut[2 * mid + 1], mid},
                                 Mat{a.lut[2 * hi], a.lut[2 * hi + 1], hi});
        }
    }
    const Kc* klut = klut_s;

    const Layout3& L = a.L;
    const int nzs = (L.mz + kseg - 1) / kseg;
    int t = blockIdx.x;
    const int ntiles = a.ntx * a.nty * nzs;
    if (a.xcd & 1) {
        // blocks are dealt to the XCDs round-robin; renumber so that XCD x
        // works on blocks [x*per, (x+1)*per), neighbours sharing its L2
        const int per = (ntiles + 7) / 8;
        t = (t & 7) * per + (t >> 3);
    }
    if (t >= ntiles) return;  // uniform over the block
    const int tx = t % a.ntx;
    t /= a.ntx;
    const int ty = t % a.nty;
    const int zs = t / a.nty;
    const int i = kOut * tx + l - 1;
    const int j0 = TY * ty + 1;
    const int j = j0 - 1 + r;
    const int kb = zs * kseg + 1;
    const int ke = min(kb + kseg - 1, L.mz);
    const bool inner = r >= 1 && r <= TY;                   // a row of the tile
    const bool store_row = inner && j <= L.my;
    const bool out_lane = l >= 2 && l < 62 && i <= L.mx;
    const long fs = L.fs;
    const bool trans = Cfg::m3(a) > 0;
    const double dx_ = a.dtdx, dy_ = a.dtdy, dz_ = a.dtdz;

    auto mat_of = [&](int id) -> Mat { return {lut[2 * id], lut[2 * id + 1], id}; };
    auto id_at = [&](int jj, int kk) -> int { return a.mat[L.offc(i, jj, kk)]; };

    // material ids (< kMaxMat <= 256) of rows j-1..j+1 at planes k-1..k+2,
    // one byte per plane: Rp[s] byte p is the id of row j+s-1, plane k-1+p
    static_assert(kMaxMat <= 256, "material ids are kept as bytes");
    unsigned Rp[3];
#pragma unroll
    for (int s = 0; s < 3; s++) {
        Rp[s] = 0u;
#pragma unroll
        for (int p = 0; p < 4; p++) Rp[s] |= (unsigned)id_at(j + s - 1, kb - 2 + p) << (8 * p);
    }
    auto R = [&](int s, int p) -> int { return (int)((Rp[s] >> (8 * p)) & 0xffu); };

    // z pipeline of the own column, ready for plane kb: interfaces kb-1/2
    // (Fa) and kb+1/2 (Fb), cqxx of kb-1/2, q(p, w) of cell kb+1
    Alpha Fa, Fb;
    Wv cqa;
    double pz1, wz1;
    {
        double pz[4], wz[4];
        int mz_[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const long c = L.offc(i, j, kb - 2 + e);
            pz[e] = a.q[c];
            wz[e] = a.q[3 * fs + c];
            mz_[e] = a.mat[c];
        }
        const Iface f0 = solve<M>(pz[0], wz[0], mat_of(mz_[0]), pz[1], wz[1], mat_of(mz_[1]));
        const Iface f1 = solve<M>(pz[1], wz[1], mat_of(mz_[1]), pz[2], wz[2], mat_of(mz_[2]));
        const Iface f2 = solve<M>(pz[2], wz[2], mat_of(mz_[2]), pz[3], wz[3], mat_of(mz_[3]));
        cqa = limited_cq<M, Cfg>(a, f1, f0, f2, dz_);
        Fa = {f1.w1.n, f1.w2.n};
        Fb = {f2.w1.n, f2.w2.n};
        pz1 = pz[3];
        wz1 = wz[3];
    }

    // q of the current plane: own cell (p, u, v, w), the row above (p, v)
    // and, for wave 0, the row below (p, v); cell k+2 of the own column (p, w)
    auto load_plane = [&](int k, double& p0, double& u0, double& v0, double& w0, double& pN,
                          double& vN, double& pS, double& vS, double& pz2, double& wz2) {
        const long c = L.offc(i, j, k);
        p0 = a.q[c];
        u0 = a.q[fs + c];
        v0 = a.q[2 * fs + c];
        w0 = a.q[3 * fs + c];
        const long cn = L.offc(i, j + 1, k);
        pN = a.q[cn];
        vN = a.q[2 * fs + cn];
        if (r == 0) {
            const long cs = L.offc(i, j - 1, k);
            pS = a.q[cs];
            vS = a.q[2 * fs + cs];
        }
        const long cz = L.offc(i, j, k + 2);
        pz2 = a.q[cz];
        wz2 = a.q[3 * fs + cz];
    };
    double p0, u0, v0, w0, pN, vN, pS = 0.0, vS = 0.0, pz2, wz2;
    load_plane(kb - 1, p0, u0, v0, w0, pN, vN, pS, vS, pz2, wz2);

    double A[3][4];  // own-row change of planes k-1, k, k+1
#pragma unroll
    for (int b = 0; b < 3; b++)
#pragma unroll
        for (int m = 0; m < 4; m++) A[b][m] = 0.0;
    double* const myu = ubox + r * kBox * 64 + l;
    double* const myd = dbox + r * kBox * 64 + l;
    double* const ybw = ybuf + (r + 1) * 4 * 64 + l;

    __syncthreads();  // lut / klut

    // transverse coefficients (h, g) of a one-material neighbourhood
    auto hg_of = [&](int m) -> Kc {
        if constexpr (KL) return klut[(m * a.nmat + m) * a.nmat + m];
        else {
            const Mat mm = mat_of(m);
            return kcoef<M>(mm, mm, mm);
        }
    };

    for (int k = kb - 1; k <= ke + 1; k++) {
        A[1][P] += p0;  // qold of plane k
        A[1][U] += u0;
        A[1][V] += v0;
        A[1][W] += w0;
        const Mat me = mat_of(R(1, 1));
        // one material over the whole wave's 3x3 neighbourhoods: rows j-1..j+1
        // and planes k-1..k+1 (x), row j (y), plane k (z)
        const int m0 = __builtin_amdgcn_readfirstlane(R(1, 1));
        // one material over the x-, y- and z-neighbourhoods: planes k-1..k+1
        // (bytes 0..2) of each row, or plane k of each row, equal to m0
        const unsigned m4 = (unsigned)m0 * 0x01010101u;
        const bool r0 = ((Rp[0] ^ m4) & 0x00ffffffu) == 0u;
        const bool r1 = ((Rp[1] ^ m4) & 0x00ffffffu) == 0u;
        const bool r2 = ((Rp[2] ^ m4) & 0x00ffffffu) == 0u;
        const bool sy = r1;
        const bool sz = ((Rp[0] ^ m4) & 0xff00u) == 0u && ((Rp[1] ^ m4) & 0xff00u) == 0u &&
                        ((Rp[2] ^ m4) & 0xff00u) == 0u;  // plane k (byte 1) of each row
        const bool sx = r0 && r1 && r2;
        const bool ux = __all(sx), uy = __all(sy), uz = __all(sz);
        Kc K0{};
        if (ux || uy || uz) K0 = hg_of(m0);

        // ---- x-slice position (j, k): t1 = y (rows), t2 = z (planes)
        {
            Mat N[3][3];
#pragma unroll
            for (int s = 0; s < 3; s++)
#pragma unroll
                for (int p = 0; p < 3; p++) N[s][p] = mat_of(R(s, p));
            const Iface f = solve<M>(lane_up(p0), lane_up(u0), mat_up(me), p0, u0, me);
            Iface fl, fr;
            fl.w1 = wv_up(f.w1);
            fl.w2 = wv_up(f.w2);
            fr.w1 = wv_dn(f.w1);
            fr.w2 = wv_dn(f.w2);
            const Wv cq = limited_cq<M, Cfg>(a, f, fl, fr, dx_);
            Iface hi;
            hi.am = wv_dn(f.am);
            const Wv cqhi = wv_dn(cq);
            // target 3(b+1) + (a+1): row offset a, plane offset b; n = u, t1 = v, t2 = w
            auto acc_x = [&](const T9& T) {
            A[1][U] += T.n;
#pragma unroll
            for (int b = 0; b < 3; b++) {
                A[b][P] += T.p[3 * b + 1];
                A[b][V] += T.t1[3 * b + 1];
                A[b][W] += T.t2[3 * b + 1];
            }
            // mailbox slots of plane offsets -1, 0, +1: (0,1,2), (3,5,6), (7,8,9) for (p, v, w)
            if (trans) {
                constexpr int slot[3][3] = {{0, 1, 2}, {3, 5, 6}, {7, 8, 9}};
#pragma unroll
                for (int b = 0; b < 3; b++) {
                    myu[slot[b][0] * 64] = T.p[3 * b + 2];
                    myu[slot[b][1] * 64] = T.t1[3 * b + 2];
                    myu[slot[b][2] * 64] = T.t2[3 * b + 2];
                    myd[slot[b][0] * 64] = T.p[3 * b];
                    myd[slot[b][1] * 64] = T.t1[3 * b];
                    myd[slot[b][2] * 64] = T.t2[3 * b];
                }
            }
            };
            T9 T;
            if (ux) {
                homog_position<M, Cfg>(a, K0.pp, K0.pt, f, cq, hi, cqhi, dx_, dy_, dz_, T);
                acc_x(T);
            } else {
                slice_position<M, Cfg, KL>(a, f, cq, hi, cqhi, N, dx_, dy_, dz_, T, klut, a.nmat);
                acc_x(T);
            }
        }

        // ---- z-slice position (j, k): t1 = x (lanes), t2 = y (rows)
        if (k >= kb && k <= ke) {
            const Mat mk1 = mat_of(R(1, 2)), mk2 = mat_of(R(1, 3));
            const Iface fc = solve<M>(pz1, wz1, mk1, pz2, wz2, mk2);  // interface k+3/2
            const Iface fa = from_alpha(Fa, mat_of(R(1, 0)), me);      // k-1/2
            const Iface fb = from_alpha(Fb, me, mk1);                   // k+1/2
            const Wv cqb = limited_cq<M, Cfg>(a, fb, fa, fc, dz_);
            Mat N[3][3];
#pragma unroll
            for (int s = 0; s < 3; s++) {
                const Mat c = mat_of(R(s, 1));
                N[1][s] = c;
                N[0][s] = mat_up(c);
                N[2][s] = mat_dn(c);
            }
            // target 3(b+1) + (a+1): lane offset a, row offset b; n = w, t1 = u, t2 = v
            auto acc_z = [&](const T9& T) {
            A[1][W] += T.n;
            A[1][P] += T.p[4] + lane_up(T.p[5]) + lane_dn(T.p[3]);
            A[1][U] += T.t1[4] + lane_up(T.t1[5]) + lane_dn(T.t1[3]);
            A[1][V] += T.t2[4] + lane_up(T.t2[5]) + lane_dn(T.t2[3]);
            if (trans) {
                myu[3 * 64] += T.p[7] + lane_up(T.p[8]) + lane_dn(T.p[6]);
                myu[4 * 64] = T.t1[7] + lane_up(T.t1[8]) + lane_dn(T.t1[6]);
                myu[5 * 64] += T.t2[7] + lane_up(T.t2[8]) + lane_dn(T.t2[6]);
                myd[3 * 64] += T.p[1] + lane_up(T.p[2]) + lane_dn(T.p[0]);
                myd[4 * 64] = T.t1[1] + lane_up(T.t1[2]) + lane_dn(T.t1[0]);
                myd[5 * 64] += T.t2[1] + lane_up(T.t2[2]) + lane_dn(T.t2[0]);
            }
            };
            T9 T;
            if (uz) {
                homog_position<M, Cfg>(a, K0.pp, K0.pt, fa, cqa, fb, cqb, dz_, dx_, dy_, T);
                acc_z(T);
            } else {
                slice_position<M, Cfg, KL>(a, fa, cqa, fb, cqb, N, dz_, dx_, dy_, T, klut, a.nmat);
                acc_z(T);
            }
            Fa = Fb;
            Fb = {fc.w1.n, fc.w2.n};
            cqa = cqb;
            pz1 = pz2;
            wz1 = wz2;
        } else if (trans) {
            myu[4 * 64] = 0.0;
            myd[4 * 64] = 0.0;
        }

        // ---- y interface j+1/2 (and j-1/2 for wave 0) at plane k
        {
            const Iface fy = solve<M>(p0, v0, me, pN, vN, mat_of(R(2, 1)));
            ybw[0] = fy.w1.p;
            ybw[64] = fy.w1.n;
            ybw[128] = fy.w2.p;
            ybw[192] = fy.w2.n;
            if (r == 0) {
                const Iface fs_ = solve<M>(pS, vS, mat_of(R(0, 1)), p0, v0, me);
                double* yb0 = ybuf + l;
                yb0[0] = fs_.w1.p;
                yb0[64] = fs_.w1.n;
                yb0[128] = fs_.w2.p;
                yb0[192] = fs_.w2.n;
            }
        }
        __syncthreads();

        // next plane's q, in flight during the rest of this one
        double np0 = 0.0, nu0 = 0.0, nv0 = 0.0, nw0 = 0.0, npN = 0.0, nvN = 0.0, npS = 0.0,
               nvS = 0.0, npz2 = 0.0, nwz2 = 0.0;
        if (k + 1 <= ke + 1)
            load_plane(k + 1, np0, nu0, nv0, nw0, npN, nvN, npS, nvS, npz2, nwz2);
        int Rn[3];
#pragma unroll
        for (int s = 0; s < 3; s++) Rn[s] = id_at(j + s - 1, k + 3);

        if (inner) {
            // ---- y-slice position (j, k): t1 = z (planes), t2 = x (lanes)
            auto raw_at = [&](int idx) -> Raw {
                const double* y = ybuf + idx * 4 * 64 + l;
                return {{y[0], y[64]}, {y[128], y[192]}};
            };
            const Raw y0 = raw_at(r - 1), y1 = raw_at(r), y2 = raw_at(r + 1), y3 = raw_at(r + 2);
            const Iface flo = from_raw(y1, mat_of(R(0, 1)), me);  // j-1/2
            const Iface fhi = from_raw(y2, me, mat_of(R(2, 1)));  // j+1/2
            const Wv cqlo = limited_cq<M, Cfg>(a, flo, waves_only(y0), waves_only(y2), dy_);
            const Wv cqhi = limited_cq<M, Cfg>(a, fhi, waves_only(y1), waves_only(y3), dy_);
            Mat N[3][3];
#pragma unroll
            for (int s = 0; s < 3; s++) {
                const Mat c = mat_of(R(1, s));
                N[s][1] = c;
                N[s][0] = mat_up(c);
                N[s][2] = mat_dn(c);
            }
            // target 3(b+1) + (a+1): plane offset a, lane offset b; n = v, t1 = w, t2 = u
            auto acc_y = [&](const T9& T) {
                A[1][V] += T.n;
#pragma unroll
                for (int e = 0; e < 3; e++) {
                    A[e][P] += T.p[3 + e] + lane_up(T.p[6 + e]) + lane_dn(T.p[e]);
                    A[e][W] += T.t1[3 + e] + lane_up(T.t1[6 + e]) + lane_dn(T.t1[e]);
                    A[e][U] += T.t2[3 + e] + lane_up(T.t2[6 + e]) + lane_dn(T.t2[e]);
                }
            };
            T9 T;
            if (uy) {
                homog_position<M, Cfg>(a, K0.pp, K0.pt, flo, cqlo, fhi, cqhi, dy_, dz_, dx_, T);
                acc_y(T);
            } else {
                slice_position<M, Cfg, KL>(a, flo, cqlo, fhi, cqhi, N, dy_, dz_, dx_, T, klut, a.nmat);
                acc_y(T);
            }
            // ---- what the rows below and above sent
            if (trans) {
                const double* fu = ubox + (r - 1) * kBox * 64 + l;
                const double* fd = dbox + (r + 1) * kBox * 64 + l;
                double g[kBox];
#pragma unroll
                for (int s = 0; s < kBox; s++) g[s] = fu[s * 64] + fd[s * 64];
                A[0][P] += g[0]; A[0][V] += g[1]; A[0][W] += g[2];
                A[1][P] += g[3]; A[1][U] += g[4]; A[1][V] += g[5]; A[1][W] += g[6];
                A[2][P] += g[7]; A[2][V] += g[8]; A[2][W] += g[9];
            }
            // ---- plane k-1 of this row is complete
            if (store_row && out_lane && k - 1 >= kb) {
                const long c = L.off(i, j, k - 1);
#pragma unroll
                for (int m = 0; m < 4; m++) a.qn[m * fs + c] = A[0][m];
            }
        }
#pragma unroll
        for (int m = 0; m < 4; m++) {
            A[0][m] = A[1][m];
            A[1][m] = A[2][m];
            A[2][m] = 0.0;
        }
#pragma unroll
        for (int s = 0; s < 3; s++) Rp[s] = (Rp[s] >> 8) | ((unsigned)Rn[s] << 24);
        p0 = np0; u0 = nu0; v0 = nv0; w0 = nw0; pN = npN; vN = nvN; pS = npS; vS = nvS;
        pz2 = npz2; wz2 = nwz2;
        __syncthreads();
    }
}

}  // namespace s3
}  // namespace clawgpu
