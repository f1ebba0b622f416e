// Learned flux limiter: features, BF16-MFMA inference, FP32 training kernels.
//
// A small network (5 -> 32 -> 32 -> 1, GELU) replaces phi(theta) of the
// wave limiter (src/2d/inlinelimiter.f90:71-97).  For every interface wave
// the pre-pass computes five features -- the upwind ratio theta (the
// reference's r), the downwind ratio, the signed Courant number s dt/dx,
// the wave strength |W| relative to the two states, and s relative to the
// fastest wave of the interface -- and the network's output is clipped to
// the Sweby TVD region 0 <= phi <= min(2 theta, 2) (phi = 0 for theta <= 0);
// wherever the clip would change the output the MC value is used instead.
// The resulting phi tables are read by the fused step kernel (limiter 6).
//
// Inference runs as batched BF16 GEMMs on the matrix cores with FP32
// accumulation: each wavefront takes 16 waves at a time and computes the
// layers transposed (hidden x waves), H1^T = W1^T X^T and H2^T = W2^T H1^T,
// with v_mfma_f32_16x16x16_bf16, so that the accumulator layout of one layer
// (lane l holds rows 4*(l/16)+r of column l%16) is already the B-operand
// layout of the next and no data leaves the registers between layers.
//
// Training (FP32): the forward/backward kernel evaluates the network per
// wave and reduces the parameter gradients per block through LDS (the
// dW2 = H1^T dZ2 outer-product sum is a small GEMM over the block's waves),
// then adds one partial per parameter per block to global memory.
#include <hip/hip_runtime.h>

#include <cmath>
#include <new>

#include "ctx.hpp"
#include "lanes.hpp"
#include "riemann2d.hpp"

namespace clawgpu {

constexpr int kNF = 8;  // features stored per wave (5 used)
constexpr int kH = 32;
constexpr int oW1 = 0, oB1 = 5 * kH, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH,
              oB3 = oW3 + kH, kNParams = oB3 + 1;  // 1281
constexpr float kThetaClip = 3.0f;

struct LearnedState {
    long ntab;        // interface slots per direction: (mx+4)*(my+4)
    long nw;          // waves per direction: ntab*mwaves
    float* feat;      // [2][nw][kNF]
    double* theta;    // [2][nw] upwind ratio
    float* g;         // [2][nw][meqn] gradient coefficients dt/dx * 0.5|s|(1-|s|dt/dx) W
    float* dphi;      // [2][nw] dL/dphi
    float* params;    // [kNParams]
    float* grad;      // [kNParams]
    double* target;   // q* in field planes
    double* lossacc;  // [1]
    bool have_params;
};

__device__ inline float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }

__device__ inline float gelu_d(float x)
{
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) +
           x * 0.3989422804014327f * expf(-0.5f * x * x);
}

__device__ inline double mc_phi(double th)
{
    double c = (1.0 + th) / 2.0;
    return fmax(0.0, fmin(c, fmin(2.0, 2.0 * th)));
}

// superbee, max(0, min(1, 2 th), min(2, th)): the upper edge of the
// second-order TVD region (src/2d/inlinelimiter.f90:78-79)
__device__ inline double superbee_phi(double th)
{
    return fmax(0.0, fmax(fmin(1.0, 2.0 * th), fmin(2.0, th)));
}

// TVD clip with MC fallback
__device__ inline double clip_phi(double raw, double th, bool* inside)
{
    double hi = th > 0.0 ? fmin(2.0 * th, 2.0) : 0.0;
    bool in = raw >= 0.0 && raw <= hi && th > 0.0;
    if (inside) *inside = in;
    return in ? raw : mc_phi(th);
}

struct FeatArgs {
    const double* q;
    Layout2 L;
    double dtdx, dtdy;
    SolverParams P;
    float* feat;
    double* theta;
    float* g;
    double* phix;  // MC fallback written here for slots without a wave
    double* phiy;
    long ntab;
};

// The five features of one wave from its interface's data: dl, dr the dot
// products of the wave with the same family's waves of the interfaces to
// the left and right along the direction, wn2 its squared norm, sp its
// speed, qn the sum of the two cells' state norms plus 1e-300, smax the
// interface's largest speed magnitude, dtd = dt/dx (or dt/dy).
template <class Mth>
__device__ inline void wave_features(double dl, double dr, double wn2, double sp, double qn,
                                     double smax, double dtd, float (&f)[5], double& th, bool& nz)
{
    nz = wn2 != 0.0;
    th = 0.0;
    for (int k = 0; k < 5; k++) f[k] = 0.0f;
    if (!nz) return;
    const double t = Mth::ratio(sp > 0.0 ? dl : dr, wn2);
    const double td = Mth::ratio(sp > 0.0 ? dr : dl, wn2);
    const double tc = fmin(fmax(t, -kThetaClip), kThetaClip);
    const double tdc = fmin(fmax(td, -kThetaClip), kThetaClip);
    if constexpr (Mth::fast) {
        // the float features to a few double ulps: products with
        // reciprocals instead of correctly rounded divisions
        constexpr double inv_clip = 1.0 / kThetaClip;
        f[0] = (float)(tc * inv_clip);
        f[1] = (float)(tdc * inv_clip);
        f[3] = (float)Mth::ratio(sqrt(wn2), qn);
        f[4] = (float)Mth::ratio(sp, smax + 1e-300);
    } else {
        f[0] = (float)(tc / kThetaClip);
        f[1] = (float)(tdc / kThetaClip);
        f[3] = (float)(sqrt(wn2) / qn);
        f[4] = (float)(sp / (smax + 1e-300));
    }
    f[2] = (float)(sp * dtd);
    th = t;
}

// Features of the MW waves of interface (i, J) along direction dir: the
// interface and its two neighbours along the direction are solved with the
// step kernel's own arithmetic.  f receives five features per wave, th the
// upwind ratio, nz whether the wave is nonzero; with g given, the gradient
// coefficients (MW x meqn floats) of the nonzero waves are written there.
template <class S>
__device__ inline void slot_features(const FeatArgs& a, int dir, int i, int J,
                                     float (&f)[S::MW][5], double (&th)[S::MW], bool (&nz)[S::MW],
                                     float* g)
{
    using Mth = typename S::Math;
    constexpr int M = S::M, MW = S::MW;
    const Layout2 L = a.L;
    // cells k-2 .. k+1 along the direction around interface k
    double qs[4][M];
    for (int c = 0; c < 4; c++) {
        int ii = dir == 0 ? i - 2 + c : i, jj = dir == 0 ? J : J - 2 + c;
        for (int m = 0; m < M; m++) qs[c][m] = a.q[m * L.fs + L.off(ii, jj)];
    }
    double w[3][MW][M], s[3][MW], am[M], ap[M];
    typename S::Cache cc;
    for (int k = 0; k < 3; k++) {
        typename S::Cell cl = S::cell(a.P, qs[k]), cr = S::cell(a.P, qs[k + 1]);
        if (dir == 0)
            S::template rpn<1, 2>(a.P, cl, cr, w[k], s[k], am, ap, cc);
        else
            S::template rpn<2, 1>(a.P, cl, cr, w[k], s[k], am, ap, cc);
    }
    const double dtd = dir == 0 ? a.dtdx : a.dtdy;
    double smax = 0.0, nl = 0.0, nr = 0.0;
    for (int mw = 0; mw < MW; mw++) smax = fmax(smax, fabs(s[1][mw]));
    for (int m = 0; m < M; m++) {
        nl += qs[1][m] * qs[1][m];
        nr += qs[2][m] * qs[2][m];
    }
    const double qn = sqrt(nl) + sqrt(nr) + 1e-300;
    for (int mw = 0; mw < MW; mw++) {
        double wn2 = 0.0, dl = 0.0, dr = 0.0;
        for (int m = 0; m < M; m++) {
            wn2 = wn2 + w[1][mw][m] * w[1][mw][m];
            dl = dl + w[0][mw][m] * w[1][mw][m];
            dr = dr + w[1][mw][m] * w[2][mw][m];
        }
        const double sp = s[1][mw];
        wave_features<Mth>(dl, dr, wn2, sp, qn, smax, dtd, f[mw], th[mw], nz[mw]);
        if (!nz[mw]) continue;
        if (g) {
            const double c = dtd * 0.5 * fabs(sp) * (1.0 - fabs(sp) * dtd);
            for (int m = 0; m < M; m++) g[mw * M + m] = (float)(c * w[1][mw][m]);
        }
    }
}

// Whether table slot (i, J) of direction dir holds an interface of the step.
__device__ inline bool slot_live(int dir, int i, int J, int mx, int my)
{
    return dir == 0 ? (i >= 1 && i <= mx + 1 && J >= 0 && J <= my + 1)
                    : (J >= 1 && J <= my + 1 && i >= 0 && i <= mx + 1);
}

// One thread per (direction, table slot): writes features, theta, gradient
// coefficients and (as a default) the MC value of phi of the slot's waves.
template <class S>
__global__ void features_kernel(FeatArgs a)
{
    constexpr int M = S::M, MW = S::MW;
    const int mx = a.L.mx, my = a.L.my;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * a.ntab) return;
    const int dir = (int)(t / a.ntab);
    const long slot = t % a.ntab;
    const int i = (int)(slot % (mx + 4)) - 1, J = (int)(slot / (mx + 4)) - 1;
    const long wbase = ((long)dir * a.ntab + slot) * MW;
    double* phit = dir == 0 ? a.phix : a.phiy;
    for (int mw = 0; mw < MW; mw++) {
        for (int k = 0; k < kNF; k++) a.feat[(wbase + mw) * kNF + k] = 0.0f;
        for (int m = 0; m < M; m++) a.g[(wbase + mw) * M + m] = 0.0f;
        a.theta[wbase + mw] = 0.0;
        phit[slot * MW + mw] = 0.0;
    }
    if (!slot_live(dir, i, J, mx, my)) return;
    float f[MW][5];
    double th[MW];
    bool nz[MW];
    slot_features<S>(a, dir, i, J, f, th, nz, a.g + wbase * M);
    for (int mw = 0; mw < MW; mw++) {
        if (!nz[mw]) continue;
        float* fo = a.feat + (wbase + mw) * kNF;
        for (int k = 0; k < 5; k++) fo[k] = f[mw][k];
        a.theta[wbase + mw] = th[mw];
        phit[slot * MW + mw] = mc_phi(th[mw]);
    }
}

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2h __attribute__((ext_vector_type(2)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

// Four floats to BF16 (round to nearest even) as an MFMA operand, two
// hardware packed conversions.
__device__ inline bf16x4 to_bf16x4(float a, float b, float c, float d)
{
    const f32x2 lo = {a, b}, hi = {c, d};
    const u32x2 r = {__builtin_bit_cast(unsigned, __builtin_convertvector(lo, bf16x2h)),
                     __builtin_bit_cast(unsigned, __builtin_convertvector(hi, bf16x2h))};
    return __builtin_bit_cast(bf16x4, r);
}

struct InferArgs {
    const float* feat;
    const double* theta;
    const float* params;
    double* phix;
    double* phiy;
    long nw;  // waves per direction
    long ntiles;
};

// GELU of the matrix-core network: the tanh form 0.5 x (1 + tanh(u)),
// u = sqrt(2/pi) (x + 0.044715 x^3), written as x / (1 + 2^(-2 u log2 e))
// with the hardware exp2 and reciprocal.  It differs from the erf form of the
// FP32 network by less than 5e-4, below the BF16 rounding of the inputs of
// every layer.  Large negative x gives exp2 = inf and the result -0.
__device__ inline float gelu_mfma(float x)
{
    const float k = 2.30220819f;  // 2 sqrt(2/pi) log2(e)
    const float z = x * fmaf(0.044715f * k, x * x, k);
    return x * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-z));
}

// The network's weights as MFMA fragments of one lane (grp = lane/16,
// col = lane%16): W1^T and W2^T in the A-operand layout (A[row=col][k =
// 4*grp+j]); the biases of the accumulator rows the lane holds (they
// initialise the accumulators) and the output weights of those rows.
struct MlpFrag {
    bf16x4 w1f[2], w2f[2][2];
    f32x4 b1v[2], b2v[2];
    float w3v[2][4], b3;
};

__device__ inline void mlp_load(MlpFrag& F, const float* P, int grp, int col)
{
    for (int t = 0; t < 2; t++) {
        float v[4];
        for (int j = 0; j < 4; j++) {
            int k = 4 * grp + j;  // input feature
            v[j] = k < 5 ? P[oW1 + k * kH + 16 * t + col] : 0.0f;
        }
        F.w1f[t] = to_bf16x4(v[0], v[1], v[2], v[3]);
        for (int kt = 0; kt < 2; kt++) {
            for (int j = 0; j < 4; j++) v[j] = P[oW2 + (16 * kt + 4 * grp + j) * kH + 16 * t + col];
            F.w2f[t][kt] = to_bf16x4(v[0], v[1], v[2], v[3]);
        }
        for (int r = 0; r < 4; r++) {
            int h = 16 * t + 4 * grp + r;
            F.b1v[t][r] = P[oB1 + h];
            F.b2v[t][r] = P[oB2 + h];
            F.w3v[t][r] = P[oW3 + h];
        }
    }
    F.b3 = P[oB3];
}

// Network output (before the clip) of a tile of 16 waves: x holds features
// 4*grp .. 4*grp+3 of wave col (zero for grp >= 2); every lane of column col
// returns wave col's output.
__device__ inline float mlp_tile(const MlpFrag& F, bf16x4 x)
{
    float h1[2][4];
    for (int t = 0; t < 2; t++) {
        f32x4 z = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(F.w1f[t], x, F.b1v[t], 0, 0, 0);
        for (int r = 0; r < 4; r++) h1[t][r] = gelu_mfma(z[r]);
    }
    bf16x4 hb[2];
    for (int kt = 0; kt < 2; kt++) hb[kt] = to_bf16x4(h1[kt][0], h1[kt][1], h1[kt][2], h1[kt][3]);
    float part = 0.0f;
    for (int t = 0; t < 2; t++) {
        f32x4 z = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(F.w2f[t][0], hb[0], F.b2v[t], 0, 0, 0);
        z = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(F.w2f[t][1], hb[1], z, 0, 0, 0);
        for (int r = 0; r < 4; r++) part += F.w3v[t][r] * gelu_mfma(z[r]);
    }
    part += __shfl_xor(part, 16);
    part += __shfl_xor(part, 32);
    return part + F.b3;
}

// BF16-MFMA inference over stored features, 16 waves per wavefront
// iteration; see the file head for the layout.  Lanes 0..15 write the 16
// results of a tile.
__global__ __launch_bounds__(256) void mlp_infer_mfma_kernel(InferArgs a)
{
    const int lane = threadIdx.x & 63;
    const int grp = lane >> 4, col = lane & 15;
    const long wave_id = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const long nwave = (long)gridDim.x * 4;
    MlpFrag F;
    mlp_load(F, a.params, grp, col);
    const long nall = 2 * a.nw;
    for (long tile = wave_id; tile < a.ntiles; tile += nwave) {
        const long n = tile * 16 + col;  // wave handled by this lane's column
        const bool nok = n < nall;
        float v[4];
        for (int j = 0; j < 4; j++) {
            int k = 4 * grp + j;
            v[j] = (nok && k < kNF) ? a.feat[n * kNF + k] : 0.0f;
        }
        const bf16x4 x = to_bf16x4(v[0], v[1], v[2], v[3]);
        const float out = mlp_tile(F, x);
        if (grp == 0 && nok) {
            const double th = a.theta[n];
            double* tab = n < a.nw ? a.phix : a.phiy;
            const long idx = n < a.nw ? n : n - a.nw;
            if (a.feat[n * kNF + 3] > 0.0f)  // nonzero wave; others keep the MC default
                tab[idx] = clip_phi((double)out, th, nullptr);
        }
    }
}

// The pre-pass of a step with limiter 6 in one kernel: features of every
// interface wave (as features_kernel) go through LDS straight into the MFMA
// network, and only the phi tables are written -- the same values as
// features_kernel followed by mlp_infer_mfma_kernel, without the feature,
// theta and gradient tables in HBM.
//
// Work: a wavefront takes a line of table slots along one direction (a
// table row for the x-interfaces, a table column for the y-interfaces; the
// same code, with the lanes stepping through rows instead of columns for y)
// and 61 consecutive slots of it.  Lane l holds the cell at position
// p = 61*chunk + l - 3 along the line and solves the interface between
// cells p-1 (from lane l-1) and p once; lanes 2..62 own the slots p, whose
// neighbouring interfaces come from lanes l-1 and l+1 by cross-lane shifts.
// So every interface is solved once (plus 3 halo lanes of 64), instead of
// three times in features_kernel.
//
// Only a wave that is nonzero, has a nonzero float strength feature and
// theta > 0 needs the network: for every other wave the clip's fallback
// gives phi = MC(theta) (which is 0 for theta <= 0) whatever the network
// says, and a zero wave keeps phi = 0.  Each wavefront compacts its waves
// that need the network into an LDS queue (ballot + prefix count) and runs
// only the queue through the MFMA tiles, 16 waves per tile; the rest are
// written directly.  On evolved Euler states a small fraction of the waves
// is nonzero (exactly constant regions give exact zero waves), so most of
// the network work is skipped; the tables are the same as without the queue.
constexpr int kLine = 61;  // slots per wavefront along its line

template <class S>
__global__ __launch_bounds__(64) void learned_line_kernel(FeatArgs a, const float* params,
                                                          int ntask)
{
    using Mth = typename S::Math;
    using Cell = typename S::Cell;
    constexpr int M = S::M, MW = S::MW, NS = 64 * MW;
    __shared__ bf16x4 xs[NS][2];
    __shared__ double qth[NS];
    __shared__ int qix[NS];  // lane * MW + mw of the queued wave
    const int lane = threadIdx.x;
    const int grp = lane >> 4, col = lane & 15;
    const Layout2 L = a.L;
    const int mx = L.mx, my = L.my, W = mx + 4;
    const int nchx = (mx + 4 + kLine - 1) / kLine, nchy = (my + 4 + kLine - 1) / kLine;
    const int nxt = (my + 4) * nchx;  // x tasks: table rows x chunks
    const SolverParams P = a.P;
    MlpFrag F;
    mlp_load(F, params, grp, col);
    // Blocks are dealt to the 8 XCDs round-robin; XCD x works through the
    // tasks [x*per, (x+1)*per) in order, so that the y lines of adjacent
    // table columns (tasks in a row, chunk-major) share its L2: their strided
    // loads and 24-byte table writes fall into the same cache lines.
    const int per = (ntask + 7) / 8;
    const int xcd = blockIdx.x & 7, nper = gridDim.x >> 3;
    for (int u = blockIdx.x >> 3; u < per; u += nper) {
        const int task = xcd * per + u;
        if (task >= ntask) break;  // uniform over the block
        const bool ydir = task >= nxt;
        const int t = ydir ? task - nxt : task;
        // x: line = table row J, row-major over chunks; y: line = table
        // column i, chunk-major over columns
        const int line = ydir ? t % (mx + 4) - 1 : t / nchx - 1;
        const int chunk = ydir ? t / (mx + 4) : t % nchx;
        const int n = ydir ? my : mx;
        const int p = kLine * chunk + lane - 3;
        const int pc = min(max(p, -1), n + 2);
        const long cell0 = ydir ? L.off(line, pc) : L.off(pc, line);
        double q[M];
#pragma unroll
        for (int m = 0; m < M; m++) q[m] = a.q[m * L.fs + cell0];
        const Cell ce = S::cell(P, q);
        double n2 = 0.0;
#pragma unroll
        for (int m = 0; m < M; m++) n2 = n2 + q[m] * q[m];
        const double cn = sqrt(n2);
        Cell cl;
        {
            const double* src = reinterpret_cast<const double*>(&ce);
            double* dst = reinterpret_cast<double*>(&cl);
#pragma unroll
            for (int k = 0; k < S::XSHIFT; k++) dst[k] = lane_up(src[k]);
        }
        double w[MW][M], s[MW], am[M], ap[M];
        typename S::Cache ch;
        if (ydir)
            S::template rpn<2, 1>(P, cl, ce, w, s, am, ap, ch);
        else
            S::template rpn<1, 2>(P, cl, ce, w, s, am, ap, ch);
        const double qn = (lane_up(cn) + cn) + 1e-300;
        double smax = 0.0;
#pragma unroll
        for (int mw = 0; mw < MW; mw++) smax = fmax(smax, fabs(s[mw]));
        const bool own = lane >= 2 && lane <= kLine + 1 && p <= n + 2;
        const bool live = own && (ydir ? (p >= 1 && p <= my + 1 && line >= 0 && line <= mx + 1)
                                       : (p >= 1 && p <= mx + 1 && line >= 0 && line <= my + 1));
        const int i = ydir ? line : p, J = ydir ? p : line;
        const long e0 = ((long)(J + 1) * W + (i + 1)) * MW;
        double* tab = ydir ? a.phiy : a.phix;
        const double dtd = ydir ? a.dtdy : a.dtdx;
        int count = 0;  // queued waves (uniform)
#pragma unroll
        for (int mw = 0; mw < MW; mw++) {
            double wn2 = 0.0, dl = 0.0;
#pragma unroll
            for (int m = 0; m < M; m++) {
                const double wl = lane_up(w[mw][m]);
                wn2 = wn2 + w[mw][m] * w[mw][m];
                dl = dl + wl * w[mw][m];
            }
            const double dr = lane_dn(dl);
            float f[5];
            double th;
            bool nz;
            wave_features<Mth>(dl, dr, wn2, s[mw], qn, smax, dtd, f, th, nz);
            const bool need = live && nz && f[3] > 0.0f && th > 0.0;
            const unsigned long long mask = __ballot(need);
            if (need) {
                const int pos = count + (int)__builtin_amdgcn_mbcnt_hi(
                                            (unsigned)(mask >> 32),
                                            __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
                xs[pos][0] = to_bf16x4(f[0], f[1], f[2], f[3]);
                xs[pos][1] = to_bf16x4(f[4], 0.0f, 0.0f, 0.0f);
                qth[pos] = th;
                qix[pos] = lane * MW + mw;
            } else if (own) {
                tab[e0 + mw] = (live && nz) ? mc_phi(th) : 0.0;
            }
            count += __popcll(mask);
        }
        if (count > 0) {
            __syncthreads();  // queue entries written by other lanes
            // slot of lane l is e(l) = e0(lane 0) + l * (MW or W*MW) along the line
            const long step = ydir ? (long)W * MW : MW;
            const long ebase = e0 - (long)lane * step;
            for (int T = 0; 16 * T < count; T++) {
                const int sidx = 16 * T + col;
                const bool ok = sidx < count;
                bf16x4 x = {0, 0, 0, 0};
                if (grp < 2 && ok) x = xs[sidx][grp];
                const float out = mlp_tile(F, x);
                if (grp == 0 && ok) {
                    const int e = qix[sidx];
                    tab[ebase + (long)(e / MW) * step + e % MW] =
                        clip_phi((double)out, qth[sidx], nullptr);
                }
            }
            __syncthreads();  // before the queue is written again
        }
    }
}

// FP32 per-wave forward pass (training path and reference for the MFMA
// kernel); returns the raw output and keeps the pre-activations.
__device__ inline float mlp_fwd_f32(const float* P, const float* f, float* z1, float* h1, float* z2,
                                    float* h2)
{
    for (int h = 0; h < kH; h++) {
        float acc = P[oB1 + h];
        for (int k = 0; k < 5; k++) acc += f[k] * P[oW1 + k * kH + h];
        z1[h] = acc;
        h1[h] = gelu(acc);
    }
    float out = P[oB3];
    for (int h = 0; h < kH; h++) {
        float acc = P[oB2 + h];
        for (int k = 0; k < kH; k++) acc += h1[k] * P[oW2 + k * kH + h];
        z2[h] = acc;
        h2[h] = gelu(acc);
        out += P[oW3 + h] * h2[h];
    }
    return out;
}

struct TrainArgs {
    const float* feat;
    const double* theta;
    const float* dphi;  // upstream dL/dphi (mode 1)
    const float* params;
    float* grad;
    double* lossacc;
    double* phix;  // mode 2: write phi tables from the FP32 forward
    double* phiy;
    long nw;
    int mode;      // 0: regress a classical phi (loss (out - phi)^2); 1: backprop dphi; 2: forward only
    int target;    // mode 0: 0 MC, 1 superbee
    float scale;   // 1 / number of samples (mode 0)
};

constexpr int kTB = 128;  // waves per training block

// FP32 forward (+ backward): thread per wave; the block's per-wave vectors go
// through LDS and every thread reduces a slice of the parameter gradient
// over the block's waves (dW2 is a 32x32 sum of outer products).
__global__ __launch_bounds__(kTB) void mlp_train_kernel(TrainArgs a)
{
    __shared__ float sP[kNParams];
    __shared__ float sF[kTB][5];
    __shared__ float sH1[kTB][kH];
    __shared__ float sD1[kTB][kH];
    __shared__ float sH2[kTB][kH];
    __shared__ float sD2[kTB][kH];
    __shared__ float sD3[kTB];
    __shared__ double sLoss[kTB];
    for (int k = threadIdx.x; k < kNParams; k += kTB) sP[k] = a.params[k];
    __syncthreads();
    const long nall = 2 * a.nw;
    const long n = (long)blockIdx.x * kTB + threadIdx.x;
    float f[5] = {0, 0, 0, 0, 0}, z1[kH], h1[kH], z2[kH], h2[kH];
    float d3 = 0.0f;
    double loss = 0.0;
    bool live = false;
    if (n < nall) {
        const float* fp = a.feat + n * kNF;
        live = fp[3] > 0.0f;  // nonzero wave
        for (int k = 0; k < 5; k++) f[k] = fp[k];
    }
    float out = mlp_fwd_f32(sP, f, z1, h1, z2, h2);
    if (live) {
        const double th = a.theta[n];
        if (a.mode == 0) {
            double e = (double)out - (a.target ? superbee_phi(th) : mc_phi(th));
            loss = e * e * a.scale;
            d3 = (float)(2.0 * e) * a.scale;
        } else if (a.mode == 1) {
            bool inside;
            clip_phi((double)out, th, &inside);
            d3 = inside ? a.dphi[n] : 0.0f;
        } else {
            double* tab = n < a.nw ? a.phix : a.phiy;
            const long idx = n < a.nw ? n : n - a.nw;
            tab[idx] = clip_phi((double)out, th, nullptr);
        }
    }
    if (a.mode == 2) return;
    // backward of this wave
    float dz2[kH], dz1[kH];
    for (int h = 0; h < kH; h++) dz2[h] = d3 * sP[oW3 + h] * gelu_d(z2[h]);
    for (int k = 0; k < kH; k++) {
        float acc = 0.0f;
        for (int h = 0; h < kH; h++) acc += sP[oW2 + k * kH + h] * dz2[h];
        dz1[k] = acc * gelu_d(z1[k]);
    }
    const int tid = threadIdx.x;
    for (int k = 0; k < 5; k++) sF[tid][k] = live ? f[k] : 0.0f;
    for (int h = 0; h < kH; h++) {
        sH1[tid][h] = h1[h];
        sD1[tid][h] = live ? dz1[h] : 0.0f;
        sH2[tid][h] = h2[h];
        sD2[tid][h] = live ? dz2[h] : 0.0f;
    }
    sD3[tid] = live ? d3 : 0.0f;
    sLoss[tid] = loss;
    __syncthreads();
    for (int p = tid; p < kNParams; p += kTB) {
        float acc = 0.0f;
        if (p < oB1) {  // W1[k][h]
            int k = p / kH, h = p % kH;
            for (int b = 0; b < kTB; b++) acc += sF[b][k] * sD1[b][h];
        } else if (p < oW2) {
            int h = p - oB1;
            for (int b = 0; b < kTB; b++) acc += sD1[b][h];
        } else if (p < oB2) {  // W2[k][h]
            int k = (p - oW2) / kH, h = (p - oW2) % kH;
            for (int b = 0; b < kTB; b++) acc += sH1[b][k] * sD2[b][h];
        } else if (p < oW3) {
            int h = p - oB2;
            for (int b = 0; b < kTB; b++) acc += sD2[b][h];
        } else if (p < oB3) {
            int h = p - oW3;
            for (int b = 0; b < kTB; b++) acc += sH2[b][h] * sD3[b];
        } else {
            for (int b = 0; b < kTB; b++) acc += sD3[b];
        }
        if (acc != 0.0f) atomicAdd(&a.grad[p], acc);
    }
    if (tid == 0) {
        double l = 0.0;
        for (int b = 0; b < kTB; b++) l += sLoss[b];
        if (l != 0.0) atomicAdd(a.lossacc, l);
    }
}

struct DphiArgs {
    const double* qnew;
    const double* target;
    Layout2 L;
    int meqn;
    int mwaves;
    const float* g;
    float* dphi;
    long ntab;
    double* lossacc;
    double scale;  // 1 / (mx * my)
    int l1;        // 0: squared error per cell; 1: absolute error per cell and field
};

// dL/dphi of the one-step loss L = scale * sum over interior cells of
// |qnew - q*|^2 (or, with l1, of sum_m |qnew_m - q*_m|), through the normal
// correction flux: a wave of x-interface i adds +g to cell (i, J) and -g to
// cell (i-1, J) per unit phi.
__global__ void dphi_kernel(DphiArgs a)
{
    const Layout2 L = a.L;
    const int mx = L.mx, my = L.my, M = a.meqn, MW = a.mwaves;
    const long t0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool on = t0 < 2 * a.ntab;  // every lane stays for the loss reduction below
    const long t = on ? t0 : 0;
    const int dir = (int)(t / a.ntab);
    const long slot = t % a.ntab;
    const int i = (int)(slot % (mx + 4)) - 1, J = (int)(slot / (mx + 4)) - 1;
    const int ir = i, jr = J;
    const int il = dir == 0 ? i - 1 : i, jl = dir == 0 ? J : J - 1;
    auto res = [&](int ii, int jj, int m) -> double {
        if (ii < 1 || ii > mx || jj < 1 || jj > my) return 0.0;
        long o = m * L.fs + L.off(ii, jj);
        return a.qnew[o] - a.target[o];
    };
    for (int mw = 0; on && mw < MW; mw++) {
        const long w = ((long)dir * a.ntab + slot) * MW + mw;
        double d = 0.0;
        if (a.l1) {
            auto sgn = [](double e) { return e > 0.0 ? 1.0 : (e < 0.0 ? -1.0 : 0.0); };
            for (int m = 0; m < M; m++)
                d += (double)a.g[w * M + m] * (sgn(res(ir, jr, m)) - sgn(res(il, jl, m)));
            a.dphi[w] = (float)(a.scale * d);
        } else {
            for (int m = 0; m < M; m++)
                d += (double)a.g[w * M + m] * (res(ir, jr, m) - res(il, jl, m));
            a.dphi[w] = (float)(2.0 * a.scale * d);
        }
    }
    // loss: one term per interior cell (dir 0 slots with i, J interior),
    // summed over the wave and added with one atomic per wave (one per cell
    // on one address would serialise at about 11 ns each)
    double l = 0.0;
    if (on && dir == 0 && i >= 1 && i <= mx && J >= 1 && J <= my) {
        for (int m = 0; m < M; m++) {
            double e = res(i, J, m);
            l += a.l1 ? fabs(e) : e * e;
        }
        l = l * a.scale;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) l += __shfl_xor(l, d);
    if ((threadIdx.x & 63) == 0 && l != 0.0) atomicAdd(a.lossacc, l);
}

}  // namespace clawgpu

using namespace clawgpu;

static LearnedState* learned_state(clawgpu_ctx* c)
{
    if (c->learned) return (LearnedState*)c->learned;
    LearnedState* s = new (std::nothrow) LearnedState();
    if (!s) return nullptr;
    s->ntab = (long)(c->mx + 4) * (c->my + 4);
    s->nw = s->ntab * c->mwaves;
    size_t tabb = (size_t)s->nw * sizeof(double);
    size_t qbytes = (size_t)c->L.fs * c->meqn * sizeof(double);
    bool ok = hipMalloc((void**)&c->phix, tabb) == hipSuccess &&
              hipMalloc((void**)&c->phiy, tabb) == hipSuccess &&
              hipMalloc((void**)&s->feat, 2 * s->nw * kNF * sizeof(float)) == hipSuccess &&
              hipMalloc((void**)&s->theta, 2 * s->nw * sizeof(double)) == hipSuccess &&
              hipMalloc((void**)&s->g, 2 * s->nw * c->meqn * sizeof(float)) == hipSuccess &&
              hipMalloc((void**)&s->dphi, 2 * s->nw * sizeof(float)) == hipSuccess &&
              hipMalloc((void**)&s->params, kNParams * sizeof(float)) == hipSuccess &&
              hipMalloc((void**)&s->grad, kNParams * sizeof(float)) == hipSuccess &&
              hipMalloc((void**)&s->target, qbytes) == hipSuccess &&
              hipMalloc((void**)&s->lossacc, sizeof(double)) == hipSuccess;
    if (!ok) {
        delete s;
        return nullptr;
    }
    s->have_params = false;
    c->learned = s;
    return s;
}

// The context's learned-limiter state, or an error: the features and the
// network are defined for the built-in acoustics and Euler solvers only.
static int learned_get(clawgpu_ctx* c, LearnedState** out)
{
    if (c->solver != CLAWGPU_SOLVER_ACOUSTICS && c->solver != CLAWGPU_SOLVER_EULER)
        return set_err(CLAWGPU_ERR_UNSUPPORTED, "the learned limiter runs with the built-in solvers only");
    *out = learned_state(c);
    if (!*out) return set_err(CLAWGPU_ERR_HIP, "learned limiter allocation failed");
    return CLAWGPU_OK;
}

void learned_release(clawgpu_ctx* c)
{
    LearnedState* s = (LearnedState*)c->learned;
    if (c->phix) (void)hipFree(c->phix);
    if (c->phiy) (void)hipFree(c->phiy);
    c->phix = c->phiy = nullptr;
    if (!s) return;
    (void)hipFree(s->feat);
    (void)hipFree(s->theta);
    (void)hipFree(s->g);
    (void)hipFree(s->dphi);
    (void)hipFree(s->params);
    (void)hipFree(s->grad);
    (void)hipFree(s->target);
    (void)hipFree(s->lossacc);
    delete s;
    c->learned = nullptr;
}

static int enqueue_features(clawgpu_ctx* c, LearnedState* s, double dx, double dy, double dt)
{
    FeatArgs a;
    a.q = c->dq[c->cur];
    a.L = c->L;
    a.dtdx = dt / dx;
    a.dtdy = dt / dy;
    a.P = c->P;
    a.feat = s->feat;
    a.theta = s->theta;
    a.g = s->g;
    a.phix = c->phix;
    a.phiy = c->phiy;
    a.ntab = s->ntab;
    int nb = (int)((2 * s->ntab + 127) / 128);
    // the same wave arithmetic as the step kernel, so that theta (and the MC
    // default of phi) are the step kernel's own values
    if (c->solver == CLAWGPU_SOLVER_ACOUSTICS) {
        if (c->arith == ARITH_FAST)
            hipLaunchKernelGGL(features_kernel<Acoustics2<FastMath>>, dim3(nb), dim3(128), 0, c->stream, a);
        else
            hipLaunchKernelGGL(features_kernel<Acoustics2<ExactMath>>, dim3(nb), dim3(128), 0, c->stream, a);
    } else {
        if (c->arith == ARITH_FAST)
            hipLaunchKernelGGL(features_kernel<Euler2<FastMath>>, dim3(nb), dim3(128), 0, c->stream, a);
        else
            hipLaunchKernelGGL(features_kernel<Euler2<ExactMath>>, dim3(nb), dim3(128), 0, c->stream, a);
    }
    HIPCHK(c, hipGetLastError());
    return CLAWGPU_OK;
}

static int enqueue_phi(clawgpu_ctx* c, LearnedState* s, int mode)
{
    if (mode == 2) return CLAWGPU_OK;  // MC: already written by the feature pass
    if (!s->have_params) return set_err(CLAWGPU_ERR_ARG, "learned limiter parameters not set");
    if (mode == 0) {
        InferArgs a;
        a.feat = s->feat;
        a.theta = s->theta;
        a.params = s->params;
        a.phix = c->phix;
        a.phiy = c->phiy;
        a.nw = s->nw;
        a.ntiles = (2 * s->nw + 15) / 16;
        long nb = (a.ntiles + 3) / 4;
        if (nb > 4096) nb = 4096;
        hipLaunchKernelGGL(mlp_infer_mfma_kernel, dim3((int)nb), dim3(256), 0, c->stream, a);
    } else {
        TrainArgs a{};
        a.feat = s->feat;
        a.theta = s->theta;
        a.params = s->params;
        a.grad = s->grad;
        a.lossacc = s->lossacc;
        a.phix = c->phix;
        a.phiy = c->phiy;
        a.nw = s->nw;
        a.mode = 2;
        int nb = (int)((2 * s->nw + kTB - 1) / kTB);
        hipLaunchKernelGGL(mlp_train_kernel, dim3(nb), dim3(kTB), 0, c->stream, a);
    }
    HIPCHK(c, hipGetLastError());
    return CLAWGPU_OK;
}

// Called by the claw2 loop before each step attempt when every wave family
// uses limiter 6: features of the current q at this dt, then phi by the
// BF16 MFMA network with the TVD clip / MC fallback.
// Features and network in one pass (learned_fused_kernel): the phi tables
// of the step, as enqueue_features followed by enqueue_phi(mode 0).
static int enqueue_fused(clawgpu_ctx* c, LearnedState* s, double dx, double dy, double dt)
{
    if (!s->have_params) return set_err(CLAWGPU_ERR_ARG, "learned limiter parameters not set");
    FeatArgs a{};
    a.q = c->dq[c->cur];
    a.L = c->L;
    a.dtdx = dt / dx;
    a.dtdy = dt / dy;
    a.P = c->P;
    a.phix = c->phix;
    a.phiy = c->phiy;
    a.ntab = s->ntab;
    const long ntask = (long)(c->my + 4) * ((c->mx + 4 + kLine - 1) / kLine) +
                       (long)(c->mx + 4) * ((c->my + 4 + kLine - 1) / kLine);
    if (ntask > 0x7fffffffL) return set_err(CLAWGPU_ERR_ARG, "grid too large");
    // one wavefront per block, a multiple of 8 blocks (one share per XCD);
    // a grid-stride loop over the lines beyond 16384 blocks (64 per CU)
    const int nb = (int)(ntask < 16384 ? (ntask + 7) / 8 * 8 : 16384);
    const dim3 g(nb), b(64);
    const float* p = s->params;
    const int nt = (int)ntask;
    if (c->solver == CLAWGPU_SOLVER_ACOUSTICS) {
        if (c->arith == ARITH_FAST)
            hipLaunchKernelGGL(learned_line_kernel<Acoustics2<FastMath>>, g, b, 0, c->stream, a, p, nt);
        else
            hipLaunchKernelGGL(learned_line_kernel<Acoustics2<ExactMath>>, g, b, 0, c->stream, a, p, nt);
    } else {
        if (c->arith == ARITH_FAST)
            hipLaunchKernelGGL(learned_line_kernel<Euler2<FastMath>>, g, b, 0, c->stream, a, p, nt);
        else
            hipLaunchKernelGGL(learned_line_kernel<Euler2<ExactMath>>, g, b, 0, c->stream, a, p, nt);
    }
    HIPCHK(c, hipGetLastError());
    return CLAWGPU_OK;
}

int learned_prepare(clawgpu_ctx* c, double dx, double dy, double dt)
{
    LearnedState* s = nullptr;
    if (const int rc_ = learned_get(c, &s)) return rc_;
    return enqueue_fused(c, s, dx, dy, dt);
}

extern "C" int clawgpu_learned_set_params(clawgpu_ctx* c, const float* params, int n)
{
    if (!c || !params || n != kNParams) return set_err(CLAWGPU_ERR_ARG, "expected 1281 parameters");
    LearnedState* s = nullptr;
    if (const int rc_ = learned_get(c, &s)) return rc_;
    HIPCHK(c, hipMemcpy(s->params, params, kNParams * sizeof(float), hipMemcpyHostToDevice));
    s->have_params = true;
    return CLAWGPU_OK;
}

extern "C" int clawgpu_learned_features(clawgpu_ctx* c, double dx, double dy, double dt)
{
    if (!c) return set_err(CLAWGPU_ERR_ARG, "null context");
    LearnedState* s = nullptr;
    if (const int rc_ = learned_get(c, &s)) return rc_;
    int rc = enqueue_features(c, s, dx, dy, dt);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLAWGPU_OK;
}

extern "C" int clawgpu_learned_phi(clawgpu_ctx* c, int mode)
{
    if (!c || !c->learned) return set_err(CLAWGPU_ERR_ARG, "features not computed");
    if (mode < 0 || mode > 2) return set_err(CLAWGPU_ERR_ARG, "mode must be 0, 1 or 2");
    int rc = enqueue_phi(c, (LearnedState*)c->learned, mode);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLAWGPU_OK;
}

// Measurement helper: reps fused pre-passes (the kernel a step with limiter
// 6 runs), then one feature pass and reps passes of the stand-alone
// BF16-MFMA inference over its tables, on the current q, timed with HIP
// events on the context's stream; averages per launch in milliseconds.
extern "C" int clawgpu_learned_time(clawgpu_ctx* c, double dx, double dy, double dt, int reps,
                                    float* fused_ms, float* infer_ms)
{
    if (!c || reps < 1 || !fused_ms || !infer_ms) return set_err(CLAWGPU_ERR_ARG, "bad argument");
    LearnedState* s = nullptr;
    if (const int rc_ = learned_get(c, &s)) return rc_;
    if (!s->have_params) return set_err(CLAWGPU_ERR_ARG, "learned limiter parameters not set");
    hipEvent_t ev[4];
    for (int k = 0; k < 4; k++) HIPCHK(c, hipEventCreate(&ev[k]));
    int rc = CLAWGPU_OK;
    HIPCHK(c, hipEventRecord(ev[0], c->stream));
    for (int r = 0; r < reps && !rc; r++) rc = enqueue_fused(c, s, dx, dy, dt);
    HIPCHK(c, hipEventRecord(ev[1], c->stream));
    if (!rc) rc = enqueue_features(c, s, dx, dy, dt);
    HIPCHK(c, hipEventRecord(ev[2], c->stream));
    for (int r = 0; r < reps && !rc; r++) rc = enqueue_phi(c, s, 0);
    HIPCHK(c, hipEventRecord(ev[3], c->stream));
    HIPCHK(c, hipEventSynchronize(ev[3]));
    float a = 0.0f, b = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&a, ev[0], ev[1]));
    HIPCHK(c, hipEventElapsedTime(&b, ev[2], ev[3]));
    for (int k = 0; k < 4; k++) (void)hipEventDestroy(ev[k]);
    if (rc) return rc;
    *fused_ms = a / reps;
    *infer_ms = b / reps;
    return CLAWGPU_OK;
}

extern "C" int clawgpu_learned_read(clawgpu_ctx* c, int what, void* out, long n)
{
    if (!c || !c->learned || !out) return set_err(CLAWGPU_ERR_ARG, "bad argument");
    LearnedState* s = (LearnedState*)c->learned;
    const void* src = nullptr;
    size_t bytes = 0;
    switch (what) {
    case 0: src = s->feat; bytes = 2 * s->nw * kNF * sizeof(float); break;
    case 1: src = c->phix; bytes = s->nw * sizeof(double); break;
    case 2: src = c->phiy; bytes = s->nw * sizeof(double); break;
    case 3: src = s->theta; bytes = 2 * s->nw * sizeof(double); break;
    case 4: src = s->g; bytes = 2 * s->nw * c->meqn * sizeof(float); break;
    case 5: src = s->dphi; bytes = 2 * s->nw * sizeof(float); break;
    default: return set_err(CLAWGPU_ERR_ARG, "unknown table");
    }
    if ((size_t)n < bytes) return set_err(CLAWGPU_ERR_ARG, "output buffer too small");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(out, src, bytes, hipMemcpyDeviceToHost));
    return CLAWGPU_OK;
}

extern "C" int clawgpu_learned_write(clawgpu_ctx* c, int what, const void* in, long n)
{
    // overwrite a table (features 0, phi 1 / 2, theta 3 or dphi 5) from the
    // host: for the gradient checks of the training kernels and for tests
    // that a pass writes every entry
    if (!c || !in) return set_err(CLAWGPU_ERR_ARG, "bad argument");
    LearnedState* s = nullptr;
    if (const int rc_ = learned_get(c, &s)) return rc_;
    void* dst = nullptr;
    size_t bytes = 0;
    switch (what) {
    case 0: dst = s->feat; bytes = 2 * s->nw * kNF * sizeof(float); break;
    case 1: dst = c->phix; bytes = s->nw * sizeof(double); break;
    case 2: dst = c->phiy; bytes = s->nw * sizeof(double); break;
    case 3: dst = s->theta; bytes = 2 * s->nw * sizeof(double); break;
    case 5: dst = s->dphi; bytes = 2 * s->nw * sizeof(float); break;
    default: return set_err(CLAWGPU_ERR_ARG, "table not writable");
    }
    if ((size_t)n != bytes) return set_err(CLAWGPU_ERR_ARG, "size mismatch");
    HIPCHK(c, hipMemcpy(dst, in, bytes, hipMemcpyHostToDevice));
    return CLAWGPU_OK;
}

extern "C" int clawgpu_learned_table_size(clawgpu_ctx* c, long* waves_per_direction)
{
    if (!c || !waves_per_direction) return set_err(CLAWGPU_ERR_ARG, "bad argument");
    *waves_per_direction = (long)(c->mx + 4) * (c->my + 4) * c->mwaves;
    return CLAWGPU_OK;
}

extern "C" int clawgpu_learned_set_target(clawgpu_ctx* c, const double* q)
{
    if (!c || !q) return set_err(CLAWGPU_ERR_ARG, "bad argument");
    LearnedState* s = nullptr;
    if (const int rc_ = learned_get(c, &s)) return rc_;
    // reuse the upload path: stage into the current buffer's spare copy
    int keep = c->cur;
    c->cur = (keep + 2) % 3;
    int rc = clawgpu_set_q(c, q);
    c->cur = keep;
    if (rc) return rc;
    HIPCHK(c, hipMemcpy(s->target, c->dq[(keep + 2) % 3], (size_t)c->L.fs * c->meqn * sizeof(double),
                        hipMemcpyDeviceToDevice));
    return CLAWGPU_OK;
}

bool learned_ready(clawgpu_ctx* c)
{
    return c->learned && ((LearnedState*)c->learned)->have_params;
}

int enqueue_step_public(clawgpu_ctx* c, double dx, double dy, double dt, const int* method,
                        const int* mthlim, int use_fwave);
int enqueue_bc_public(clawgpu_ctx* c, const int* mthbc);

// Gradient of a training loss with respect to the network parameters.
//   mode 0: regression onto MC, L = sum over live waves of (out - phi_MC)^2
//   mode 2: as mode 0 with superbee in place of MC
//   mode 1: one-step loss of the unsplit step (method, mthlim ignored: all
//           waves use the network through the FP32 forward, clipped) from the
//           device q against the target set by clawgpu_learned_set_target,
//           L = mean over interior cells of |q_new - q*|^2; the gradient
//           flows through the normal correction fluxes.
//   mode 3: as mode 1 with L = mean over interior cells of sum_m |q_new - q*|
// grad (host, 1281 floats) and loss receive the results; q is unchanged.
extern "C" int clawgpu_learned_grad(clawgpu_ctx* c, double dx, double dy, double dt,
                                    const int* method, const int* mthbc, int mode, float* grad,
                                    double* loss)
{
    if (!c || !grad || !loss || mode < 0 || mode > 3 || !method || !mthbc)
        return set_err(CLAWGPU_ERR_ARG, "bad argument");
    LearnedState* s = nullptr;
    if (const int rc_ = learned_get(c, &s)) return rc_;
    if (!s->have_params) return set_err(CLAWGPU_ERR_ARG, "parameters not set");
    HIPCHK(c, hipMemsetAsync(s->grad, 0, kNParams * sizeof(float), c->stream));
    HIPCHK(c, hipMemsetAsync(s->lossacc, 0, sizeof(double), c->stream));
    int rc;
    if (mode == 1 || mode == 3) {
        rc = enqueue_bc_public(c, mthbc);
        if (rc) return rc;
        rc = enqueue_features(c, s, dx, dy, dt);
        if (rc) return rc;
        rc = enqueue_phi(c, s, 1);  // FP32 forward, clipped
        if (rc) return rc;
        int lim[kMaxWaves];
        for (int k = 0; k < kMaxWaves; k++) lim[k] = kLimTable;
        int keep = c->cur;
        rc = enqueue_step_public(c, dx, dy, dt, method, lim, 0);
        if (rc) return rc;
        int res = (keep + 1) % 3;
        c->cur = keep;  // q is not advanced
        DphiArgs d;
        d.qnew = c->dq[res];
        d.target = s->target;
        d.L = c->L;
        d.meqn = c->meqn;
        d.mwaves = c->mwaves;
        d.g = s->g;
        d.dphi = s->dphi;
        d.ntab = s->ntab;
        d.lossacc = s->lossacc;
        d.scale = 1.0 / ((double)c->mx * c->my);
        d.l1 = mode == 3;
        int nb = (int)((2 * s->ntab + 127) / 128);
        hipLaunchKernelGGL(dphi_kernel, dim3(nb), dim3(128), 0, c->stream, d);
        HIPCHK(c, hipGetLastError());
    }
    TrainArgs a{};
    a.feat = s->feat;
    a.theta = s->theta;
    a.dphi = s->dphi;
    a.params = s->params;
    a.grad = s->grad;
    a.lossacc = s->lossacc;
    a.nw = s->nw;
    a.mode = mode == 2 ? 0 : (mode == 3 ? 1 : mode);
    a.target = mode == 2 ? 1 : 0;
    if (mode != 1) {
        // mean over live waves: count them on the host from theta/features
        a.scale = 1.0f;
    }
    int nb = (int)((2 * s->nw + kTB - 1) / kTB);
    hipLaunchKernelGGL(mlp_train_kernel, dim3(nb), dim3(kTB), 0, c->stream, a);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(grad, s->grad, kNParams * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(loss, s->lossacc, sizeof(double), hipMemcpyDeviceToHost));
    return CLAWGPU_OK;
}
