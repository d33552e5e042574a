// Device helpers shared by the quartet kernels (series.hip) and their
// adjoints (adjoint.hip): FP64 sin/cos and reciprocal, the closed-form a(l,p)
// coefficients of the j0-derivative expansion, and the launch shapes and
// Bessel regimes both passes must agree on.
#pragma once

#include <math.h>

#include "thistle_internal.hpp"

#define WAVE 64

#define XDPH SERIES_XDPH  // q h from which dphrec's start-100 recurrence is
                          // followed step by step (it departs from the exact
                          // j_beta near 76)

namespace {

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

__device__ __forceinline__ void kahan(double& s, double& c, double v) {
#pragma clang fp contract(off)
    double y = v - c;
    double t = s + y;
    c = (t - s) - y;
    s = t;
}

__device__ __forceinline__ double rcp64(double x) {
    // hardware reciprocal estimate refined by two Newton steps
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    r = fma(fma(-x, r, 1.0), r, r);
    return r;
}

// Cody-Waite reduction by pi/2 split in three parts (33, 33 and 53 bits),
// then Taylor polynomials on |r| <= pi/4.  Valid for 0 <= x < 2^20.
__device__ __forceinline__ void sincos_cw(double x, double& s, double& c) {
    const double C1 = 0x1.921fb54400000p+0, C2 = 0x1.0b4611a600000p-34,
                 C3 = 0x1.3198a2e037073p-69;
    double kf = rint(x * 0.6366197723675814);
    double r = fma(-kf, C1, x);
    r = fma(-kf, C2, r);
    r = fma(-kf, C3, r);
    double r2 = r * r;
    // sin r = r (1 - r^2/3! + ... - r^16/17!)
    double ps = -1.0 / 355687428096000.0;   // -1/17!
    ps = fma(ps, r2, 1.0 / 1307674368000.0);
    ps = fma(ps, r2, -1.0 / 6227020800.0);
    ps = fma(ps, r2, 1.0 / 39916800.0);
    ps = fma(ps, r2, -1.0 / 362880.0);
    ps = fma(ps, r2, 1.0 / 5040.0);
    ps = fma(ps, r2, -1.0 / 120.0);
    ps = fma(ps, r2, 1.0 / 6.0);
    double sr = fma(-r * r2, ps, r);
    // cos r = 1 - r^2/2! + ... + r^20/20!
    double pc = 1.0 / 2432902008176640000.0;  // 1/20!
    pc = fma(pc, r2, -1.0 / 6402373705728000.0);
    pc = fma(pc, r2, 1.0 / 20922789888000.0);
    pc = fma(pc, r2, -1.0 / 87178291200.0);
    pc = fma(pc, r2, 1.0 / 479001600.0);
    pc = fma(pc, r2, -1.0 / 3628800.0);
    pc = fma(pc, r2, 1.0 / 40320.0);
    pc = fma(pc, r2, -1.0 / 720.0);
    pc = fma(pc, r2, 1.0 / 24.0);
    pc = fma(pc, r2, -0.5);
    double cr = fma(pc, r2, 1.0);
    int k = (int)kf & 3;
    double ss = (k & 1) ? cr : sr;
    double cc = (k & 1) ? sr : cr;
    s = (k & 2) ? -ss : ss;
    c = ((k + 1) & 2) ? -cc : cc;
}

// FP32 polynomial after an FP64 reduction (mixed precision).
__device__ __forceinline__ void sincos_mixed(double x, float& s, float& c) {
    const double C1 = 0x1.921fb54400000p+0, C2 = 0x1.0b4611a600000p-34,
                 C3 = 0x1.3198a2e037073p-69;
    double kf = rint(x * 0.6366197723675814);
    double rd = fma(-kf, C1, x);
    rd = fma(-kf, C2, rd);
    rd = fma(-kf, C3, rd);
    float r = (float)rd, r2 = r * r;
    float ps = 1.0f / 362880.0f;
    ps = fmaf(ps, r2, -1.0f / 5040.0f);
    ps = fmaf(ps, r2, 1.0f / 120.0f);
    ps = fmaf(ps, r2, -1.0f / 6.0f);
    float sr = fmaf(r * r2, ps, r);
    float pc = -1.0f / 3628800.0f;
    pc = fmaf(pc, r2, 1.0f / 40320.0f);
    pc = fmaf(pc, r2, -1.0f / 720.0f);
    pc = fmaf(pc, r2, 1.0f / 24.0f);
    pc = fmaf(pc, r2, -0.5f);
    float cr = fmaf(pc, r2, 1.0f);
    int k = (int)kf & 3;
    float ss = (k & 1) ? cr : sr;
    float cc = (k & 1) ? sr : cr;
    s = (k & 2) ? -ss : ss;
    c = ((k + 1) & 2) ? -cc : cc;
}

// a(l,p) of rrdj0 in closed form: coefficient table built at compile time
struct ACoef {
    double v[17][17];
};
constexpr ACoef make_acoef() {
    ACoef t{};
    for (int l = 0; l <= 16; ++l)
        for (int p = 0; p <= l; ++p) {
            if ((l - p) & 1) continue;
            int k = (l - p) / 2;
            double c = 1.0;
            for (int i = 2; i <= l; ++i) c *= i;
            for (int i = 2; i <= k; ++i) c /= i;
            for (int i = 2; i <= p; ++i) c /= i;
            for (int i = 0; i < k; ++i) c *= 0.5;
            t.v[l][p] = (((l + p) / 2) & 1) ? -c : c;
        }
    return t;
}
__constant__ ACoef kAcoef = make_acoef();

}  // namespace

// G lanes per quartet: small classes (a few GTO quadruples and exponent
// tuples per quartet) share one wave between 64/G quartets; LL = 9 quartets
// take two waves and from LL = 10 on, with thousands of quadruples, four
// (XY6: LL 9 on one wave 7.82 ms against 7.50, on four 7.55; LL 8 on four
// 7.92; LL 10+ on two 8.11, on eight 8.10).  With the pipelined quadruple
// loop the small classes gained from narrower groups: 4 lanes up to LL = 2,
// 8 for LL 3-4, 16 for LL 5, 32 for LL 6, a wave for LL 7-8 (XY6 step 7.12
// against 7.35 ms with 4, 8, 16 and 32 lanes up to LL 1, 2, 3 and 6; LL 6
// on 16 lanes or LL 7 on 32: slower).
template <int LL>
struct ContractShape {
    static constexpr int CT = LL < 9 ? 64 : LL == 9 ? 128 : 256;
    static constexpr int G = LL <= 2 ? 4 : LL <= 4 ? 8 : LL == 5 ? 16 : LL == 6 ? 32 : CT;
    static constexpr int NG = CT / G;
    // row stride of the px/py/pz tables: odd, so that rows gathered by the
    // lanes of a wave spread over all LDS banks (an even stride of 8 doubles
    // maps every row onto two bank positions)
    static constexpr int DP = (LL + 1) | 1;
};

// The contraction adjoint's lanes per quartet (its tables are four times
// the forward's, so its groups stay wider).
template <int LL>
struct AdjointShape {
    static constexpr int CT = LL < 9 ? 64 : LL == 9 ? 128 : 256;
    static constexpr int G = LL <= 1 ? 4 : LL == 2 ? 8 : LL == 3 ? 16 : LL <= 6 ? 32 : CT;
    static constexpr int NG = CT / G;
};

// Upward recurrence from j0, j1 (f_up64's form, on u_b = j_b (q/h)^b) keeps
// every g_b = j_b / x^b (b <= LL) within 2e-14 of its envelope (|g_b| below
// the turning point x = sqrt(b (b + 1)), the Hankel modulus / x^b above it)
// for x >= xup = LL - 1 (0.5 at LL = 1, 1.2 at LL = 2).  Measured on 12000
// points per LL, the plain j recurrence stays within 1e-14 from x = LL - 1.5
// (LL >= 6; 0.2, 0.86, 1.68, 2.70, 3.69 for LL = 1..5), so xup leaves a
// margin of at least 0.3; a unit lower the error exceeds 2e-14.  Below xup
// Miller's backward recurrence is used.
//
// Miller's start.  Started M(LL) orders above LL from (0, 1) and
// normalised by j0 or j1, whichever is larger in magnitude (j0 alone loses
// up to 5e-12 of the envelope next to its zeros, x = pi, 2 pi, ...), the
// backward recurrence gives every g_b = j_b / x^b, b <= LL, within 4e-15 of
// its envelope (|g_b| below the turning point x = sqrt(b (b + 1)), the
// Hankel modulus above it) for x < xup, checked against a long-double
// recurrence started 80 orders up on 4000 points of [0, xup).  The least M
// that does so is 5, 7, 8, 9, 10, 11, 11, 12, 13, 13, 14, 14, 14, 15, 15, 16
// for LL = 1..16; M(LL) adds two orders to it, at most 16.  (Chosen per wave
// from the slice's largest x instead, the extra code copies cost more than
// the shorter recurrences save.)
constexpr int miller_orders(int LL) {
    return LL <= 1 ? 7 : LL == 2 ? 9 : LL == 3 ? 10 : LL == 4 ? 11 : LL == 5 ? 12
         : LL <= 7 ? 13 : LL == 8 ? 14 : LL <= 10 ? 15 : 16;
}
template <int LL>
struct SeriesRegime {
    static constexpr double xup = LL == 0 ? 0.0 : LL == 1 ? 0.5 : LL == 2 ? 1.2 : LL - 1.0;
    // the start for any x < xup
    static constexpr int nstart = LL + miller_orders(LL);
};

// Z of one quartet.  Total: the lower-triangle elements (max(xb,xk),
// min(xb,xk)) of bra element xb and ket element xk all lie in one block of
// the npp x npp matrix, so an element is base[row + ldz * col] with 32-bit
// block-local indices (npp < 2^24).  Elastic: the GTO starts of the four
// shells in the ngto x ngto one-particle Z.
struct ZView {
    const double* base;
    int ldz;
    int mode;            // total: 0 ket on rows, 1 bra on rows, 2 same pair
    int sa, sb, sc, sd;  // elastic
};

__device__ __forceinline__ ZView zview(const SeriesArgs& A, int pb, int pk, int64_t offb,
                                       int64_t offk, const GroupInfo& Gi, const GroupInfo& Gj,
                                       const GroupInfo& Gk, const GroupInfo& Gr) {
    ZView V;
    V.ldz = (int)A.ldz;
    V.sa = Gi.start;
    V.sb = Gj.start;
    V.sc = Gk.start;
    V.sd = Gr.start;
    if (A.elastic) {
        V.base = A.Z;
        V.mode = 0;
    } else if (pk > pb) {
        V.base = A.Z + offk + A.ldz * offb;
        V.mode = 0;
    } else if (pk < pb) {
        V.base = A.Z + offb + A.ldz * offk;
        V.mode = 1;
    } else {
        V.base = A.Z + offb + A.ldz * offb;
        V.mode = 2;
    }
    return V;
}

// ztot of one GTO quadruple from the quadruple-table entry w (pair-block
// offsets in bits 0-15, component indices in bits 46-61).
__device__ __forceinline__ double zget(const SeriesArgs& A, const ZView& V, uint64_t w) {
    if (A.elastic) {
        const int ia = V.sa + (int)((w >> 46) & 15u), ib = V.sb + (int)((w >> 50) & 15u),
                  ic = V.sc + (int)((w >> 54) & 15u), id = V.sd + (int)((w >> 58) & 15u);
        return V.base[ic + (int64_t)V.ldz * id] * V.base[ia + (int64_t)V.ldz * ib];
    }
    const int bl = (int)(w & 255u), kl = (int)((w >> 8) & 255u);
    const bool swap = V.mode == 1 || (V.mode == 2 && bl > kl);
    const int r = swap ? bl : kl, c = swap ? kl : bl;
    return V.base[(unsigned)r + __umul24((unsigned)V.ldz, (unsigned)c)];
}

