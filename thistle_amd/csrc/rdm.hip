// createtwordm on the GPU: maxcoincidence, createtwordm's accumulation and its
// selection loop (src/twordms.f90:8-556), with the reference's floating-point
// result.
//
// The reference adds every contribution of every (c1, c2) determinant pair
// into a dense twordm, in (c1, c2, statement) order; the values, and so the
// rows that clear the cutoff, depend on that order.  The device path keeps it:
//   1. connected pairs (ndiff <= 4) per c1, c2 increasing: popcount of the
//      XOR of the determinants' slot masks, one pass to count and one to write
//      at the scanned offsets;
//   2. each pair's number of contributions, scanned into global positions;
//   3. every contribution (element, value) written at its position, i.e. in
//      the reference's global order;
//   4. a stable radix sort by element keeps that order inside each element;
//   5. one thread per element adds its contributions in that order, from 0;
//   6. the selection loop (p, q, r, s) as a flag scan and an ordered emit.
//
// maxc_individual (src/twordms.f90:83-172) aligns c2's occupied slots onto
// c1's by swaps, one sign flip per swap.  With the slots unique to c1, a
// (ascending), and to c2, b (ascending), and C the common slots:
//   - ndiff 2: d2 = b at a's place, sign (-1)^#C strictly between a and b;
//   - ndiff 4: b goes to a's places in ascending order, except when all of a
//     lie above (or below) all of b with an odd number of C between the two
//     groups, where the two b are exchanged; the sign is
//     (-1)^(#C between a0 and its b + #C between a1 and its b + exchanged).
// (Both follow from the sign being the parity of the permutation the swaps
// build; checked against the swap procedure itself on random slot sets.)
//
// Slots are 0-based bits j: orbital j/2 + 1, spin (j + 1) % 2 (the reference's
// 1-based slot modulo 2).  Determinants must be canonical (value 1 on even
// slots, 2 on odd ones, as the punch reader writes them).
#include "thistle_internal.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>

namespace {

constexpr int RW = 64;  // wave

__device__ __forceinline__ uint64_t below(int k) {
    return k <= 0 ? 0ull : (k >= 64 ? ~0ull : ((1ull << k) - 1ull));
}

template <int NW>
struct DMask {
    uint64_t w[NW];
};

template <int NW>
__device__ __forceinline__ DMask<NW> load_mask(const uint64_t* __restrict__ m, int c) {
    DMask<NW> d;
#pragma unroll
    for (int w = 0; w < NW; ++w) d.w[w] = m[(size_t)c * NW + w];
    return d;
}

// set bits of m strictly between slots x and y
template <int NW>
__device__ __forceinline__ int between(const DMask<NW>& m, int x, int y) {
    const int lo = (x < y ? x : y) + 1, hi = x < y ? y : x;
    int c = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) c += __popcll(m.w[w] & below(hi - 64 * w) & ~below(lo - 64 * w));
    return c;
}

// the k-th (0-based) set bit of m
template <int NW>
__device__ __forceinline__ int nth_bit(const DMask<NW>& m, int k) {
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        uint64_t v = m.w[w];
        const int c = __popcll(v);
        if (k < c) {
            for (int i = 0; i < k; ++i) v &= v - 1;
            return 64 * w + __builtin_ctzll(v);
        }
        k -= c;
    }
    return -1;
}

__device__ __forceinline__ int orb_of(int j) { return j / 2 + 1; }
__device__ __forceinline__ int spin_of(int j) { return (j + 1) & 1; }

// slots of spin sp: sp = 1 on even bits, 0 on odd bits
__device__ __forceinline__ uint64_t spin_bits(int sp) {
    return sp ? 0x5555555555555555ull : 0xAAAAAAAAAAAAAAAAull;
}

struct Exc {
    int e;
    int d1[2], d2[2], s1[2], s2[2];
};

template <int NW>
__device__ Exc excitation(const DMask<NW>& m1, const DMask<NW>& m2, int nd) {
    DMask<NW> A, B, C;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        A.w[w] = m1.w[w] & ~m2.w[w];
        B.w[w] = m2.w[w] & ~m1.w[w];
        C.w[w] = m1.w[w] & m2.w[w];
    }
    Exc x{};
    if (nd == 2) {
        const int a = nth_bit(A, 0), b = nth_bit(B, 0);
        x.e = (between(C, a, b) & 1) ? -1 : 1;
        x.d1[0] = orb_of(a);
        x.d2[0] = orb_of(b);
        x.s1[0] = spin_of(a);
        x.s2[0] = spin_of(b);
        return x;
    }
    const int a0 = nth_bit(A, 0), a1 = nth_bit(A, 1), b0 = nth_bit(B, 0), b1 = nth_bit(B, 1);
    int sw = 0;
    if (a0 > b1) sw = between(C, b1, a0) & 1;
    else if (a1 < b0) sw = between(C, a1, b0) & 1;
    const int c0 = sw ? b1 : b0, c1 = sw ? b0 : b1;
    x.e = ((between(C, a0, c0) + between(C, a1, c1) + sw) & 1) ? -1 : 1;
    x.d1[0] = orb_of(a0);
    x.d1[1] = orb_of(a1);
    x.d2[0] = orb_of(c0);
    x.d2[1] = orb_of(c1);
    x.s1[0] = spin_of(a0);
    x.s1[1] = spin_of(a1);
    x.s2[0] = spin_of(c0);
    x.s2[1] = spin_of(c1);
    return x;
}

__global__ void k_rdm_pack(const int32_t* __restrict__ confs, int nconf, int lconfs, int nw,
                           uint64_t* __restrict__ masks) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nconf) return;
    for (int w = 0; w < nw; ++w) {
        uint64_t v = 0;
        for (int b = 0; b < 64 && 64 * w + b < lconfs; ++b)
            if (confs[(size_t)c + (size_t)nconf * (64 * w + b)] != 0) v |= 1ull << b;
        masks[(size_t)c * nw + w] = v;
    }
}

// Connected pairs: per group of CB determinants c1 (one wave), every c2 with
// ndiff <= maxdiff in increasing order.  off == nullptr: counts only.
constexpr int CB = 4;
template <int NW>
__global__ void __launch_bounds__(256) k_rdm_pairs(const uint64_t* __restrict__ masks, int nconf,
                                                   int maxdiff, int64_t* __restrict__ cnt,
                                                   const int64_t* __restrict__ off,
                                                   int32_t* __restrict__ c1s,
                                                   int32_t* __restrict__ c2s,
                                                   int8_t* __restrict__ nds) {
    const int lane = threadIdx.x & (RW - 1);
    const int ngrp = (nconf + CB - 1) / CB;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (int g = blockIdx.x * (blockDim.x / RW) + threadIdx.x / RW; g < ngrp;
         g += gridDim.x * (blockDim.x / RW)) {
        DMask<NW> m1[CB];
        int64_t run[CB];
#pragma unroll
        for (int k = 0; k < CB; ++k) {
            const int c1 = g * CB + k;
            m1[k] = load_mask<NW>(masks, c1 < nconf ? c1 : 0);
            run[k] = off ? off[c1 < nconf ? c1 : 0] : 0;
        }
        for (int c2b = 0; c2b < nconf; c2b += RW) {
            const int c2 = c2b + lane;
            const DMask<NW> m2 = load_mask<NW>(masks, c2 < nconf ? c2 : 0);
#pragma unroll
            for (int k = 0; k < CB; ++k) {
                int nd = 0;
#pragma unroll
                for (int w = 0; w < NW; ++w) nd += __popcll(m1[k].w[w] ^ m2.w[w]);
                const bool ok = c2 < nconf && g * CB + k < nconf && nd <= maxdiff;
                const uint64_t bal = __ballot(ok);
                if (off && ok) {
                    const int64_t pos = run[k] + __popcll(bal & lt);
                    c1s[pos] = g * CB + k;
                    c2s[pos] = c2;
                    nds[pos] = (int8_t)nd;
                }
                run[k] += __popcll(bal);
            }
        }
        if (!off && lane == 0)
#pragma unroll
            for (int k = 0; k < CB; ++k)
                if (g * CB + k < nconf) cnt[g * CB + k] = run[k];
    }
}

// contributions of one pair, in createtwordm's statement order; emit(key, v)
template <int NW, class F>
__device__ __forceinline__ void pair_contribs(const DMask<NW>& m1, const DMask<NW>& m2, int nd,
                                              double cc, int N, F&& emit) {
    auto TW = [N](int p, int r, int s, int q) -> uint32_t {
        return (uint32_t)((p - 1) + N * ((r - 1) + N * ((s - 1) + N * (q - 1))));
    };
    if (nd == 4) {
        const Exc x = excitation(m1, m2, 4);
        const double base = cc * (double)x.e;
        if (x.s2[0] == x.s1[0] && x.s2[1] == x.s1[1])
            emit(TW(x.d1[1], x.d1[0], x.d2[0], x.d2[1]), base * 1.0);
        if (x.s2[0] == x.s1[1] && x.s2[1] == x.s1[0])
            emit(TW(x.d1[0], x.d1[1], x.d2[0], x.d2[1]), base * -1.0);
        if (x.s1[0] == x.s2[1] && x.s2[0] == x.s1[1])
            emit(TW(x.d1[1], x.d1[0], x.d2[1], x.d2[0]), base * -1.0);
        if (x.s1[1] == x.s2[1] && x.s2[0] == x.s1[0])
            emit(TW(x.d1[0], x.d1[1], x.d2[1], x.d2[0]), base * 1.0);
    } else if (nd == 2) {
        const Exc x = excitation(m1, m2, 2);
        const double base = cc * (double)x.e;
        const int d1 = x.d1[0], d2 = x.d2[0], sp1 = x.s1[0], sp2 = x.s2[0];
        // the slots occupied in both, in slot order, four times
        for (int loop = 0; loop < 4; ++loop) {
            if ((loop == 0 || loop == 3) && sp2 != sp1) continue;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                uint64_t v = m1.w[w] & m2.w[w];
                if (loop == 1 || loop == 2) v &= (sp1 == sp2) ? spin_bits(sp1) : 0ull;
                while (v) {
                    const int o = orb_of(64 * w + __builtin_ctzll(v));
                    v &= v - 1;
                    if (loop == 0) emit(TW(d1, o, o, d2), base * 1.0);
                    else if (loop == 1) emit(TW(o, d1, o, d2), base * -1.0);
                    else if (loop == 2) emit(TW(d1, o, d2, o), base * -1.0);
                    else emit(TW(o, d1, d2, o), base * 1.0);
                }
            }
        }
    } else if (nd == 0) {
        const double base = cc * 1.0;
        for (int w1 = 0; w1 < NW; ++w1) {
            uint64_t v1 = m1.w[w1];
            while (v1) {
                const int i1 = 64 * w1 + __builtin_ctzll(v1);
                v1 &= v1 - 1;
                const int so = orb_of(i1);
                for (int w2 = 0; w2 < NW; ++w2) {
                    uint64_t v2 = m2.w[w2];
                    while (v2) {
                        const int i2 = 64 * w2 + __builtin_ctzll(v2);
                        v2 &= v2 - 1;
                        if (i2 == i1) continue;
                        const int qo = orb_of(i2);
                        if (spin_of(i1) == spin_of(i2)) emit(TW(so, qo, so, qo), base * -1.0);
                        emit(TW(qo, so, so, qo), base * 1.0);
                    }
                }
            }
        }
    }
}

template <int NW>
__global__ void k_rdm_ccount(const uint64_t* __restrict__ masks, const int32_t* __restrict__ c1s,
                             const int32_t* __restrict__ c2s, const int8_t* __restrict__ nds,
                             int64_t np, int64_t* __restrict__ cnt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const DMask<NW> m1 = load_mask<NW>(masks, c1s[i]), m2 = load_mask<NW>(masks, c2s[i]);
    const int nd = nds[i];
    int64_t n = 0;
    if (nd == 4) {
        const Exc x = excitation(m1, m2, 4);
        n = (x.s2[0] == x.s1[0] && x.s2[1] == x.s1[1]) + (x.s2[0] == x.s1[1] && x.s2[1] == x.s1[0]) +
            (x.s1[0] == x.s2[1] && x.s2[0] == x.s1[1]) + (x.s1[1] == x.s2[1] && x.s2[0] == x.s1[0]);
    } else if (nd == 2) {
        const Exc x = excitation(m1, m2, 2);
        if (x.s1[0] == x.s2[0]) {
            int nocc = 0, same = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                const uint64_t c = m1.w[w] & m2.w[w];
                nocc += __popcll(c);
                same += __popcll(c & spin_bits(x.s1[0]));
            }
            n = 2 * nocc + 2 * same;
        }
    } else if (nd == 0) {
        int na = 0, nb = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            na += __popcll(m1.w[w] & spin_bits(1));
            nb += __popcll(m1.w[w] & spin_bits(0));
        }
        const int64_t nt = na + nb;
        n = nt * (nt - 1) + (int64_t)na * (na - 1) + (int64_t)nb * (nb - 1);
    }
    cnt[i] = n;
}

template <int NW>
__global__ void k_rdm_cwrite(const uint64_t* __restrict__ masks, const int32_t* __restrict__ c1s,
                             const int32_t* __restrict__ c2s, const int8_t* __restrict__ nds,
                             int64_t np, const double* __restrict__ c1v,
                             const double* __restrict__ c2v, const int64_t* __restrict__ coff,
                             int N, uint32_t* __restrict__ keys, double* __restrict__ vals) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const int c1 = c1s[i], c2 = c2s[i];
    const DMask<NW> m1 = load_mask<NW>(masks, c1), m2 = load_mask<NW>(masks, c2);
    int64_t pos = coff[i];
    const int64_t end = coff[i + 1];  // k_rdm_ccount's count for this pair
    const double cc = c1v[c1] * c2v[c2];
    pair_contribs<NW>(m1, m2, nds[i], cc, N, [&](uint32_t k, double v) {
        if (pos < end) {
            keys[pos] = k;
            vals[pos] = v;
        }
        ++pos;
    });
}

// createonerdm (src/one_rdm.f90:109-222) on the same pair list (ndiff <= 2):
// an ndiff-2 pair adds c1 c2 ep at (d1, d2), the orbitals of c1's and c2's
// differing slots; an ndiff-0 pair adds c1v[c1] c2v[c1] at (o, o) for every
// occupied slot, in slot order (the reference's loop reads c2v at c1).
template <int NW>
__global__ void k_rdm1_ccount(const uint64_t* __restrict__ masks, const int32_t* __restrict__ c1s,
                              const int8_t* __restrict__ nds, int64_t np,
                              int64_t* __restrict__ cnt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const int nd = nds[i];
    int64_t n = 0;
    if (nd == 2) {
        n = 1;
    } else if (nd == 0) {
        const DMask<NW> m1 = load_mask<NW>(masks, c1s[i]);
#pragma unroll
        for (int w = 0; w < NW; ++w) n += __popcll(m1.w[w]);
    }
    cnt[i] = n;
}

template <int NW>
__global__ void k_rdm1_cwrite(const uint64_t* __restrict__ masks, const int32_t* __restrict__ c1s,
                              const int32_t* __restrict__ c2s, const int8_t* __restrict__ nds,
                              int64_t np, const double* __restrict__ c1v,
                              const double* __restrict__ c2v, const int64_t* __restrict__ coff,
                              int N, uint32_t* __restrict__ keys, double* __restrict__ vals) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const int c1 = c1s[i], c2 = c2s[i], nd = nds[i];
    int64_t pos = coff[i];
    const int64_t end = coff[i + 1];  // k_rdm1_ccount's count for this pair
    const DMask<NW> m1 = load_mask<NW>(masks, c1);
    if (nd == 2) {
        const DMask<NW> m2 = load_mask<NW>(masks, c2);
        const Exc x = excitation(m1, m2, 2);
        if (pos < end) {
            keys[pos] = (uint32_t)((x.d1[0] - 1) + N * (x.d2[0] - 1));
            vals[pos] = c1v[c1] * c2v[c2] * (double)x.e;
        }
    } else if (nd == 0) {
        const double v = c1v[c1] * c2v[c1] * 1.0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            uint64_t b = m1.w[w];
            while (b) {
                const int j = 64 * w + __builtin_ctzll(b);
                b &= b - 1;
                const int o = orb_of(j) - 1;
                if (pos < end) {
                    keys[pos] = (uint32_t)(o + N * o);
                    vals[pos] = v;
                }
                ++pos;
            }
        }
    }
}

// start of each element's run in the sorted keys (lower bound)
__global__ void k_rdm_bounds(const uint32_t* __restrict__ keys, int64_t n, int64_t nel,
                             int64_t* __restrict__ start) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e > nel) return;
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)keys[mid] < e) lo = mid + 1;
        else hi = mid;
    }
    start[e] = lo;
}

// twordm(e) = (((0 + v_0) + v_1) + ...) over the element's contributions in
// the reference's order
__global__ void k_rdm_esum(const double* __restrict__ vals, const int64_t* __restrict__ start,
                           int64_t nel, double* __restrict__ tw) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nel) return;
    double s = 0.0;
    for (int64_t i = start[e]; i < start[e + 1]; ++i) s = s + vals[i];
    tw[e] = s;
}

// selection loop p, q, r, s (p outermost): flag of loop position L
__global__ void k_rdm_flags(const double* __restrict__ tw, int N, double cutoff,
                            uint8_t* __restrict__ flags) {
    const int64_t L = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n4 = (int64_t)N * N * N * N;
    if (L >= n4) return;
    const int s = (int)(L % N), r = (int)((L / N) % N), q = (int)((L / N / N) % N),
              p = (int)(L / N / N / N);
    const double v = tw[p + (int64_t)N * (q + (int64_t)N * (r + (int64_t)N * s))];
    flags[L] = fabs(v) >= cutoff ? 1 : 0;
}

// rows (p, s, q, r) and values twordm(p, q, r, s) at their scanned positions
__global__ void k_rdm_emit(const double* __restrict__ tw, int N, const uint8_t* __restrict__ flags,
                           const int64_t* __restrict__ pos, int32_t* __restrict__ rows,
                           double* __restrict__ out) {
    const int64_t L = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n4 = (int64_t)N * N * N * N;
    if (L >= n4 || !flags[L]) return;
    const int s = (int)(L % N), r = (int)((L / N) % N), q = (int)((L / N / N) % N),
              p = (int)(L / N / N / N);
    const int64_t o = pos[L];
    rows[4 * o + 0] = p + 1;
    rows[4 * o + 1] = s + 1;
    rows[4 * o + 2] = q + 1;
    rows[4 * o + 3] = r + 1;
    out[o] = tw[p + (int64_t)N * (q + (int64_t)N * (r + (int64_t)N * s))];
}

// ---- bit-string builders (src/one_rdm.f90) -------------------------------
// A determinant is two 64-bit words, w[0] alpha and w[1] beta, orbital k on
// bit k-1 (the reference's integer bit strings).  The item list is the n
// determinants themselves (the no-difference pass, run first for every
// determinant) followed by the connected ordered pairs of k_rdm_pairs, so
// scanning per-item counts gives every contribution its place in the
// reference's global order.

// orbitals strictly between 1-based orbitals x and y
__device__ __forceinline__ uint64_t bit_between(int x, int y) {
    const int lo = x < y ? x : y, hi = x < y ? y : x;
    return below(hi - 1) & ~below(lo);
}

__global__ void k_bit_pack(const int64_t* __restrict__ alpha, const int64_t* __restrict__ beta,
                           int n, uint64_t* __restrict__ masks) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    masks[2 * (size_t)i] = (uint64_t)alpha[i];
    masks[2 * (size_t)i + 1] = (uint64_t)beta[i];
}

// the determinants' own items (c1 = c2 = i) at the head of the list
__global__ void k_bit_diag(int n, int32_t* __restrict__ c1s, int32_t* __restrict__ c2s,
                           int8_t* __restrict__ nds) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    c1s[i] = i;
    c2s[i] = i;
    nds[i] = 0;
}

// createtwordm_bit's contributions of item i (src/one_rdm.f90:1009-1300),
// in statement order; emit(key, value)
template <class F>
__device__ __forceinline__ void bit2_contribs(const DMask<2>& m1, const DMask<2>& m2, int nd,
                                              bool diag, double c1i, double c2j, int N,
                                              F&& emit) {
    auto TW = [N](int p, int q, int r, int s) -> uint32_t {
        return (uint32_t)((p - 1) + N * ((q - 1) + N * ((r - 1) + N * (s - 1))));
    };
    if (diag) {
        // no differences: every ordered choice of two occupied slots
        const double c = c1i * c2j;
        for (int sp = 0; sp < 2; ++sp) {
            uint64_t b1 = m1.w[sp];
            while (b1) {
                const int so = __builtin_ctzll(b1) + 1;
                b1 &= b1 - 1;
                for (int sp2 = 0; sp2 < 2; ++sp2) {
                    uint64_t b2 = m1.w[sp2];
                    while (b2) {
                        const int qo = __builtin_ctzll(b2) + 1;
                        b2 &= b2 - 1;
                        if (qo == so && sp == sp2) continue;
                        if (sp == sp2) emit(TW(so, qo, so, qo), -c);
                        emit(TW(qo, so, so, qo), c);
                    }
                }
            }
        }
        return;
    }
    if (nd == 4) {
        int hol[2][2] = {{0, 0}, {0, 0}}, par[2][2] = {{0, 0}, {0, 0}};
        int nperm = 0, same = -1;
        for (int sp = 0; sp < 2; ++sp) {
            const uint64_t x = m1.w[sp] ^ m2.w[sp];
            uint64_t h = x & m1.w[sp], pt = x & m2.w[sp];
            int nh = 0, np_ = 0;
            while (pt && np_ < 2) {
                par[sp][np_++] = __builtin_ctzll(pt) + 1;
                pt &= pt - 1;
            }
            while (h && nh < 2) {
                hol[sp][nh++] = __builtin_ctzll(h) + 1;
                h &= h - 1;
            }
            for (int k = 0; k < nh; ++k)
                nperm += __popcll(m1.w[sp] & bit_between(hol[sp][k], par[sp][k]));
            if (nh == 2) {
                same = sp;
                const int a = min(hol[sp][0], par[sp][0]), b = max(hol[sp][0], par[sp][0]);
                const int c = min(hol[sp][1], par[sp][1]), d = max(hol[sp][1], par[sp][1]);
                if (c > a && c < b && d > b) ++nperm;
                break;
            }
        }
        const double c = c1i * c2j * ((nperm & 1) ? -1.0 : 1.0);
        if (same < 0) {
            const int so = par[0][0], qo = par[1][0], po = hol[1][0], ro = hol[0][0];
            emit(TW(po, ro, so, qo), c);
            emit(TW(ro, po, qo, so), c);
        } else {
            const int so = par[same][0], qo = par[same][1], po = hol[same][1], ro = hol[same][0];
            emit(TW(po, ro, so, qo), c);
            emit(TW(ro, po, so, qo), -c);
            emit(TW(po, ro, qo, so), -c);
            emit(TW(ro, po, qo, so), c);
        }
    } else if (nd == 2) {
        const int sp = (m1.w[0] ^ m2.w[0]) ? 0 : 1;
        const uint64_t x = m1.w[sp] ^ m2.w[sp];
        const int po = __builtin_ctzll(x & m1.w[sp]) + 1, qo = __builtin_ctzll(x & m2.w[sp]) + 1;
        const int nperm = __popcll(m1.w[sp] & bit_between(po, qo));
        const double c = c1i * c2j * ((nperm & 1) ? -1.0 : 1.0);
        uint64_t b = m1.w[sp] & ~(1ull << (po - 1));
        while (b) {
            const int ro = __builtin_ctzll(b) + 1;
            b &= b - 1;
            emit(TW(po, ro, ro, qo), c);
            emit(TW(ro, po, qo, ro), c);
            emit(TW(ro, po, ro, qo), -c);
            emit(TW(po, ro, qo, ro), -c);
        }
        b = m1.w[1 - sp];
        while (b) {
            const int ro = __builtin_ctzll(b) + 1;
            b &= b - 1;
            emit(TW(po, ro, ro, qo), c);
            emit(TW(ro, po, qo, ro), c);
        }
    }
}

__device__ __forceinline__ int64_t bit2_count(const DMask<2>& m1, const DMask<2>& m2, int nd,
                                              bool diag) {
    const int64_t na = __popcll(m1.w[0]), nb = __popcll(m1.w[1]);
    if (diag) return 2 * (na * (na - 1) + nb * (nb - 1)) + 2 * na * nb;
    if (nd == 4) return (__popcll(m1.w[0] ^ m2.w[0]) == 4 || __popcll(m1.w[1] ^ m2.w[1]) == 4) ? 4 : 2;
    if (nd == 2) {
        const int sp = (m1.w[0] ^ m2.w[0]) ? 0 : 1;
        return 4 * (__popcll(m1.w[sp]) - 1) + 2 * (int64_t)__popcll(m1.w[1 - sp]);
    }
    return 0;
}

__global__ void k_bit2_ccount(const uint64_t* __restrict__ masks, const int32_t* __restrict__ c1s,
                              const int32_t* __restrict__ c2s, const int8_t* __restrict__ nds,
                              int64_t nitem, int ndiag, int64_t* __restrict__ cnt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nitem) return;
    const DMask<2> m1 = load_mask<2>(masks, c1s[i]), m2 = load_mask<2>(masks, c2s[i]);
    cnt[i] = bit2_count(m1, m2, nds[i], i < ndiag);
}

__global__ void k_bit2_cwrite(const uint64_t* __restrict__ masks, const int32_t* __restrict__ c1s,
                              const int32_t* __restrict__ c2s, const int8_t* __restrict__ nds,
                              int64_t nitem, int ndiag, const double* __restrict__ c1v,
                              const double* __restrict__ c2v, const int64_t* __restrict__ coff,
                              int N, uint32_t* __restrict__ keys, double* __restrict__ vals) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nitem) return;
    const int c1 = c1s[i], c2 = c2s[i];
    const DMask<2> m1 = load_mask<2>(masks, c1), m2 = load_mask<2>(masks, c2);
    int64_t pos = coff[i];
    const int64_t end = coff[i + 1];  // k_bit2_ccount's count for this item
    bit2_contribs(m1, m2, nds[i], i < ndiag, c1v[c1], c2v[c2], N, [&](uint32_t k, double v) {
        if (pos < end) {
            keys[pos] = k;
            vals[pos] = v;
        }
        ++pos;
    });
}

// one_rdm_bit (src/one_rdm.f90:1440-1516): a determinant adds civs^2 on the
// diagonal of each occupied orbital, alpha then beta; a pair c1 < c2 with one
// difference adds civs(c1) civs(c2) ep at (p, q) and at (q, p).
__global__ void k_bit1_ccount(const uint64_t* __restrict__ masks, const int32_t* __restrict__ c1s,
                              const int32_t* __restrict__ c2s, const int8_t* __restrict__ nds,
                              int64_t nitem, int ndiag, int64_t* __restrict__ cnt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nitem) return;
    int64_t n = 0;
    if (i < ndiag) {
        const DMask<2> m1 = load_mask<2>(masks, c1s[i]);
        n = __popcll(m1.w[0]) + __popcll(m1.w[1]);
    } else if (nds[i] == 2 && c2s[i] > c1s[i]) {
        n = 2;
    }
    cnt[i] = n;
}

__global__ void k_bit1_cwrite(const uint64_t* __restrict__ masks, const int32_t* __restrict__ c1s,
                              const int32_t* __restrict__ c2s, const int8_t* __restrict__ nds,
                              int64_t nitem, int ndiag, const double* __restrict__ civ,
                              const int64_t* __restrict__ coff, int N,
                              uint32_t* __restrict__ keys, double* __restrict__ vals) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nitem) return;
    int64_t pos = coff[i];
    const int64_t end = coff[i + 1];  // k_bit1_ccount's count for this item
    const int c1 = c1s[i], c2 = c2s[i];
    const DMask<2> m1 = load_mask<2>(masks, c1);
    if (i < ndiag) {
        const double v = civ[c1] * civ[c1];
        for (int sp = 0; sp < 2; ++sp) {
            uint64_t b = m1.w[sp];
            while (b) {
                const int j = __builtin_ctzll(b);
                b &= b - 1;
                if (pos < end) {
                    keys[pos] = (uint32_t)(j + N * j);
                    vals[pos] = v;
                }
                ++pos;
            }
        }
    } else if (nds[i] == 2 && c2 > c1 && pos + 2 <= end) {
        const DMask<2> m2 = load_mask<2>(masks, c2);
        const int sp = (m1.w[0] ^ m2.w[0]) ? 0 : 1;
        const uint64_t x = m1.w[sp] ^ m2.w[sp];
        const int po = __builtin_ctzll(m1.w[sp] & x) + 1, qo = __builtin_ctzll(m2.w[sp] & x) + 1;
        const int ep = __popcll(m1.w[sp] & bit_between(po, qo));
        const double v = civ[c1] * civ[c2] * ((ep & 1) ? -1.0 : 1.0);
        keys[pos] = (uint32_t)((po - 1) + N * (qo - 1));
        vals[pos] = v;
        keys[pos + 1] = (uint32_t)((qo - 1) + N * (po - 1));
        vals[pos + 1] = v;
    }
}

// mcci_to_bit's sign (src/one_rdm.f90:1920-1972): the parity of the
// reordering of a determinant's slots from orbital order (alpha before beta
// on each orbital) to alpha-then-beta order, i.e. of the number of
// (beta j, alpha k) pairs with k > j.  Coefficients are multiplied by it.
__global__ void k_mcci_phase(const uint64_t* __restrict__ masks, int n,
                             const double* __restrict__ c1in, const double* __restrict__ c2in,
                             double* __restrict__ c1out, double* __restrict__ c2out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t a = masks[2 * (size_t)i], b0 = masks[2 * (size_t)i + 1];
    int inv = 0;
    uint64_t b = b0;
    while (b) {
        const int j = __builtin_ctzll(b);
        b &= b - 1;
        inv += __popcll(a & ~below(j + 1));
    }
    const double ph = (inv & 1) ? -1.0 : 1.0;
    c1out[i] = ph * c1in[i];
    c2out[i] = ph * c2in[i];
}

// createtwordm_bit_fci (src/one_rdm.f90:390-875).  One item per file
// determinant c1: its no-difference pass, then for every single excitation
// A' of its alpha string A (added orbital j increasing, removed k increasing;
// strings not among the unique alpha strings, or with no determinant,
// skipped) the single contributions when (A', B) is a determinant, and for
// every single excitation B' of its beta string (taken from the same
// unique-alpha table) the alpha-beta double contributions when (A', B') is a
// determinant; then every alpha-alpha double excitation (removed pair,
// added pair, both increasing).  Determinants are found through the
// (alpha, beta) keys sorted stably, i.e. the first in file order, which is
// findloc's choice within a string's contiguous block.
struct FciView {
    const uint64_t* masks;   // (alpha, beta) per determinant
    const double* civ;
    const uint64_t* uniq;    // unique alpha strings, increasing
    const int64_t* irep;     // determinants per unique string
    const uint64_t* keys;    // alpha * 2^32 + beta, sorted stably
    const int32_t* kidx;     // file position of each sorted key
    int n, nuniq, N;
};

__device__ __forceinline__ bool fci_has_string(const FciView& V, uint64_t s) {
    int lo = 0, hi = V.nuniq;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (V.uniq[mid] < s) lo = mid + 1;
        else hi = mid;
    }
    return lo < V.nuniq && V.uniq[lo] == s && V.irep[lo] != 0;
}

__device__ __forceinline__ int fci_find_det(const FciView& V, uint64_t a, uint64_t b) {
    const uint64_t key = (a << 32) | b;
    int lo = 0, hi = V.n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (V.keys[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return (lo < V.n && V.keys[lo] == key) ? V.kidx[lo] : -1;
}

template <class F>
__device__ void fci_contribs(const FciView& V, int c1, F&& emit) {
    const int N = V.N;
    auto TW = [N](int p, int q, int r, int s) -> uint32_t {
        return (uint32_t)((p - 1) + N * ((q - 1) + N * ((r - 1) + N * (s - 1))));
    };
    const DMask<2> m = load_mask<2>(V.masks, c1);
    const uint64_t A = m.w[0], B = m.w[1];
    const uint64_t comp = N >= 64 ? ~0ull : ((1ull << N) - 1ull);
    const double civ1 = V.civ[c1];
    bit2_contribs(m, m, 0, true, civ1, civ1, N, emit);
    // alpha singles, each followed by its alpha-beta doubles
    uint64_t bj = A ^ comp;
    while (bj) {
        const int j = __builtin_ctzll(bj);
        bj &= bj - 1;
        uint64_t bk = A;
        while (bk) {
            const int k = __builtin_ctzll(bk);
            bk &= bk - 1;
            const uint64_t A1 = (A | (1ull << j)) & ~(1ull << k);
            if (!fci_has_string(V, A1)) continue;
            const int ep = __popcll(A & bit_between(j + 1, k + 1));
            const int d = fci_find_det(V, A1, B);
            if (d >= 0) {
                const double ct = ((ep & 1) ? -1.0 : 1.0) * 2.0 * civ1 * V.civ[d];
                const int po = k + 1, qo = j + 1;
                uint64_t b = A & ~(1ull << (po - 1));
                while (b) {
                    const int ro = __builtin_ctzll(b) + 1;
                    b &= b - 1;
                    emit(TW(po, ro, ro, qo), ct);
                    emit(TW(ro, po, qo, ro), ct);
                    emit(TW(ro, po, ro, qo), -ct);
                    emit(TW(po, ro, qo, ro), -ct);
                }
                b = B;
                while (b) {
                    const int ro = __builtin_ctzll(b) + 1;
                    b &= b - 1;
                    emit(TW(po, ro, ro, qo), ct);
                    emit(TW(ro, po, qo, ro), ct);
                }
            }
            uint64_t cj = B ^ comp;
            while (cj) {
                const int jb = __builtin_ctzll(cj);
                cj &= cj - 1;
                uint64_t ck = B;
                while (ck) {
                    const int kb = __builtin_ctzll(ck);
                    ck &= ck - 1;
                    const uint64_t B1 = (B | (1ull << jb)) & ~(1ull << kb);
                    if (!fci_has_string(V, B1)) continue;
                    const int d2 = fci_find_det(V, A1, B1);
                    if (d2 < 0) continue;
                    const int e2 = ep + __popcll(B & bit_between(jb + 1, kb + 1));
                    const double ct = 1.0 * ((e2 & 1) ? -1.0 : 1.0) * civ1 * V.civ[d2];
                    const int ro = k + 1, so = j + 1, po = kb + 1, qo = jb + 1;
                    emit(TW(po, ro, so, qo), ct);
                    emit(TW(ro, po, qo, so), ct);
                }
            }
        }
    }
    // alpha-alpha doubles
    int occ[64], unocc[64], no = 0, nu = 0;
    for (int k = 0; k < N; ++k) {
        if ((A >> k) & 1) occ[no++] = k;
        else unocc[nu++] = k;
    }
    for (int a = 0; a + 1 < no; ++a)
        for (int b2 = a + 1; b2 < no; ++b2)
            for (int k = 0; k + 1 < nu; ++k)
                for (int l = k + 1; l < nu; ++l) {
                    const uint64_t A2 = (((A & ~(1ull << occ[a])) & ~(1ull << occ[b2])) |
                                         (1ull << unocc[k])) | (1ull << unocc[l]);
                    if (!fci_has_string(V, A2)) continue;
                    const int d = fci_find_det(V, A2, B);
                    if (d < 0) continue;
                    const int l1 = min(occ[a], unocc[k]) + 1, h1 = max(occ[a], unocc[k]) + 1;
                    const int l2 = min(occ[b2], unocc[l]) + 1, h2 = max(occ[b2], unocc[l]) + 1;
                    int ep = __popcll(A & bit_between(l1, h1)) + __popcll(A & bit_between(l2, h2));
                    if (l2 > l1 && l2 < h1 && h2 > h1) ++ep;
                    const double ct = ((ep & 1) ? -1.0 : 1.0) * 2.0 * civ1 * V.civ[d];
                    const int so = unocc[k] + 1, qo = unocc[l] + 1, po = occ[b2] + 1,
                              ro = occ[a] + 1;
                    emit(TW(po, ro, so, qo), ct);
                    emit(TW(ro, po, so, qo), -ct);
                    emit(TW(po, ro, qo, so), -ct);
                    emit(TW(ro, po, qo, so), ct);
                }
}

__global__ void k_fci_keys(const uint64_t* __restrict__ masks, int n, uint64_t* __restrict__ keys,
                           int32_t* __restrict__ idx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = (masks[2 * (size_t)i] << 32) | masks[2 * (size_t)i + 1];
    idx[i] = i;
}

__global__ void k_fci_ccount(FciView V, int64_t* __restrict__ cnt) {
    const int c1 = blockIdx.x * blockDim.x + threadIdx.x;
    if (c1 >= V.n) return;
    int64_t k = 0;
    fci_contribs(V, c1, [&](uint32_t, double) { ++k; });
    cnt[c1] = k;
}

__global__ void k_fci_cwrite(FciView V, const int64_t* __restrict__ coff,
                             uint32_t* __restrict__ keys, double* __restrict__ vals) {
    const int c1 = blockIdx.x * blockDim.x + threadIdx.x;
    if (c1 >= V.n) return;
    int64_t pos = coff[c1];
    const int64_t end = coff[c1 + 1];  // k_fci_ccount's count for this determinant
    fci_contribs(V, c1, [&](uint32_t k, double v) {
        if (pos < end) {
            keys[pos] = k;
            vals[pos] = v;
        }
        ++pos;
    });
}

inline unsigned nb64(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

hipError_t rdm_pack(const int32_t* confs, int nconf, int lconfs, int nw, uint64_t* masks,
                    hipStream_t s) {
    hipLaunchKernelGGL(k_rdm_pack, dim3(nb64(nconf, 128)), dim3(128), 0, s, confs, nconf, lconfs,
                       nw, masks);
    return hipGetLastError();
}

#define RDM_NW(X) \
    X(1) X(2) X(3) X(4)

hipError_t rdm_pairs(const uint64_t* masks, int nconf, int nw, int maxdiff, int64_t* cnt,
                     const int64_t* off, int32_t* c1s, int32_t* c2s, int8_t* nds, hipStream_t s) {
    const int64_t ngrp = (nconf + CB - 1) / CB;
    const unsigned blocks = (unsigned)std::min<int64_t>(nb64(ngrp, 4), 65536);
    switch (nw) {
#define P_(k)                                                                               \
    case k:                                                                                 \
        hipLaunchKernelGGL(k_rdm_pairs<k>, dim3(blocks), dim3(256), 0, s, masks, nconf, maxdiff, \
                           cnt, off, c1s, c2s, nds);                                         \
        break;
        RDM_NW(P_)
#undef P_
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t rdm_ccount(const uint64_t* masks, int nw, const int32_t* c1s, const int32_t* c2s,
                      const int8_t* nds, int64_t np, int64_t* cnt, hipStream_t s) {
    if (np <= 0) return hipSuccess;
    switch (nw) {
#define P_(k)                                                                                 \
    case k:                                                                                   \
        hipLaunchKernelGGL(k_rdm_ccount<k>, dim3(nb64(np, 256)), dim3(256), 0, s, masks, c1s, c2s, \
                           nds, np, cnt);                                                      \
        break;
        RDM_NW(P_)
#undef P_
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t rdm_cwrite(const uint64_t* masks, int nw, const int32_t* c1s, const int32_t* c2s,
                      const int8_t* nds, int64_t np, const double* c1v, const double* c2v,
                      const int64_t* coff, int N, uint32_t* keys, double* vals, hipStream_t s) {
    if (np <= 0) return hipSuccess;
    switch (nw) {
#define P_(k)                                                                                 \
    case k:                                                                                   \
        hipLaunchKernelGGL(k_rdm_cwrite<k>, dim3(nb64(np, 256)), dim3(256), 0, s, masks, c1s, c2s, \
                           nds, np, c1v, c2v, coff, N, keys, vals);                            \
        break;
        RDM_NW(P_)
#undef P_
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
hipError_t rdm1_ccount(const uint64_t* masks, int nw, const int32_t* c1s, const int8_t* nds,
                       int64_t np, int64_t* cnt, hipStream_t s) {
    if (np <= 0) return hipSuccess;
    switch (nw) {
#define P_(k)                                                                                  \
    case k:                                                                                    \
        hipLaunchKernelGGL(k_rdm1_ccount<k>, dim3(nb64(np, 256)), dim3(256), 0, s, masks, c1s,  \
                           nds, np, cnt);                                                       \
        break;
        RDM_NW(P_)
#undef P_
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t rdm1_cwrite(const uint64_t* masks, int nw, const int32_t* c1s, const int32_t* c2s,
                       const int8_t* nds, int64_t np, const double* c1v, const double* c2v,
                       const int64_t* coff, int N, uint32_t* keys, double* vals, hipStream_t s) {
    if (np <= 0) return hipSuccess;
    switch (nw) {
#define P_(k)                                                                                  \
    case k:                                                                                    \
        hipLaunchKernelGGL(k_rdm1_cwrite<k>, dim3(nb64(np, 256)), dim3(256), 0, s, masks, c1s,  \
                           c2s, nds, np, c1v, c2v, coff, N, keys, vals);                        \
        break;
        RDM_NW(P_)
#undef P_
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
#undef RDM_NW

hipError_t rdm_scan_i64(void* temp, size_t& bytes, const int64_t* in, int64_t* out, int64_t n,
                        hipStream_t s) {
    return hipcub::DeviceScan::ExclusiveSum(temp, bytes, in, out, n, s);
}

hipError_t rdm_sort(void* temp, size_t& bytes, const uint32_t* kin, uint32_t* kout,
                    const double* vin, double* vout, int64_t n, int keybits, hipStream_t s) {
    return hipcub::DeviceRadixSort::SortPairs(temp, bytes, kin, kout, vin, vout, n, 0, keybits, s);
}

hipError_t class_sort(void* temp, size_t& bytes, const uint32_t* kin, uint32_t* kout,
                      const int32_t* vin, int32_t* vout, int64_t n, hipStream_t s) {
    return hipcub::DeviceRadixSort::SortPairs(temp, bytes, kin, kout, vin, vout, n, 0, 5, s);
}

hipError_t rdm_esum(const uint32_t* keys, const double* vals, int64_t n, int64_t nel,
                    int64_t* start, double* tw, hipStream_t s) {
    hipLaunchKernelGGL(k_rdm_bounds, dim3(nb64(nel + 1, 256)), dim3(256), 0, s, keys, n, nel, start);
    hipLaunchKernelGGL(k_rdm_esum, dim3(nb64(nel, 256)), dim3(256), 0, s, vals, start, nel, tw);
    return hipGetLastError();
}

hipError_t rdm_flags(const double* tw, int N, double cutoff, uint8_t* flags, hipStream_t s) {
    const int64_t n4 = (int64_t)N * N * N * N;
    hipLaunchKernelGGL(k_rdm_flags, dim3(nb64(n4, 256)), dim3(256), 0, s, tw, N, cutoff, flags);
    return hipGetLastError();
}

hipError_t rdm_emit(const double* tw, int N, const uint8_t* flags, const int64_t* pos,
                    int32_t* rows, double* out, hipStream_t s) {
    const int64_t n4 = (int64_t)N * N * N * N;
    hipLaunchKernelGGL(k_rdm_emit, dim3(nb64(n4, 256)), dim3(256), 0, s, tw, N, flags, pos, rows,
                       out);
    return hipGetLastError();
}

hipError_t bit_pack(const int64_t* alpha, const int64_t* beta, int n, uint64_t* masks,
                    hipStream_t s) {
    hipLaunchKernelGGL(k_bit_pack, dim3(nb64(n, 256)), dim3(256), 0, s, alpha, beta, n, masks);
    return hipGetLastError();
}

hipError_t bit_diag(int n, int32_t* c1s, int32_t* c2s, int8_t* nds, hipStream_t s) {
    hipLaunchKernelGGL(k_bit_diag, dim3(nb64(n, 256)), dim3(256), 0, s, n, c1s, c2s, nds);
    return hipGetLastError();
}

hipError_t bit2_ccount(const uint64_t* masks, const int32_t* c1s, const int32_t* c2s,
                       const int8_t* nds, int64_t nitem, int ndiag, int64_t* cnt, hipStream_t s) {
    if (nitem <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bit2_ccount, dim3(nb64(nitem, 256)), dim3(256), 0, s, masks, c1s, c2s, nds,
                       nitem, ndiag, cnt);
    return hipGetLastError();
}

hipError_t bit2_cwrite(const uint64_t* masks, const int32_t* c1s, const int32_t* c2s,
                       const int8_t* nds, int64_t nitem, int ndiag, const double* c1v,
                       const double* c2v, const int64_t* coff, int N, uint32_t* keys, double* vals,
                       hipStream_t s) {
    if (nitem <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bit2_cwrite, dim3(nb64(nitem, 256)), dim3(256), 0, s, masks, c1s, c2s, nds,
                       nitem, ndiag, c1v, c2v, coff, N, keys, vals);
    return hipGetLastError();
}

hipError_t bit1_ccount(const uint64_t* masks, const int32_t* c1s, const int32_t* c2s,
                       const int8_t* nds, int64_t nitem, int ndiag, int64_t* cnt, hipStream_t s) {
    if (nitem <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bit1_ccount, dim3(nb64(nitem, 256)), dim3(256), 0, s, masks, c1s, c2s, nds,
                       nitem, ndiag, cnt);
    return hipGetLastError();
}

hipError_t bit1_cwrite(const uint64_t* masks, const int32_t* c1s, const int32_t* c2s,
                       const int8_t* nds, int64_t nitem, int ndiag, const double* civ,
                       const int64_t* coff, int N, uint32_t* keys, double* vals, hipStream_t s) {
    if (nitem <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bit1_cwrite, dim3(nb64(nitem, 256)), dim3(256), 0, s, masks, c1s, c2s, nds,
                       nitem, ndiag, civ, coff, N, keys, vals);
    return hipGetLastError();
}

hipError_t mcci_phase(const uint64_t* masks, int n, const double* c1in, const double* c2in,
                      double* c1out, double* c2out, hipStream_t s) {
    hipLaunchKernelGGL(k_mcci_phase, dim3(nb64(n, 256)), dim3(256), 0, s, masks, n, c1in, c2in,
                       c1out, c2out);
    return hipGetLastError();
}

hipError_t fci_keys(const uint64_t* masks, int n, uint64_t* keys, int32_t* idx, hipStream_t s) {
    hipLaunchKernelGGL(k_fci_keys, dim3(nb64(n, 256)), dim3(256), 0, s, masks, n, keys, idx);
    return hipGetLastError();
}

hipError_t fci_key_sort(void* temp, size_t& bytes, const uint64_t* kin, uint64_t* kout,
                        const int32_t* vin, int32_t* vout, int64_t n, hipStream_t s) {
    return hipcub::DeviceRadixSort::SortPairs(temp, bytes, kin, kout, vin, vout, n, 0, 64, s);
}

hipError_t fci_ccount(const FciArgs& a, int64_t* cnt, hipStream_t s) {
    const FciView V{a.masks, a.civ, a.uniq, a.irep, a.keys, a.kidx, a.n, a.nuniq, a.N};
    hipLaunchKernelGGL(k_fci_ccount, dim3(nb64(a.n, 64)), dim3(64), 0, s, V, cnt);
    return hipGetLastError();
}

hipError_t fci_cwrite(const FciArgs& a, const int64_t* coff, uint32_t* keys, double* vals,
                      hipStream_t s) {
    const FciView V{a.masks, a.civ, a.uniq, a.irep, a.keys, a.kidx, a.n, a.nuniq, a.N};
    hipLaunchKernelGGL(k_fci_cwrite, dim3(nb64(a.n, 64)), dim3(64), 0, s, V, coff, keys, vals);
    return hipGetLastError();
}
