// Quartet kernels of the total / elastic j0 path (gfx950).
//
// Two kernels per LLmax class (LL = L_i + L_j + L_k + L_r of the quartet's
// four primitive shells), launched over the class's quartet positions:
//
//   k_contract<LL>  one 256-thread workgroup per quartet.  Contracts the
//                   quartet's 2-RDM block with the Hermite expansions and the
//                   j0-derivative coefficients into h(beta), beta = 0..LL
//                   (h_saved of tot_integral_k_ijkr / integral_k_ijkr and the
//                   P0 sum of the h = 0 branch; src/Integrals.f90:4751-4863,
//                   4955-5226,5730-5901).  FP64.
//   k_series<LL,M>  one wave per (quartet slice, 64-point q slice).  Evaluates
//                   F(q) = sum_beta h(beta) (q/h)^beta j_beta(q h)
//                   (bessels0rr / dphrec, src/Integrals.f90:6662-6739,
//                   7147-7184) or the h = 0 polynomial (h(beta) carrying the
//                   quartet's multiplicity mult), weights it with
//                   e12(q,i,j) * e12(q,k,r) and accumulates S(q) with
//                   compensated sums.  M = true runs the large-argument
//                   Bessel recursion in FP32 (mixed precision).
//
// Factorisation used by k_contract.  The reference forms, per GTO quadruple,
//   sum_{t,u,v,t',u',v'} (-1)^{t+u+v} E^x_ab(t) E^y_ab(u) E^z_ab(v)
//                         E^x_cd(t') E^y_cd(u') E^z_cd(v') h_pre2(beta,T,U,V)
// with T = t+t' etc. and h_pre2(beta,T,U,V) = sum a(T,i) b(U,j) c(V,k) over
// beta = ceil((T+U+V-i-j-k)/2)+i+j+k (BesselDeriv).  a(T,i) vanishes unless
// T-i is even, so beta = (T+i)/2 + (U+j)/2 + (V+k)/2 and the sum factorises:
//   h(beta) = sum_{mx+my+mz=beta} px(mx) py(my) pz(mz),
//   px(m)   = sum_T c_x(T) a(T, 2m-T),  c_x(T) = sum_t (-1)^t E^x_ab(t) E^x_cd(T-t).
// px depends only on the x exponents (la,lb,lc,ld) of the quadruple, so each
// workgroup tabulates px/py/pz per exponent tuple in LDS and every quadruple
// is a gather plus a triple convolution.  a(l,p) has the closed form
//   a(l,p) = (-1)^((l+p)/2) l! / (k! p! 2^k) x^p,  k = (l-p)/2,
// which is what rrdj0's recurrence (src/Integrals.f90:7062-7091) produces.
#include "series_math.hpp"

#include <algorithm>
#include <mutex>


// ------------------------------------------------------------ contraction
// J2 (total j2, tot_integral_k_ijkr_j2, src/Integrals.f90:5230-5420): besides
// h(beta) the kernel accumulates g(beta) = sum z sum px py (mz pz), and writes
//   series: h2(b) = c (eta h(b-2) + 2(b-1) h(b-1) - 6 g(b-1)),  b <= LL+2,
//           c = sqrt(5/pi)/4, eta = 3 hz^2 - h^2 (BesselDeriv2j with rrdj2a/b
//           factorised like the j0 case: the rrdj2b weight 2 ceil((l+p)/2)
//           is 2 m per axis, and -2 on z)
//   h = 0:  c0(b) = sqrt(5/pi)/2 (b h(b) - 3 g(b)),  b <= LL  (P0_j2, :3034-3170)
// Quartet groups g0, g0 + gstep, ... (NG quartets each) of a class list
// (nidx headers) whose records go to hsb.
template <int LL, bool J2>
__device__ __forceinline__ void contract_groups(const SeriesArgs& A, double* __restrict__ hsb,
                                                const QHdr* __restrict__ hdr, int64_t nidx,
                                                int tmax, int64_t g0, int64_t gstep) {
    constexpr int D = LL + 1;
    constexpr int NA = J2 ? 2 : 1;  // accumulated vectors: h, and g for j2
    constexpr int CT = ContractShape<LL>::CT, G = ContractShape<LL>::G, NG = ContractShape<LL>::NG;
    constexpr int NW = CT / WAVE;
    extern __shared__ double sm[];
    __shared__ int anyw[NW];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    const int grp = tid / G, lg = tid % G;
    constexpr int DP = ContractShape<LL>::DP;
    double* PX = sm + (size_t)grp * 3 * tmax * DP;  // [3][T][DP] of this group's quartet
    double* red = sm + (size_t)NG * 3 * tmax * DP;  // [NA][NW][D] when a quartet spans waves
    const int L1 = A.maxl + 1, T1 = 2 * A.maxl + 1;

    // grid-stride over the class list: neighbouring quartets go to
    // neighbouring workgroups, hence to different XCDs (measured faster on
    // XY6 than contiguous runs per workgroup or per XCD)
    // The quartet's header (and, for the elastic kinds, its groups) are
    // loaded before the loop for the first group and at the end of an
    // iteration for the next (XY6 step 6.86 -> 6.74 ms against loading them
    // at the top of the loop body, with fewer VGPRs: 100 against 106 at LL
    // 4).  Loading the next group's at the start of an iteration, beside the
    // current one's table build, with 4096-16384 workgroups looping over the
    // groups of LL <= 6: 6.72-6.77 ms against 6.68 on one box.  The header
    // carries the pair centres' difference, the pair block offsets and the
    // shells' L, so the total kinds read no pair or group record.
    const int64_t stride = gstep * NG;
    QHdr Hn;
    GroupInfo Gin{}, Gjn{}, Gkn{}, Grn{};
    auto fetch_h = [&](int64_t b) {
        const int64_t k = b + grp;
        Hn = hdr[k < nidx ? k : b];
    };
    auto fetch_pg = [&]() {
        if (!A.elastic) return;
        const uint64_t en = QH_ENT(Hn.w);
        Gin = A.G[QE_GI(en)];
        Gjn = A.G[QE_GJ(en)];
        Gkn = A.G[QE_GK(en)];
        Grn = A.G[QE_GR(en)];
    };
    if (g0 * NG < nidx) {
        fetch_h(g0 * NG);
        fetch_pg();
    }
    for (int64_t base = g0 * NG; base < nidx; base += stride) {
        const int64_t kq = base + grp;
        const bool valid = kq < nidx;
        const bool more = base + stride < nidx;
        const QHdr H = Hn;
        const GroupInfo Gi = Gin, Gj = Gjn, Gk = Gkn, Gr = Grn;
        const int64_t e = H.e;
        const uint64_t ent = QH_ENT(H.w);
        const int gi = QE_GI(ent), gj = QE_GJ(ent), gk = QE_GK(ent), gr = QE_GR(ent);
        const int br = QE_BR(ent);
        const int pb = pair_index(gi, gj, A.ng), pk = pair_index(gk, gr, A.ng);
        const int Li = QH_LA(H.ob), Lj = QH_LB(H.ob), Lk = QH_LA(H.ok), Lr = QH_LB(H.ok);
        const int T = valid ? (Li + 1) * (Lj + 1) * (Lk + 1) * (Lr + 1) : 0;
        const double rx = H.rx, ry = H.ry, rz = H.rz;
        const double hx = br ? 0.0 : rx, hy = br ? 0.0 : ry, hz = br ? 0.0 : rz;

        // The quadruple loop is software-pipelined: the lane's first table
        // entry is requested here, before the table build, and its Z before
        // the build's closing barrier; each step of the loop then requests
        // the next step's Z and the entry after that before it contracts its
        // own quadruple, so that a quartet waits for one chain of dependent
        // loads (entry, then Z) once, not once per step (XY6 step 8.15 ->
        // 7.57 ms; two quadruples per step: 7.83)
        const uint32_t nq4 = valid ? QH_NQ4(H.w) : 0u;
        const uint64_t* qt = A.quadtab + H.qoff;
        uint64_t wc = lg < nq4 ? qt[lg] : 0ull;

        __syncthreads();  // the previous quartets are done with the tables
        // px(m) = sum_T c(T) a(T, 2m - T) per axis and exponent tuple, with
        // a(l,p) = kAcoef(l,p) h^p
        // w = ax T + t, t = ((la (Lj+1) + lb) (Lk+1) + lc) (Lr+1) + ld, decoded
        // with 16-bit multiply-shift quotients: floor(t m / 2^16), m =
        // ceil(2^16 / d), is t / d exactly for t < 2^16 / d (every shell has
        // L <= 4: t < 625, d <= 5); offsets in 24-bit products (XY6 step 7.51 -> 7.38 ms
        // against the compiler's division sequences and 64-bit offsets)
        const int rowax = L1 * L1 * T1;
        const double* Ebb = A.E + (int64_t)pb * (3 * rowax);
        const double* Ekb = A.E + (int64_t)pk * (3 * rowax);
        const unsigned mr = (65536u + Lr) / (unsigned)(Lr + 1);
        const unsigned mk = (65536u + Lk) / (unsigned)(Lk + 1);
        const unsigned mj = (65536u + Lj) / (unsigned)(Lj + 1);
        for (int w = lg; w < 3 * T; w += G) {
            const int ax = (w >= T) + (w >= 2 * T);
            const int t = w - ax * T;
            const int t1 = (int)(__umul24((unsigned)t, mr) >> 16);
            const int ld = t - t1 * (Lr + 1);
            const int t2 = (int)(__umul24((unsigned)t1, mk) >> 16);
            const int lc = t1 - t2 * (Lk + 1);
            const int la = (int)(__umul24((unsigned)t2, mj) >> 16);
            const int lb = t2 - la * (Lj + 1);
            const double* Eb = Ebb + (__mul24(ax, rowax) + __mul24(__mul24(la, L1) + lb, T1));
            const double* Ek = Ekb + (__mul24(ax, rowax) + __mul24(__mul24(lc, L1) + ld, T1));
            const int lab = la + lb, lcd = lc + ld;
            double eb[D], ek[D];
#pragma unroll
            for (int s = 0; s < D; ++s) {
                double vb = s <= lab ? Eb[s] : 0.0;
                eb[s] = (s & 1) ? -vb : vb;
                ek[s] = s <= lcd ? Ek[s] : 0.0;
            }
            double cx[D];
#pragma unroll
            for (int T2 = 0; T2 < D; ++T2) {
                double s = 0.0;
#pragma unroll
                for (int t3 = 0; t3 <= T2; ++t3) s += eb[t3] * ek[T2 - t3];
                cx[T2] = s;
            }
            const double hax = ax == 0 ? hx : (ax == 1 ? hy : hz);
            double hp[D];
            hp[0] = 1.0;
#pragma unroll
            for (int k = 1; k < D; ++k) hp[k] = hp[k - 1] * hax;
            double* out = PX + ((int64_t)ax * T + t) * DP;
#pragma unroll
            for (int m = 0; m < D; ++m) {
                double s = 0.0;
#pragma unroll
                for (int T2 = m; T2 <= 2 * m && T2 < D; ++T2)
                    s += cx[T2] * (kAcoef.v[T2][2 * m - T2] * hp[2 * m - T2]);
                out[m] = s;
            }
        }
        const ZView ZV = zview(A, pb, pk, QH_OFF(H.ob), QH_OFF(H.ok), Gi, Gj, Gk, Gr);
        double zc = lg < nq4 ? zget(A, ZV, wc) : 0.0;
        uint64_t wn = lg + G < nq4 ? qt[lg + G] : 0ull;
        __syncthreads();

        double h[NA][D];
#pragma unroll
        for (int k = 0; k < NA; ++k)
#pragma unroll
            for (int b = 0; b < D; ++b) h[k][b] = 0.0;
        int any = 0;
        // one GTO quadruple: h(beta) += z * (px * py * pz)(beta)
        auto quad = [&](uint64_t w, double z) {
            const int tx = (int)((w >> 16) & 1023u), ty = (int)((w >> 26) & 1023u),
                      tz = (int)((w >> 36) & 1023u);
            const double* px = PX + tx * DP;
            const double* py = PX + (T + ty) * DP;
            const double* pz = PX + (2 * T + tz) * DP;
            double vy[D], vz[D], vx[D];
#pragma unroll
            for (int m = 0; m < D; ++m) {
                vx[m] = px[m];
                vy[m] = py[m];
                vz[m] = pz[m];
            }
            double yz[D];
#pragma unroll
            for (int k2 = 0; k2 < D; ++k2) {
                double s = 0.0;
#pragma unroll
                for (int my = 0; my <= k2; ++my) s += vy[my] * vz[k2 - my];
                yz[k2] = s;
            }
#pragma unroll
            for (int b = 0; b < D; ++b) {
                double s = 0.0;
#pragma unroll
                for (int mx = 0; mx <= b; ++mx) s += vx[mx] * yz[b - mx];
                h[0][b] = fma(z, s, h[0][b]);
            }
            if constexpr (J2) {
                double yzm[D];
#pragma unroll
                for (int k2 = 0; k2 < D; ++k2) {
                    double s = 0.0;
#pragma unroll
                    for (int my = 0; my < k2; ++my) s += vy[my] * ((double)(k2 - my) * vz[k2 - my]);
                    yzm[k2] = s;
                }
#pragma unroll
                for (int b = 0; b < D; ++b) {
                    double s = 0.0;
#pragma unroll
                    for (int mx = 0; mx <= b; ++mx) s += vx[mx] * yzm[b - mx];
                    h[NA - 1][b] = fma(z, s, h[NA - 1][b]);
                }
            }
        };
        // the lane's quadruples u = lg, lg + G, ... in increasing u
        for (uint32_t u = lg; u < nq4; u += G) {
            // every load issued so far has arrived (the step's entry and Z):
            // said explicitly, so that the loads issued below are not waited
            // for before the contraction (without it the compiler's wait
            // before the first use of zc also covered them: 8.02 against
            // 7.57 ms).  vmcnt(0); expcnt and lgkmcnt at their maxima.
            __builtin_amdgcn_s_waitcnt(0x0F70);
            const double zn = u + G < nq4 ? zget(A, ZV, wn) : 0.0;
            const uint64_t w2 = u + 2 * G < nq4 ? qt[u + 2 * G] : 0ull;
            if (!(fabs(zc) < A.cutoffz)) {
                any = 1;
                quad(wc, zc);
            }
            wc = wn;
            zc = zn;
            wn = w2;
        }
        // the values of h(beta) (j0) or of the j2 series / h = 0 polynomial
        // coefficients (DO of them), given the reduced sums in hr / gr
        constexpr int DO = J2 ? D + 2 : D;
        double* rec = hsb + kq * HS_ROW(DO);
        // the series kernel's per-quartet scalars, written by one lane
        auto tail = [&]() {
            const double hq = sqrt(rx * rx + ry * ry + rz * rz);
            rec[DO] = hq;
            rec[DO + 1] = br ? 0.0 : rcp64(hq);
            rec[DO + 2] = __longlong_as_double((long long)HS_META(pb, pk, QE_MULT(ent), br));
        };
        auto coef = [&](int b, const double* hr, const double* grr) -> double {
            if constexpr (!J2) {
                return hr[b];
            } else {
                const double r5 = 1.2615662610100802;  // sqrt(5/pi)
                if (br) return b < D ? 0.5 * r5 * ((double)b * hr[b] - 3.0 * grr[b]) : 0.0;
                const double eta = 3.0 * rz * rz - (rx * rx + ry * ry + rz * rz);
                double v = 0.0;
                if (b >= 2) v = eta * hr[b - 2];
                if (b >= 1 && b - 1 < D) v += 2.0 * (double)(b - 1) * hr[b - 1] - 6.0 * grr[b - 1];
                return 0.25 * r5 * v;
            }
        };
        if constexpr (G <= WAVE) {
            // butterfly over the group's lanes; every lane ends with the sums
#pragma unroll
            for (int k = 0; k < NA; ++k)
#pragma unroll
                for (int b = 0; b < D; ++b) {
                    double v = h[k][b];
#pragma unroll
                    for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
                    h[k][b] = v;
                }
            const uint64_t bal = __ballot(any);
            const uint64_t gmask =
                G == WAVE ? ~0ull : (((1ull << G) - 1ull) << ((lane / G) * G));
            if (valid) {
                // lane lg writes the coefficients lg, lg + G, ... (j2 records
                // of the 4-lane classes have more than G)
#pragma unroll
                for (int b0 = 0; b0 < DO; b0 += G) {
                    double mine = 0.0;
#pragma unroll
                    for (int b = b0; b < DO && b < b0 + G; ++b)
                        if (lg == b - b0) mine = coef(b, h[0], h[NA - 1]);
                    if (b0 + lg < DO) rec[b0 + lg] = mine * (double)QE_MULT(ent);
                }
                if (lg == 0) {
                    tail();
                    A.flags[e] = (bal & gmask) ? 1 : 0;
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < NA; ++k)
#pragma unroll
                for (int b = 0; b < D; ++b) {
                    double v = wsum(h[k][b]);
                    if (lane == 0) red[(k * NW + wid) * D + b] = v;
                }
            int wa = __any(any);
            if (lane == 0) anyw[wid] = wa;
            __syncthreads();
            if (valid && tid < DO) {
                double hr[D], grr[D];
#pragma unroll
                for (int b = 0; b < D; ++b) {
                    double s0 = red[b], s1 = red[((NA - 1) * NW) * D + b];
#pragma unroll
                    for (int w = 1; w < NW; ++w) {
                        s0 += red[w * D + b];
                        s1 += red[((NA - 1) * NW + w) * D + b];
                    }
                    hr[b] = s0;
                    grr[b] = s1;
                }
                double mine = 0.0;
#pragma unroll
                for (int b = 0; b < DO; ++b)
                    if (tid == b) mine = coef(b, hr, grr);
                rec[tid] = mine * (double)QE_MULT(ent);
            }
            if (valid && tid == 0) {
                tail();
                int f = 0;
#pragma unroll
                for (int w = 0; w < NW; ++w) f |= anyw[w];
                A.flags[e] = f ? 1 : 0;
            }
        }
        if (more) {
            fetch_h(base + stride);
            fetch_pg();
        }
    }
}

// The header of each quartet of a class list: its entry, GTO-quadruple
// count (at most 15^4 < 2^16: a group holds at most 15 components) and
// quadruple-table start.
__global__ void __launch_bounds__(256)
    k_qhdr(const int32_t* __restrict__ cls, int64_t n, const uint64_t* __restrict__ qlist,
           const GroupInfo* __restrict__ G, const PairInfo* __restrict__ P, int ng,
           const int64_t* __restrict__ comboff, int S, QHdr* __restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const int32_t e = cls[k];
    const uint64_t ent = qlist[e];
    const int gi = QE_GI(ent), gj = QE_GJ(ent), gk = QE_GK(ent), gr = QE_GR(ent);
    const GroupInfo Gi = G[gi], Gj = G[gj], Gk = G[gk], Gr = G[gr];
    const PairInfo PB = P[pair_index(gi, gj, ng)], PK = P[pair_index(gk, gr, ng)];
    const uint64_t nq4 = (uint64_t)(Gi.count * Gj.count * Gk.count * Gr.count);
    QHdr h;
    h.w = ent | (nq4 << 45);
    h.e = e;
    h.qoff = (uint32_t)comboff[((Gi.pad * S + Gj.pad) * S + Gk.pad) * S + Gr.pad];
    h.rx = PK.Px - PB.Px;
    h.ry = PK.Py - PB.Py;
    h.rz = PK.Pz - PB.Pz;
    // (pair offsets below 2^24: checked where Z is allocated; the elastic
    // kinds read only the L fields)
    h.ob = ((uint32_t)PB.off & 0xFFFFFFu) | ((uint32_t)Gi.L << 24) | ((uint32_t)Gj.L << 28);
    h.ok = ((uint32_t)PK.off & 0xFFFFFFu) | ((uint32_t)Gk.L << 24) | ((uint32_t)Gr.L << 28);
    out[k] = h;
}

hipError_t launch_qhdr(const int32_t* cls, int64_t n, const uint64_t* qlist, const GroupInfo* G,
                       const PairInfo* P, int ng, const int64_t* comboff, int nsig, QHdr* out,
                       hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_qhdr, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, cls, n, qlist,
                       G, P, ng, comboff, nsig, out);
    return hipGetLastError();
}

template <int LL, bool J2>
__global__ void __launch_bounds__(ContractShape<LL>::CT)
    k_contract(SeriesArgs A, const QHdr* __restrict__ hdr, int64_t nidx, int tmax) {
    contract_groups<LL, J2>(A, A.hs, hdr, nidx, tmax, blockIdx.x, gridDim.x);
}

// The entry of a cross-frame work table that holds block b: the last one
// whose first block is <= b (entries in increasing first block).
template <class W>
__device__ __forceinline__ const W& find_work(const W* __restrict__ work, int nwork, int64_t b) {
    int lo = 0, hi = nwork - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (work[mid].blk0 <= b) lo = mid;
        else hi = mid - 1;
    }
    return work[lo];
}

// One class of many frames in one launch: each frame's quartet groups take
// consecutive blocks (one group per block), so a block's frame is uniform.
template <int LL, bool J2>
__global__ void __launch_bounds__(ContractShape<LL>::CT)
    k_contract_multi(const SeriesArgs* __restrict__ args, const ClassWork* __restrict__ work,
                     int nwork, int tmax) {
    const ClassWork& W = find_work(work, nwork, blockIdx.x);
    contract_groups<LL, J2>(args[W.frame], W.hs, W.hdr, W.n, tmax,
                            (int64_t)blockIdx.x - W.blk0, (int64_t)1 << 40);
}

// ------------------------------------------------------------------ series
template <int LL>
__device__ __forceinline__ double f_pzero(const double (&hs)[LL + 1], double q) {
    double q2 = q * q;
    double dfi = 1.0;
#pragma unroll
    for (int b = 1; b <= LL; ++b) dfi /= (double)(2 * b + 1);
    double F = hs[LL] * dfi;
#pragma unroll
    for (int b = LL - 1; b >= 0; --b) {
        dfi *= (double)(2 * b + 3);
        F = fma(F, q2, hs[b] * dfi);
    }
    return F;
}

// dphrec's divisors 4 LU^2 - 1 (LU = 0..99, exact in FP64) and their
// correctly rounded reciprocals.  XX / d is evaluated as q0 = XX c followed
// by one FMA-exact residual correction, q0 + (XX - q0 d) c: the correctly
// rounded quotient (identical to the division for every LU here and x from
// 70 to 5000 on 1.96e8 sampled arguments), at three operations instead of a
// full FP64 division.
struct DphTable {
    double d[100], c[100];
};
constexpr DphTable make_dph_table() {
    DphTable t{};
    for (int i = 0; i < 100; ++i) {
        t.d[i] = 4.0 * (double)i * (double)i - 1.0;
        t.c[i] = 1.0 / t.d[i];
    }
    return t;
}
__constant__ const DphTable kDph = make_dph_table();

__device__ __forceinline__ double dph_quot(double XX, int LU) {
    const double d = kDph.d[LU], c = kDph.c[LU];
    const double q0 = XX * c;
    return fma(fma(-q0, d, XX), c, q0);
}

// dphrec with start order 100 followed literally, then bessels0rr's sum.
// Out of line and rarely taken; it reads the quartet's h(beta) row from
// global memory, so that the caller's h(beta) registers are never spilled
// to scratch to pass them.
template <int LL>
__device__ __noinline__ double f_dphrec(const double* __restrict__ hs, double q, double h,
                                        double x, double sx, double cx) {
#pragma clang fp contract(off)
    const double XX = x * x;
    double r2 = 1.0, r1 = 1.0;  // R[i+2], R[i+1] with R[99] = R[98] = 1
    int i = 97;
    for (; i > LL + 1; --i) {
        double r0 = r1 - dph_quot(XX, i + 2) * r2;
        r2 = r1;
        r1 = r0;
    }
    double R[LL + 2];
#pragma unroll
    for (int k = LL + 1; k >= 0; --k) {
        double r0 = r1 - dph_quot(XX, k + 2) * r2;
        r2 = r1;
        r1 = r0;
        R[k] = r0;
    }
    double Aw = R[0] * (cx + x * sx) - XX / 3.0 * R[1] * cx;
    double CU = 1.0 / x;
    double F = sx / x * hs[0];
    double invp = q / h, rb = 1.0;
#pragma unroll
    for (int K = 1; K <= LL + 1; ++K) {
        double W = 2.0 * (K - 1);
        CU = x / (W + 1.0) * CU;
        if (K >= 2) {
            double jb = CU * (R[K - 1] / Aw);
            rb = rb * invp;
            F = F + jb * rb * hs[K - 1];
        }
    }
    return F;
}

// Miller backward recurrence for g_b = j_b(x)/x^b started M orders above LL,
// normalised by g_0 = j0 or by x g_1 = j1, whichever of j0, j1 is larger in
// magnitude: F = sum_b h(b) q^(2b) g_b.  FP64 (small arguments, where the
// terms cancel); fixed trip count so that a wave does not diverge.
template <int LL, int M>
__device__ __forceinline__ double f_miller(const double (&hs)[LL + 1], double q, double x,
                                           double j0, double j1) {
    const double x2 = x * x, q2 = q * q;
    double vp1 = 0.0, v = 1.0;
#pragma unroll
    for (int b = LL + M; b > LL; --b) {
        double vm = fma((double)(2 * b + 1), v, -x2 * vp1);
        vp1 = v;
        v = vm;
    }
    double acc = hs[LL] * v;
#pragma unroll
    for (int b = LL; b >= 1; --b) {
        double vm = fma((double)(2 * b + 1), v, -x2 * vp1);
        vp1 = v;
        v = vm;
        acc = fma(acc, q2, hs[b - 1] * v);
    }
    const bool by0 = fabs(j0) >= fabs(j1);
    return acc * ((by0 ? j0 : j1) * rcp64(by0 ? v : x * vp1));
}

// Upward recurrence on u_b = j_b(x) (q/h)^b, so that the series needs no
// power of q/h: since (q/h)/x = 1/h^2,
//   u_{b+1} = (2b+1)/h^2 u_b - (q/h)^2 u_{b-1},  u_0 = j_0,  u_1 = (j_0 - cos x)/h^2.
// cb[b] = (2b+1)/h^2 is the quartet's (uniform over q).  F = sum_b h(b) u_b
// is summed by Clenshaw's recurrence on the same coefficients,
//   y_b = h(b) + cb[b] y_{b+1} - (q/h)^2 y_{b+2}  (b = LL .. 1, y_{LL+1} = y_{LL+2} = 0),
//   F = h(0) u_0 + y_1 u_1 - (q/h)^2 y_2 u_0,
// two FMAs per order instead of the three of recurrence plus sum; against a
// long-double sum, both stay within 4e-15 of sum_b |h(b)| |u_b|_env (the
// Hankel-modulus envelope) for x from LL - 1 to 74, LL = 2..16.
template <int LL>
__device__ __forceinline__ double f_up64(const double (&hs)[LL + 1], const double (&cb)[LL + 1],
                                         double q, double invh, double j0, double cx) {
    if constexpr (LL == 0) {
        return j0 * hs[0];
    } else {
        const double r = q * invh;
        const double nr2 = -(r * r);
        double y1 = hs[LL], y2 = 0.0;  // y_{b+1}, y_{b+2}
        if constexpr (LL >= 2) {
            y2 = y1;
            y1 = fma(cb[LL - 1], y2, hs[LL - 1]);
#pragma unroll
            for (int b = LL - 2; b >= 1; --b) {
                const double y0 = fma(cb[b], y1, fma(nr2, y2, hs[b]));
                y2 = y1;
                y1 = y0;
            }
        }
        const double u1 = (j0 - cx) * cb[0];
        if constexpr (LL == 1) return fma(y1, u1, j0 * hs[0]);
        return fma(y1, u1, j0 * fma(nr2, y2, hs[0]));
    }
}

// Mixed precision: the upward recurrence of f_up64 in FP32, on the same
// u_b = j_b(x) (q/h)^b with the quartet's coefficients hf and cbf = (2b+1)/h^2
// rounded once per record.  sin, cos x come from the FP64 phase (an FP32
// argument x would carry an absolute error of 6e-8 x into them).  Fa
// receives sum_b |hw_b u_b|, hw_b = (b + 2) |h_b|: the rounding error of u_b
// grows about one unit per recurrence step, so this is the scale of the
// term's FP32 error.
template <int LL>
__device__ __forceinline__ float f_up32(const float (&hf)[LL + 1], const float (&hw)[LL + 1],
                                        const float (&cbf)[LL + 1], float q, float invh,
                                        float x, float sx, float cx, float& Fa) {
    const float inv = __builtin_amdgcn_rcpf(x);
    const float j0 = sx * inv;
    float F = j0 * hf[0];
    Fa = fabsf(j0) * hw[0];
    if (LL >= 1) {
        const float r = q * invh;
        const float nr2 = -(r * r);
        float um = j0, uc = (j0 - cx) * cbf[0];
        F = fmaf(uc, hf[1], F);
        Fa = fmaf(fabsf(uc), hw[1], Fa);
#pragma unroll
        for (int b = 1; b < LL; ++b) {
            const float un = fmaf(cbf[b], uc, nr2 * um);
            um = uc;
            uc = un;
            F = fmaf(uc, hf[b + 1], F);
            Fa = fmaf(fabsf(uc), hw[b + 1], Fa);
        }
    }
    return F;
}

// q slices (of 64 points) per wave: SQ = 4 on the full 256-point groups of a
// chunk, 1 or 2 on its tail, so that short grids do not pad to 256 points

// h = 0 polynomial of total j2: sum_b c(b) q^(2b) / (2b+3)!!
template <int LL>
__device__ __forceinline__ double f_pzero_j2(const double (&hs)[LL + 1], double q) {
    double q2 = q * q;
    double dfi = 1.0;
#pragma unroll
    for (int b = 0; b <= LL; ++b) dfi /= (double)(2 * b + 3);
    double F = hs[LL] * dfi;
#pragma unroll
    for (int b = LL - 1; b >= 0; --b) {
        dfi *= (double)(2 * b + 5);
        F = fma(F, q2, hs[b] * dfi);
    }
    return F;
}

// qinv = 1/q (0 at q = 0), so that 1/x = invh * qinv needs no reciprocal.
// sx, cx = sin, cos of x = q h, computed by the caller.  Points with
// x >= XDPH are left to k_series_far (0 here): the dphrec call would cost
// this kernel about 50 registers per lane for a regime most workloads never
// reach.  MIXED: x > LL + 6 runs in FP32 (hf, cbf: the record's coefficients
// rounded to FP32) and Fa receives the scale of its rounding error; Fa is 0
// on every FP64 branch.
template <int LL, bool MIXED>
__device__ __forceinline__ double f_series(const double (&hs)[LL + 1], const double (&cb)[LL + 1],
                                           const float (&hf)[LL + 1], const float (&hw)[LL + 1],
                                           const float (&cbf)[LL + 1],
                                           double q, double qinv, double h, double invh,
                                           double sx, double cx, double& Fa) {
    const double x = q * h;
    Fa = 0.0;
    if (x >= XDPH) return 0.0;
    if (MIXED && x > (double)LL + 6.0) {
        float fa;
        const float F = f_up32<LL>(hf, hw, cbf, (float)q, (float)invh, (float)x, (float)sx,
                                   (float)cx, fa);
        Fa = (double)fa;
        return (double)F;
    }
    const double inv = invh * qinv;
    // q = 0 (x = 0): the j0 limit 1 instead of 0/0
    const double j0 = x > 0.0 ? sx * inv : 1.0;
    if (LL == 0 || x >= SeriesRegime<LL>::xup) return f_up64<LL>(hs, cb, q, invh, j0, cx);
    return f_miller<LL, SeriesRegime<LL>::nstart - LL>(hs, q, x, j0, (j0 - cx) * inv);
}

// F(q) of one quartet.  J2: the series divided by q^2 (bessels0rr then
// F/mu**2, src/Integrals.f90:5415-5418), whose q -> 0 limit is c(1)/3 (the
// reference's 0/0), or the j2 h = 0 polynomial.
template <int LL, bool MIXED, bool J2>
__device__ __forceinline__ double f_eval(const double (&hs)[LL + 1], const double (&cb)[LL + 1],
                                         const float (&hf)[LL + 1], const float (&hw)[LL + 1],
                                         const float (&cbf)[LL + 1],
                                         double q, double qinv, double h, double invh, int br,
                                         double sx, double cx, double& Fa) {
    if constexpr (J2) {
        if (br) {
            Fa = 0.0;
            return f_pzero_j2<LL>(hs, q);
        }
        const double F =
            f_series<LL, MIXED>(hs, cb, hf, hw, cbf, q, qinv, h, invh, sx, cx, Fa);
        if (q > 0.0) {
            Fa = Fa / (q * q);
            return F / (q * q);
        }
        return LL >= 1 ? hs[LL >= 1 ? 1 : 0] * (1.0 / 3.0) : 0.0;
    } else {
        if (br) {
            Fa = 0.0;
            return f_pzero<LL>(hs, q);
        }
        return f_series<LL, MIXED>(hs, cb, hf, hw, cbf, q, qinv, h, invh, sx, cx, Fa);
    }
}

// lane l's value of v, wave-uniform
__device__ __forceinline__ double rdlane(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
// Row pb of e12 ([pair][nq] doubles) as a raw buffer resource (gfx9 word 3:
// 32-bit dword data, no swizzle); loads past its nq values return 0.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t e12_row(const SeriesArgs& A, int pb) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)(A.e12 + (int64_t)pb * A.nq), (short)0,
                                             A.nq * 8, 0x00020000);
}
__device__ __forceinline__ double buf_ld64(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, byte_off, 0, 0);
    return __longlong_as_double((long long)(((unsigned long long)v.y << 32) | v.x));
}

// Chunk `chunk` of nchunk (the records of one wave) of a class whose records
// are hsb (nidx of them), on q slice group ys of the points [qs0, qs1) of
// the chunk of q; the sums go to partial row `chunk`.
template <int LL, bool MIXED, bool J2, int SQ>
__device__ __forceinline__ void series_chunk(const SeriesArgs& A, const double* __restrict__ hsb,
                                             int64_t nidx, int first, int64_t chunk,
                                             int64_t nchunk, int ys, int qs0, int qs1) {
    constexpr int R = HS_ROW(LL + 1);
    constexpr int NB = WAVE;  // records per batch (one phase step per lane)
    const int lane = threadIdx.x;
    int iqs[SQ];
    uint32_t qbyte[SQ];
    double qv[SQ], qiv[SQ], S[SQ], Sc[SQ], Sa2[SQ];
#pragma unroll
    for (int s = 0; s < SQ; ++s) {
        const int iql = qs0 + (ys * SQ + s) * WAVE + lane;
        iqs[s] = A.q0 + (iql < qs1 ? iql : qs0);
        qbyte[s] = (uint32_t)iqs[s] * 8u;
        qv[s] = A.q[iqs[s]];
        qiv[s] = qv[s] > 0.0 ? rcp64(qv[s]) : 0.0;
        S[s] = Sc[s] = Sa2[s] = 0.0;
    }
    // a contiguous chunk of the class's records per wave.  The records are
    // wave-uniform: read by scalar loads (through the constant address space;
    // the contraction wrote them in an earlier kernel), the next record's
    // requested before the current one is evaluated, so that no vector
    // register, LDS broadcast or readfirstlane carries them (XY6 series alone
    // 3.04 -> 2.91 ms against records staged through LDS)
    typedef const __attribute__((address_space(4))) double* crec_p;
    const int64_t k0 = chunk * nidx / nchunk, k1 = (chunk + 1) * nidx / nchunk;
    for (int64_t kb = k0; kb < k1; kb += NB) {
        const int nb = (int)(k1 - kb < NB ? k1 - kb : NB);
        // the phase step sin, cos(64 dq h) between a lane's q slices on a
        // uniform grid, lane j for record kb + j (0 on the h = 0 branch,
        // which uses no sine); once per record and wave rather than per lane
        double psd = 0.0, pcd = 0.0;
        if (SQ > 1 && A.qlin && lane < nb) {
            const double* r = hsb + (kb + lane) * R;
            const uint64_t m = (uint64_t)__double_as_longlong(r[LL + 3]);
            if (!((m >> 44) & 1u)) sincos_cw((double)WAVE * A.dq * r[LL + 1], psd, pcd);
        }
        double nrec[R];  // the next record
        {
            const crec_p rn = (crec_p)(hsb + kb * R);
#pragma unroll
            for (int b = 0; b < R; ++b) nrec[b] = rn[b];
        }
        // plain sums over the batch's records, added to S with compensation
        // once per batch
        double Tb[SQ];
#pragma unroll
        for (int s = 0; s < SQ; ++s) Tb[s] = 0.0;
        for (int j = 0; j < nb; ++j) {
            double hs[LL + 1];
#pragma unroll
            for (int b = 0; b <= LL; ++b) hs[b] = nrec[b];
            const double h = nrec[LL + 1], invh = nrec[LL + 2];
            const uint64_t meta = (uint64_t)__double_as_longlong(nrec[LL + 3]);
            if (j + 1 < nb) {
                const crec_p rn = (crec_p)(hsb + (kb + j + 1) * R);
#pragma unroll
                for (int b = 0; b < R; ++b) nrec[b] = rn[b];
            }
            const int pb = (int)(meta & 0xFFFFFu), pk = (int)((meta >> 20) & 0xFFFFFu);
            const int br = (int)((meta >> 44) & 1u);
            // both pair factors of the lane's q points, issued before the
            // series so that their latency hides behind it (loading them one
            // record ahead measured slower: more registers live; XY6 series
            // alone 3.19 against 2.91 ms with two records in scalar registers)
            // (buffer loads: the pair's row as a scalar resource plus the
            // lane's byte offset, no per-record vector address arithmetic)
            const __amdgpu_buffer_rsrc_t eb = e12_row(A, pb), ek = e12_row(A, pk);
            double ebv[SQ], ekv[SQ];
#pragma unroll
            for (int s = 0; s < SQ; ++s) {
                ebv[s] = buf_ld64(eb, qbyte[s]);
                ekv[s] = buf_ld64(ek, qbyte[s]);
            }
            const double ih2 = invh * invh;
            double cb[LL + 1];
#pragma unroll
            for (int b = 0; b <= LL; ++b) cb[b] = (double)(2 * b + 1) * ih2;
            // mixed: the FP32 regime's coefficients, rounded once per record
            float hf[LL + 1], hw[LL + 1], cbf[LL + 1];
#pragma unroll
            for (int b = 0; b <= LL; ++b) {
                hf[b] = MIXED ? (float)hs[b] : 0.0f;
                hw[b] = MIXED ? (float)(b + 2) * fabsf(hf[b]) : 0.0f;
                cbf[b] = MIXED ? (float)cb[b] : 0.0f;
            }
            // sin, cos of q h on the lane's SQ q points.  On a uniform grid the
            // points are x0 + s D, D = 64 dq h, so one sincos of x0, one of D and
            // rotations give them all; a q[0] off the grid gets its own.
            double sxs[SQ], cxs[SQ];
#pragma unroll
            for (int s = 0; s < SQ; ++s) sxs[s] = cxs[s] = 0.0;
            if (!br) {
                if (SQ > 1 && A.qlin) {
                    const bool off = A.q0off && iqs[0] == 0;
                    double sb, cbb;
                    sincos_cw(off ? A.qa * h : qv[0] * h, sb, cbb);
                    sxs[0] = sb;
                    cxs[0] = cbb;
                    if (off) sincos_cw(qv[0] * h, sxs[0], cxs[0]);
                    const double sd = rdlane(psd, j), cd = rdlane(pcd, j);
#pragma unroll
                    for (int s = 1; s < SQ; ++s) {
                        const double ns = fma(sb, cd, cbb * sd), nc = fma(cbb, cd, -(sb * sd));
                        sb = ns;
                        cbb = nc;
                        sxs[s] = sb;
                        cxs[s] = cbb;
                    }
                } else {
#pragma unroll
                    for (int s = 0; s < SQ; ++s) sincos_cw(qv[s] * h, sxs[s], cxs[s]);
                }
            }
#pragma unroll
            for (int s = 0; s < SQ; ++s) {
                double Fa;
                const double F = f_eval<LL, MIXED, J2>(hs, cb, hf, hw, cbf, qv[s], qiv[s], h,
                                                       invh, br, sxs[s], cxs[s], Fa);
                double v = F * ebv[s];
                v = v * ekv[s];
                Tb[s] += v;
                // error scale of the FP32 terms: |weight| sum_b |h_b u_b|
                if (MIXED) {
                    const double e = fabs(ebv[s] * ekv[s]) * Fa;
                    Sa2[s] = fma(e, e, Sa2[s]);
                }
            }
        }
#pragma unroll
        for (int s = 0; s < SQ; ++s) kahan(S[s], Sc[s], Tb[s]);
    }
#pragma unroll
    for (int s = 0; s < SQ; ++s) {
        const int iql = qs0 + (ys * SQ + s) * WAVE + lane;
        if (iql >= qs1) continue;
        int64_t o = chunk * A.nqc + iql;
        if (MIXED) A.parta[o] = first ? Sa2[s] : A.parta[o] + Sa2[s];
        if (first) {
            A.part[o] = S[s];
            A.partc[o] = Sc[s];
        } else {
            double ps = A.part[o], pc = A.partc[o];
            kahan(ps, pc, S[s]);
            kahan(ps, pc, -Sc[s]);
            A.part[o] = ps;
            A.partc[o] = pc;
        }
    }
}

template <int LL, bool MIXED, bool J2, int SQ>
__global__ void __launch_bounds__(WAVE) k_series(SeriesArgs A, int64_t nidx, int first) {
    series_chunk<LL, MIXED, J2, SQ>(A, A.hs, nidx, first, blockIdx.x, gridDim.x, blockIdx.y,
                                    A.qs0, A.qs1);
}

// One class of many frames: each work entry (a frame's class and q range)
// takes nchunk x ny consecutive blocks.
template <int LL, bool J2, int SQ>
__global__ void __launch_bounds__(WAVE)
    k_series_multi(const SeriesArgs* __restrict__ args, const SeriesWork* __restrict__ work,
                   int nwork) {
    const SeriesWork& W = find_work(work, nwork, blockIdx.x);
    const int64_t local = (int64_t)blockIdx.x - W.blk0;
    series_chunk<LL, false, J2, SQ>(args[W.frame], W.hs, W.n, 0, local / W.ny, W.nchunk,
                                    (int)(local % W.ny), W.qs0, W.qs1);
}

// The x = q h >= XDPH points of one class (dphrec), with k_series's launch
// shape and class partition; each wave adds its sums to the partial rows the
// series kernel wrote.  A wave skips every quartet whose h keeps its q points
// below XDPH, so the kernel is a read of the records where the regime is
// rare.
template <int LL, bool J2, int SQ>
__device__ __forceinline__ void far_chunk(const SeriesArgs& A, const double* __restrict__ hsb,
                                          int64_t nidx, int64_t chunk, int64_t nchunk, int ys,
                                          int qs0, int qs1) {
    constexpr int R = HS_ROW(LL + 1);
    const int lane = threadIdx.x;
    int iqs[SQ];
    double qv[SQ], S[SQ], Sc[SQ];
    double qmx = 0.0;
#pragma unroll
    for (int s = 0; s < SQ; ++s) {
        const int iql = qs0 + (ys * SQ + s) * WAVE + lane;
        iqs[s] = A.q0 + (iql < qs1 ? iql : qs0);
        qv[s] = A.q[iqs[s]];
        qmx = fmax(qmx, qv[s]);
        S[s] = Sc[s] = 0.0;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) qmx = fmax(qmx, __shfl_xor(qmx, off, WAVE));
    const int64_t k0 = chunk * nidx / nchunk, k1 = (chunk + 1) * nidx / nchunk;
    for (int64_t kq = k0; kq < k1; ++kq) {
        const double* rc = hsb + kq * R;
        const double h = rc[LL + 1];
        const uint64_t meta = (uint64_t)__double_as_longlong(rc[LL + 3]);
        if (((meta >> 44) & 1u) || qmx * h < XDPH) continue;
        const int pb = (int)(meta & 0xFFFFFu), pk = (int)((meta >> 20) & 0xFFFFFu);
#pragma unroll
        for (int s = 0; s < SQ; ++s) {
            const double x = qv[s] * h;
            if (x < XDPH) continue;
            double sx, cx;
            sincos_cw(x, sx, cx);
            double F = f_dphrec<LL>(rc, qv[s], h, x, sx, cx);
            if constexpr (J2) F = F / (qv[s] * qv[s]);
            double v = F * A.e12[(int64_t)pb * A.nq + iqs[s]];
            v = v * A.e12[(int64_t)pk * A.nq + iqs[s]];
            kahan(S[s], Sc[s], v);
        }
    }
#pragma unroll
    for (int s = 0; s < SQ; ++s) {
        const int iql = qs0 + (ys * SQ + s) * WAVE + lane;
        if (iql >= qs1) continue;
        const int64_t o = chunk * A.nqc + iql;
        double ps = A.part[o], pc = A.partc[o];
        kahan(ps, pc, S[s]);
        kahan(ps, pc, -Sc[s]);
        A.part[o] = ps;
        A.partc[o] = pc;
    }
}

template <int LL, bool J2, int SQ>
__global__ void __launch_bounds__(WAVE) k_series_far(SeriesArgs A, int64_t nidx) {
    far_chunk<LL, J2, SQ>(A, A.hs, nidx, blockIdx.x, gridDim.x, blockIdx.y, A.qs0, A.qs1);
}

template <int LL, bool J2, int SQ>
__global__ void __launch_bounds__(WAVE)
    k_series_far_multi(const SeriesArgs* __restrict__ args, const SeriesWork* __restrict__ work,
                       int nwork) {
    const SeriesWork& W = find_work(work, nwork, blockIdx.x);
    const int64_t local = (int64_t)blockIdx.x - W.blk0;
    far_chunk<LL, J2, SQ>(args[W.frame], W.hs, W.n, local / W.ny, W.nchunk, (int)(local % W.ny),
                          W.qs0, W.qs1);
}

// Mixed precision, second pass: the flagged q points (A.qsel[A.pbase + k],
// k < min(*A.nsel - A.pbase, nqc)) evaluated in FP64 in every regime,
// dphrec's included, from the class's records.  The lanes of a wave split
// into NF point lanes (NF the power of two covering the points, at most 64)
// times 64 / NF quartet lanes; each lane walks its quartets in order, and
// the quartet lanes of one point are combined in a fixed butterfly, so the
// sums are deterministic.  Row w of part / partc holds wave w's sums.
template <int LL, bool J2>
__global__ void __launch_bounds__(WAVE) k_series_pts(SeriesArgs A, int64_t nidx) {
    constexpr int R = HS_ROW(LL + 1);
    const int lane = threadIdx.x;
    const int left = *A.nsel - A.pbase;
    const int np = left < A.nqc ? left : A.nqc;
    if (np <= 0) return;
    int NF = 1;
    while (NF < np && NF < WAVE) NF <<= 1;
    const int NJ = WAVE / NF, k = lane & (NF - 1), jl = lane / NF;
    const int64_t k0 = (int64_t)blockIdx.x * nidx / gridDim.x,
                  k1 = ((int64_t)blockIdx.x + 1) * nidx / gridDim.x;
    const float zf[LL + 1] = {};
    for (int p0 = 0; p0 < np; p0 += NF) {
        const int pt = p0 + k;
        const bool live = pt < np;
        const int iq = A.qsel[A.pbase + (live ? pt : 0)];
        const double q = A.q[iq], qinv = q > 0.0 ? rcp64(q) : 0.0;
        double S = 0.0, Sc = 0.0;
        for (int64_t kq = k0 + jl; kq < k1; kq += NJ) {
            const double* rc = A.hs + kq * R;
            double hs[LL + 1];
#pragma unroll
            for (int b = 0; b <= LL; ++b) hs[b] = rc[b];
            const double h = rc[LL + 1], invh = rc[LL + 2];
            const uint64_t meta = (uint64_t)__double_as_longlong(rc[LL + 3]);
            const int pb = (int)(meta & 0xFFFFFu), pk = (int)((meta >> 20) & 0xFFFFFu);
            const int br = (int)((meta >> 44) & 1u);
            const double ih2 = invh * invh;
            double cb[LL + 1];
#pragma unroll
            for (int b = 0; b <= LL; ++b) cb[b] = (double)(2 * b + 1) * ih2;
            double F, Fa;
            const double x = q * h;
            if (!br && x >= XDPH) {
                double sx, cx;
                sincos_cw(x, sx, cx);
                F = f_dphrec<LL>(rc, q, h, x, sx, cx);
                if constexpr (J2) F = F / (q * q);
            } else {
                double sx = 0.0, cx = 0.0;
                if (!br) sincos_cw(x, sx, cx);
                F = f_eval<LL, false, J2>(hs, cb, zf, zf, zf, q, qinv, h, invh, br, sx, cx, Fa);
            }
            double v = F * A.e12[(int64_t)pb * A.nq + iq];
            v = v * A.e12[(int64_t)pk * A.nq + iq];
            kahan(S, Sc, v);
        }
#pragma unroll
        for (int off = 1; off < WAVE; off <<= 1) {
            if (off < NF) continue;
            const double os = __shfl_xor(S, off, WAVE), oc = __shfl_xor(Sc, off, WAVE);
            kahan(S, Sc, os);
            kahan(S, Sc, -oc);
        }
        if (jl == 0 && live) {
            const int64_t o = (int64_t)blockIdx.x * A.nqc + pt;
            double ps = A.part[o], pc = A.partc[o];
            kahan(ps, pc, S);
            kahan(ps, pc, -Sc);
            A.part[o] = ps;
            A.partc[o] = pc;
        }
    }
}

// One workgroup: the q points with |S(q)| < kappa sqrt(A2(q)), in
// increasing q order (a NaN S(q) is flagged as well).
#define QF_T 1024
__global__ void __launch_bounds__(QF_T) k_qflag(const double* __restrict__ S,
                                                const double* __restrict__ Aq, int nq,
                                                double kappa, int32_t* __restrict__ qsel,
                                                int32_t* __restrict__ nsel) {
    __shared__ int wsum_[QF_T / WAVE];
    __shared__ int base;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    if (tid == 0) base = 0;
    __syncthreads();
    for (int q0 = 0; q0 < nq; q0 += QF_T) {
        const int iq = q0 + tid;
        const bool f = iq < nq && Aq[iq] > 0.0 && !(fabs(S[iq]) >= kappa * sqrt(Aq[iq]));
        const uint64_t bal = __ballot(f);
        const int below = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wsum_[wid] = __popcll(bal);
        __syncthreads();
        int off = base;
        for (int w = 0; w < wid; ++w) off += wsum_[w];
        if (f) qsel[off + below] = iq;
        __syncthreads();
        if (tid == 0) {
            int t = 0;
            for (int w = 0; w < QF_T / WAVE; ++w) t += wsum_[w];
            base += t;
        }
        __syncthreads();
    }
    if (tid == 0) *nsel = base;
}

__global__ void k_qscatter(const double* __restrict__ Sfix, const int32_t* __restrict__ qsel,
                           const int32_t* __restrict__ nsel, int pbase, int nqc,
                           double* __restrict__ S) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int left = *nsel - pbase;
    if (k < nqc && k < left) S[qsel[pbase + k]] = Sfix[k];
}

// --------------------------------------------- electron-nucleus (types 5/6)
// integral_k_ij / integral_ij_pzero for the group pair (gi, gj) and each
// nucleus (src/Integrals.f90:6192-6326,4865-4951), summed as
// nuclei_electron_integration does (:2540-2590):
//   h(beta) = sum_{a,b} Z(a,b) (px * py * pz)(beta),
//   px(m)   = sum_T (-1)^T E^x_ab(T) a(T, 2m-T) hx^(2m-T),  h = R_nucleus - P,
// the factorisation of BesselDeriv with Hermite_like_coeffs (which equals
// rrdj0's a(l,p)).  F(q) follows BesselSum (:6329-6368) literally, upward
// recurrence at every q h as the reference does, or the h = 0 polynomial
// for h < cutoffcentre.  The reference visits (gi,gj) and (gj,gi) with the
// same value; a pair gi < gj counts twice.
template <int LP>
__global__ void __launch_bounds__(WAVE) k_nuc(NucArgs A, const int32_t* __restrict__ pairs,
                                              int row0) {
    constexpr int D = LP + 1;
    const int lane = threadIdx.x;
    const int p = pairs[blockIdx.x];
    const PairInfo PP = A.P[p];
    const GroupInfo Gi = A.G[PP.gi], Gj = A.G[PP.gj];
    const double w = PP.gi == PP.gj ? 1.0 : 2.0;
    const int L1 = A.maxl + 1, T1 = 2 * A.maxl + 1;
    double* row = A.part + (int64_t)(row0 + blockIdx.x) * A.nq;
    const double* e12 = A.e12 + (int64_t)p * A.nq;
    for (int iq = lane; iq < A.nq; iq += WAVE) row[iq] = 0.0;
    const int nab = Gi.count * Gj.count;
    for (int na = 0; na < A.natom; ++na) {
        const double rx = A.geom[na] - PP.Px, ry = A.geom[na + A.natom] - PP.Py,
                     rz = A.geom[na + 2 * A.natom] - PP.Pz;
        const double h = sqrt(rx * rx + ry * ry + rz * rz);
        const bool br = h < A.cutoffcentre;
        const double hv3[3] = {br ? 0.0 : rx, br ? 0.0 : ry, br ? 0.0 : rz};
        double hb[D];
#pragma unroll
        for (int b = 0; b < D; ++b) hb[b] = 0.0;
        for (int u = lane; u < nab; u += WAVE) {
            const int a = u / Gj.count, b = u - a * Gj.count;
            const int ia = Gi.start + a, ib = Gj.start + b;
            const double z = A.Z[ia + (int64_t)A.ngto * ib];
            const char4 ca = A.lxyz[ia], cb = A.lxyz[ib];
            const int la[3] = {ca.x, ca.y, ca.z}, lb[3] = {cb.x, cb.y, cb.z};
            double v[3][D];
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                const double* Er = A.E + (((int64_t)p * 3 + ax) * L1 * L1 + la[ax] * L1 + lb[ax]) * T1;
                const int tmax = la[ax] + lb[ax];
                double hp[D];
                hp[0] = 1.0;
#pragma unroll
                for (int k = 1; k < D; ++k) hp[k] = hp[k - 1] * hv3[ax];
#pragma unroll
                for (int m = 0; m < D; ++m) {
                    double s = 0.0;
#pragma unroll
                    for (int T = m; T <= 2 * m && T < D; ++T) {
                        if (T > tmax) continue;
                        const double c = (T & 1) ? -Er[T] : Er[T];
                        s += c * (kAcoef.v[T][2 * m - T] * hp[2 * m - T]);
                    }
                    v[ax][m] = s;
                }
            }
#pragma unroll
            for (int b2 = 0; b2 < D; ++b2) {
                double s = 0.0;
#pragma unroll
                for (int mx = 0; mx <= b2; ++mx) {
                    double t = 0.0;
#pragma unroll
                    for (int my = 0; my <= b2 - mx; ++my) t += v[1][my] * v[2][b2 - mx - my];
                    s += v[0][mx] * t;
                }
                hb[b2] = fma(z, s, hb[b2]);
            }
        }
#pragma unroll
        for (int b2 = 0; b2 < D; ++b2) hb[b2] = wsum(hb[b2]);
        const double zn = A.Zn[na];
        for (int iq = lane; iq < A.nq; iq += WAVE) {
            const double q = A.q[iq];
            double F;
            if (br) {
                F = f_pzero<LP>(hb, q);
            } else if (q == 0.0) {
                F = hb[0];  // the q -> 0 limit, where BesselSum divides 0 by 0
            } else {
                // BesselSum
                const double Pmu = h * q;
                double sx, cx;
                sincos_cw(Pmu, sx, cx);
                double h0 = sx / Pmu;
                F = h0 * hb[0];
                if (LP == 1) {
                    const double h1 = (sx / (Pmu * Pmu) - cx / Pmu) * q / h;
                    F = F + hb[LP >= 1 ? 1 : 0] * h1;
                } else if (LP > 1) {
                    const double muOH = q / h;
                    double h1 = (sx / (Pmu * Pmu) - cx / Pmu) * muOH;
                    F = F + hb[1] * h1;
#pragma unroll
                    for (int ra = 2; ra <= LP; ++ra) {
                        const double hr = ((double)(2 * ra - 1) / Pmu * h1 - h0 * muOH) * muOH;
                        F = F + hr * hb[ra];
                        h0 = h1;
                        h1 = hr;
                    }
                }
            }
            row[iq] = row[iq] + w * (zn * F * e12[iq]);
        }
    }
}

__global__ void k_rowsum(const double* __restrict__ part, int rows, int nq,
                         double* __restrict__ out) {
    const int iq = blockIdx.x * blockDim.x + threadIdx.x;
    if (iq >= nq) return;
    double s = 0.0;
    for (int r = 0; r < rows; ++r) s += part[(int64_t)r * nq + iq];
    out[iq] = s;
}

hipError_t launch_nuc(const NucArgs& a, int LP, const int32_t* pairs, int npairs, int row0,
                      hipStream_t s) {
    if (npairs <= 0) return hipSuccess;
    switch (LP) {
#define N_(k)                                                                                  \
    case k:                                                                                    \
        hipLaunchKernelGGL(k_nuc<k>, dim3(npairs), dim3(WAVE), 0, s, a, pairs, row0);          \
        break;
        N_(0) N_(1) N_(2) N_(3) N_(4) N_(5) N_(6) N_(7) N_(8)
#undef N_
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_rowsum(const double* part, int rows, int nq, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_rowsum, dim3((nq + 127) / 128), dim3(128), 0, s, part, rows, nq, out);
    return hipGetLastError();
}

// ----------------------------------------------------------------- launchers
namespace {
template <int LL>
size_t contract_lds_for(int Tmax) {
    constexpr int D = LL + 1, DP = ContractShape<LL>::DP, NG = ContractShape<LL>::NG,
                  NW = ContractShape<LL>::CT / WAVE;
    // one quartet per block: h and (for j2) g partial sums per wave
    return sizeof(double) * ((size_t)NG * 3 * Tmax * DP + (NG == 1 ? (size_t)2 * NW * D : 0));
}
}  // namespace

#define ALL_LL(X) X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) \
    X(14) X(15) X(16)

size_t contract_lds_bytes(int LL, int Tmax) {
    switch (LL) {
#define C_(k) \
    case k: return contract_lds_for<k>(Tmax);
        ALL_LL(C_)
#undef C_
        default: return 0;
    }
}

hipError_t launch_contract(const SeriesArgs& a, int LL, int Tmax, const QHdr* hdr,
                           int64_t nidx, int max_blocks, hipStream_t s) {
    if (nidx <= 0) return hipSuccess;
    size_t lds = contract_lds_bytes(LL, Tmax);
    switch (LL) {
#define C_(k)                                                                                  \
    case k: {                                                                                  \
        constexpr int NG = ContractShape<k>::NG;                                               \
        int blocks = (int)std::min<int64_t>(max_blocks, (nidx + NG - 1) / NG);                 \
        if (a.j2)                                                                              \
            hipLaunchKernelGGL((k_contract<k, true>), dim3(blocks), dim3(ContractShape<k>::CT),  \
                               lds, s, a, hdr, nidx, Tmax);                                    \
        else                                                                                   \
            hipLaunchKernelGGL((k_contract<k, false>), dim3(blocks), dim3(ContractShape<k>::CT), \
                               lds, s, a, hdr, nidx, Tmax);                                    \
        break;                                                                                 \
    }
        ALL_LL(C_)
#undef C_
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_contract_multi(const SeriesArgs* args, const ClassWork* work, int nwork,
                                 int64_t nblocks, int LL, int tmax, bool j2, hipStream_t s) {
    if (nwork <= 0 || nblocks <= 0) return hipSuccess;
    const size_t lds = contract_lds_bytes(LL, tmax);
    switch (LL) {
#define C_(k)                                                                                   \
    case k:                                                                                     \
        if (j2)                                                                                 \
            hipLaunchKernelGGL((k_contract_multi<k, true>), dim3((unsigned)nblocks),            \
                               dim3(ContractShape<k>::CT), lds, s, args, work, nwork, tmax);    \
        else                                                                                    \
            hipLaunchKernelGGL((k_contract_multi<k, false>), dim3((unsigned)nblocks),           \
                               dim3(ContractShape<k>::CT), lds, s, args, work, nwork, tmax);    \
        break;
        ALL_LL(C_)
#undef C_
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int contract_groups_per_block(int LL) {
    switch (LL) {
#define C_(k) \
    case k: return ContractShape<k>::NG;
        ALL_LL(C_)
#undef C_
        default: return 1;
    }
}

// Raises (never lowers) the dynamic LDS limit of the contraction kernels:
// plans with different needs share the kernels.
hipError_t set_contract_lds_limit(size_t bytes) {
    static std::mutex mu;
    static size_t cur = 0;
    std::lock_guard<std::mutex> lock(mu);
    if (bytes <= cur) return hipSuccess;
#define C_(k)                                                                            \
    {                                                                                    \
        hipError_t e = hipFuncSetAttribute((const void*)k_contract<k, false>,            \
                                           hipFuncAttributeMaxDynamicSharedMemorySize,   \
                                           (int)bytes);                                  \
        if (e != hipSuccess) return e;                                                   \
        e = hipFuncSetAttribute((const void*)k_contract<k, true>,                        \
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes); \
        if (e != hipSuccess) return e;                                                   \
        e = hipFuncSetAttribute((const void*)k_contract_multi<k, false>,                 \
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes); \
        if (e != hipSuccess) return e;                                                   \
        e = hipFuncSetAttribute((const void*)k_contract_multi<k, true>,                  \
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes); \
        if (e != hipSuccess) return e;                                                   \
    }
    ALL_LL(C_)
#undef C_
    cur = bytes;
    return hipSuccess;
}

namespace {
// The sub-ranges [qs0, qs1) of a q chunk and their slices per wave: the
// 256-point groups at SQ = 4 (one sincos and three rotations per lane and
// record), the tail at the SQ that pads it least, so that a 64-point grid
// does not run 256-point waves.  Mixed precision keeps SQ = 4 throughout.
struct QSplit {
    int n, qs0[2], qs1[2], sq[2];
};
QSplit q_split(int nqc, bool only4) {
    QSplit r{};
    const int full = only4 ? nqc : (nqc / 256) * 256;
    int k = 0;
    if (full > 0) {
        r.qs0[k] = 0;
        r.qs1[k] = full;
        r.sq[k] = 4;
        ++k;
    }
    if (full < nqc) {
        const int tail = nqc - full;
        r.qs0[k] = full;
        r.qs1[k] = nqc;
        r.sq[k] = tail <= 64 ? 1 : tail <= 128 ? 2 : tail <= 192 ? 1 : 4;
        ++k;
    }
    r.n = k;
    return r;
}

template <int LL, int SQV>
void series_go(const SeriesArgs& a, int64_t nidx, int waves, int first, hipStream_t s) {
    const dim3 grid(waves, (a.qs1 - a.qs0 + SQV * WAVE - 1) / (SQV * WAVE));
    if constexpr (SQV == 4) {
        if (a.mixed) {
            if (a.j2)
                hipLaunchKernelGGL((k_series<LL, true, true, 4>), grid, dim3(WAVE), 0, s, a, nidx,
                                   first);
            else
                hipLaunchKernelGGL((k_series<LL, true, false, 4>), grid, dim3(WAVE), 0, s, a,
                                   nidx, first);
            return;
        }
    }
    if (a.j2)
        hipLaunchKernelGGL((k_series<LL, false, true, SQV>), grid, dim3(WAVE), 0, s, a, nidx,
                           first);
    else
        hipLaunchKernelGGL((k_series<LL, false, false, SQV>), grid, dim3(WAVE), 0, s, a, nidx,
                           first);
}

template <int LL, int SQV>
void series_far_go(const SeriesArgs& a, int64_t nidx, int waves, hipStream_t s) {
    const dim3 grid(waves, (a.qs1 - a.qs0 + SQV * WAVE - 1) / (SQV * WAVE));
    if (a.j2)
        hipLaunchKernelGGL((k_series_far<LL, true, SQV>), grid, dim3(WAVE), 0, s, a, nidx);
    else
        hipLaunchKernelGGL((k_series_far<LL, false, SQV>), grid, dim3(WAVE), 0, s, a, nidx);
}

template <int LL>
void series_ll(SeriesArgs a, int64_t nidx, int waves, int first, bool far, hipStream_t s) {
    const QSplit sp = q_split(a.nqc, a.mixed);
    for (int k = 0; k < sp.n; ++k) {
        a.qs0 = sp.qs0[k];
        a.qs1 = sp.qs1[k];
        if (far) {
            if (sp.sq[k] == 1) series_far_go<LL, 1>(a, nidx, waves, s);
            else if (sp.sq[k] == 2) series_far_go<LL, 2>(a, nidx, waves, s);
            else series_far_go<LL, 4>(a, nidx, waves, s);
        } else {
            if (sp.sq[k] == 1) series_go<LL, 1>(a, nidx, waves, first, s);
            else if (sp.sq[k] == 2) series_go<LL, 2>(a, nidx, waves, first, s);
            else series_go<LL, 4>(a, nidx, waves, first, s);
        }
    }
}

hipError_t series_any(const SeriesArgs& a, int LL, int64_t nidx, int waves, int first, bool far,
                      hipStream_t s) {
    if (nidx <= 0) return hipSuccess;
    switch (LL) {
#define CASE(k)                                       \
    case k:                                           \
        series_ll<k>(a, nidx, waves, first, far, s);  \
        break;
        CASE(0) CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
        CASE(9) CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
#undef CASE
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
}  // namespace

int series_q_split(int nqc, bool only4, int* qs0, int* qs1, int* sq) {
    const QSplit sp = q_split(nqc, only4);
    for (int k = 0; k < sp.n; ++k) {
        qs0[k] = sp.qs0[k];
        qs1[k] = sp.qs1[k];
        sq[k] = sp.sq[k];
    }
    return sp.n;
}

namespace {
template <int LL, int SQV>
void series_multi_go(const SeriesArgs* args, const SeriesWork* work, int nwork, int64_t nblocks,
                     bool j2, bool far, hipStream_t s) {
    const dim3 grid((unsigned)nblocks);
    if (far) {
        if (j2)
            hipLaunchKernelGGL((k_series_far_multi<LL, true, SQV>), grid, dim3(WAVE), 0, s, args,
                               work, nwork);
        else
            hipLaunchKernelGGL((k_series_far_multi<LL, false, SQV>), grid, dim3(WAVE), 0, s, args,
                               work, nwork);
    } else {
        if (j2)
            hipLaunchKernelGGL((k_series_multi<LL, true, SQV>), grid, dim3(WAVE), 0, s, args, work,
                               nwork);
        else
            hipLaunchKernelGGL((k_series_multi<LL, false, SQV>), grid, dim3(WAVE), 0, s, args,
                               work, nwork);
    }
}
}  // namespace

hipError_t launch_series_multi(const SeriesArgs* args, const SeriesWork* work, int nwork,
                               int64_t nblocks, int LL, int sq, bool j2, bool far,
                               hipStream_t s) {
    if (nwork <= 0 || nblocks <= 0) return hipSuccess;
    switch (LL) {
#define CASE(k)                                                                      \
    case k:                                                                          \
        if (sq == 1) series_multi_go<k, 1>(args, work, nwork, nblocks, j2, far, s);  \
        else if (sq == 2) series_multi_go<k, 2>(args, work, nwork, nblocks, j2, far, s); \
        else series_multi_go<k, 4>(args, work, nwork, nblocks, j2, far, s);          \
        break;
        CASE(0) CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
        CASE(9) CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
#undef CASE
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_series_far(const SeriesArgs& a, int LL, int64_t nidx, int waves,
                             hipStream_t s) {
    return series_any(a, LL, nidx, waves, 0, true, s);
}

hipError_t launch_series_pts(const SeriesArgs& a, int LL, int64_t nidx, int waves, int first,
                             hipStream_t s) {
    (void)first;
    if (nidx <= 0) return hipSuccess;
    switch (LL) {
#define CASE(k)                                                                              \
    case k:                                                                                  \
        if (a.j2)                                                                            \
            hipLaunchKernelGGL((k_series_pts<k, true>), dim3(waves), dim3(WAVE), 0, s, a, nidx); \
        else                                                                                 \
            hipLaunchKernelGGL((k_series_pts<k, false>), dim3(waves), dim3(WAVE), 0, s, a, nidx); \
        break;
        CASE(0) CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
        CASE(9) CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
#undef CASE
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_qflag(const double* S, const double* Aq, int nq, double kappa, int32_t* qsel,
                        int32_t* nsel, hipStream_t s) {
    hipLaunchKernelGGL(k_qflag, dim3(1), dim3(QF_T), 0, s, S, Aq, nq, kappa, qsel, nsel);
    return hipGetLastError();
}

hipError_t launch_qscatter(const double* Sfix, const int32_t* qsel, const int32_t* nsel,
                           int pbase, int nqc, double* S, hipStream_t s) {
    hipLaunchKernelGGL(k_qscatter, dim3((nqc + 255) / 256), dim3(256), 0, s, Sfix, qsel, nsel,
                       pbase, nqc, S);
    return hipGetLastError();
}

hipError_t launch_series(const SeriesArgs& a, int LL, int64_t nidx, int waves, int first,
                         hipStream_t s) {
    return series_any(a, LL, nidx, waves, first, false, s);
}
