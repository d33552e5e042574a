// Adjoint of the total j0 path: dL/dS(q) -> dL/d(primitive-shell centres)
// and dL/d(2-RDM values).  No reference counterpart (the reference has no
// derivatives); the forward it differentiates is the one of series.hip.
//
// With w(q) = dL/dS(q), one quartet contributes
//   L_Q = sum_q w mult e12_b e12_k F(q),  F = sum_beta h(beta) B_beta(q),
//   B_beta = (q/h)^beta j_beta(qh) = q^(2 beta) g_beta(qh)   (h = 0: q^(2b)/(2b+1)!!)
// and every derivative goes through
//   G_beta = sum_q w mult e12_b e12_k B_beta(q),   beta = 0 .. LL+1
// (k_series_adj; dB_beta/dh = -h B_(beta+1) gives the Bessel part of the
// h derivative).  k_contract_adj recomputes the px tables of k_contract
// together with their derivatives with respect to the three components of
// h = P_k - P_b (through a(l,p) h^p) and of the pair separations A-B, C-D
// (through dE/dX_AB), and forms per GTO quadruple
//   dL/dz = sum_m px(m) Yx(m),  Yx(m) = sum_k yz(k) G(m+k)
// (the same for y and z), which is dL/dZ for the 2-RDM gradient, and
// z sum_m dpx(m) Yx(m) for the geometry.  The e12 factors add
// -2 a b/p (A-B) sum_beta h(beta) G_beta.  Per quartet the derivatives are
// mapped to the four shell centres through P = (a A + b B)/p and reduced per
// shell in a fixed order.
#include "series_math.hpp"

#include <algorithm>

#define GS 18  // stride of the G_beta buffer (beta <= LL+1 <= 17)

// ------------------------------------------------------- dE / dX_AB tables
// Forward-mode derivative of md_slice's recursion (kernels.hip) with
// X_PA = -(b/p) X_AB and X_PB = (a/p) X_AB.
__device__ void md_slice_d(double PA, double PB, double dPA, double dPB, double a, int l1,
                           int l2, double* dout) {
    double cur[17], dc[17];
    for (int t = 0; t < 17; ++t) cur[t] = dc[t] = 0.0;
    cur[0] = 1.0;
    int top = 0;
    for (int pass = 0; pass < l1 + l2; ++pass) {
        const double PX = pass < l1 ? PA : PB, dPX = pass < l1 ? dPA : dPB;
        double nx[17], nd[17];
        for (int t = 0; t <= top + 1; ++t) {
            double em1 = t > 0 ? cur[t - 1] : 0.0, e0 = t <= top ? cur[t] : 0.0,
                   ep1 = t + 1 <= top ? cur[t + 1] : 0.0;
            double dm1 = t > 0 ? dc[t - 1] : 0.0, d0 = t <= top ? dc[t] : 0.0,
                   dp1 = t + 1 <= top ? dc[t + 1] : 0.0;
            nx[t] = a * em1 + PX * e0 + (double)(t + 1) * ep1;
            nd[t] = a * dm1 + PX * d0 + dPX * e0 + (double)(t + 1) * dp1;
        }
        top += 1;
        for (int t = 0; t <= top; ++t) {
            cur[t] = nx[t];
            dc[t] = nd[t];
        }
    }
    for (int t = 0; t <= l1 + l2; ++t) dout[t] = dc[t];
}

// thread per (pair, axis, la, lb); same layout as the E tables
__global__ void k_pair_dE(const GroupInfo* __restrict__ G, const PairInfo* __restrict__ P,
                          int npair, int maxl, double* __restrict__ dE) {
    const int L1 = maxl + 1, T1 = 2 * maxl + 1;
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)npair * 3 * L1 * L1) return;
    int lb = (int)(idx % L1);
    int64_t t = idx / L1;
    int la = (int)(t % L1);
    t /= L1;
    int ax = (int)(t % 3);
    int p = (int)(t / 3);
    double* out = dE + idx * T1;
    for (int k = 0; k < T1; ++k) out[k] = 0.0;
    PairInfo pi = P[p];
    GroupInfo A = G[pi.gi], B = G[pi.gj];
    if (la > A.L || lb > B.L) return;
    double xa = ax == 0 ? A.x : (ax == 1 ? A.y : A.z);
    double xb = ax == 0 ? B.x : (ax == 1 ? B.y : B.z);
    double gaP = A.ga + B.ga;
    double Px = (A.ga * xa + B.ga * xb) / gaP;
    double dPA = -B.ga / gaP, dPB = A.ga / gaP;
    if (la >= lb)
        md_slice_d(Px - xa, Px - xb, dPA, dPB, 0.5 / gaP, la, lb, out);
    else
        md_slice_d(Px - xb, Px - xa, dPB, dPA, 0.5 / gaP, lb, la, out);
}

// ------------------------------------------------------------ series adjoint
// j_beta for beta <= M by dphrec's start-100 recurrence followed literally
// (the forward's x >= XDPH regime), times (q/h)^beta.
template <int M>
__device__ __noinline__ void b_dphrec(double (&B)[M + 1], double q, double h, double x, double sx,
                                      double cx) {
#pragma clang fp contract(off)
    const double XX = x * x;
    double r2 = 1.0, r1 = 1.0;
    for (int i = 97; i > M + 1; --i) {
        double LU = (double)(i + 2);
        double r0 = r1 - XX / (4.0 * LU * LU - 1.0) * r2;
        r2 = r1;
        r1 = r0;
    }
    double R[M + 2];
#pragma unroll
    for (int k = M + 1; k >= 0; --k) {
        double LU = (double)(k + 2);
        double r0 = r1 - XX / (4.0 * LU * LU - 1.0) * r2;
        r2 = r1;
        r1 = r0;
        R[k] = r0;
    }
    double Aw = R[0] * (cx + x * sx) - XX / 3.0 * R[1] * cx;
    double CU = 1.0 / x;
    B[0] = sx / x;
    double invp = q / h, rb = 1.0;
#pragma unroll
    for (int K = 1; K <= M + 1; ++K) {
        double W = 2.0 * (K - 1);
        CU = x / (W + 1.0) * CU;
        if (K >= 2) {
            rb = rb * invp;
            B[K - 1] = CU * (R[K - 1] / Aw) * rb;
        }
    }
}

// B_beta(q), beta = 0..M, in the regimes of the forward series.
template <int M>
__device__ __forceinline__ void b_all(double (&B)[M + 1], double q, double h, double invh,
                                      int br) {
    if (br) {
        const double q2 = q * q;
        double t = 1.0;
        B[0] = 1.0;
#pragma unroll
        for (int b = 1; b <= M; ++b) {
            t = t * q2 * (1.0 / (double)(2 * b + 1));
            B[b] = t;
        }
        return;
    }
    const double x = q * h;
    double sx, cx;
    sincos_cw(x, sx, cx);
    const double inv = rcp64(x);
    const double j0 = x > 0.0 ? sx * inv : 1.0;
    if (x >= XDPH) {
        b_dphrec<M>(B, q, h, x, sx, cx);
        return;
    }
    if (x >= SeriesRegime<M>::xup) {
        const double r = q * invh;
        double jm = j0, jc = (j0 - cx) * inv, rb = r;
        B[0] = j0;
        B[1] = jc * rb;
#pragma unroll
        for (int b = 1; b < M; ++b) {
            double jn = fma((double)(2 * b + 1) * inv, jc, -jm);
            jm = jc;
            jc = jn;
            rb *= r;
            B[b + 1] = jc * rb;
        }
        return;
    }
    // Miller: g_b = j_b / x^b, B_b = q^(2b) g_b
    const double x2 = x * x;
    constexpr int N = SeriesRegime<M>::nstart;
    double vp1 = 0.0, v = 1.0;
#pragma unroll
    for (int b = N; b > M; --b) {
        double vm = fma((double)(2 * b + 1), v, -x2 * vp1);
        vp1 = v;
        v = vm;
    }
    double gg[M + 1];
    gg[M] = v;
#pragma unroll
    for (int b = M; b >= 1; --b) {
        double vm = fma((double)(2 * b + 1), v, -x2 * vp1);
        vp1 = v;
        v = vm;
        gg[b - 1] = v;
    }
    // normalised by j0 or j1 = x g_1, whichever is larger in magnitude (j0
    // alone loses accuracy next to its zeros)
    const double j1 = x > 0.0 ? (j0 - cx) * inv : 0.0;
    const bool by0 = fabs(j0) >= fabs(j1);
    const double nrm = (by0 ? j0 : j1) * rcp64(by0 ? v : x * vp1);
    const double q2 = q * q;
    double qp = 1.0;
#pragma unroll
    for (int b = 0; b <= M; ++b) {
        B[b] = gg[b] * nrm * qp;
        qp *= q2;
    }
}

// one wave per contiguous chunk of the class list; lanes over q
template <int LL>
__global__ void __launch_bounds__(WAVE) k_series_adj(AdjointArgs J, const int32_t* __restrict__ idx,
                                                     int64_t nidx) {
    constexpr int K = LL + 2;
    const SeriesArgs& A = J.s;
    const int lane = threadIdx.x;
    const int64_t per = (nidx + gridDim.x - 1) / gridDim.x;
    const int64_t k0 = (int64_t)blockIdx.x * per;
    const int64_t k1 = k0 + per < nidx ? k0 + per : nidx;
    for (int64_t kq = k0; kq < k1; ++kq) {
        const int64_t e = idx[kq];
        if (!A.flags[e]) continue;  // every block below cutoffz: no gradient
        const uint64_t ent = A.qlist[e];
        const int gi = QE_GI(ent), gj = QE_GJ(ent), gk = QE_GK(ent), gr = QE_GR(ent);
        const int br = QE_BR(ent);
        const double mult = (double)QE_MULT(ent);
        const int pb = pair_index(gi, gj, A.ng), pk = pair_index(gk, gr, A.ng);
        const double hx = A.P[pk].Px - A.P[pb].Px, hy = A.P[pk].Py - A.P[pb].Py,
                     hz = A.P[pk].Pz - A.P[pb].Pz;
        const double h = sqrt(hx * hx + hy * hy + hz * hz);
        const double invh = br ? 0.0 : rcp64(h);
        const double* eb = A.e12 + (int64_t)pb * A.nq;
        const double* ek = A.e12 + (int64_t)pk * A.nq;
        double g[K];
#pragma unroll
        for (int b = 0; b < K; ++b) g[b] = 0.0;
        for (int iq = lane; iq < A.nq; iq += WAVE) {
            const double q = A.q[iq];
            const double wt = J.w[iq] * mult * eb[iq] * ek[iq];
            double B[K];
            b_all<K - 1>(B, q, h, invh, br);
#pragma unroll
            for (int b = 0; b < K; ++b) g[b] = fma(wt, B[b], g[b]);
        }
        double mine = 0.0;
#pragma unroll
        for (int b = 0; b < K; ++b) {
            double v = wsum(g[b]);
            if (lane == b) mine = v;
        }
        if (lane < K) J.gadj[e * GS + lane] = mine;
    }
}

// --------------------------------------------------------- contraction adjoint
// Per group and axis: PX (forward), PXH (d/dh), PXB (d/dX_AB), PXK (d/dX_CD);
// each [3][T][D].  In LDS when they fit, else in a global scratch slice.
template <int LL, bool GLOB>
__global__ void __launch_bounds__(AdjointShape<LL>::CT)
    k_contract_adj(AdjointArgs J, const int32_t* __restrict__ idx, int64_t nidx, int tmax) {
    constexpr int D = LL + 1;
    constexpr int CT = AdjointShape<LL>::CT, G = AdjointShape<LL>::G, NG = AdjointShape<LL>::NG;
    constexpr int NW = CT / WAVE;
    constexpr int NV = 9;  // dh (3), dX_AB (3), dX_CD (3)
    const SeriesArgs& A = J.s;
    extern __shared__ double sm[];
    __shared__ double redw[NW][NV];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    const int grp = tid / G, lg = tid % G;
    const size_t per = (size_t)4 * 3 * tmax * D;
    double* tab = GLOB ? J.scratch + ((size_t)blockIdx.x * NG + grp) * per : sm + (size_t)grp * per;
    const int L1 = A.maxl + 1, T1 = 2 * A.maxl + 1;

    for (int64_t base = (int64_t)blockIdx.x * NG; base < nidx; base += (int64_t)gridDim.x * NG) {
        const int64_t kq = base + grp;
        const int64_t e = idx[kq < nidx ? kq : base];
        const bool valid = kq < nidx;
        const bool live = valid && A.flags[e];
        const uint64_t ent = A.qlist[e];
        const int gi = QE_GI(ent), gj = QE_GJ(ent), gk = QE_GK(ent), gr = QE_GR(ent);
        const int br = QE_BR(ent);
        const int pb = pair_index(gi, gj, A.ng), pk = pair_index(gk, gr, A.ng);
        const PairInfo PB = A.P[pb], PK = A.P[pk];
        const GroupInfo Gi = A.G[gi], Gj = A.G[gj], Gk = A.G[gk], Gr = A.G[gr];
        const int Li = Gi.L, Lj = Gj.L, Lk = Gk.L, Lr = Gr.L;
        const int T = live ? (Li + 1) * (Lj + 1) * (Lk + 1) * (Lr + 1) : 0;
        const double hv[3] = {PK.Px - PB.Px, PK.Py - PB.Py, PK.Pz - PB.Pz};
        double Gb[D + 1];
#pragma unroll
        for (int b = 0; b <= D; ++b) Gb[b] = live ? J.gadj[e * GS + b] : 0.0;

        __syncthreads();
        for (int w = lg; w < 3 * T; w += G) {
            int ax = w / T, t = w % T;
            int ld = t % (Lr + 1), t2 = t / (Lr + 1);
            int lc = t2 % (Lk + 1);
            t2 /= (Lk + 1);
            int lb = t2 % (Lj + 1), la = t2 / (Lj + 1);
            const int64_t ob = (((int64_t)pb * 3 + ax) * L1 * L1 + la * L1 + lb) * T1;
            const int64_t ok = (((int64_t)pk * 3 + ax) * L1 * L1 + lc * L1 + ld) * T1;
            const int lab = la + lb, lcd = lc + ld;
            double eb[D], ek[D], deb[D], dek[D];
#pragma unroll
            for (int s = 0; s < D; ++s) {
                double vb = s <= lab ? A.E[ob + s] : 0.0, dvb = s <= lab ? J.dE[ob + s] : 0.0;
                eb[s] = (s & 1) ? -vb : vb;
                deb[s] = (s & 1) ? -dvb : dvb;
                ek[s] = s <= lcd ? A.E[ok + s] : 0.0;
                dek[s] = s <= lcd ? J.dE[ok + s] : 0.0;
            }
            double cx[D], cb[D], ck[D];
#pragma unroll
            for (int T2 = 0; T2 < D; ++T2) {
                double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
                for (int t3 = 0; t3 <= T2; ++t3) {
                    s0 += eb[t3] * ek[T2 - t3];
                    s1 += deb[t3] * ek[T2 - t3];
                    s2 += eb[t3] * dek[T2 - t3];
                }
                cx[T2] = s0;
                cb[T2] = s1;
                ck[T2] = s2;
            }
            const double hax = br ? 0.0 : (ax == 0 ? hv[0] : (ax == 1 ? hv[1] : hv[2]));
            double hp[D + 1];
            hp[0] = 1.0;
#pragma unroll
            for (int k = 1; k <= D; ++k) hp[k] = hp[k - 1] * hax;
            const size_t row = ((size_t)ax * T + t) * D;
            double* o0 = tab + row;
            double* o1 = tab + (size_t)3 * T * D + row;
            double* o2 = tab + (size_t)6 * T * D + row;
            double* o3 = tab + (size_t)9 * T * D + row;
#pragma unroll
            for (int m = 0; m < D; ++m) {
                double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
                for (int T2 = m; T2 <= 2 * m && T2 < D; ++T2) {
                    const int p = 2 * m - T2;
                    const double ac = kAcoef.v[T2][p];
                    s0 += cx[T2] * (ac * hp[p]);
                    s2 += cb[T2] * (ac * hp[p]);
                    s3 += ck[T2] * (ac * hp[p]);
                    if (p >= 1) s1 += cx[T2] * (ac * (double)p * hp[p - 1]);
                }
                o0[m] = s0;
                o1[m] = br ? 0.0 : s1;
                o2[m] = s2;
                o3[m] = s3;
            }
        }
        if (GLOB) __threadfence_block();
        __syncthreads();

        const uint32_t nq4 = live ? (uint32_t)(Gi.count * Gj.count * Gk.count * Gr.count) : 0u;
        const int S = A.nsig;
        const uint64_t* qt =
            A.quadtab + A.comboff[((Gi.pad * S + Gj.pad) * S + Gk.pad) * S + Gr.pad];
        double acc[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] = 0.0;
        const ZView ZV = zview(A, pb, pk, PB.off, PK.off, Gi, Gj, Gk, Gr);
        for (uint32_t u = lg; u < nq4; u += G) {
            const uint64_t w = qt[u];
            const int64_t xb = PB.off + (int)(w & 255u);
            const int64_t xk = PK.off + (int)((w >> 8) & 255u);
            const double z = zget(A, ZV, w);
            if (fabs(z) < A.cutoffz) continue;
            const int tt[3] = {(int)((w >> 16) & 1023u), (int)((w >> 26) & 1023u),
                               (int)((w >> 36) & 1023u)};
            double v3[3][D];
#pragma unroll
            for (int ax = 0; ax < 3; ++ax)
#pragma unroll
                for (int m = 0; m < D; ++m) v3[ax][m] = tab[((size_t)ax * T + tt[ax]) * D + m];
            // Y_ax(m) = sum_k conv(other two axes)(k) G(m + k)
            double Y[3][D];
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                const int a1 = ax == 0 ? 1 : 0, a2 = ax == 2 ? 1 : 2;
                double cv[D];
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    double s = 0.0;
#pragma unroll
                    for (int m1 = 0; m1 <= k; ++m1) s += v3[a1][m1] * v3[a2][k - m1];
                    cv[k] = s;
                }
#pragma unroll
                for (int m = 0; m < D; ++m) {
                    double s = 0.0;
#pragma unroll
                    for (int k = 0; k + m < D; ++k) s += cv[k] * Gb[m + k];
                    Y[ax][m] = s;
                }
            }
            double dz = 0.0;
#pragma unroll
            for (int m = 0; m < D; ++m) dz += v3[0][m] * Y[0][m];
            J.Wz[xk + J.ldw * xb] = dz;
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                const double* r1 = tab + ((size_t)(3 + ax) * T + tt[ax]) * D;
                const double* r2 = tab + ((size_t)(6 + ax) * T + tt[ax]) * D;
                const double* r3 = tab + ((size_t)(9 + ax) * T + tt[ax]) * D;
                double s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
                for (int m = 0; m < D; ++m) {
                    s1 += r1[m] * Y[ax][m];
                    s2 += r2[m] * Y[ax][m];
                    s3 += r3[m] * Y[ax][m];
                }
                acc[ax] = fma(z, s1, acc[ax]);
                acc[3 + ax] = fma(z, s2, acc[3 + ax]);
                acc[6 + ax] = fma(z, s3, acc[6 + ax]);
            }
        }
        // reduce the nine derivatives over the quartet's lanes
        if constexpr (G <= WAVE) {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                double s = acc[v];
#pragma unroll
                for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, WAVE);
                acc[v] = s;
            }
        } else {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                double s = wsum(acc[v]);
                if (lane == 0) redw[wid][v] = s;
            }
            __syncthreads();
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                double s = redw[0][v];
#pragma unroll
                for (int w2 = 1; w2 < NW; ++w2) s += redw[w2][v];
                acc[v] = s;
            }
        }
        if (valid && lg == 0) {
            double gh[3] = {acc[0], acc[1], acc[2]}, gb[3] = {acc[3], acc[4], acc[5]},
                   gk2[3] = {acc[6], acc[7], acc[8]};
            if (live) {
                double Cb = 0.0, Ch = 0.0;
#pragma unroll
                for (int b = 0; b < D; ++b) {
                    // the record holds mult h(beta); 1 / mult (1, 2, 4 or 8) is exact
                    const double hb = A.hs[kq * HS_ROW(D) + b] * (1.0 / (double)QE_MULT(ent));
                    Cb = fma(hb, Gb[b], Cb);
                    Ch = fma(hb, Gb[b + 1], Ch);
                }
                const double db[3] = {Gi.x - Gj.x, Gi.y - Gj.y, Gi.z - Gj.z};
                const double dk[3] = {Gk.x - Gr.x, Gk.y - Gr.y, Gk.z - Gr.z};
                const double fb = -2.0 * Gi.ga * Gj.ga / PB.gaP * Cb;
                const double fk = -2.0 * Gk.ga * Gr.ga / PK.gaP * Cb;
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) {
                    if (!br) gh[ax] -= hv[ax] * Ch;
                    gb[ax] += fb * db[ax];
                    gk2[ax] += fk * dk[ax];
                }
            }
            double* o = J.qgrad + e * 12;
            const double ai = Gi.ga / PB.gaP, aj = Gj.ga / PB.gaP, ak = Gk.ga / PK.gaP,
                         ar = Gr.ga / PK.gaP;
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                o[ax] = gb[ax] - ai * gh[ax];
                o[3 + ax] = -gb[ax] - aj * gh[ax];
                o[6 + ax] = gk2[ax] + ak * gh[ax];
                o[9 + ax] = -gk2[ax] + ar * gh[ax];
            }
        }
    }
}

// -------------------------------------------------------- per-shell reduction
// One block per shell; entries e*4+slot of the quartets touching the shell,
// summed in a fixed order.
__global__ void __launch_bounds__(256) k_group_grad(const double* __restrict__ qgrad,
                                                    const int64_t* __restrict__ goff,
                                                    const int64_t* __restrict__ gent,
                                                    double* __restrict__ out) {
    __shared__ double red[256][3];
    const int g = blockIdx.x, tid = threadIdx.x;
    double s[3] = {0.0, 0.0, 0.0};
    for (int64_t k = goff[g] + tid; k < goff[g + 1]; k += 256) {
        const int64_t v = gent[k];
        const double* q = qgrad + (v >> 2) * 12 + (v & 3) * 3;
        s[0] += q[0];
        s[1] += q[1];
        s[2] += q[2];
    }
    for (int c = 0; c < 3; ++c) red[tid][c] = s[c];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w)
            for (int c = 0; c < 3; ++c) red[tid][c] += red[tid + w][c];
        __syncthreads();
    }
    if (tid < 3) out[g * 3 + tid] = red[0][tid];
}

// dL/dtotal(n) = (W(P,Q) + W(Q,P)) / 8 for row n = (p,s,q,r), P = {p,s},
// Q = {q,r}: the adjoint of Gs = (g + g^T)/8 and of the V scatter.
__global__ void k_dtotal(const int32_t* __restrict__ mat, int64_t nmat, int norb,
                         const double* __restrict__ W, int npmo, double* __restrict__ dtot) {
    int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nmat) return;
    int p = mat[n] - 1, s = mat[n + nmat] - 1, q = mat[n + 2 * nmat] - 1, r = mat[n + 3 * nmat] - 1;
    int a = p < s ? p : s, b = p < s ? s : p, c = q < r ? q : r, d = q < r ? r : q;
    int P = pair_index(a, b, norb), Q = pair_index(c, d, norb);
    dtot[n] = (W[P + (int64_t)npmo * Q] + W[Q + (int64_t)npmo * P]) * 0.125;
}

// ----------------------------------------------------------------- launchers
#define ALL_LL(X) X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) \
    X(14) X(15) X(16)

hipError_t launch_pair_dE(const GroupInfo* G, const PairInfo* P, int npair, int maxl, double* dE,
                          hipStream_t s) {
    const int L1 = maxl + 1;
    int64_t n = (int64_t)npair * 3 * L1 * L1;
    hipLaunchKernelGGL(k_pair_dE, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, G, P, npair,
                       maxl, dE);
    return hipGetLastError();
}

hipError_t launch_series_adj(const AdjointArgs& j, int LL, const int32_t* idx, int64_t nidx,
                             int waves, hipStream_t s) {
    if (nidx <= 0) return hipSuccess;
    int w = (int)std::min<int64_t>(waves, nidx);
    switch (LL) {
#define C_(k)                                                                          \
    case k:                                                                            \
        hipLaunchKernelGGL((k_series_adj<k>), dim3(w), dim3(WAVE), 0, s, j, idx, nidx); \
        break;
        ALL_LL(C_)
#undef C_
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

size_t contract_adj_table_doubles(int LL, int Tmax) {
    return (size_t)4 * 3 * Tmax * (LL + 1);
}

int contract_adj_groups(int LL) {
    switch (LL) {
#define C_(k) \
    case k: return AdjointShape<k>::NG;
        ALL_LL(C_)
#undef C_
        default: return 1;
    }
}

hipError_t launch_contract_adj(const AdjointArgs& j, int LL, int Tmax, const int32_t* idx,
                               int64_t nidx, int blocks, bool glob, size_t lds_limit,
                               hipStream_t s) {
    if (nidx <= 0) return hipSuccess;
    const int NG = contract_adj_groups(LL);
    const size_t lds = glob ? 0 : sizeof(double) * NG * contract_adj_table_doubles(LL, Tmax);
    if (!glob && lds > lds_limit) return hipErrorInvalidValue;
    switch (LL) {
#define C_(k)                                                                             \
    case k:                                                                               \
        if (glob)                                                                         \
            hipLaunchKernelGGL((k_contract_adj<k, true>), dim3(blocks),                   \
                               dim3(AdjointShape<k>::CT), 0, s, j, idx, nidx, Tmax);     \
        else                                                                              \
            hipLaunchKernelGGL((k_contract_adj<k, false>), dim3(blocks),                  \
                               dim3(AdjointShape<k>::CT), lds, s, j, idx, nidx, Tmax);   \
        break;
        ALL_LL(C_)
#undef C_
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t set_contract_adj_lds_limit(size_t bytes) {
#define C_(k)                                                                          \
    {                                                                                  \
        hipError_t e = hipFuncSetAttribute((const void*)k_contract_adj<k, false>,      \
                                           hipFuncAttributeMaxDynamicSharedMemorySize, \
                                           (int)bytes);                                \
        if (e != hipSuccess) return e;                                                 \
    }
    ALL_LL(C_)
#undef C_
    return hipSuccess;
}

hipError_t launch_group_grad(const double* qgrad, const int64_t* goff, const int64_t* gent,
                             int ng, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_group_grad, dim3(ng), dim3(256), 0, s, qgrad, goff, gent, out);
    return hipGetLastError();
}

hipError_t launch_dtotal(const int32_t* mat, int64_t nmat, int norb, const double* W, int npmo,
                         double* dtot, hipStream_t s) {
    if (nmat <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dtotal, dim3((unsigned)((nmat + 255) / 256)), dim3(256), 0, s, mat, nmat,
                       norb, W, npmo, dtot);
    return hipGetLastError();
}
