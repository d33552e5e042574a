// Device kernels of the MI355X total / elastic j0 scattering path (gfx950).
//
// Stages (reference routine each restates in parentheses):
//   pair tables   -- McMurchie-Davidson E tables, pair centres, e12
//                    (fill_md_table / variables_total, src/Variables.f90)
//   enumeration   -- the five quartet nests of tot_integration
//                    (src/Integrals.f90:79-82,178-180,258-260,340-341,416)
//   transform     -- unique 2-RDM rows, MO-pair matrix, T matrix
//                    (unique_total, z11/z22, src/Variables.f90:1000-1061)
//   series        -- per quartet: 2-RDM block x Hermite contraction -> h(beta),
//                    then the j_beta series over the q grid
//                    (tot_integral_k_ijkr / _pzero, BesselDeriv, rrdj0,
//                     bessels0rr, dphrec; src/Integrals.f90:4708-5226,
//                     6662-6739,7062-7184)
//   reduction     -- fixed-order compensated sum of per-wave partial S(q)
#include "thistle_internal.hpp"

#include <math.h>

#include <algorithm>
#include <climits>
#include <cstdlib>

#define WAVE 64

// pi = acos(-1.0000) in single precision, widened (src/Variables.f90:994)
__constant__ double kPiSp = 3.1415927410125732421875;

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

__device__ __forceinline__ void kahan_add(double& s, double& c, double v) {
#pragma clang fp contract(off)
    double y = v - c;
    double t = s + y;
    c = (t - s) - y;
    s = t;
}

// ------------------------------------------------------------ pair tables
// One top-level E^{l1,l2}_t slice: raise A to l1 with PA, then B to l2 with
// PB, in the operation order of fill_md_table (src/Variables.f90:31-944).
__device__ void md_slice(double PA, double PB, double a, int l1, int l2,
                         double* out) {
#pragma clang fp contract(off)
    double cur[17], nxt[17];
    for (int t = 0; t < 17; ++t) cur[t] = 0.0;
    cur[0] = 1.0;
    int top = 0;
    for (int pass = 0; pass < l1 + l2; ++pass) {
        double PX = pass < l1 ? PA : PB;
        for (int t = 0; t <= top + 1; ++t) {
            double em1 = t > 0 ? cur[t - 1] : 0.0;
            double e0 = t <= top ? cur[t] : 0.0;
            double ep1 = t + 1 <= top ? cur[t + 1] : 0.0;
            double v;
            if (t == 0)
                v = PX * e0 + ep1;
            else if (t == top + 1)
                v = a * em1;
            else if (t == top)
                v = a * em1 + PX * e0;
            else
                v = a * em1 + PX * e0 + (double)(t + 1) * ep1;
            nxt[t] = v;
        }
        top += 1;
        for (int t = 0; t <= top; ++t) cur[t] = nxt[t];
    }
    for (int t = 0; t <= l1 + l2; ++t) out[t] = cur[t];
}

__global__ void k_pair_info(const GroupInfo* __restrict__ G,
                            PairInfo* __restrict__ P, int npair) {
#pragma clang fp contract(off)
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npair) return;
    PairInfo pi = P[p];
    GroupInfo A = G[pi.gi], B = G[pi.gj];
    double gaP = A.ga + B.ga;
    pi.gaP = gaP;
    pi.Px = (A.ga * A.x + B.ga * B.x) / gaP;
    pi.Py = (A.ga * A.y + B.ga * B.y) / gaP;
    pi.Pz = (A.ga * A.z + B.ga * B.z) / gaP;
    double dx = A.x - B.x, dy = A.y - B.y, dz = A.z - B.z;
    double r2 = dx * dx + dy * dy + dz * dz;
    pi.pre = pow(kPiSp / gaP, 1.5);
    pi.kab = exp(-(A.ga * B.ga) / gaP * r2);
    P[p] = pi;
}

// thread per (pair, axis, la, lb)
__global__ void k_pair_E(const GroupInfo* __restrict__ G,
                         const PairInfo* __restrict__ P, int npair, int maxl,
                         double* __restrict__ E) {
#pragma clang fp contract(off)
    const int L1 = maxl + 1, T1 = 2 * maxl + 1;
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t total = (int64_t)npair * 3 * L1 * L1;
    if (idx >= total) return;
    int lb = (int)(idx % L1);
    int64_t t = idx / L1;
    int la = (int)(t % L1);
    t /= L1;
    int ax = (int)(t % 3);
    int p = (int)(t / 3);
    double* out = E + idx * T1;
    for (int k = 0; k < T1; ++k) out[k] = 0.0;
    PairInfo pi = P[p];
    GroupInfo A = G[pi.gi], B = G[pi.gj];
    if (la > A.L || lb > B.L) return;
    double xa = ax == 0 ? A.x : (ax == 1 ? A.y : A.z);
    double xb = ax == 0 ? B.x : (ax == 1 ? B.y : B.z);
    double gaP = A.ga + B.ga;
    double Px = (A.ga * xa + B.ga * xb) / gaP;
    double a = 0.5 / gaP;
    // the higher-l member is raised first; ties go to the lower GTO index,
    // which is the A side because gi <= gj
    if (la >= lb)
        md_slice(Px - xa, Px - xb, a, la, lb, out);
    else
        md_slice(Px - xb, Px - xa, a, lb, la, out);
}

// e12(q,i,j) = (pi_sp/p)^1.5 exp(-q^2/(4p)) exp(-ab/p |AB|^2)
__global__ void k_e12(const PairInfo* __restrict__ P, int npair,
                      const double* __restrict__ q, int nq,
                      double* __restrict__ e12, float* __restrict__ e12f) {
#pragma clang fp contract(off)
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)npair * nq) return;
    int p = (int)(idx / nq), iq = (int)(idx % nq);
    PairInfo pi = P[p];
    double v = pi.pre * exp(-(q[iq] * q[iq]) * 0.25 / pi.gaP) * pi.kab;
    e12[idx] = v;
    if (e12f) e12f[idx] = (float)v;
}

// ------------------------------------------------------------ enumeration
// One 64-thread block per inner loop; entries written in reference order.
__global__ void k_enumerate(const EnumRow* __restrict__ rows,
                            const PairInfo* __restrict__ P, int ng,
                            double cutoffcentre, uint64_t* __restrict__ out) {
#pragma clang fp contract(off)
    EnumRow R = rows[blockIdx.x];
    for (int u = threadIdx.x; u < R.count; u += blockDim.x) {
        int g[4];
        const int mult = row_quartet(R, R.r0 + u, g);
        const int gi = g[0], gj = g[1], gk = g[2], gr = g[3];
        int br = 1;
        if (R.nest != 5) {
            const PairInfo& b = P[pair_index(gi, gj, ng)];
            const PairInfo& k = P[pair_index(gk, gr, ng)];
            double hx = k.Px - b.Px, hy = k.Py - b.Py, hz = k.Pz - b.Pz;
            double h = sqrt(hx * hx + hy * hy + hz * hz);
            br = h < cutoffcentre ? 1 : 0;
        }
        out[R.out + u] = (uint64_t)gi | ((uint64_t)gj << 10) | ((uint64_t)gk << 20) |
                         ((uint64_t)gr << 30) | ((uint64_t)mult << 40) |
                         ((uint64_t)br << 44);
    }
}

// LL class of every quartet position (one 64-thread block per enumeration
// row, as k_enumerate): key[e] = L_i + L_j + L_k + L_r and pos[e] = e for a
// stable sort by class; per block, the class histogram and the largest
// exponent-tuple count prod (L+1) per class (merged with one atomic per
// class); the shell-signature combinations that occur.
__global__ void k_classify(const EnumRow* __restrict__ rows, const int8_t* __restrict__ gL,
                           const int8_t* __restrict__ gsig, int S, uint32_t* __restrict__ key,
                           int32_t* __restrict__ pos, unsigned long long* __restrict__ cnt,
                           int* __restrict__ tmax, uint8_t* __restrict__ used) {
    __shared__ unsigned int hc[18];
    __shared__ int ht[18];
    if (threadIdx.x < 18) {
        hc[threadIdx.x] = 0;
        ht[threadIdx.x] = 0;
    }
    __syncthreads();
    const EnumRow R = rows[blockIdx.x];
    for (int u = threadIdx.x; u < R.count; u += blockDim.x) {
        int g[4];
        row_quartet(R, R.r0 + u, g);
        const int L0 = gL[g[0]], L1 = gL[g[1]], L2 = gL[g[2]], L3 = gL[g[3]];
        const int LL = L0 + L1 + L2 + L3;
        const int T = (L0 + 1) * (L1 + 1) * (L2 + 1) * (L3 + 1);
        const int64_t e = R.out + u;
        key[e] = (uint32_t)LL;
        pos[e] = (int32_t)e;
        atomicAdd(&hc[LL], 1u);
        atomicMax(&ht[LL], T);
        used[((gsig[g[0]] * S + gsig[g[1]]) * S + gsig[g[2]]) * S + gsig[g[3]]] = 1;
    }
    __syncthreads();
    if (threadIdx.x < 18 && hc[threadIdx.x]) {
        atomicAdd(&cnt[threadIdx.x], (unsigned long long)hc[threadIdx.x]);
        atomicMax(&tmax[threadIdx.x], ht[threadIdx.x]);
    }
}

__global__ void k_unpack_quartets(const uint64_t* __restrict__ in, int64_t n,
                                  int32_t* __restrict__ out) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    uint64_t v = in[e];
    int32_t* o = out + 6 * e;
    o[0] = QE_GI(v) + 1;
    o[1] = QE_GJ(v) + 1;
    o[2] = QE_GK(v) + 1;
    o[3] = QE_GR(v) + 1;
    o[4] = QE_MULT(v);
    o[5] = QE_BR(v);
}

// ------------------------------------------------------------------- scan
// Exclusive scan of 0/1 bytes into int64 (three passes, fixed order).
#define SCAN_T 256
#define SCAN_I 8
__global__ void k_scan_blocks(const uint8_t* __restrict__ in, int64_t n,
                              int64_t* __restrict__ out,
                              int64_t* __restrict__ bsum) {
    __shared__ int64_t s[SCAN_T];
    int64_t base = (int64_t)blockIdx.x * SCAN_T * SCAN_I + (int64_t)threadIdx.x * SCAN_I;
    int64_t v[SCAN_I];
    int64_t run = 0;
    for (int k = 0; k < SCAN_I; ++k) {
        int64_t i = base + k;
        v[k] = run;
        run += (i < n) ? (int64_t)in[i] : 0;
    }
    s[threadIdx.x] = run;
    __syncthreads();
    for (int off = 1; off < SCAN_T; off <<= 1) {
        int64_t t = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    int64_t excl = s[threadIdx.x] - run;
    for (int k = 0; k < SCAN_I; ++k) {
        int64_t i = base + k;
        if (i < n) out[i] = v[k] + excl;
    }
    if (threadIdx.x == SCAN_T - 1) bsum[blockIdx.x] = s[threadIdx.x];
}

// Exclusive scan of the nb block sums in place, total to *total.  One
// workgroup of SCAN_SUMS_T threads: a contiguous run of block sums per thread,
// then a scan of the run totals.
#define SCAN_SUMS_T 1024
__global__ void __launch_bounds__(SCAN_SUMS_T) k_scan_sums(int64_t* __restrict__ bsum, int64_t nb,
                                                          int64_t* __restrict__ total) {
    __shared__ int64_t s[SCAN_SUMS_T];
    const int t = threadIdx.x;
    const int64_t b0 = (int64_t)t * nb / SCAN_SUMS_T, b1 = ((int64_t)t + 1) * nb / SCAN_SUMS_T;
    int64_t run = 0;
    for (int64_t b = b0; b < b1; ++b) run += bsum[b];
    s[t] = run;
    __syncthreads();
    for (int off = 1; off < SCAN_SUMS_T; off <<= 1) {
        const int64_t v = t >= off ? s[t - off] : 0;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    int64_t acc = s[t] - run;  // exclusive prefix of this thread's run
    for (int64_t b = b0; b < b1; ++b) {
        const int64_t v = bsum[b];
        bsum[b] = acc;
        acc += v;
    }
    if (t == SCAN_SUMS_T - 1) *total = s[t];
}

__global__ void k_scan_add(int64_t* __restrict__ out, int64_t n,
                           const int64_t* __restrict__ bsum) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] += bsum[i / (SCAN_T * SCAN_I)];
}

__global__ void k_compact(const uint8_t* __restrict__ flags,
                          const uint64_t* __restrict__ in, int64_t n,
                          const int64_t* __restrict__ pos,
                          uint64_t* __restrict__ out) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    if (flags[e]) out[pos[e]] = in[e];
}

// ----------------------------------------------------------------- unique
static __device__ __forceinline__ uint64_t mix64(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

static __device__ __forceinline__ uint64_t row_key(const int32_t* m, int64_t nrow,
                                                   int ncol, int64_t r) {
    uint64_t key = 0;
    int bits = ncol <= 4 ? 16 : 8;
    for (int c = 0; c < ncol; ++c)
        key |= ((uint64_t)(uint32_t)m[r + nrow * c] & ((1ull << bits) - 1)) << (bits * c);
    return key;
}

#define HEMPTY 0xffffffffffffffffull

__global__ void k_hash_init(uint64_t* keys, int32_t* minrow, int32_t* cnt,
                            int64_t hsize) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hsize) return;
    keys[i] = HEMPTY;
    minrow[i] = 0x7fffffff;
    cnt[i] = 0;
}

__global__ void k_hash_insert(const int32_t* __restrict__ m, int64_t nrow,
                              int ncol, uint64_t* keys, int32_t* minrow,
                              int32_t* cnt, int32_t* slot, int64_t hsize) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrow) return;
    uint64_t key = row_key(m, nrow, ncol, r);
    uint64_t h = mix64(key) & (uint64_t)(hsize - 1);
    while (true) {
        unsigned long long prev = atomicCAS((unsigned long long*)&keys[h],
                                            (unsigned long long)HEMPTY,
                                            (unsigned long long)key);
        if (prev == HEMPTY || prev == key) {
            atomicMin(&minrow[h], (int32_t)r);
            atomicAdd(&cnt[h], 1);
            slot[r] = (int32_t)h;
            return;
        }
        h = (h + 1) & (uint64_t)(hsize - 1);
    }
}

__global__ void k_first_flags(const int32_t* __restrict__ slot,
                              const int32_t* __restrict__ minrow, int64_t nrow,
                              uint8_t* __restrict__ flag) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrow) return;
    flag[r] = minrow[slot[r]] == (int32_t)r ? 1 : 0;
}

// rtot(u) = sum of total over the rows equal to the u-th unique row, in row
// order (sum(same*total) of src/unique_module.f90:126-131).
__global__ void k_unique_out(const int32_t* __restrict__ m, int64_t nrow,
                             int ncol, const double* __restrict__ total,
                             const int32_t* __restrict__ slot,
                             const int32_t* __restrict__ minrow,
                             const int32_t* __restrict__ cnt,
                             const int64_t* __restrict__ pos, int32_t* rmat,
                             int64_t ldr, double* rtot, int32_t* iuni,
                             int32_t* irec) {
#pragma clang fp contract(off)
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrow) return;
    int s = slot[r];
    int64_t first = minrow[s];
    if (irec) irec[r] = (int32_t)(pos[first] + 1);
    if (first != r) return;
    int64_t u = pos[r];
    if (iuni) iuni[u] = (int32_t)(r + 1);
    if (rmat)
        for (int c = 0; c < ncol; ++c) rmat[u + ldr * c] = m[r + nrow * c];
    if (rtot) {
        double acc = 0.0;
        int left = cnt[s];
        for (int64_t k = r; k < nrow && left > 0; ++k)
            if (slot[k] == s) {
                acc = acc + total[k];
                --left;
            }
        rtot[u] = acc;
    }
}

// MO-pair 2-RDM: V[p,s,q,r] from unique rows, then
// Gs(P,Q) = (G(P,Q) + G(Q,P))/8 with G summed over the orderings of each
// unordered pair (P = {m1,m2}, Q = {m3,m4}).
__global__ void k_scatter_v(const int32_t* __restrict__ rmat,
                            const double* __restrict__ rtot, int64_t nu,
                            const int64_t* __restrict__ nu_dev, int64_t ldr,
                            int norb, double* __restrict__ V) {
    int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nu || u >= *nu_dev) return;
    int64_t N = norb;
    int p = rmat[u] - 1, s = rmat[u + ldr] - 1, q = rmat[u + 2 * ldr] - 1,
        r = rmat[u + 3 * ldr] - 1;
    V[((p * N + s) * N + q) * N + r] = rtot[u];
}

__device__ __forceinline__ double g_unordered(const double* V, int N, int p,
                                              int s, int q, int r) {
#pragma clang fp contract(off)
    auto at = [&](int a, int b, int c, int d) {
        return V[(((int64_t)a * N + b) * N + c) * N + d];
    };
    double v = at(p, s, q, r);
    if (p != s) v = v + at(s, p, q, r);
    if (q != r) {
        v = v + at(p, s, r, q);
        if (p != s) v = v + at(s, p, r, q);
    }
    return v;
}

__global__ void k_gs(const double* __restrict__ V, int norb,
                     const int2* __restrict__ mopairs, int npmo,
                     double* __restrict__ Gs) {
#pragma clang fp contract(off)
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)npmo * npmo) return;
    int P = (int)(idx % npmo), Q = (int)(idx / npmo);
    int2 a = mopairs[P], b = mopairs[Q];
    double g1 = g_unordered(V, norb, a.x, a.y, b.x, b.y);
    double g2 = g_unordered(V, norb, b.x, b.y, a.x, a.y);
    Gs[P + (int64_t)npmo * Q] = (g1 + g2) * 0.125;
}

// T(P, x) = C(p,a)C(s,b) + C(p,b)C(s,a) for pair-space element x = (a,b);
// column-major with leading dimension npmo; 0 for the padding elements
// (a = b = -1) at the end of the pair space.
__global__ void k_tmat(const double* __restrict__ mmod, int nmo,
                       const int2* __restrict__ mopairs, int npmo,
                       const int32_t* __restrict__ xa,
                       const int32_t* __restrict__ xb, int64_t npp,
                       double* __restrict__ T) {
#pragma clang fp contract(off)
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)npmo * npp) return;
    int P = (int)(idx % npmo);
    int64_t x = idx / npmo;
    int2 ps = mopairs[P];
    int a = xa[x], b = xb[x];
    if (a < 0) {
        T[idx] = 0.0;
        return;
    }
    double ca_p = mmod[ps.x + (int64_t)nmo * a], cb_s = mmod[ps.y + (int64_t)nmo * b];
    double cb_p = mmod[ps.x + (int64_t)nmo * b], ca_s = mmod[ps.y + (int64_t)nmo * a];
    T[idx] = ca_p * cb_s + cb_p * ca_s;
}

// Z(ii,jj) of variables_elastic, src/Variables.f90:1199-1216
__global__ void k_zel(const double* __restrict__ mmod, int nmo, int norb,
                      const double* __restrict__ g, int ngto,
                      double* __restrict__ Z) {
#pragma clang fp contract(off)
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)ngto * ngto) return;
    int ii = (int)(idx % ngto), jj = (int)(idx / ngto);
    double temp = 0.0;
    for (int cc = 0; cc < norb; ++cc)
        for (int cc2 = 0; cc2 < norb; ++cc2) {
            double gv = g[cc + (int64_t)norb * cc2];
            temp = temp + gv * mmod[cc + (int64_t)nmo * ii] * mmod[cc2 + (int64_t)nmo * jj];
            temp = temp + gv * mmod[cc + (int64_t)nmo * jj] * mmod[cc2 + (int64_t)nmo * ii];
        }
    Z[idx] = temp / 2.0;
}

// clears and copies as kernels, so that captured graphs hold no memset nodes
// C(0:norb, a) padded with zeros to NO doubles per GTO a, so that the
// factorised transform stages whole columns into LDS with contiguous loads.
__global__ void k_cpad(const double* __restrict__ C, int nmo, int norb, int ngto, int NO,
                       double* __restrict__ Cp) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ngto * NO) return;
    const int a = i / NO, p = i - a * NO;
    Cp[i] = p < norb ? C[p + (int64_t)nmo * a] : 0.0;
}

// The MO-pair transform factorises because T((p,s),(a,b)) is a sum of
// products of MO coefficients C_pa C_sb (the diagonal p == s counted twice,
// as the ordered pairs (p,s) and (s,p) both reach it).  With M^T(i, x) =
// sum_y T(y, i) Gs(x, y):
//   W3(p, b, i) = sum_s Mo(p, s, i) C_sb,  Mo(p, s, i) = M^T(i, pair(p, s)) (x2 if p == s)
//   Z(i, j)     = sum_p C_{p a_j} W3(p, b_j, i)      (j = primitive pair (a_j, b_j))
// W3 is laid out [p][b][i] (i fastest).  One thread per i, one block row per
// p: Mo(p, :, i) stays in registers while the C columns b are staged through
// LDS ZF_CH at a time (measured faster than scalar loads of each column).
// W3 and Z are far larger than the caches and read back by the next kernel
// from HBM, so both kernels store them non-temporally (XY6's transform 1.70
// -> 1.63 ms).
#define ZF_T 256  // threads per block of the transform kernels
#define ZF_CH 16  // C columns per LDS stage
template <int NO>
__global__ void __launch_bounds__(ZF_T) k_w3(const double* __restrict__ Mt,
                                             const double* __restrict__ Cp, int norb, int ngto,
                                             int64_t npp, double* __restrict__ W3) {
    __shared__ double cs[ZF_CH * NO];
    const int p = blockIdx.y, tid = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * ZF_T + tid;
    const bool live = i < npp;
    double m[NO];
#pragma unroll
    for (int s = 0; s < NO; ++s) {
        m[s] = 0.0;
        if (s < norb && live) {
            const int x = p < s ? pair_index(p, s, norb) : pair_index(s, p, norb);
            const double v = Mt[i + npp * x];
            m[s] = s == p ? 2.0 * v : v;
        }
    }
    double* out = W3 + i + npp * (int64_t)ngto * p;
    for (int b0 = 0; b0 < ngto; b0 += ZF_CH) {
        const int nb = ngto - b0 < ZF_CH ? ngto - b0 : ZF_CH;
        __syncthreads();  // the previous stage is consumed
        for (int t = tid; t < nb * NO; t += ZF_T) cs[t] = Cp[(int64_t)b0 * NO + t];
        __syncthreads();
        for (int bb = 0; bb < nb; ++bb) {
            const double* c = cs + bb * NO;
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
            for (int s = 0; s < NO; s += 4) {
                s0 = fma(m[s], c[s], s0);
                s1 = fma(m[s + 1], c[s + 1], s1);
                s2 = fma(m[s + 2], c[s + 2], s2);
                s3 = fma(m[s + 3], c[s + 3], s3);
            }
            if (live) __builtin_nontemporal_store((s0 + s1) + (s2 + s3), &out[npp * (int64_t)(b0 + bb)]);
        }
    }
}

typedef const __attribute__((address_space(4))) double* cdouble_p;

// Lower triangle Z(i, j), i >= j, column-major with leading dimension npp.
// Block (i tile, GTO b): W3(:, b, i) stays in registers while the block walks
// the columns j = jtab[b][a] (a < aend[b]; j increases with a) of the pairs
// whose second GTO is b.  Writes are contiguous along i.
template <int NO>
__global__ void __launch_bounds__(ZF_T) k_zform(const double* __restrict__ W3,
                                                const double* __restrict__ Cp,
                                                const int32_t* __restrict__ jtab,
                                                const int32_t* __restrict__ aend_of, int norb,
                                                int ngto, int64_t npp, double* __restrict__ Z) {
    const int b = blockIdx.y;
    const int64_t i0 = (int64_t)blockIdx.x * ZF_T;
    const int64_t ilast = i0 + ZF_T - 1;
    const int32_t* jb = jtab + (int64_t)b * ngto;
    if (jb[0] > ilast) return;  // the whole tile lies above the diagonal
    const int64_t i = i0 + threadIdx.x;
    const bool live = i < npp;
    double w[NO];
#pragma unroll
    for (int p = 0; p < NO; ++p)
        w[p] = (live && p < norb) ? W3[i + npp * (b + (int64_t)ngto * p)] : 0.0;
    const int aend = aend_of[b];
    // C through the constant address space: uniform reads become scalar
    // loads even where the compiler sinks the sums into the divergent store
    // (measured faster here than staging C through LDS)
    const cdouble_p Cc = (cdouble_p)Cp;
    for (int a = 0; a < aend; ++a) {
        const int64_t j = jb[a];
        if (j > ilast) break;
        const cdouble_p c = Cc + (int64_t)a * NO;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
        for (int p = 0; p < NO; p += 4) {
            s0 = fma(w[p], c[p], s0);
            s1 = fma(w[p + 1], c[p + 1], s1);
            s2 = fma(w[p + 2], c[p + 2], s2);
            s3 = fma(w[p + 3], c[p + 3], s3);
        }
        if (live && i >= j) __builtin_nontemporal_store((s0 + s1) + (s2 + s3), &Z[i + npp * j]);
    }
}

// The same two contractions fused on the matrix cores, W3 never leaving
// the CU.  Persistent blocks of ZmShape::W waves, one per CU (the LDS holds C whole
// plus one W chunk); each block walks tiles of ZM_TI primitive pairs i,
// largest i first.  Per tile, Mo(p, :, i) of each wave's NO / W MO rows p
// stays in registers as MFMA operands, and per chunk of ZM_BC GTOs b:
//   stage 1  W(p, b, i)  = sum_s C_sb Mo(p, s, i)     (D[b][i] per p)
//   stage 2  Z(i, j(a,b)) = sum_p C_pa W(p, b, i)     (D[a][i] per b)
// both as v_mfma_f64_16x16x4 products (A[m][k] from lane (m = l % 16,
// k = l / 16), B[k][n] from lane (k = l / 16, n = l % 16), D[m][n] in item
// m / 4 of lane n + 16 (m % 4), as measured on gfx950).  Chunks whose first
// pair column lies beyond the tile are skipped: the triangle i >= j bounds
// both stages.
#define ZM_TI 16  // primitive pairs i per tile
#define ZM_BC 16  // GTOs b per chunk
#define ZM_NGMAX 320  // GTOs at most (the LDS bounds ngto below this anyway)
#define ZM_WMAX 16  // waves per block at NO = 32 (8 measured 1.75 against 1.47 ms)
typedef double zdbl4 __attribute__((ext_vector_type(4)));

template <int NO>
struct ZmShape {
    static constexpr int W = NO >= 32 ? ZM_WMAX : NO / 2;  // waves per block
    static constexpr int CS = NO + 2;               // C row stride in LDS (doubles)
    static constexpr int WB = NO * ZM_TI + 16;      // W chunk stride per b (doubles)
    static constexpr int PW = NO / W;               // MO rows p per wave
    static constexpr int KS = NO / 4;               // MFMA k steps
};

static inline size_t zm_lds_bytes(int NO, int ngto) {
    const int ng16 = (ngto + 15) & ~15;
    return sizeof(double) * ((size_t)ng16 * (NO + 2) + (size_t)ZM_BC * (NO * ZM_TI + 16)) +
           sizeof(int32_t) * (size_t)ZM_BC * ng16;
}

template <int NO>
__global__ void __launch_bounds__(ZmShape<NO>::W * 64) k_zmfma(const double* __restrict__ Mt,
                                                     const double* __restrict__ Cp,
                                                     const int32_t* __restrict__ jtab,
                                                     int norb, int ngto, int64_t npp, int ntile,
                                                     double* __restrict__ Z) {
    using S = ZmShape<NO>;
    extern __shared__ double zsm[];
    const int ng16 = (ngto + 15) & ~15;
    double* Cs = zsm;                      // [ng16][CS]: C(p, a) at a * CS + p
    double* Wc = zsm + (size_t)ng16 * S::CS;  // [ZM_BC][WB]: W(p, b, i) at b * WB + p * ZM_TI + i
    // [ZM_BC][ng16]: the chunk's pair columns j(a, b), INT_MAX past aend[b]
    int32_t* Jt = (int32_t*)(Wc + ZM_BC * S::WB);
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int l16 = lane & 15, l4 = lane >> 4;
    for (int t = tid; t < ng16 * S::CS; t += S::W * 64) {
        const int a = t / S::CS, p = t - a * S::CS;
        Cs[t] = (a < ngto && p < NO) ? Cp[(int64_t)a * NO + p] : 0.0;
    }
    __syncthreads();
    for (int tile = ntile - 1 - (int)blockIdx.x; tile >= 0; tile -= (int)gridDim.x) {
        // npp < 2^24 (checked at plan set-up): pair indices fit 32 bits
        const int i0 = tile * ZM_TI, ilast = i0 + ZM_TI - 1;
        const int i = i0 + l16;
        const bool live = i < npp;
        // stage-1 B operands: Mo(p, s, i), k = s
        double mo[S::PW][S::KS];
#pragma unroll
        for (int pr = 0; pr < S::PW; ++pr) {
            const int p = wid * S::PW + pr;
#pragma unroll
            for (int ks = 0; ks < S::KS; ++ks) {
                const int s = l4 + 4 * ks;
                double v = 0.0;
                if (live && p < norb && s < norb) {
                    const int x = p < s ? pair_index(p, s, norb) : pair_index(s, p, norb);
                    v = Mt[i + npp * (int64_t)x];
                    if (s == p) v *= 2.0;
                }
                mo[pr][ks] = v;
            }
        }
        // chunks that reach the triangle: the first column of a chunk's first
        // GTO grows with b, so they are a prefix (one load per lane, a ballot)
        int nchunk = 0;
        for (int c0 = 0; c0 * ZM_BC < ngto; c0 += 64) {
            const int c = c0 + lane;
            const bool in = c * ZM_BC < ngto && jtab[(int64_t)c * ZM_BC * ngto] <= ilast;
            nchunk += __popcll(__ballot(in));
        }
        for (int b0 = 0; b0 < nchunk * ZM_BC; b0 += ZM_BC) {
            {
                double cb[S::KS];
#pragma unroll
                for (int ks = 0; ks < S::KS; ++ks) cb[ks] = Cs[(b0 + l16) * S::CS + l4 + 4 * ks];
#pragma unroll
                for (int pr = 0; pr < S::PW; ++pr) {
                    zdbl4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int ks = 0; ks < S::KS; ++ks)
                        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(cb[ks], mo[pr][ks], acc, 0, 0, 0);
                    const int p = wid * S::PW + pr;
#pragma unroll
                    for (int r = 0; r < 4; ++r) Wc[(4 * r + l4) * S::WB + p * ZM_TI + l16] = acc[r];
                }
            }
            // the chunk's pair columns, loaded while the MFMAs run
            // (jtab holds -1 past aend[b]); independent loads, one row per wave
            for (int bl = wid; bl < ZM_BC; bl += S::W) {
                const int b = b0 + bl;
                const int32_t* row = jtab + (int64_t)b * ngto;
                int v[(ZM_NGMAX + 63) / 64];
#pragma unroll
                for (int k = 0; k < (ZM_NGMAX + 63) / 64; ++k) {
                    const int a = lane + 64 * k;
                    v[k] = (b < ngto && a < ngto) ? row[a] : -1;
                }
#pragma unroll
                for (int k = 0; k < (ZM_NGMAX + 63) / 64; ++k) {
                    const int a = lane + 64 * k;
                    if (a < ng16) Jt[bl * ng16 + a] = v[k] < 0 ? INT_MAX : v[k];
                }
            }
            __syncthreads();
            for (int bl = wid; bl < ZM_BC; bl += S::W) {
                const int32_t* jb = Jt + bl * ng16;
                // a tiles whose first column reaches the triangle (a prefix)
                int nat = 0;
                for (int c0 = 0; c0 * 16 < ng16; c0 += 64) {
                    const int c = c0 + lane;
                    nat += __popcll(__ballot(c * 16 < ng16 && jb[c * 16] <= ilast));
                }
                nat = __builtin_amdgcn_readfirstlane(nat);
                if (nat == 0) continue;
                double wv[S::KS];
#pragma unroll
                for (int ks = 0; ks < S::KS; ++ks)
                    wv[ks] = Wc[bl * S::WB + (l4 + 4 * ks) * ZM_TI + l16];
                for (int at = 0; at < nat; ++at) {
                    const int a0 = 16 * at;
                    double ca[S::KS];
#pragma unroll
                    for (int ks = 0; ks < S::KS; ++ks) ca[ks] = Cs[(a0 + l16) * S::CS + l4 + 4 * ks];
                    int jv[4];
#pragma unroll
                    for (int r = 0; r < 4; ++r) jv[r] = jb[a0 + 4 * r + l4];
                    zdbl4 c0 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int ks = 0; ks < S::KS; ++ks)
                        c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(ca[ks], wv[ks], c0, 0, 0, 0);
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        if (live && i >= jv[r])
                            __builtin_nontemporal_store(
                                c0[r], Z + (uint64_t)(uint32_t)jv[r] * (uint64_t)npp + i);
                }
            }
            __syncthreads();  // the chunk's W is consumed
        }
    }
}

// the fused transform applies while C and one W chunk fit the LDS
bool zfact_fused(int norb, int ngto) {
    if (const char* ev = std::getenv("THISTLE_ZMFMA"))
        if (ev[0] == '0') return false;
    const int NO = zfact_width(norb);
    return NO >= 8 && NO <= 32 && ngto <= ZM_NGMAX && zm_lds_bytes(NO, ngto) <= 160 * 1024;
}

hipError_t zfact_fused_prepare(int norb, int ngto) {
    const int NO = zfact_width(norb);
    const int lds = (int)zm_lds_bytes(NO, ngto);
    switch (NO) {
        case 8: return hipFuncSetAttribute((const void*)k_zmfma<8>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        case 16: return hipFuncSetAttribute((const void*)k_zmfma<16>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        case 32: return hipFuncSetAttribute((const void*)k_zmfma<32>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        default: return hipErrorInvalidValue;
    }
}

int zfact_width(int norb) {
    return norb <= 8 ? 8 : norb <= 16 ? 16 : norb <= 32 ? 32 : norb <= 48 ? 48 : norb <= 64 ? 64 : 0;
}

hipError_t launch_zfact(const double* Mt, const double* C, int nmo, int norb, int ngto,
                        int64_t npp, const int32_t* jtab, const int32_t* aend_of, double* Cp,
                        double* W3, double* Z, int zgrid, hipStream_t s) {
    const int NO = zfact_width(norb);
    if (NO == 0 || npp <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_cpad, dim3((ngto * NO + 255) / 256), dim3(256), 0, s, C, nmo, norb,
                       ngto, NO, Cp);
    if (zgrid > 0) {
        if (NO > 32 || ngto > ZM_NGMAX || zm_lds_bytes(NO, ngto) > 160 * 1024)
            return hipErrorInvalidValue;
        const size_t lds = zm_lds_bytes(NO, ngto);
        const int ntile = (int)((npp + ZM_TI - 1) / ZM_TI);
        const unsigned grid = (unsigned)std::min(ntile, zgrid);
#define ZMF(NOV)                                                                             \
    case NOV:                                                                                \
        hipLaunchKernelGGL(k_zmfma<NOV>, dim3(grid), dim3(ZmShape<NOV>::W * 64), lds, s, Mt, Cp, \
                           jtab, norb, ngto, npp, ntile, Z);                                 \
        break;
        switch (NO) {
            ZMF(8) ZMF(16) ZMF(32)
            default: return hipErrorInvalidValue;
        }
#undef ZMF
        return hipGetLastError();
    }
    const unsigned nt = (unsigned)((npp + ZF_T - 1) / ZF_T);
#define ZF(NOV)                                                                              \
    case NOV:                                                                                \
        hipLaunchKernelGGL(k_w3<NOV>, dim3(nt, (unsigned)norb), dim3(ZF_T), 0, s, Mt, Cp,      \
                           norb, ngto, npp, W3);                                             \
        hipLaunchKernelGGL(k_zform<NOV>, dim3(nt, (unsigned)ngto), dim3(ZF_T), 0, s, W3, Cp,   \
                           jtab, aend_of, norb, ngto, npp, Z);                               \
        break;
    switch (NO) {
        ZF(8) ZF(16) ZF(32) ZF(48) ZF(64)
    }
#undef ZF
    return hipGetLastError();
}

__global__ void k_fill(double* __restrict__ p, int64_t n, double v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_copy(const double* __restrict__ in, double* __restrict__ out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) out[i] = in[i];
}

__global__ void k_to_float(const double* __restrict__ in, float* __restrict__ out,
                           int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

// S(q) from the per-wave compensated partial sums, in a fixed order: the
// waves are split into RED_SEG fixed segments, each summed (Kahan) by one
// thread, and the segment sums are combined in segment order.
#define RED_SEG 32
#define RED_Q 8
__device__ __forceinline__ void reduce_block(const double* __restrict__ part,
                                             const double* __restrict__ partc, int waves,
                                             int nqc, double* __restrict__ out, int64_t qb) {
    __shared__ double ss[RED_SEG][RED_Q], sc[RED_SEG][RED_Q];
    const int ql = threadIdx.x % RED_Q, seg = threadIdx.x / RED_Q;
    const int iq = (int)(qb * RED_Q) + ql;
    double s = 0.0, c = 0.0;
    if (iq < nqc) {
        const int w0 = (int)((int64_t)seg * waves / RED_SEG),
                  w1 = (int)((int64_t)(seg + 1) * waves / RED_SEG);
        for (int w = w0; w < w1; ++w) {
            kahan_add(s, c, part[(int64_t)w * nqc + iq]);
            kahan_add(s, c, -partc[(int64_t)w * nqc + iq]);
        }
    }
    ss[seg][ql] = s;
    sc[seg][ql] = c;
    __syncthreads();
    if (seg == 0 && iq < nqc) {
        double S = 0.0, C = 0.0;
        for (int g = 0; g < RED_SEG; ++g) {
            kahan_add(S, C, ss[g][ql]);
            kahan_add(S, C, -sc[g][ql]);
        }
        out[iq] = S;
    }
}

__global__ void __launch_bounds__(RED_SEG * RED_Q) k_reduce(const double* __restrict__ part,
                                                           const double* __restrict__ partc,
                                                           int waves, int nqc,
                                                           double* __restrict__ out) {
    reduce_block(part, partc, waves, nqc, out, blockIdx.x);
}

template <class W>
__device__ __forceinline__ const W& find_entry(const W* __restrict__ work, int nwork, int64_t b) {
    int lo = 0, hi = nwork - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (work[mid].blk0 <= b) lo = mid;
        else hi = mid - 1;
    }
    return work[lo];
}

// k_reduce of many frames in one launch (the same order of additions)
__global__ void __launch_bounds__(RED_SEG * RED_Q)
    k_reduce_multi(const RedWork* __restrict__ work, int nwork) {
    const RedWork& W = find_entry(work, nwork, blockIdx.x);
    reduce_block(W.part, W.partc, W.waves, W.nqc, W.out, (int64_t)blockIdx.x - W.blk0);
}

// zero many buffers in one launch (256 doubles per block)
__global__ void k_fill_multi(const FillWork* __restrict__ work, int nwork) {
    const FillWork& W = find_entry(work, nwork, blockIdx.x);
    const int64_t i = ((int64_t)blockIdx.x - W.blk0) * 256 + threadIdx.x;
    if (i < W.n) W.p[i] = 0.0;
}

// ---------------------------------------------------------------- launchers
static inline unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

hipError_t launch_pair_info(const GroupInfo* G, PairInfo* P, int npair, hipStream_t s) {
    hipLaunchKernelGGL(k_pair_info, dim3(nblk(npair, 128)), dim3(128), 0, s, G, P, npair);
    return hipGetLastError();
}

hipError_t launch_pair_E(const GroupInfo* G, const PairInfo* P, int npair, int maxl,
                         double* E, hipStream_t s) {
    int64_t n = (int64_t)npair * 3 * (maxl + 1) * (maxl + 1);
    hipLaunchKernelGGL(k_pair_E, dim3(nblk(n, 128)), dim3(128), 0, s, G, P, npair, maxl, E);
    return hipGetLastError();
}

hipError_t launch_e12(const PairInfo* P, int npair, const double* q, int nq, double* e12,
                      float* e12f, hipStream_t s) {
    int64_t n = (int64_t)npair * nq;
    hipLaunchKernelGGL(k_e12, dim3(nblk(n, 256)), dim3(256), 0, s, P, npair, q, nq, e12, e12f);
    return hipGetLastError();
}

hipError_t launch_enumerate(const EnumRow* rows, int nrows, const PairInfo* P, int ng,
                            double cutoffcentre, uint64_t* out, hipStream_t s) {
    if (nrows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_enumerate, dim3(nrows), dim3(64), 0, s, rows, P, ng, cutoffcentre, out);
    return hipGetLastError();
}

hipError_t launch_classify(const EnumRow* rows, int nrows, const int8_t* gL, const int8_t* gsig,
                           int S, uint32_t* key, int32_t* pos, unsigned long long* cnt, int* tmax,
                           uint8_t* used, hipStream_t s) {
    if (nrows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_classify, dim3(nrows), dim3(64), 0, s, rows, gL, gsig, S, key, pos, cnt,
                       tmax, used);
    return hipGetLastError();
}

hipError_t launch_unpack_quartets(const uint64_t* in, int64_t n, int32_t* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_unpack_quartets, dim3(nblk(n, 256)), dim3(256), 0, s, in, n, out);
    return hipGetLastError();
}

hipError_t launch_reduce(const double* part, const double* partc, int waves, int nqc,
                         double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce, dim3(nblk(nqc, RED_Q)), dim3(RED_SEG * RED_Q), 0, s, part, partc,
                       waves, nqc, out);
    return hipGetLastError();
}

// out[q] = sum of part[w][q] over the waves in order (the FP32 error scale of
// mixed mode: a magnitude, summed without compensation)
__global__ void k_reduce_plain(const double* __restrict__ part, int waves, int nqc,
                               double* __restrict__ out) {
    const int iq = blockIdx.x * blockDim.x + threadIdx.x;
    if (iq >= nqc) return;
    double s = 0.0;
    for (int w = 0; w < waves; ++w) s += part[(int64_t)w * nqc + iq];
    out[iq] = s;
}

hipError_t launch_reduce_plain(const double* part, int waves, int nqc, double* out,
                               hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_plain, dim3(nblk(nqc, 128)), dim3(128), 0, s, part, waves, nqc,
                       out);
    return hipGetLastError();
}

hipError_t launch_exclusive_scan_u8(const uint8_t* in, int64_t n, int64_t* out,
                                    int64_t* total, hipStream_t s) {
    // bsum lives right after `out` (caller allocates n + nb + 1 int64)
    int64_t nb = (n + SCAN_T * SCAN_I - 1) / (SCAN_T * SCAN_I);
    if (nb == 0) nb = 1;
    int64_t* bsum = out + n;
    hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(SCAN_T), 0, s, in, n, out, bsum);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_SUMS_T), 0, s, bsum, nb, total);
    if (n > 0)
        hipLaunchKernelGGL(k_scan_add, dim3(nblk(n, 256)), dim3(256), 0, s, out, n, bsum);
    return hipGetLastError();
}

hipError_t launch_compact(const uint8_t* flags, const uint64_t* in, int64_t n,
                          int64_t* scratch, uint64_t* out, int64_t* count_dev,
                          hipStream_t s) {
    hipError_t err = launch_exclusive_scan_u8(flags, n, scratch, count_dev, s);
    if (err != hipSuccess) return err;
    if (n > 0)
        hipLaunchKernelGGL(k_compact, dim3(nblk(n, 256)), dim3(256), 0, s, flags, in, n, scratch, out);
    return hipGetLastError();
}

hipError_t launch_unique(const int32_t* mat, int64_t nrow, int ncol, const double* total,
                         UniqueWork& w, int32_t* rmat, double* rtot, int32_t* iuni,
                         int32_t* irec, int64_t* count_dev, hipStream_t s) {
    hipLaunchKernelGGL(k_hash_init, dim3(nblk(w.hsize, 256)), dim3(256), 0, s, w.keys,
                       w.minrow, w.cnt, w.hsize);
    if (nrow > 0)
        hipLaunchKernelGGL(k_hash_insert, dim3(nblk(nrow, 256)), dim3(256), 0, s, mat, nrow,
                           ncol, w.keys, w.minrow, w.cnt, w.slot, w.hsize);
    uint8_t* flags = reinterpret_cast<uint8_t*>(w.keys);  // keys are not needed any more
    if (nrow > 0)
        hipLaunchKernelGGL(k_first_flags, dim3(nblk(nrow, 256)), dim3(256), 0, s, w.slot,
                           w.minrow, nrow, flags);
    hipError_t err = launch_exclusive_scan_u8(flags, nrow, w.scan, count_dev, s);
    if (err != hipSuccess) return err;
    if (nrow > 0)
        hipLaunchKernelGGL(k_unique_out, dim3(nblk(nrow, 256)), dim3(256), 0, s, mat, nrow, ncol,
                           total, w.slot, w.minrow, w.cnt, w.scan, rmat, nrow, rtot, iuni, irec);
    return hipGetLastError();
}

hipError_t launch_gs(const int32_t* rmat, const double* rtot, int64_t nu,
                     const int64_t* nu_dev, int64_t ldr, int norb, double* V, hipStream_t s) {
    if (nu > 0)
        hipLaunchKernelGGL(k_scatter_v, dim3(nblk(nu, 256)), dim3(256), 0, s, rmat, rtot, nu,
                           nu_dev, ldr, norb, V);
    return hipGetLastError();
}

hipError_t launch_gs_matrix(const double* V, int norb, const int2* mopairs, int npmo,
                            double* Gs, hipStream_t s) {
    int64_t n = (int64_t)npmo * npmo;
    hipLaunchKernelGGL(k_gs, dim3(nblk(n, 256)), dim3(256), 0, s, V, norb, mopairs, npmo, Gs);
    return hipGetLastError();
}

hipError_t launch_tmat2(const double* mmod, int nmo, const int2* mopairs, int npmo,
                        const int32_t* xa, const int32_t* xb, int64_t npp, double* T,
                        hipStream_t s) {
    int64_t n = (int64_t)npmo * npp;
    hipLaunchKernelGGL(k_tmat, dim3(nblk(n, 256)), dim3(256), 0, s, mmod, nmo, mopairs, npmo,
                       xa, xb, npp, T);
    return hipGetLastError();
}

hipError_t launch_zel(const double* mmod, int nmo, int norb, const double* onerdm, int ngto,
                      double* Z, hipStream_t s) {
    int64_t n = (int64_t)ngto * ngto;
    hipLaunchKernelGGL(k_zel, dim3(nblk(n, 256)), dim3(256), 0, s, mmod, nmo, norb, onerdm,
                       ngto, Z);
    return hipGetLastError();
}

hipError_t launch_fill(double* p, int64_t n, double v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_copy(const double* in, double* out, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, s, in, out, n);
    return hipGetLastError();
}

hipError_t launch_to_float(const double* in, float* out, int64_t n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_to_float, dim3(nblk(n, 256)), dim3(256), 0, s, in, out, n);
    return hipGetLastError();
}

hipError_t launch_reduce_multi(const RedWork* work, int nwork, int64_t nblocks, hipStream_t s) {
    if (nwork <= 0 || nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduce_multi, dim3((unsigned)nblocks), dim3(RED_SEG * RED_Q), 0, s, work,
                       nwork);
    return hipGetLastError();
}

int reduce_q_per_block() { return RED_Q; }

hipError_t launch_fill_multi(const FillWork* work, int nwork, int64_t nblocks, hipStream_t s) {
    if (nwork <= 0 || nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_multi, dim3((unsigned)nblocks), dim3(256), 0, s, work, nwork);
    return hipGetLastError();
}
