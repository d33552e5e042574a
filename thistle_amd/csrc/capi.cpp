// C ABI and execution plan of the MI355X scattering path.
//
// A plan owns every device buffer of one (basis, 2-RDM, q grid) problem.
// thistle_plan_execute runs the device stages on the plan's stream:
//   1. pair tables   (k_pair_info, k_pair_E, k_e12)
//   2. enumeration   (k_enumerate: the reference's five nests)
//   3. transform     (unique rows -> MO-pair matrix Gs -> T -> Z = T^T Gs T
//                     with a rocBLAS GEMM and a symmetric rank-k update that
//                     writes Z's lower triangle, on the matrix cores), or the
//                     elastic Z of the 1-RDM
//   4. quartet kernels, per LL class: k_contract (h(beta) of each quartet)
//                     then k_series (Bessel series times e12 weights into
//                     per-wave compensated partial sums of S(q))
//   5. reduction of the partial sums + stable compaction of the screened
//                     quartet list
#include <rocblas/rocblas.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "thistle_internal.hpp"

#define HCHECK(x)                                                     \
    do {                                                              \
        hipError_t e_ = (x);                                          \
        if (e_ != hipSuccess) {                                       \
            std::fprintf(stderr, "thistle: %s failed: %s (%s:%d)\n", #x, \
                         hipGetErrorString(e_), __FILE__, __LINE__);  \
            return THISTLE_EDEVICE;                                   \
        }                                                             \
    } while (0)

#define BCHECK(x)                                                       \
    do {                                                                \
        rocblas_status s_ = (x);                                        \
        if (s_ != rocblas_status_success) {                             \
            std::fprintf(stderr, "thistle: %s failed: %d (%s:%d)\n", #x, \
                         (int)s_, __FILE__, __LINE__);                  \
            return THISTLE_EDEVICE;                                     \
        }                                                               \
    } while (0)

namespace {

// Released device blocks of 1 MiB and more are kept, per device, for the
// next allocation of about the same size (repeated calls on one basis reuse
// every plan buffer: hipMalloc and hipFree of the 5.6 GB XY6 Z and the rest
// took 8.8 ms of an 18 ms call).  A block enters the cache after a device
// synchronisation, the one hipFree would have done, so no kernel still uses
// it; the cache holds at most kCacheCap bytes (oldest blocks freed first).
// THISTLE_DEVICE_CACHE=0 frees blocks instead.
constexpr size_t kCacheMin = (size_t)1 << 20;
constexpr size_t kCacheGrain = (size_t)2 << 20;
constexpr size_t kCacheCap = (size_t)64 << 30;
struct BlockCache {
    struct Blk {
        int dev;
        size_t cap;
        void* p;
    };
    std::mutex mu;
    std::vector<Blk> blks;  // oldest first
    size_t held = 0;
};
BlockCache& block_cache() {
    static BlockCache* c = new BlockCache();  // outlives static destruction
    return *c;
}
bool cache_on() {
    static const bool on = [] {
        const char* ev = std::getenv("THISTLE_DEVICE_CACHE");
        return !(ev && ev[0] == '0');
    }();
    return on;
}
size_t cache_round(size_t bytes) {
    return bytes < kCacheMin ? bytes : (bytes + kCacheGrain - 1) / kCacheGrain * kCacheGrain;
}
void cache_release(void* p, size_t cap) {
    int dev = 0;
    if (!cache_on() || cap < kCacheMin || hipGetDevice(&dev) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess) {
        (void)hipFree(p);
        return;
    }
    BlockCache& c = block_cache();
    std::lock_guard<std::mutex> lock(c.mu);
    while (!c.blks.empty() && c.held + cap > kCacheCap) {
        (void)hipFree(c.blks.front().p);
        c.held -= c.blks.front().cap;
        c.blks.erase(c.blks.begin());
    }
    c.blks.push_back({dev, cap, p});
    c.held += cap;
}
// a cached block of this device with cap in [want, want + want / 4], or null
void* cache_take(size_t want, size_t& cap) {
    int dev = 0;
    if (!cache_on() || want < kCacheMin || hipGetDevice(&dev) != hipSuccess) return nullptr;
    BlockCache& c = block_cache();
    std::lock_guard<std::mutex> lock(c.mu);
    int best = -1;
    for (int k = 0; k < (int)c.blks.size(); ++k) {
        const auto& b = c.blks[k];
        if (b.dev == dev && b.cap >= want && b.cap <= want + want / 4 &&
            (best < 0 || b.cap < c.blks[best].cap))
            best = k;
    }
    if (best < 0) return nullptr;
    void* p = c.blks[best].p;
    cap = c.blks[best].cap;
    c.held -= cap;
    c.blks.erase(c.blks.begin() + best);
    return p;
}

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;  // requested
    size_t cap = 0;    // allocated
    ~DevBuf() {
        if (p) cache_release(p, cap);
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

int dalloc(DevBuf& b, size_t bytes) {
    if (b.p) cache_release(b.p, b.cap);
    b.p = nullptr;
    b.bytes = bytes;
    const size_t want = cache_round(bytes == 0 ? 16 : bytes);
    b.p = cache_take(want, b.cap);
    if (b.p) return THISTLE_OK;
    b.cap = want;
    if (hipMalloc(&b.p, want) != hipSuccess) {
        b.p = nullptr;
        b.cap = 0;
        return THISTLE_ENOMEM;
    }
    return THISTLE_OK;
}

// rocBLAS handles, kept per device across plans (creating one costs ms)
struct BlasCache {
    std::mutex mu;
    std::vector<std::pair<int, rocblas_handle>> free;
};
BlasCache& blas_cache() {
    static BlasCache* c = new BlasCache();
    return *c;
}
rocblas_handle blas_take() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    {
        BlasCache& c = blas_cache();
        std::lock_guard<std::mutex> lock(c.mu);
        for (size_t k = 0; k < c.free.size(); ++k)
            if (c.free[k].first == dev) {
                rocblas_handle h = c.free[k].second;
                c.free.erase(c.free.begin() + (long)k);
                return h;
            }
    }
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) return nullptr;
    return h;
}
void blas_give(rocblas_handle h, int dev) {
    BlasCache& c = blas_cache();
    std::lock_guard<std::mutex> lock(c.mu);
    c.free.push_back({dev, h});
}

const ThistleOpts kDefaultOpts = {THISTLE_FP64, THISTLE_RDM_CUTOFF, nullptr, 0, 0};
static_assert(THISTLE_RDM_CUTOFF == (double)1e-30f, "1E-30 as a default real");

// Mixed precision: a q point is re-evaluated in FP64 when |S(q)| <
// kMixedKappa * sqrt(A2(q)), A2(q) = sum over the FP32-evaluated terms of
// e^2, e = |weight| sum_b (b + 2) |h_b u_b| (the term's FP32 error scale).
// The terms' rounding errors are independent, so their sum has the size of
// 2^-24 sqrt(A2); kMixedKappa = 4 * 2^-24 / 1e-6 flags every point where four
// times that could exceed 1e-6 |S(q)|.
constexpr double kMixedKappa = 4.0 * 5.9604644775390625e-08 / 1e-6;

}  // namespace

struct ThistlePlan {
    int type = 1, precision = 0;
    int ngto = 0, ng = 0, maxl = 0, nq = 0, nmo = 0, norb = 0;
    int npair = 0, npmo = 0, waves = 0, waves_alloc = 0, nrows = 0;
    bool zfact = false;  // total Z through the MO factorisation of T (else one dsyrkx)
    int zgrid = 0;       // > 0: that factorisation fused on the matrix cores, this many blocks
    bool far = false;    // some q h may reach the dphrec regime (k_series_far)
    int qlin = 0, q0off = 0;  // q[k] = qa + k dq for k >= 1; q[0] apart
    double qa = 0.0, dq = 0.0;
    int64_t npp = 0, nquart = 0, nmat = 0, units = 0;
    int64_t nlisted = 0;  // quartets in the class lists (nquart, or a subset)
    double cutoffz = 0, cutoffmd = 0, cutoffcentre = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int device = 0;
    rocblas_handle blas = nullptr;
    std::vector<double> h_q;
    // device buffers
    DevBuf G, P, E, e12, q, lxyz, rows, qlist, flags, scan, qcomp, count, qhdr;
    DevBuf mmod, mat, total, rmat, rtot, hkeys, hmin, hcnt, hslot, uscan;
    DevBuf V, Gs, T, M, Z, Cp, W3, jtab, aend, mopairs, xa, xb, onerdm;
    DevBuf part, partc, S, cls;
    // mixed precision: FP32 error scale per wave / per q, the q points
    // re-evaluated in FP64 and their count, the re-evaluated sums
    DevBuf parta, Aq, qsel, nsel, Sfix;
    double kappa = 0.0;
    int32_t last_nsel = 0;
    std::vector<int> gstart, gcount, gL;  // per primitive shell (host)
    int64_t cls_off[18] = {0};  // per-LL ranges in `cls` (positions, ref order)
    int64_t cls_hoff[18] = {0}; // per-LL start of the class's records in `hs` (doubles)
    int cls_tmax[17] = {0};     // largest exponent-tuple count per LL class
    DevBuf hs;                  // h(beta) per quartet
    DevBuf quadtab, comboff;    // GTO-quadruple tables per shell-signature combination
    int nsig = 0;
    hipEvent_t ev[6] = {};
    double ms[5] = {0, 0, 0, 0, 0};
    // contraction kernels (plan stream) and series kernels (series stream)
    // of the last timed execution: first launch to last launch
    hipEvent_t kev[4] = {};
    double kms[2] = {0, 0};
    // series kernels on a second stream, each behind its class's contraction
    hipStream_t aux = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr, cev[17] = {};
    void drop_overlap() {
        for (auto& e : cev)
            if (e) (void)hipEventDestroy(e), e = nullptr;
        if (fork_ev) (void)hipEventDestroy(fork_ev), fork_ev = nullptr;
        if (join_ev) (void)hipEventDestroy(join_ev), join_ev = nullptr;
        if (aux) (void)hipStreamDestroy(aux), aux = nullptr;
    }
    ~ThistlePlan() {
        drop_overlap();
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : kev)
            if (e) (void)hipEventDestroy(e);
        if (blas) {
            // back to the per-device cache (its stream is set again on use)
            if (stream) (void)hipStreamSynchronize(stream);
            (void)rocblas_set_stream(blas, nullptr);
            blas_give(blas, device);
        }
        if (own_stream && stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// Shell signature: the ordered (l,m,n) list of a group's GTOs.  Quartets
// whose four groups have the same signatures share one GTO-quadruple table.
int check_basis(const ThistleBasis* b, std::vector<GroupInfo>& groups, int& maxl_eff,
                std::vector<std::vector<int32_t>>* sigs = nullptr) {
    if (!b || b->ngto <= 0 || b->ng <= 0 || b->nq <= 0 || !b->ga || !b->l || !b->m ||
        !b->n || !b->xx || !b->yy || !b->zz || !b->group || !b->q)
        return THISTLE_EINVAL;
    if (b->ng > 1024) return THISTLE_EINVAL;
    // |q| grid: the series is even in q, so the reference's sin(pmu)/pmu
    // stays right for q < 0 while x = q h < 0 would not; reject it
    for (int k = 0; k < b->nq; ++k)
        if (!std::isfinite(b->q[k]) || b->q[k] < 0.0) return THISTLE_EINVAL;
    for (double c : {b->cutoffz, b->cutoffmd, b->cutoffcentre})
        if (!std::isfinite(c) || c < 0.0) return THISTLE_EINVAL;
    if (b->cutoffmd > THISTLE_CUTOFFMD_MAX) return THISTLE_ECUTOFF;
    int maxv = 0;
    for (int j = 0; j < b->ngto; ++j) {
        if (b->l[j] < 0 || b->m[j] < 0 || b->n[j] < 0) return THISTLE_EINVAL;
        maxv = std::max({maxv, (int)b->l[j], (int)b->m[j], (int)b->n[j]});
        if (b->group[j] < 1 || b->group[j] > b->ng) return THISTLE_EGROUP;
    }
    if (maxv > 4) return THISTLE_ELMAX;
    maxl_eff = std::max(maxv, (int)b->maxl);
    groups.assign(b->ng, GroupInfo{});
    std::vector<int> seen(b->ng, 0);
    for (int g = 0; g < b->ng; ++g) groups[g].start = -1;
    for (int j = 0; j < b->ngto; ++j) {
        int g = b->group[j] - 1;
        if (groups[g].start < 0) {
            groups[g].start = j;
            groups[g].x = b->xx[j];
            groups[g].y = b->yy[j];
            groups[g].z = b->zz[j];
            groups[g].ga = b->ga[j];
            groups[g].L = b->l[j] + b->m[j] + b->n[j];
        }
        groups[g].count++;
    }
    for (int g = 0; g < b->ng; ++g) {
        if (groups[g].start < 0) return THISTLE_EGROUP;
        for (int j = groups[g].start; j < groups[g].start + groups[g].count; ++j)
            if (b->group[j] != g + 1) return THISTLE_EGROUP;
    }
    if (sigs) {
        sigs->clear();
        for (int g = 0; g < b->ng; ++g) {
            std::vector<int32_t> key;
            for (int j = groups[g].start; j < groups[g].start + groups[g].count; ++j)
                key.push_back((b->l[j] << 8) | (b->m[j] << 4) | b->n[j]);
            int id = -1;
            for (size_t k = 0; k < sigs->size(); ++k)
                if ((*sigs)[k] == key) id = (int)k;
            if (id < 0) {
                if (sigs->size() >= 32) return THISTLE_EGROUP;
                id = (int)sigs->size();
                sigs->push_back(key);
            }
            groups[g].pad = id;
        }
    }
    // LLmax + 1 > 20 for the largest quartet (src/Integrals.f90:5117-5120)
    int Lmax = 0;
    for (auto& g : groups) Lmax = std::max(Lmax, g.L);
    if (4 * Lmax + 1 > 20) return THISTLE_ELLMAX;
    // one primitive shell per group: at most 15 cartesian components of one L
    for (int g = 0; g < b->ng; ++g) {
        if (groups[g].count > 15) return THISTLE_EGROUP;
        for (int j = groups[g].start; j < groups[g].start + groups[g].count; ++j)
            if (b->l[j] + b->m[j] + b->n[j] != groups[g].L) return THISTLE_EGROUP;
    }
    return THISTLE_OK;
}

// Rows of the five nests in reference order (src/Integrals.f90:79-416).
// Rows of the reference's quartet nests in reference order; ej2 selects
// elastic_integration_j2's six nests (src/Integrals.f90:2004-2325).
std::vector<EnumRow> enum_rows(int ng, int64_t& total, bool ej2 = false) {
    std::vector<EnumRow> rows;
    int64_t off = 0;
    auto add = [&](int nest, int i, int j, int k, int r0, int cnt) {
        if (cnt <= 0) return;
        rows.push_back(EnumRow{nest, i, j, k, r0, cnt, off});
        off += cnt;
    };
    for (int i = 0; i < ng; ++i)
        for (int j = i + 1; j < ng; ++j)
            for (int k = i + 1; k < ng; ++k) add(1, i, j, k, k + 1, ng - 1 - k);
    if (ej2) {
        for (int i = 0; i < ng; ++i)
            for (int j = i + 1; j < ng; ++j) add(6, i, j, -1, j + 1, ng - 1 - j);
        for (int i = 0; i < ng; ++i) add(7, i, -1, -1, i + 1, ng - 1 - i);
    } else {
        for (int i = 0; i < ng; ++i)
            for (int j = i + 1; j < ng; ++j) add(2, i, j, -1, i + 1, ng - 1 - i);
    }
    for (int i = 0; i < ng; ++i)
        for (int k = 0; k < ng; ++k) add(3, i, -1, k, k + 1, ng - 1 - k);
    for (int i = 0; i < ng; ++i) add(4, i, -1, -1, i + 1, ng - 1 - i);
    add(5, -1, -1, -1, 0, ng);
    total = off;
    return rows;
}

// Group pairs gi <= gj (pair_index order) with their blocks in primitive
// pair space, and the GTOs (a, b) of every primitive pair.  Returns npp.
int64_t build_pairs(const std::vector<GroupInfo>& groups, std::vector<PairInfo>& pairs,
                    std::vector<int32_t>& xa, std::vector<int32_t>& xb) {
    const int ng = (int)groups.size();
    pairs.assign((size_t)ng * (ng + 1) / 2, PairInfo{});
    xa.clear();
    xb.clear();
    int64_t off = 0;
    for (int gi = 0; gi < ng; ++gi)
        for (int gj = gi; gj < ng; ++gj) {
            PairInfo& p = pairs[pair_index(gi, gj, ng)];
            std::memset(&p, 0, sizeof p);
            p.gi = gi;
            p.gj = gj;
            p.off = (int32_t)off;
            p.nab = groups[gi].count * groups[gj].count;
            for (int a = 0; a < groups[gi].count; ++a)
                for (int c = 0; c < groups[gj].count; ++c) {
                    xa.push_back(groups[gi].start + a);
                    xb.push_back(groups[gj].start + c);
                }
            off += p.nab;
        }
    return off;
}

// GTO-quadruple tables of the signature combinations in use.  Entry u of
// a combination (d fastest): bits 0-7 a*cj+b, 8-15 c*cr+d (offsets in the
// bra / ket pair blocks), 16-25 / 26-35 / 36-45 the x / y / z exponent
// tuple index, 46-61 a, b, c, d (4 bits each).  The ket offset runs
// fastest, so that a wave's Z reads fall on few cache lines.  (A copy
// with the bra offset fastest for the quartets whose bra elements are
// Z's rows measured slower on XY6: 8.93 against 8.83 ms.)
void build_quad_tables(const std::vector<std::vector<int32_t>>& sigs,
                       const std::vector<uint8_t>& used, std::vector<int64_t>& comboff,
                       std::vector<uint64_t>& qtab) {
    const int S = (int)sigs.size();
    comboff.assign((size_t)S * S * S * S, 0);
    qtab.clear();
    for (int si = 0; si < S; ++si)
        for (int sj = 0; sj < S; ++sj)
            for (int sk = 0; sk < S; ++sk)
                for (int sr = 0; sr < S; ++sr) {
                    size_t c = ((size_t)(si * S + sj) * S + sk) * S + sr;
                    if (!used[c]) continue;
                    comboff[c] = (int64_t)qtab.size();
                    const std::vector<int32_t>* sg[4] = {&sigs[si], &sigs[sj], &sigs[sk],
                                                         &sigs[sr]};
                    int cnt4[4], L4[4];
                    for (int k = 0; k < 4; ++k) {
                        cnt4[k] = (int)sg[k]->size();
                        int v = (*sg[k])[0];
                        L4[k] = (v >> 8) + ((v >> 4) & 15) + (v & 15);
                    }
                    auto entry = [&](int a, int bb, int cc, int d) {
                        int e4[4] = {(*sg[0])[a], (*sg[1])[bb], (*sg[2])[cc], (*sg[3])[d]};
                        uint64_t t3[3];
                        for (int ax = 0; ax < 3; ++ax) {
                            int sh = 8 - 4 * ax, t = 0;
                            for (int k = 0; k < 4; ++k)
                                t = t * (L4[k] + 1) + ((e4[k] >> sh) & 15);
                            t3[ax] = (uint64_t)t;
                        }
                        return (uint64_t)(a * cnt4[1] + bb) |
                               ((uint64_t)(cc * cnt4[3] + d) << 8) | (t3[0] << 16) |
                               (t3[1] << 26) | (t3[2] << 36) | ((uint64_t)a << 46) |
                               ((uint64_t)bb << 50) | ((uint64_t)cc << 54) |
                               ((uint64_t)d << 58);
                    };
                    for (int a = 0; a < cnt4[0]; ++a)
                        for (int bb = 0; bb < cnt4[1]; ++bb)
                            for (int cc = 0; cc < cnt4[2]; ++cc)
                                for (int d = 0; d < cnt4[3]; ++d)
                                    qtab.push_back(entry(a, bb, cc, d));
                }
}

int setup_common(ThistlePlan* pl, const ThistleBasis* b, const ThistleOpts* o) {
    std::vector<GroupInfo> groups;
    std::vector<std::vector<int32_t>> sigs;
    int maxl = 0;
    int rc = check_basis(b, groups, maxl, &sigs);
    if (rc) return rc;
    const int S = (int)sigs.size();
    pl->precision = o->precision;
    for (const auto& g : groups) {
        pl->gstart.push_back(g.start);
        pl->gcount.push_back(g.count);
        pl->gL.push_back(g.L);
    }
    pl->ngto = b->ngto;
    pl->ng = b->ng;
    pl->maxl = maxl;
    pl->nq = b->nq;
    {
        // q h >= SERIES_XDPH needs the far-argument kernel only if some q
        // times some pair-centre distance gets there; a pair centre lies
        // between its two group centres, so the widest centre distance
        // bounds every h
        std::vector<std::array<double, 3>> c;
        for (const auto& g : groups) {
            std::array<double, 3> v{g.x, g.y, g.z};
            if (std::find(c.begin(), c.end(), v) == c.end()) c.push_back(v);
        }
        double d2 = 0.0, qmax = 0.0;
        for (size_t i = 0; i < c.size(); ++i)
            for (size_t j = i + 1; j < c.size(); ++j) {
                const double dx = c[i][0] - c[j][0], dy = c[i][1] - c[j][1],
                             dz = c[i][2] - c[j][2];
                d2 = std::max(d2, dx * dx + dy * dy + dz * dz);
            }
        for (int k = 0; k < b->nq; ++k) qmax = std::max(qmax, std::fabs(b->q[k]));
        pl->far = qmax * std::sqrt(d2) * (1.0 + 1e-9) >= SERIES_XDPH;
    }
    pl->cutoffz = b->cutoffz;
    pl->cutoffmd = b->cutoffmd;
    pl->cutoffcentre = b->cutoffcentre;
    pl->device = o->device;
    int ng = b->ng;
    pl->npair = ng * (ng + 1) / 2;
    std::vector<PairInfo> pairs;
    std::vector<int32_t> xa, xb;
    pl->npp = build_pairs(groups, pairs, xa, xb);
    if (pl->npp > 0x7fffffff) return THISTLE_EINVAL;
    // primitive-pair space padded to a multiple of 16 with elements of no GTO
    // pair (T, M and Z are zero there): the leading dimension of M and Z then
    // keeps every 16-row tile column of the transform on whole 128-byte
    // lines (a row count of 8 mod 16 measured 2.2x slower in k_zmfma)
    while (pl->npp % 16) {
        xa.push_back(-1);
        xb.push_back(-1);
        ++pl->npp;
    }
    std::vector<char4> lx(b->ngto);
    for (int j = 0; j < b->ngto; ++j)
        lx[j] = make_char4((char)b->l[j], (char)b->m[j], (char)b->n[j],
                           (char)(b->l[j] + b->m[j] + b->n[j]));
    int64_t Q = 0;
    std::vector<EnumRow> rows = enum_rows(ng, Q, pl->type == 8);
    pl->nquart = Q;
    pl->nlisted = Q;
    pl->nrows = (int)rows.size();
    if (Q > 0x7fffffff) return THISTLE_EINVAL;

    // ---- device from here on
    HCHECK(hipSetDevice(o->device));
    if (o->stream) {
        pl->stream = (hipStream_t)o->stream;
    } else {
        HCHECK(hipStreamCreate(&pl->stream));
        pl->own_stream = true;
    }
    for (auto& e : pl->ev) HCHECK(hipEventCreate(&e));
    for (auto& e : pl->kev) HCHECK(hipEventCreate(&e));
    // THISTLE_OVERLAP=0 keeps every stage on the one stream
    const char* ov = std::getenv("THISTLE_OVERLAP");
    if (!(ov && ov[0] == '0')) {
        HCHECK(hipStreamCreateWithFlags(&pl->aux, hipStreamNonBlocking));
        HCHECK(hipEventCreateWithFlags(&pl->fork_ev, hipEventDisableTiming));
        HCHECK(hipEventCreateWithFlags(&pl->join_ev, hipEventDisableTiming));
        for (auto& e : pl->cev) HCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    if ((rc = dalloc(pl->rows, sizeof(EnumRow) * rows.size()))) return rc;
    if ((rc = dalloc(pl->cls, sizeof(int32_t) * Q))) return rc;
    HCHECK(hipMemcpyAsync(pl->rows.p, rows.data(), sizeof(EnumRow) * rows.size(),
                          hipMemcpyHostToDevice, pl->stream));
    // positions of each LL = L_i+L_j+L_k+L_r class in reference order: a
    // class key per quartet on the GPU, then a stable radix sort by key
    std::vector<uint8_t> used((size_t)S * S * S * S, 0);
    {
        std::vector<int8_t> gL8(ng), gs8(ng);
        for (int g = 0; g < ng; ++g) {
            gL8[g] = (int8_t)groups[g].L;
            gs8[g] = (int8_t)groups[g].pad;
        }
        DevBuf dgL, dgs, key, key2, pos, cnt, tmx, dused, tmp;
        if ((rc = dalloc(dgL, ng)) || (rc = dalloc(dgs, ng)) ||
            (rc = dalloc(key, sizeof(uint32_t) * Q)) || (rc = dalloc(key2, sizeof(uint32_t) * Q)) ||
            (rc = dalloc(pos, sizeof(int32_t) * Q)) || (rc = dalloc(cnt, 18 * 8)) ||
            (rc = dalloc(tmx, 18 * 4)) || (rc = dalloc(dused, used.size())))
            return rc;
        HCHECK(hipMemcpyAsync(dgL.p, gL8.data(), ng, hipMemcpyHostToDevice, pl->stream));
        HCHECK(hipMemcpyAsync(dgs.p, gs8.data(), ng, hipMemcpyHostToDevice, pl->stream));
        HCHECK(hipMemsetAsync(cnt.p, 0, 18 * 8, pl->stream));
        HCHECK(hipMemsetAsync(tmx.p, 0, 18 * 4, pl->stream));
        HCHECK(hipMemsetAsync(dused.p, 0, used.size(), pl->stream));
        HCHECK(launch_classify(pl->rows.as<EnumRow>(), pl->nrows, dgL.as<int8_t>(),
                               dgs.as<int8_t>(), S, key.as<uint32_t>(), pos.as<int32_t>(),
                               cnt.as<unsigned long long>(), tmx.as<int>(), dused.as<uint8_t>(),
                               pl->stream));
        size_t tb = 0;
        HCHECK(class_sort(nullptr, tb, key.as<uint32_t>(), key2.as<uint32_t>(), pos.as<int32_t>(),
                          pl->cls.as<int32_t>(), Q, pl->stream));
        if ((rc = dalloc(tmp, std::max<size_t>(tb, 16)))) return rc;
        HCHECK(class_sort(tmp.p, tb, key.as<uint32_t>(), key2.as<uint32_t>(), pos.as<int32_t>(),
                          pl->cls.as<int32_t>(), Q, pl->stream));
        unsigned long long hc[18];
        int ht[18];
        HCHECK(hipMemcpyAsync(hc, cnt.p, sizeof hc, hipMemcpyDeviceToHost, pl->stream));
        HCHECK(hipMemcpyAsync(ht, tmx.p, sizeof ht, hipMemcpyDeviceToHost, pl->stream));
        HCHECK(hipMemcpyAsync(used.data(), dused.p, used.size(), hipMemcpyDeviceToHost,
                              pl->stream));
        HCHECK(hipStreamSynchronize(pl->stream));
        for (int k = 0; k < 17; ++k) pl->cls_tmax[k] = ht[k];
        pl->cls_off[0] = 0;
        pl->cls_hoff[0] = 0;
        // records of class LL: DO = LL + 1 series coefficients (LL + 3 for
        // j2) plus the per-quartet words, HS_ROW(DO) doubles each
        const int dj = (pl->type == 7 || pl->type == 8) ? 3 : 1;
        for (int k = 0; k < 17; ++k) {
            pl->cls_off[k + 1] = pl->cls_off[k] + (int64_t)hc[k];
            pl->cls_hoff[k + 1] = pl->cls_hoff[k] + (int64_t)hc[k] * HS_ROW(k + dj);
        }
        if (pl->cls_off[17] != Q) return THISTLE_EDEVICE;
    }
    {
        // the contraction kernel keeps per-exponent-tuple tables in LDS
        hipDeviceProp_t prop;
        HCHECK(hipGetDeviceProperties(&prop, o->device));
        size_t need = 0;
        for (int LL = 0; LL <= 16; ++LL)
            if (pl->cls_off[LL + 1] > pl->cls_off[LL])
                need = std::max(need, contract_lds_bytes(LL, pl->cls_tmax[LL]));
        if (need > prop.sharedMemPerBlock) {
            std::fprintf(stderr, "thistle: quartet tables need %zu B of LDS (device: %zu B)\n",
                         need, (size_t)prop.sharedMemPerBlock);
            return THISTLE_ELMAX;
        }
        if (need > 65536) HCHECK(set_contract_lds_limit(need));
    }
    std::vector<int64_t> comboff;
    std::vector<uint64_t> qtab;
    build_quad_tables(sigs, used, comboff, qtab);
    // quartet headers hold 32-bit table starts (a larger table, 32 GB, would
    // not be held anyway)
    if (qtab.size() > 0xffffffffull) return THISTLE_ENOMEM;
    pl->nsig = S;
    int L1 = maxl + 1, T1 = 2 * maxl + 1;
    if ((rc = dalloc(pl->G, sizeof(GroupInfo) * ng))) return rc;
    if ((rc = dalloc(pl->P, sizeof(PairInfo) * pl->npair))) return rc;
    if ((rc = dalloc(pl->E, sizeof(double) * (size_t)pl->npair * 3 * L1 * L1 * T1))) return rc;
    if ((rc = dalloc(pl->e12, sizeof(double) * (size_t)pl->npair * b->nq))) return rc;
    if ((rc = dalloc(pl->q, sizeof(double) * b->nq))) return rc;
    if ((rc = dalloc(pl->lxyz, sizeof(char4) * b->ngto))) return rc;
    if ((rc = dalloc(pl->qlist, sizeof(uint64_t) * Q))) return rc;
    if ((rc = dalloc(pl->flags, Q))) return rc;
    if ((rc = dalloc(pl->qhdr, sizeof(QHdr) * Q))) return rc;
    int64_t nb = (Q + 2047) / 2048 + 1;
    if ((rc = dalloc(pl->scan, sizeof(int64_t) * (Q + nb + 1)))) return rc;
    if ((rc = dalloc(pl->qcomp, sizeof(uint64_t) * Q))) return rc;
    if ((rc = dalloc(pl->count, sizeof(int64_t) * 2))) return rc;
    if ((rc = dalloc(pl->xa, sizeof(int32_t) * pl->npp))) return rc;
    if ((rc = dalloc(pl->xb, sizeof(int32_t) * pl->npp))) return rc;
    if ((rc = dalloc(pl->hs, sizeof(double) * (size_t)std::max<int64_t>(pl->cls_hoff[17], 1))))
        return rc;
    if ((rc = dalloc(pl->quadtab, sizeof(uint64_t) * qtab.size()))) return rc;
    if ((rc = dalloc(pl->comboff, sizeof(int64_t) * comboff.size()))) return rc;
    HCHECK(hipMemcpyAsync(pl->quadtab.p, qtab.data(), sizeof(uint64_t) * qtab.size(),
                          hipMemcpyHostToDevice, pl->stream));
    HCHECK(hipMemcpyAsync(pl->comboff.p, comboff.data(), sizeof(int64_t) * comboff.size(),
                          hipMemcpyHostToDevice, pl->stream));
    HCHECK(hipStreamSynchronize(pl->stream));
    pl->h_q.assign(b->q, b->q + b->nq);
    // a uniform grid from q[1] on (q[0] may be set apart, as 1e-10 is by the
    // reference's driver) lets the series kernel rotate sin/cos across q
    if (b->nq >= 3) {
        const double* q = b->q;
        const int n = b->nq;
        const double dq = (q[n - 1] - q[1]) / (n - 2), qa = q[1] - dq;
        double qmax = 0.0;
        for (int k = 0; k < n; ++k) qmax = std::max(qmax, std::fabs(q[k]));
        const double tol = 8.0 * 2.220446049250313e-16 * qmax;
        bool lin = dq > 0.0 && std::isfinite(dq);
        for (int k = 1; lin && k < n; ++k) lin = std::fabs(q[k] - (qa + k * dq)) <= tol;
        pl->qlin = lin ? 1 : 0;
        pl->q0off = lin && std::fabs(q[0] - qa) > tol ? 1 : 0;
        pl->qa = qa;
        pl->dq = dq;
    }
    HCHECK(hipMemcpyAsync(pl->G.p, groups.data(), sizeof(GroupInfo) * ng, hipMemcpyHostToDevice, pl->stream));
    HCHECK(hipMemcpyAsync(pl->P.p, pairs.data(), sizeof(PairInfo) * pl->npair, hipMemcpyHostToDevice, pl->stream));
    HCHECK(hipMemcpyAsync(pl->q.p, b->q, sizeof(double) * b->nq, hipMemcpyHostToDevice, pl->stream));
    HCHECK(hipMemcpyAsync(pl->lxyz.p, lx.data(), sizeof(char4) * b->ngto, hipMemcpyHostToDevice, pl->stream));
    HCHECK(hipMemcpyAsync(pl->xa.p, xa.data(), sizeof(int32_t) * pl->npp, hipMemcpyHostToDevice, pl->stream));
    HCHECK(hipMemcpyAsync(pl->xb.p, xb.data(), sizeof(int32_t) * pl->npp, hipMemcpyHostToDevice, pl->stream));
    int waves = o->waves;
    if (waves <= 0) {
        hipDeviceProp_t prop;
        HCHECK(hipGetDeviceProperties(&prop, o->device));
        // series waves per CU: 8 (XY6 quartet stage 7.03 ms; 6: 7.04, 12: 7.06
        // with the classes in increasing LL)
        waves = prop.multiProcessorCount * 8;
    }
    pl->waves = (int)std::max<int64_t>(1, std::min<int64_t>(waves, Q));
    pl->waves_alloc = pl->waves;
    int nqc = std::min(b->nq, 512);
    if ((rc = dalloc(pl->part, sizeof(double) * (size_t)pl->waves * nqc))) return rc;
    if ((rc = dalloc(pl->partc, sizeof(double) * (size_t)pl->waves * nqc))) return rc;
    if ((rc = dalloc(pl->S, sizeof(double) * b->nq))) return rc;
    if (pl->precision == THISTLE_MIXED) {
        if ((rc = dalloc(pl->parta, sizeof(double) * (size_t)pl->waves * nqc))) return rc;
        if ((rc = dalloc(pl->Aq, sizeof(double) * (size_t)b->nq))) return rc;
        if ((rc = dalloc(pl->qsel, sizeof(int32_t) * b->nq))) return rc;
        if ((rc = dalloc(pl->nsel, sizeof(int32_t) * 2))) return rc;
        if ((rc = dalloc(pl->Sfix, sizeof(double) * nqc))) return rc;
        pl->kappa = kMixedKappa;
        if (const char* ev = std::getenv("THISTLE_MIXED_KAPPA")) pl->kappa = std::atof(ev);
    }
    if (b->nmo > 0 && b->mmod) {
        pl->nmo = b->nmo;
        if ((rc = dalloc(pl->mmod, sizeof(double) * (size_t)b->nmo * b->ngto))) return rc;
        HCHECK(hipMemcpyAsync(pl->mmod.p, b->mmod, sizeof(double) * (size_t)b->nmo * b->ngto,
                              hipMemcpyHostToDevice, pl->stream));
    }
    // the host tables above go out of scope here
    HCHECK(hipStreamSynchronize(pl->stream));
    return THISTLE_OK;
}

int run_tables_and_enum(ThistlePlan* pl, bool timing) {
    HCHECK(launch_pair_info(pl->G.as<GroupInfo>(), pl->P.as<PairInfo>(), pl->npair, pl->stream));
    HCHECK(launch_pair_E(pl->G.as<GroupInfo>(), pl->P.as<PairInfo>(), pl->npair, pl->maxl,
                         pl->E.as<double>(), pl->stream));
    HCHECK(launch_e12(pl->P.as<PairInfo>(), pl->npair, pl->q.as<double>(), pl->nq,
                      pl->e12.as<double>(), nullptr, pl->stream));
    if (timing) HCHECK(hipEventRecord(pl->ev[1], pl->stream));
    HCHECK(launch_enumerate(pl->rows.as<EnumRow>(), pl->nrows, pl->P.as<PairInfo>(), pl->ng,
                            pl->cutoffcentre, pl->qlist.as<uint64_t>(), pl->stream));
    HCHECK(launch_qhdr(pl->cls.as<int32_t>(), pl->cls_off[17], pl->qlist.as<uint64_t>(),
                       pl->G.as<GroupInfo>(), pl->P.as<PairInfo>(), pl->ng,
                       pl->comboff.as<int64_t>(), pl->nsig, pl->qhdr.as<QHdr>(), pl->stream));
    if (timing) HCHECK(hipEventRecord(pl->ev[2], pl->stream));
    return THISTLE_OK;
}

int run_transform_total(ThistlePlan* pl) {
    const int norb = pl->norb, npmo = pl->npmo;
    UniqueWork w{pl->hkeys.as<uint64_t>(), pl->hmin.as<int32_t>(), pl->hcnt.as<int32_t>(),
                 pl->hslot.as<int32_t>(), pl->uscan.as<int64_t>(),
                 (int64_t)(pl->hkeys.bytes / sizeof(uint64_t))};
    HCHECK(launch_unique(pl->mat.as<int32_t>(), pl->nmat, 4, pl->total.as<double>(), w,
                         pl->rmat.as<int32_t>(), pl->rtot.as<double>(), nullptr, nullptr,
                         pl->count.as<int64_t>() + 1, pl->stream));
    HCHECK(launch_fill(pl->V.as<double>(), (int64_t)(pl->V.bytes / sizeof(double)), 0.0,
                       pl->stream));
    // only the unique prefix of rmat/rtot is valid; its length stays on the
    // device and guards the scatter
    HCHECK(launch_gs(pl->rmat.as<int32_t>(), pl->rtot.as<double>(), pl->nmat,
                     pl->count.as<int64_t>() + 1, pl->nmat, norb, pl->V.as<double>(),
                     pl->stream));
    HCHECK(launch_gs_matrix(pl->V.as<double>(), norb, pl->mopairs.as<int2>(), npmo,
                            pl->Gs.as<double>(), pl->stream));
    HCHECK(launch_tmat2(pl->mmod.as<double>(), pl->nmo, pl->mopairs.as<int2>(), npmo,
                        pl->xa.as<int32_t>(), pl->xb.as<int32_t>(), pl->npp, pl->T.as<double>(),
                        pl->stream));
    const double one = 1.0, zero = 0.0;
    const rocblas_int npp = (rocblas_int)pl->npp;
    if (pl->zfact) {
        // M^T = T^T Gs^T  (npp x npmo), then Z through the MO factorisation:
        // 2 npp norb ngto (norb + ngto / 2) flops instead of npp^2 npmo
        BCHECK(rocblas_dgemm(pl->blas, rocblas_operation_transpose, rocblas_operation_transpose,
                             npp, npmo, npmo, &one, pl->T.as<double>(), npmo,
                             pl->Gs.as<double>(), npmo, &zero, pl->M.as<double>(), npp));
        HCHECK(launch_zfact(pl->M.as<double>(), pl->mmod.as<double>(), pl->nmo, norb, pl->ngto,
                            pl->npp, pl->jtab.as<int32_t>(), pl->aend.as<int32_t>(),
                            pl->Cp.as<double>(), pl->W3.as<double>(), pl->Z.as<double>(),
                            pl->zgrid, pl->stream));
        return THISTLE_OK;
    }
    // M = Gs * T   (npmo x npp)
    BCHECK(rocblas_dgemm(pl->blas, rocblas_operation_none, rocblas_operation_none, npmo, npp,
                         npmo, &one, pl->Gs.as<double>(), npmo, pl->T.as<double>(), npmo, &zero,
                         pl->M.as<double>(), npmo));
    // Z = T^T * M  (npp x npp, symmetric): lower triangle only
    BCHECK(rocblas_dsyrkx(pl->blas, rocblas_fill_lower, rocblas_operation_transpose, npp, npmo,
                          &one, pl->T.as<double>(), npmo, pl->M.as<double>(), npmo, &zero,
                          pl->Z.as<double>(), npp));
    return THISTLE_OK;
}

SeriesArgs series_args(const ThistlePlan* pl) {
    const bool elastic = pl->type == 2 || pl->type == 8;
    SeriesArgs a{};
    a.qlist = pl->qlist.as<uint64_t>();
    a.nquart = pl->nquart;
    a.G = pl->G.as<GroupInfo>();
    a.P = pl->P.as<PairInfo>();
    a.E = pl->E.as<double>();
    a.maxl = pl->maxl;
    a.ng = pl->ng;
    a.ngto = pl->ngto;
    a.lxyz = pl->lxyz.as<char4>();
    a.q = pl->q.as<double>();
    a.nq = pl->nq;
    a.e12 = pl->e12.as<double>();
    a.Z = pl->Z.as<double>();
    a.ldz = elastic ? pl->ngto : pl->npp;
    a.elastic = elastic ? 1 : 0;
    a.mixed = pl->precision == THISTLE_MIXED;
    a.j2 = (pl->type == 7 || pl->type == 8) ? 1 : 0;
    a.cutoffz = pl->cutoffz;
    a.part = pl->part.as<double>();
    a.partc = pl->partc.as<double>();
    a.parta = pl->parta.as<double>();
    a.qsel = pl->qsel.as<int32_t>();
    a.nsel = pl->nsel.as<int32_t>();
    a.flags = pl->flags.as<uint8_t>();
    a.hs = pl->hs.as<double>();
    a.quadtab = pl->quadtab.as<uint64_t>();
    a.comboff = pl->comboff.as<int64_t>();
    a.nsig = pl->nsig;
    a.nqc = std::min(512, pl->nq);
    a.qs0 = 0;
    a.qs1 = a.nqc;
    a.qlin = pl->qlin;
    a.q0off = pl->q0off;
    a.qa = pl->qa;
    a.dq = pl->dq;
    return a;
}

// The contraction (gather- and latency-bound) and the series (FP64 VALU
// bound) of different classes run concurrently: contractions on the plan's
// stream, series on the auxiliary stream, each series launch behind its own
// class's contraction through an event.  The series launches stay in one
// stream order, so the partial rows accumulate exactly as on one stream.
int run_series(ThistlePlan* pl, bool timing) {
    SeriesArgs a = series_args(pl);
    const bool ov = pl->aux != nullptr;
    hipStream_t ss = ov ? pl->aux : pl->stream;
    if (ov) {
        HCHECK(hipEventRecord(pl->fork_ev, pl->stream));
        HCHECK(hipStreamWaitEvent(pl->aux, pl->fork_ev, 0));
    }
    // h(beta) of every quartet (q independent).  Classes in increasing LL:
    // the low classes' contractions are short and their series long, the
    // high classes' the other way round (Johnson's two-stage order), so the
    // series stream has work while the heavy contractions run and little is
    // left after the last one (XY6 quartet stage 7.30 ms in decreasing order)
    if (timing) HCHECK(hipEventRecord(pl->kev[0], pl->stream));
    for (int LL = 0; LL <= 16; ++LL) {
        int64_t n = pl->cls_off[LL + 1] - pl->cls_off[LL];
        if (n <= 0) continue;
        // one workgroup per quartet group up to 2^20 groups (grid-stride
        // beyond): XY6's quartet stage 8.0 ms against 8.15 with 16384
        int blocks = (int)std::min<int64_t>(1 << 20, n);
        a.hs = pl->hs.as<double>() + pl->cls_hoff[LL];
        HCHECK(launch_contract(a, LL, pl->cls_tmax[LL], pl->qhdr.as<QHdr>() + pl->cls_off[LL],
                               n, blocks, pl->stream));
        if (ov) HCHECK(hipEventRecord(pl->cev[LL], pl->stream));
    }
    if (timing) HCHECK(hipEventRecord(pl->kev[1], pl->stream));
    bool first_series = true;
    for (int q0 = 0; q0 < pl->nq; q0 += 512) {
        a.q0 = q0;
        a.nqc = std::min(512, pl->nq - q0);
        const int64_t rows = (int64_t)pl->waves * a.nqc;
        HCHECK(launch_fill(pl->part.as<double>(), rows, 0.0, ss));
        HCHECK(launch_fill(pl->partc.as<double>(), rows, 0.0, ss));
        a.pstride = rows;
        if (a.mixed) HCHECK(launch_fill(pl->parta.as<double>(), rows, 0.0, ss));
        for (int LL = 0; LL <= 16; ++LL) {
            int64_t n = pl->cls_off[LL + 1] - pl->cls_off[LL];
            if (n <= 0) continue;
            if (ov && q0 == 0) HCHECK(hipStreamWaitEvent(ss, pl->cev[LL], 0));
            if (timing && first_series) HCHECK(hipEventRecord(pl->kev[2], ss));
            first_series = false;
            int w = (int)std::min<int64_t>(pl->waves, n);
            a.hs = pl->hs.as<double>() + pl->cls_hoff[LL];
            // j2 carries its series two orders further (bessels0rr(LLmax+2))
            HCHECK(launch_series(a, a.j2 ? LL + 2 : LL, n, w, 0, ss));
            if (pl->far) HCHECK(launch_series_far(a, a.j2 ? LL + 2 : LL, n, w, ss));
        }
        HCHECK(launch_reduce(a.part, a.partc, pl->waves, a.nqc, pl->S.as<double>() + q0, ss));
        if (a.mixed)
            HCHECK(launch_reduce_plain(a.parta, pl->waves, a.nqc, pl->Aq.as<double>() + q0, ss));
    }
    if (timing) {
        if (first_series) HCHECK(hipEventRecord(pl->kev[2], ss));
        HCHECK(hipEventRecord(pl->kev[3], ss));
    }
    if (a.mixed) {
        // second pass: the q points whose FP32 error scale is not small
        // against |S(q)|, in FP64 (every regime) from the same records
        HCHECK(launch_qflag(pl->S.as<double>(), pl->Aq.as<double>(), pl->nq, pl->kappa,
                            pl->qsel.as<int32_t>(), pl->nsel.as<int32_t>(), ss));
        const int nqc = std::min(512, pl->nq);
        a.nqc = nqc;
        for (int pbase = 0; pbase < pl->nq; pbase += nqc) {
            a.pbase = pbase;
            const int64_t rows = (int64_t)pl->waves * nqc;
            HCHECK(launch_fill(pl->part.as<double>(), rows, 0.0, ss));
            HCHECK(launch_fill(pl->partc.as<double>(), rows, 0.0, ss));
            for (int LL = 0; LL <= 16; ++LL) {
                int64_t n = pl->cls_off[LL + 1] - pl->cls_off[LL];
                if (n <= 0) continue;
                int w = (int)std::min<int64_t>(pl->waves, n);
                a.hs = pl->hs.as<double>() + pl->cls_hoff[LL];
                HCHECK(launch_series_pts(a, a.j2 ? LL + 2 : LL, n, w, 0, ss));
            }
            HCHECK(launch_reduce(a.part, a.partc, pl->waves, nqc, pl->Sfix.as<double>(), ss));
            HCHECK(launch_qscatter(pl->Sfix.as<double>(), pl->qsel.as<int32_t>(),
                                   pl->nsel.as<int32_t>(), pbase, nqc, pl->S.as<double>(), ss));
        }
    }
    if (ov) {
        HCHECK(hipEventRecord(pl->join_ev, pl->aux));
        HCHECK(hipStreamWaitEvent(pl->stream, pl->join_ev, 0));
    }
    if (timing) {
        // stage 4 is empty since the screened quartet list is compacted on
        // demand (thistle_quartet_list_screened) from the flags kept here
        HCHECK(hipEventRecord(pl->ev[4], pl->stream));
        HCHECK(hipEventRecord(pl->ev[5], pl->stream));
    }
    pl->units = pl->nlisted * (int64_t)pl->nq;
    return THISTLE_OK;
}

// Class lists of a subset of the enumeration (positions strictly
// increasing, so every class keeps reference order), or of every quartet
// when pos is null.
int set_quartets(ThistlePlan* pl, const int64_t* pos, int64_t n) {
    int64_t Q = 0;
    std::vector<EnumRow> rows = enum_rows(pl->ng, Q, pl->type == 8);
    if (Q != pl->nquart) return THISTLE_EINVAL;
    std::vector<int64_t> all;
    if (!pos) {
        n = Q;
    } else {
        if (n < 0 || n > Q) return THISTLE_EINVAL;
        for (int64_t k = 0; k < n; ++k)
            if (pos[k] < 0 || pos[k] >= Q || (k > 0 && pos[k] <= pos[k - 1]))
                return THISTLE_EINVAL;
    }
    std::vector<int8_t> LLof((size_t)Q);
    for (const auto& R : rows)
        for (int u = 0; u < R.count; ++u) {
            int g[4];
            row_quartet(R, R.r0 + u, g);
            LLof[(size_t)(R.out + u)] =
                (int8_t)(pl->gL[g[0]] + pl->gL[g[1]] + pl->gL[g[2]] + pl->gL[g[3]]);
        }
    std::vector<int64_t> cnt(18, 0);
    for (int64_t k = 0; k < n; ++k) cnt[LLof[(size_t)(pos ? pos[k] : k)]]++;
    const int dj = (pl->type == 7 || pl->type == 8) ? 3 : 1;
    pl->cls_off[0] = 0;
    pl->cls_hoff[0] = 0;
    for (int k = 0; k < 17; ++k) {
        pl->cls_off[k + 1] = pl->cls_off[k] + cnt[k];
        pl->cls_hoff[k + 1] = pl->cls_hoff[k] + cnt[k] * HS_ROW(k + dj);
    }
    std::vector<int32_t> cls((size_t)std::max<int64_t>(n, 1));
    std::vector<int64_t> fill(pl->cls_off, pl->cls_off + 17);
    for (int64_t k = 0; k < n; ++k) {
        const int64_t e = pos ? pos[k] : k;
        cls[(size_t)fill[LLof[(size_t)e]]++] = (int32_t)e;
    }
    pl->nlisted = n;
    if (n > 0)
        HCHECK(hipMemcpyAsync(pl->cls.p, cls.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice,
                              pl->stream));
    HCHECK(hipStreamSynchronize(pl->stream));
    return THISTLE_OK;
}

// Every device stage of one execution, enqueued on the plan's stream with no
// host synchronisation, memset or (unless timing) event: safe under stream
// capture.  S_dst (device, optional) receives a copy of S(q).
int enqueue_execute(ThistlePlan* pl, bool timing, double* S_dst, bool quartets = true) {
    if (timing) HCHECK(hipEventRecord(pl->ev[0], pl->stream));
    int rc = run_tables_and_enum(pl, timing);
    if (rc) return rc;
    if (pl->type == 1 || pl->type == 7) {
        rc = run_transform_total(pl);
        if (rc) return rc;
    } else {
        HCHECK(launch_zel(pl->mmod.as<double>(), pl->nmo, pl->norb, pl->onerdm.as<double>(),
                          pl->ngto, pl->Z.as<double>(), pl->stream));
    }
    if (timing) HCHECK(hipEventRecord(pl->ev[3], pl->stream));
    if (!quartets) return THISTLE_OK;  // a batch runs the quartet stage across frames
    rc = run_series(pl, timing);
    if (rc) return rc;
    if (S_dst) HCHECK(launch_copy(pl->S.as<double>(), S_dst, pl->nq, pl->stream));
    return THISTLE_OK;
}

int plan_create_impl(int32_t type, const ThistleBasis* b, int64_t nmat, const int32_t* mat,
                     const double* total, int32_t norb, const double* onerdm,
                     const ThistleOpts* o, ThistlePlan** out) {
    // every host-side argument check comes before the first HIP call
    if (type != 1 && type != 2 && type != 7 && type != 8) return THISTLE_ETYPE;
    if (!b || !b->mmod || b->nmo <= 0) return THISTLE_EINVAL;
    if ((type == 7 || type == 8) && b->l && b->m && b->n) {
        // the j2 series runs to order LLmax + 2; the kernels stop at 16
        int lmax = 0;
        for (int j = 0; j < b->ngto; ++j) lmax = std::max(lmax, b->l[j] + b->m[j] + b->n[j]);
        if (4 * lmax + 2 > 16) return THISTLE_ELLMAX;
    }
    int nrb = 0;
    if (type == 1 || type == 7) {
        if (nmat < 0 || (nmat > 0 && (!mat || !total))) return THISTLE_EINVAL;
        for (int64_t k = 0; k < 4 * nmat; ++k) {
            if (mat[k] < 1 || mat[k] > b->nmo || mat[k] > 65535) return THISTLE_EINVAL;
            nrb = std::max(nrb, (int)mat[k]);
        }
    } else if (norb <= 0 || !onerdm || norb > b->nmo) {
        return THISTLE_EINVAL;
    }
    std::unique_ptr<ThistlePlan> pl(new ThistlePlan());
    pl->type = type;
    int rc = setup_common(pl.get(), b, o);
    if (rc) return rc;
    if (type == 1 || type == 7) {
        pl->norb = std::max(nrb, 1);
        pl->nmat = nmat;
        pl->npmo = pl->norb * (pl->norb + 1) / 2;
        std::vector<int2> mop;
        for (int p = 0; p < pl->norb; ++p)
            for (int s = p; s < pl->norb; ++s) mop.push_back(make_int2(p, s));
        int64_t hs = 16;
        while (hs < 2 * std::max<int64_t>(nmat, 1)) hs <<= 1;
        size_t N4 = (size_t)pl->norb * pl->norb * pl->norb * pl->norb;
        if ((rc = dalloc(pl->mat, sizeof(int32_t) * 4 * std::max<int64_t>(nmat, 1)))) return rc;
        if ((rc = dalloc(pl->total, sizeof(double) * std::max<int64_t>(nmat, 1)))) return rc;
        if ((rc = dalloc(pl->rmat, sizeof(int32_t) * 4 * std::max<int64_t>(nmat, 1)))) return rc;
        if ((rc = dalloc(pl->rtot, sizeof(double) * std::max<int64_t>(nmat, 1)))) return rc;
        if ((rc = dalloc(pl->hkeys, sizeof(uint64_t) * hs))) return rc;
        if ((rc = dalloc(pl->hmin, sizeof(int32_t) * hs))) return rc;
        if ((rc = dalloc(pl->hcnt, sizeof(int32_t) * hs))) return rc;
        if ((rc = dalloc(pl->hslot, sizeof(int32_t) * std::max<int64_t>(nmat, 1)))) return rc;
        int64_t nb = (nmat + 2047) / 2048 + 1;
        if ((rc = dalloc(pl->uscan, sizeof(int64_t) * (nmat + nb + 1)))) return rc;
        if ((rc = dalloc(pl->V, sizeof(double) * N4))) return rc;
        if ((rc = dalloc(pl->Gs, sizeof(double) * (size_t)pl->npmo * pl->npmo))) return rc;
        if ((rc = dalloc(pl->T, sizeof(double) * (size_t)pl->npmo * pl->npp))) return rc;
        if ((rc = dalloc(pl->M, sizeof(double) * (size_t)pl->npmo * pl->npp))) return rc;
        // block-local Z indices are 24-bit products (npp^2 doubles could not
        // be allocated beyond that anyway)
        if (pl->npp >= ((int64_t)1 << 24)) return THISTLE_ENOMEM;
        if ((rc = dalloc(pl->Z, sizeof(double) * (size_t)pl->npp * pl->npp))) return rc;
        // Z through the MO factorisation of T while W3(p, :, :) fits in
        // registers (norb <= 64); THISTLE_ZFACT=0|1 overrides the choice
        pl->zfact = zfact_width(pl->norb) > 0;
        if (const char* ev = std::getenv("THISTLE_ZFACT")) {
            if (ev[0] == '0') pl->zfact = false;
            if (ev[0] == '1') pl->zfact = zfact_width(pl->norb) > 0;
        }
        if (pl->zfact) {
            const int ngt = b->ngto;
            if (zfact_fused(pl->norb, ngt)) {
                // one persistent block per CU
                HCHECK(hipDeviceGetAttribute(&pl->zgrid, hipDeviceAttributeMultiprocessorCount,
                                             o->device));
                HCHECK(zfact_fused_prepare(pl->norb, ngt));
            }
            if (pl->zgrid <= 0 &&
                (rc = dalloc(pl->W3, sizeof(double) * (size_t)pl->norb * ngt * pl->npp)))
                return rc;
            if ((rc = dalloc(pl->Cp, sizeof(double) * (size_t)ngt * zfact_width(pl->norb))))
                return rc;
            // per GTO b: the GTOs a that pair with it as (a, b) (groups up to
            // b's own) and the primitive pair index of each
            std::vector<int32_t> ae(ngt), jt((size_t)ngt * ngt, -1);
            std::vector<int64_t> poff((size_t)pl->ng * pl->ng, 0);
            int64_t o = 0;
            for (int gi = 0; gi < pl->ng; ++gi)
                for (int gj = gi; gj < pl->ng; ++gj) {
                    poff[(size_t)gi * pl->ng + gj] = o;
                    o += (int64_t)pl->gcount[gi] * pl->gcount[gj];
                }
            for (int gb = 0; gb < pl->ng; ++gb)
                for (int bb = pl->gstart[gb]; bb < pl->gstart[gb] + pl->gcount[gb]; ++bb) {
                    ae[bb] = pl->gstart[gb] + pl->gcount[gb];
                    for (int ga = 0; ga <= gb; ++ga)
                        for (int aa = pl->gstart[ga]; aa < pl->gstart[ga] + pl->gcount[ga]; ++aa)
                            jt[(size_t)bb * ngt + aa] = (int32_t)(
                                poff[(size_t)ga * pl->ng + gb] +
                                (int64_t)(aa - pl->gstart[ga]) * pl->gcount[gb] +
                                (bb - pl->gstart[gb]));
                }
            if ((rc = dalloc(pl->aend, sizeof(int32_t) * ngt))) return rc;
            if ((rc = dalloc(pl->jtab, sizeof(int32_t) * jt.size()))) return rc;
            HCHECK(hipMemcpyAsync(pl->aend.p, ae.data(), sizeof(int32_t) * ngt,
                                  hipMemcpyHostToDevice, pl->stream));
            HCHECK(hipMemcpyAsync(pl->jtab.p, jt.data(), sizeof(int32_t) * jt.size(),
                                  hipMemcpyHostToDevice, pl->stream));
            HCHECK(hipStreamSynchronize(pl->stream));
        }
        if ((rc = dalloc(pl->mopairs, sizeof(int2) * mop.size()))) return rc;
        HCHECK(hipMemcpyAsync(pl->mopairs.p, mop.data(), sizeof(int2) * mop.size(),
                              hipMemcpyHostToDevice, pl->stream));
        if (nmat > 0) {
            HCHECK(hipMemcpyAsync(pl->mat.p, mat, sizeof(int32_t) * 4 * nmat, hipMemcpyHostToDevice,
                                  pl->stream));
            HCHECK(hipMemcpyAsync(pl->total.p, total, sizeof(double) * nmat, hipMemcpyHostToDevice,
                                  pl->stream));
        }
        if (!(pl->blas = blas_take())) return THISTLE_EDEVICE;
        BCHECK(rocblas_set_stream(pl->blas, pl->stream));
    } else {
        pl->norb = norb;
        if ((rc = dalloc(pl->onerdm, sizeof(double) * norb * norb))) return rc;
        if ((rc = dalloc(pl->Z, sizeof(double) * (size_t)b->ngto * b->ngto))) return rc;
        HCHECK(hipMemcpyAsync(pl->onerdm.p, onerdm, sizeof(double) * norb * norb,
                              hipMemcpyHostToDevice, pl->stream));
    }
    HCHECK(hipStreamSynchronize(pl->stream));
    *out = pl.release();
    return THISTLE_OK;
}

// Electron-nucleus term (nuclei_electron_integration) on an elastic plan whose
// pair tables, e12 and Z are computed: out_h[q] (host).
int run_nuclear(ThistlePlan* pl, int natom, const double* Zn, const double* geom,
                double* out_h) {
    if (natom <= 0 || !Zn || !geom || !out_h) return THISTLE_EINVAL;
    std::vector<std::vector<int32_t>> byL(9);
    for (int gi = 0; gi < pl->ng; ++gi)
        for (int gj = gi; gj < pl->ng; ++gj) {
            const int LP = pl->gL[gi] + pl->gL[gj];
            if (LP > 8) return THISTLE_ELLMAX;
            byL[LP].push_back(pair_index(gi, gj, pl->ng));
        }
    std::vector<int32_t> all;
    std::vector<int> off(10, 0);
    for (int LP = 0; LP <= 8; ++LP) {
        off[LP] = (int)all.size();
        all.insert(all.end(), byL[LP].begin(), byL[LP].end());
    }
    off[9] = (int)all.size();
    DevBuf pairs, part, zn, geo, out;
    int rc;
    if ((rc = dalloc(pairs, sizeof(int32_t) * all.size()))) return rc;
    if ((rc = dalloc(part, sizeof(double) * all.size() * (size_t)pl->nq))) return rc;
    if ((rc = dalloc(zn, sizeof(double) * natom))) return rc;
    if ((rc = dalloc(geo, sizeof(double) * 3 * (size_t)natom))) return rc;
    if ((rc = dalloc(out, sizeof(double) * pl->nq))) return rc;
    HCHECK(hipMemcpyAsync(pairs.p, all.data(), sizeof(int32_t) * all.size(),
                          hipMemcpyHostToDevice, pl->stream));
    HCHECK(hipMemcpyAsync(zn.p, Zn, sizeof(double) * natom, hipMemcpyHostToDevice, pl->stream));
    HCHECK(hipMemcpyAsync(geo.p, geom, sizeof(double) * 3 * natom, hipMemcpyHostToDevice,
                          pl->stream));
    NucArgs a{};
    a.P = pl->P.as<PairInfo>();
    a.G = pl->G.as<GroupInfo>();
    a.E = pl->E.as<double>();
    a.maxl = pl->maxl;
    a.ngto = pl->ngto;
    a.nq = pl->nq;
    a.lxyz = pl->lxyz.as<char4>();
    a.Z = pl->Z.as<double>();
    a.e12 = pl->e12.as<double>();
    a.q = pl->q.as<double>();
    a.Zn = zn.as<double>();
    a.geom = geo.as<double>();
    a.natom = natom;
    a.cutoffcentre = pl->cutoffcentre;
    a.part = part.as<double>();
    for (int LP = 0; LP <= 8; ++LP)
        HCHECK(launch_nuc(a, LP, pairs.as<int32_t>() + off[LP], off[LP + 1] - off[LP], off[LP],
                          pl->stream));
    HCHECK(launch_rowsum(part.as<double>(), (int)all.size(), pl->nq, out.as<double>(),
                         pl->stream));
    HCHECK(hipMemcpyAsync(out_h, out.p, sizeof(double) * pl->nq, hipMemcpyDeviceToHost,
                          pl->stream));
    HCHECK(hipStreamSynchronize(pl->stream));
    return THISTLE_OK;
}

// Sum_ij Zn_i Zn_j sinc(q |R_i - R_j|) in the reference's loop order
// (src/main_calculation.f90:303-312, sinc :853-862)
void nuclear_nn(int natom, const double* Zn, const double* geom, const double* q, int nq,
                double* Inn) {
    for (int k = 0; k < nq; ++k) Inn[k] = 0.0;
    for (int i = 0; i < natom; ++i)
        for (int j = 0; j < natom; ++j)
            for (int k = 0; k < nq; ++k) {
                double s2 = 0.0;
                for (int c = 0; c < 3; ++c) {
                    const double d = geom[i + (size_t)natom * c] - geom[j + (size_t)natom * c];
                    s2 = s2 + d * d;
                }
                const double a = q[k] * std::fabs(std::sqrt(s2));
                const double sinc = std::fabs(a) < 1.0e-10 ? 1.0 : std::sin(a) / a;
                Inn[k] = Inn[k] + Zn[i] * Zn[j] * sinc;
            }
}

// Adjoint of one executed type-1 plan: w = dL/dS(q) -> dL/d(shell centre)
// [ng][3] (gg_host, may be null) and dL/dtotal [nmat] (dtot_host, may be null).
int run_adjoint(ThistlePlan* pl, const double* w_host, double* gg_host, double* dtot_host) {
    HCHECK(hipSetDevice(pl->device));
    const int L1 = pl->maxl + 1, T1 = 2 * pl->maxl + 1;
    DevBuf dE, w, gadj, Wz, qgrad, goff, gent, gg, dtot, scratch, Wg;
    int rc;
    if ((rc = dalloc(dE, sizeof(double) * (size_t)pl->npair * 3 * L1 * L1 * T1))) return rc;
    if ((rc = dalloc(w, sizeof(double) * pl->nq))) return rc;
    if ((rc = dalloc(gadj, sizeof(double) * 18 * (size_t)pl->nquart))) return rc;
    if ((rc = dalloc(qgrad, sizeof(double) * 12 * (size_t)pl->nquart))) return rc;
    if ((rc = dalloc(Wz, sizeof(double) * (size_t)pl->npp * pl->npp))) return rc;
    HCHECK(hipMemcpyAsync(w.p, w_host, sizeof(double) * pl->nq, hipMemcpyHostToDevice,
                          pl->stream));
    HCHECK(hipMemsetAsync(Wz.p, 0, Wz.bytes, pl->stream));
    HCHECK(launch_pair_dE(pl->G.as<GroupInfo>(), pl->P.as<PairInfo>(), pl->npair, pl->maxl,
                          dE.as<double>(), pl->stream));
    AdjointArgs j{};
    j.s = series_args(pl);
    j.dE = dE.as<double>();
    j.w = w.as<double>();
    j.gadj = gadj.as<double>();
    j.Wz = Wz.as<double>();
    j.ldw = pl->npp;
    j.qgrad = qgrad.as<double>();
    for (int LL = 16; LL >= 0; --LL) {
        int64_t n = pl->cls_off[LL + 1] - pl->cls_off[LL];
        HCHECK(launch_series_adj(j, LL, pl->cls.as<int32_t>() + pl->cls_off[LL], n, pl->waves,
                                 pl->stream));
    }
    // px tables and their derivatives: LDS up to 64 KB per workgroup, else a
    // global slice per quartet group
    const size_t lds_cap = 65536;
    size_t scratch_doubles = 0;
    auto shape = [&](int LL, int64_t n, bool& glob, int& blocks) {
        const int NG = contract_adj_groups(LL);
        const size_t per = contract_adj_table_doubles(LL, pl->cls_tmax[LL]) * NG;
        glob = per * sizeof(double) > lds_cap;
        blocks = (int)std::min<int64_t>(glob ? 1024 : (LL < 10 ? 16384 : 2048), (n + NG - 1) / NG);
        return per;
    };
    for (int LL = 0; LL <= 16; ++LL) {
        int64_t n = pl->cls_off[LL + 1] - pl->cls_off[LL];
        if (n <= 0) continue;
        bool glob;
        int blocks;
        size_t per = shape(LL, n, glob, blocks);
        if (glob) scratch_doubles = std::max(scratch_doubles, per * blocks);
    }
    if (scratch_doubles) {
        if ((rc = dalloc(scratch, sizeof(double) * scratch_doubles))) return rc;
        j.scratch = scratch.as<double>();
    }
    for (int LL = 16; LL >= 0; --LL) {
        int64_t n = pl->cls_off[LL + 1] - pl->cls_off[LL];
        if (n <= 0) continue;
        bool glob;
        int blocks;
        shape(LL, n, glob, blocks);
        j.s.hs = pl->hs.as<double>() + pl->cls_hoff[LL];
        HCHECK(launch_contract_adj(j, LL, pl->cls_tmax[LL], pl->cls.as<int32_t>() + pl->cls_off[LL],
                                   n, blocks, glob, lds_cap, pl->stream));
    }
    if (gg_host) {
        // quartet -> shell entries (e*4 + slot), grouped by shell, in list order
        int64_t Q = 0;
        std::vector<EnumRow> rows = enum_rows(pl->ng, Q, pl->type == 8);
        std::vector<int64_t> cnt(pl->ng + 1, 0);
        auto quad = [](const EnumRow& R, int x, int g[4]) { row_quartet(R, x, g); };
        for (const auto& R : rows)
            for (int u = 0; u < R.count; ++u) {
                int g[4];
                quad(R, R.r0 + u, g);
                for (int k = 0; k < 4; ++k) cnt[g[k] + 1]++;
            }
        for (int g = 0; g < pl->ng; ++g) cnt[g + 1] += cnt[g];
        std::vector<int64_t> ent(4 * (size_t)Q), fill(cnt.begin(), cnt.end() - 1);
        for (const auto& R : rows)
            for (int u = 0; u < R.count; ++u) {
                int g[4];
                quad(R, R.r0 + u, g);
                for (int k = 0; k < 4; ++k) ent[fill[g[k]]++] = (R.out + u) * 4 + k;
            }
        if ((rc = dalloc(goff, sizeof(int64_t) * cnt.size()))) return rc;
        if ((rc = dalloc(gent, sizeof(int64_t) * ent.size()))) return rc;
        if ((rc = dalloc(gg, sizeof(double) * 3 * pl->ng))) return rc;
        HCHECK(hipMemcpyAsync(goff.p, cnt.data(), sizeof(int64_t) * cnt.size(),
                              hipMemcpyHostToDevice, pl->stream));
        HCHECK(hipMemcpyAsync(gent.p, ent.data(), sizeof(int64_t) * ent.size(),
                              hipMemcpyHostToDevice, pl->stream));
        HCHECK(launch_group_grad(qgrad.as<double>(), goff.as<int64_t>(), gent.as<int64_t>(),
                                 pl->ng, gg.as<double>(), pl->stream));
        HCHECK(hipMemcpyAsync(gg_host, gg.p, sizeof(double) * 3 * pl->ng, hipMemcpyDeviceToHost,
                              pl->stream));
        // the host vectors above must outlive the copies
        HCHECK(hipStreamSynchronize(pl->stream));
    }
    if (dtot_host && pl->nmat > 0) {
        const int npmo = pl->npmo;
        const double one = 1.0, zero = 0.0;
        if ((rc = dalloc(Wg, sizeof(double) * (size_t)npmo * npmo))) return rc;
        if ((rc = dalloc(dtot, sizeof(double) * pl->nmat))) return rc;
        // Y = T Wz (Wz stored transposed: element (xk, xb) holds dL/dZ(xb, xk))
        BCHECK(rocblas_dgemm(pl->blas, rocblas_operation_none, rocblas_operation_transpose, npmo,
                             (rocblas_int)pl->npp, (rocblas_int)pl->npp, &one, pl->T.as<double>(),
                             npmo, Wz.as<double>(), (rocblas_int)pl->npp, &zero,
                             pl->M.as<double>(), npmo));
        // W = Y T^T  (npmo x npmo) = dL/dGs
        BCHECK(rocblas_dgemm(pl->blas, rocblas_operation_none, rocblas_operation_transpose, npmo,
                             npmo, (rocblas_int)pl->npp, &one, pl->M.as<double>(), npmo,
                             pl->T.as<double>(), npmo, &zero, Wg.as<double>(), npmo));
        HCHECK(launch_dtotal(pl->mat.as<int32_t>(), pl->nmat, pl->norb, Wg.as<double>(), npmo,
                             dtot.as<double>(), pl->stream));
        HCHECK(hipMemcpyAsync(dtot_host, dtot.p, sizeof(double) * pl->nmat,
                              hipMemcpyDeviceToHost, pl->stream));
    }
    HCHECK(hipStreamSynchronize(pl->stream));
    return THISTLE_OK;
}

}  // namespace

// Many independent frames executed together: per-frame plans spread over a
// few streams (fork/join through events), captured once into one hipGraph
// and replayed.  The first execution tunes the series kernel's waves per CU.
struct ThistleBatch {
    static constexpr int NS = 4;
    int device = 0, cus = 0, waves_per_cu = 0;
    hipStream_t streams[NS] = {};
    hipEvent_t fork = nullptr, join[NS] = {}, t0 = nullptr, t1 = nullptr;
    std::vector<std::unique_ptr<ThistlePlan>> plans;
    std::vector<int64_t> soff;
    DevBuf S;
    int64_t nqtot = 0, units = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool tuned = false, graph_ok = false;
    float ms = 0.f;
    // Cross-frame quartet stage (FP64, every q grid within one 512-point
    // chunk).  The frames form a few groups of about equal work; per group
    // and LL class one contraction launch and at most three series and
    // three far-regime launches over all its frames, driven by work tables,
    // and one zeroing and one reduction launch.  A group's per-frame set-up
    // stages (streams 2 and 3) overlap the previous group's quartet stage
    // (contractions on stream 0, series and reduction on stream 1).
    bool merged = false;
    int rows = 0, ngroups = 0;  // partial rows (series chunks) per frame, groups: tuned
    struct SLaunch {
        int LL, sq;
        bool far;
        int64_t off, n, blocks;
    };
    struct Group {
        int f0 = 0, f1 = 0;
        DevBuf dcw, dsw, dfill, dred;
        int64_t cw_off[18] = {}, cw_blocks[17] = {};
        int cls_tmax[17] = {};
        std::vector<SLaunch> slaunch;
        int nfill = 0, nred = 0;
        int64_t fill_blocks = 0, red_blocks = 0;
        hipEvent_t cev[17] = {}, joina[2] = {}, fill_ev = nullptr;
        ~Group() {
            for (auto& e : cev)
                if (e) (void)hipEventDestroy(e);
            for (auto& e : joina)
                if (e) (void)hipEventDestroy(e);
            if (fill_ev) (void)hipEventDestroy(fill_ev);
        }
    };
    DevBuf dargs;
    std::vector<std::unique_ptr<Group>> groups;
    hipEvent_t ser_ev = nullptr;
    ~ThistleBatch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        plans.clear();  // plans use the batch streams
        for (auto& e : join)
            if (e) (void)hipEventDestroy(e);
        groups.clear();
        if (ser_ev) (void)hipEventDestroy(ser_ev);
        if (fork) (void)hipEventDestroy(fork);
        if (t0) (void)hipEventDestroy(t0);
        if (t1) (void)hipEventDestroy(t1);
        for (auto& st : streams)
            if (st) (void)hipStreamDestroy(st);
    }
};

namespace {

void batch_set_waves(ThistleBatch* b, int per_cu) {
    b->waves_per_cu = per_cu;
    for (auto& p : b->plans)
        p->waves = (int)std::max<int64_t>(
            1, std::min<int64_t>({(int64_t)per_cu * b->cus, p->waves_alloc, p->nquart}));
}

// Work tables of the cross-frame quartet stage for `rows` partial rows per
// frame (a frame's class of n records is split into min(rows, n) chunks,
// as a single-frame execution with waves = rows splits it).
int group_tables(ThistleBatch* b, ThistleBatch::Group* g) {
    std::vector<ClassWork> cw;
    std::vector<SeriesWork> sw;
    std::vector<FillWork> fw;
    std::vector<RedWork> rw;
    int64_t fb = 0, rb = 0;
    const int rq = reduce_q_per_block();
    for (int f = g->f0; f < g->f1; ++f) {
        ThistlePlan* p = b->plans[f].get();
        const int64_t len = (int64_t)p->waves * p->nq;
        fw.push_back({fb, p->part.as<double>(), len});
        fb += (len + 255) / 256;
        fw.push_back({fb, p->partc.as<double>(), len});
        fb += (len + 255) / 256;
        rw.push_back({rb, p->part.as<double>(), p->partc.as<double>(),
                      b->S.as<double>() + b->soff[f], p->waves, p->nq});
        rb += (p->nq + rq - 1) / rq;
    }
    g->slaunch.clear();
    for (int LL = 0; LL <= 16; ++LL) {
        g->cw_off[LL] = (int64_t)cw.size();
        const int ngq = contract_groups_per_block(LL);
        int64_t blocks = 0;
        int tmax = 0;
        for (int f = g->f0; f < g->f1; ++f) {
            ThistlePlan* p = b->plans[f].get();
            const int64_t n = p->cls_off[LL + 1] - p->cls_off[LL];
            if (n <= 0) continue;
            cw.push_back({blocks, n, p->qhdr.as<QHdr>() + p->cls_off[LL],
                          p->hs.as<double>() + p->cls_hoff[LL], f});
            blocks += (n + ngq - 1) / ngq;
            tmax = std::max(tmax, p->cls_tmax[LL]);
        }
        g->cw_blocks[LL] = blocks;
        g->cls_tmax[LL] = tmax;
        // series launches in the order a single frame issues them: the
        // 256-point groups (q slices of 4), the tail, then the far regime
        for (int far = 0; far < 2; ++far)
            for (int sq : {4, 2, 1}) {
                const int64_t off = (int64_t)sw.size();
                int64_t sb = 0;
                for (int f = g->f0; f < g->f1; ++f) {
                    ThistlePlan* p = b->plans[f].get();
                    const int64_t n = p->cls_off[LL + 1] - p->cls_off[LL];
                    if (n <= 0 || (far && !p->far)) continue;
                    int qs0[2], qs1[2], sqs[2];
                    const int np = series_q_split(p->nq, false, qs0, qs1, sqs);
                    for (int k = 0; k < np; ++k) {
                        if (sqs[k] != sq) continue;
                        const int nchunk = (int)std::min<int64_t>(p->waves, n);
                        const int ny = (qs1[k] - qs0[k] + sq * 64 - 1) / (sq * 64);
                        sw.push_back({sb, n, p->hs.as<double>() + p->cls_hoff[LL], f, nchunk, ny,
                                      qs0[k], qs1[k], 0});
                        sb += (int64_t)nchunk * ny;
                    }
                }
                if (sb > 0)
                    g->slaunch.push_back({LL, sq, far != 0, off, (int64_t)sw.size() - off, sb});
            }
    }
    g->cw_off[17] = (int64_t)cw.size();
    g->nfill = (int)fw.size();
    g->fill_blocks = fb;
    g->nred = (int)rw.size();
    g->red_blocks = rb;
    int rc;
    if ((rc = dalloc(g->dcw, sizeof(ClassWork) * std::max<size_t>(cw.size(), 1))) ||
        (rc = dalloc(g->dsw, sizeof(SeriesWork) * std::max<size_t>(sw.size(), 1))) ||
        (rc = dalloc(g->dfill, sizeof(FillWork) * std::max<size_t>(fw.size(), 1))) ||
        (rc = dalloc(g->dred, sizeof(RedWork) * std::max<size_t>(rw.size(), 1))))
        return rc;
    if (!cw.empty())
        HCHECK(hipMemcpy(g->dcw.p, cw.data(), sizeof(ClassWork) * cw.size(),
                         hipMemcpyHostToDevice));
    if (!sw.empty())
        HCHECK(hipMemcpy(g->dsw.p, sw.data(), sizeof(SeriesWork) * sw.size(),
                         hipMemcpyHostToDevice));
    if (!fw.empty())
        HCHECK(hipMemcpy(g->dfill.p, fw.data(), sizeof(FillWork) * fw.size(),
                         hipMemcpyHostToDevice));
    if (!rw.empty())
        HCHECK(hipMemcpy(g->dred.p, rw.data(), sizeof(RedWork) * rw.size(),
                         hipMemcpyHostToDevice));
    return THISTLE_OK;
}

int batch_build_tables(ThistleBatch* b, int rows) {
    b->rows = rows;
    const int nf = (int)b->plans.size();
    std::vector<SeriesArgs> args((size_t)nf);
    for (int f = 0; f < nf; ++f) {
        ThistlePlan* p = b->plans[f].get();
        p->waves = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)rows, p->waves_alloc,
                                                                p->nquart}));
        args[f] = series_args(p);
        args[f].q0 = 0;
        args[f].nqc = p->nq;
    }
    int rc = dalloc(b->dargs, sizeof(SeriesArgs) * std::max<size_t>(args.size(), 1));
    if (rc) return rc;
    HCHECK(hipMemcpy(b->dargs.p, args.data(), sizeof(SeriesArgs) * args.size(),
                     hipMemcpyHostToDevice));
    for (auto& g : b->groups)
        if ((rc = group_tables(b, g.get()))) return rc;
    return THISTLE_OK;
}

// Frames split into contiguous groups of about equal work (quartets x q
// points), with their events.
int batch_make_groups(ThistleBatch* b, int ngroups) {
    const int nf = (int)b->plans.size();
    ngroups = std::max(1, std::min(ngroups, nf));
    int64_t total = 0;
    for (auto& p : b->plans) total += p->nquart * (int64_t)p->nq;
    b->groups.clear();
    int f = 0;
    int64_t acc = 0;
    for (int k = 0; k < ngroups; ++k) {
        std::unique_ptr<ThistleBatch::Group> g(new ThistleBatch::Group());
        g->f0 = f;
        const int64_t goal = total * (k + 1) / ngroups;
        while (f < nf && (k == ngroups - 1 || acc < goal || f == g->f0)) {
            acc += b->plans[f]->nquart * (int64_t)b->plans[f]->nq;
            ++f;
        }
        g->f1 = f;
        if (g->f1 == g->f0) continue;
        for (auto& e : g->cev) HCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        for (auto& e : g->joina) HCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        HCHECK(hipEventCreateWithFlags(&g->fill_ev, hipEventDisableTiming));
        b->groups.push_back(std::move(g));
    }
    return THISTLE_OK;
}

// Rows per frame and number of frame groups.  One group: the per-frame
// stages spread over all four streams; several: over streams 2 and 3, so
// that they overlap the previous group's quartet stage on streams 0 and 1.
int batch_configure(ThistleBatch* b, int rows, int ngroups) {
    int rc = batch_make_groups(b, ngroups);
    if (rc) return rc;
    const bool split = b->groups.size() > 1;
    for (size_t f = 0; f < b->plans.size(); ++f) {
        ThistlePlan* p = b->plans[f].get();
        p->stream = split ? b->streams[2 + (f & 1)] : b->streams[f % ThistleBatch::NS];
        if (p->blas) BCHECK(rocblas_set_stream(p->blas, p->stream));
    }
    b->ngroups = (int)b->groups.size();
    return batch_build_tables(b, rows);
}

int batch_enqueue_merged(ThistleBatch* b) {
    hipStream_t s0 = b->streams[0], s1 = b->streams[1];
    HCHECK(hipEventRecord(b->fork, s0));
    for (int s = 1; s < ThistleBatch::NS; ++s) HCHECK(hipStreamWaitEvent(b->streams[s], b->fork, 0));
    const SeriesArgs* args = b->dargs.as<SeriesArgs>();
    for (auto& gp : b->groups) {
        ThistleBatch::Group* g = gp.get();
        for (int f = g->f0; f < g->f1; ++f) {
            int rc = enqueue_execute(b->plans[f].get(), false, nullptr, false);
            if (rc) return rc;
        }
        if (b->groups.size() > 1) {
            HCHECK(hipEventRecord(g->joina[0], b->streams[2]));
            HCHECK(hipEventRecord(g->joina[1], b->streams[3]));
            HCHECK(hipStreamWaitEvent(s0, g->joina[0], 0));
            HCHECK(hipStreamWaitEvent(s0, g->joina[1], 0));
        } else {
            for (int s = 1; s < ThistleBatch::NS; ++s) {
                HCHECK(hipEventRecord(b->join[s], b->streams[s]));
                HCHECK(hipStreamWaitEvent(s0, b->join[s], 0));
            }
        }
        HCHECK(launch_fill_multi(g->dfill.as<FillWork>(), g->nfill, g->fill_blocks, s0));
        HCHECK(hipEventRecord(g->fill_ev, s0));
        HCHECK(hipStreamWaitEvent(s1, g->fill_ev, 0));
        for (int LL = 0; LL <= 16; ++LL) {
            const int nw = (int)(g->cw_off[LL + 1] - g->cw_off[LL]);
            if (nw <= 0) continue;
            HCHECK(launch_contract_multi(args, g->dcw.as<ClassWork>() + g->cw_off[LL], nw,
                                         g->cw_blocks[LL], LL, g->cls_tmax[LL], false, s0));
            HCHECK(hipEventRecord(g->cev[LL], s0));
        }
        int waited = -1;
        for (const auto& L : g->slaunch) {
            if (L.LL != waited) {
                HCHECK(hipStreamWaitEvent(s1, g->cev[L.LL], 0));
                waited = L.LL;
            }
            HCHECK(launch_series_multi(args, g->dsw.as<SeriesWork>() + L.off, (int)L.n, L.blocks,
                                       L.LL, L.sq, false, L.far, s1));
        }
        HCHECK(launch_reduce_multi(g->dred.as<RedWork>(), g->nred, g->red_blocks, s1));
    }
    HCHECK(hipEventRecord(b->ser_ev, s1));
    HCHECK(hipStreamWaitEvent(s0, b->ser_ev, 0));
    for (int s = 2; s < ThistleBatch::NS; ++s) {
        HCHECK(hipEventRecord(b->join[s], b->streams[s]));
        HCHECK(hipStreamWaitEvent(s0, b->join[s], 0));
    }
    return THISTLE_OK;
}

int batch_enqueue(ThistleBatch* b) {
    if (b->merged) return batch_enqueue_merged(b);
    HCHECK(hipEventRecord(b->fork, b->streams[0]));
    for (int s = 1; s < ThistleBatch::NS; ++s) HCHECK(hipStreamWaitEvent(b->streams[s], b->fork, 0));
    for (size_t f = 0; f < b->plans.size(); ++f) {
        int rc = enqueue_execute(b->plans[f].get(), false, b->S.as<double>() + b->soff[f]);
        if (rc) return rc;
    }
    for (int s = 1; s < ThistleBatch::NS; ++s) {
        HCHECK(hipEventRecord(b->join[s], b->streams[s]));
        HCHECK(hipStreamWaitEvent(b->streams[0], b->join[s], 0));
    }
    return THISTLE_OK;
}

int batch_timed(ThistleBatch* b, float& ms) {
    HCHECK(hipEventRecord(b->t0, b->streams[0]));
    int rc = batch_enqueue(b);
    if (rc) return rc;
    HCHECK(hipEventRecord(b->t1, b->streams[0]));
    HCHECK(hipStreamSynchronize(b->streams[0]));
    HCHECK(hipEventElapsedTime(&ms, b->t0, b->t1));
    return THISTLE_OK;
}

// Autotune on eager executions, then capture the chosen configuration.
// Per-frame launches: series waves per CU.  Cross-frame launches: partial
// rows per frame, i.e. records per series wave (fewer rows, longer chunks
// and a shorter reduction; more rows, more waves in flight).
int batch_prepare(ThistleBatch* b) {
    // merged: (rows per frame, frame groups) = c / 16, c % 16
    const std::vector<int> cand =
        b->merged ? std::vector<int>{512 * 16 + 1, 128 * 16 + 1, 2048 * 16 + 1, 512 * 16 + 2,
                                     512 * 16 + 4}
                  : std::vector<int>{4, 8, 12};
    float best = 0.f;
    int pick = cand[0];
    for (int c : cand) {
        if (b->merged) {
            int rc = batch_configure(b, c / 16, c % 16);
            if (rc) return rc;
        } else {
            batch_set_waves(b, c);
        }
        float ms = 0.f;
        int rc = batch_timed(b, ms);  // the first candidate also warms up rocBLAS
        if (rc) return rc;
        rc = batch_timed(b, ms);
        if (rc) return rc;
        if (c == cand[0] || ms < best) {
            best = ms;
            pick = c;
        }
    }
    if (b->merged) {
        int rc = batch_configure(b, pick / 16, pick % 16);
        if (rc) return rc;
    } else {
        batch_set_waves(b, pick);
    }
    b->tuned = true;
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(b->streams[0], hipStreamCaptureModeGlobal) != hipSuccess)
        return THISTLE_OK;  // eager fallback
    int rc = batch_enqueue(b);
    hipError_t ec = hipStreamEndCapture(b->streams[0], &g);
    if (rc || ec != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return THISTLE_OK;
    }
    hipGraphExec_t ex = nullptr;
    if (hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return THISTLE_OK;
    }
    b->graph = g;
    b->exec = ex;
    b->graph_ok = true;
    return THISTLE_OK;
}

}  // namespace

extern "C" {

int32_t thistle_batch_create(int32_t nframes, const ThistleBasis* bases, const int64_t* nmat,
                             const int32_t* const* mats, const double* const* totals,
                             const ThistleOpts* opts, ThistleBatch** out) {
    if (!out) return THISTLE_EINVAL;
    *out = nullptr;
    if (nframes <= 0 || !bases || !nmat || !mats || !totals) return THISTLE_EINVAL;
    const ThistleOpts* o = opts ? opts : &kDefaultOpts;
    std::unique_ptr<ThistleBatch> b(new ThistleBatch());
    b->device = o->device;
    HCHECK(hipSetDevice(o->device));
    hipDeviceProp_t prop;
    HCHECK(hipGetDeviceProperties(&prop, o->device));
    b->cus = prop.multiProcessorCount;
    for (auto& st : b->streams) HCHECK(hipStreamCreate(&st));
    HCHECK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    for (auto& e : b->join) HCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HCHECK(hipEventCreate(&b->t0));
    HCHECK(hipEventCreate(&b->t1));
    for (int f = 0; f < nframes; ++f) {
        ThistleOpts of = *o;
        of.stream = b->streams[f % ThistleBatch::NS];
        of.waves = 12 * b->cus;
        ThistlePlan* raw = nullptr;
        int rc = thistle_plan_create(1, &bases[f], nmat[f], mats[f], totals[f], 0, nullptr, &of,
                                     &raw);
        if (rc) return rc;
        raw->drop_overlap();  // the frames already overlap across the batch streams
        b->plans.emplace_back(raw);
        b->soff.push_back(b->nqtot);
        b->nqtot += bases[f].nq;
        b->units += raw->nquart * (int64_t)raw->nq;
    }
    int rc = dalloc(b->S, sizeof(double) * b->nqtot);
    if (rc) return rc;
    // the cross-frame quartet stage: FP64 frames whose q grids fit one chunk
    b->merged = o->precision == THISTLE_FP64;
    for (int f = 0; f < nframes; ++f) b->merged = b->merged && bases[f].nq <= 512;
    if (b->merged) {
        HCHECK(hipEventCreateWithFlags(&b->ser_ev, hipEventDisableTiming));
        if ((rc = batch_configure(b.get(), 512, 1))) return rc;
    }
    *out = b.release();
    return THISTLE_OK;
}

int32_t thistle_batch_execute(ThistleBatch* b, double* S_out) {
    if (!b) return THISTLE_EINVAL;
    HCHECK(hipSetDevice(b->device));
    if (!b->tuned) {
        int rc = batch_prepare(b);
        if (rc) return rc;
    }
    HCHECK(hipEventRecord(b->t0, b->streams[0]));
    if (b->exec) {
        HCHECK(hipGraphLaunch(b->exec, b->streams[0]));
    } else {
        int rc = batch_enqueue(b);
        if (rc) return rc;
    }
    HCHECK(hipEventRecord(b->t1, b->streams[0]));
    if (S_out)
        HCHECK(hipMemcpyAsync(S_out, b->S.p, sizeof(double) * b->nqtot, hipMemcpyDeviceToHost,
                              b->streams[0]));
    HCHECK(hipStreamSynchronize(b->streams[0]));
    HCHECK(hipEventElapsedTime(&b->ms, b->t0, b->t1));
    return THISTLE_OK;
}

int32_t thistle_batch_info(ThistleBatch* b, double* ms_last, int64_t* units, int32_t* waves,
                           int32_t* graph) {
    if (!b) return THISTLE_EINVAL;
    if (ms_last) *ms_last = b->ms;
    if (units) *units = b->units;
    if (waves) *waves = b->merged ? b->rows : b->waves_per_cu * b->cus;
    if (graph) *graph = b->graph_ok ? 1 : 0;
    return THISTLE_OK;
}

int32_t thistle_batch_tuning(ThistleBatch* b, int32_t* rows, int32_t* groups) {
    if (!b) return THISTLE_EINVAL;
    if (rows) *rows = b->merged ? b->rows : 0;
    if (groups) *groups = b->merged ? b->ngroups : 0;
    return THISTLE_OK;
}

void thistle_batch_destroy(ThistleBatch* b) { delete b; }

int32_t thistle_total_scattering_adjoint(const ThistleBasis* basis, int64_t nmat,
                                         const int32_t* mat, const double* total, int32_t natom,
                                         const double* geom, const double* dL_dS,
                                         const ThistleOpts* opts, double* S_out, double* dL_dR,
                                         double* dL_dtotal) {
    if (!basis || !dL_dS) return THISTLE_EINVAL;
    const ThistleOpts* o = opts ? opts : &kDefaultOpts;
    // shell -> atom by centre (host, before any device call)
    std::vector<int> gatom;
    if (dL_dR) {
        if (natom <= 0 || !geom) return THISTLE_EINVAL;
        std::vector<GroupInfo> groups;
        int maxl = 0;
        int rc = check_basis(basis, groups, maxl);
        if (rc) return rc;
        gatom.assign(basis->ng, -1);
        for (int g = 0; g < basis->ng; ++g) {
            const double c[3] = {groups[g].x, groups[g].y, groups[g].z};
            for (int a = 0; a < natom && gatom[g] < 0; ++a) {
                bool same = true;
                for (int k = 0; k < 3; ++k)
                    same = same && std::fabs(geom[a + (int64_t)natom * k] - c[k]) <=
                                       1e-10 * (1.0 + std::fabs(c[k]));
                if (same) gatom[g] = a;
            }
            if (gatom[g] < 0) return THISTLE_EINVAL;
        }
    }
    ThistlePlan* raw = nullptr;
    int rc = thistle_plan_create(1, basis, nmat, mat, total, 0, nullptr, o, &raw);
    if (rc) return rc;
    std::unique_ptr<ThistlePlan> pl(raw);
    rc = thistle_plan_execute(pl.get(), S_out);
    if (rc) return rc;
    std::vector<double> gg(dL_dR ? 3 * (size_t)basis->ng : 0);
    rc = run_adjoint(pl.get(), dL_dS, dL_dR ? gg.data() : nullptr, dL_dtotal);
    if (rc) return rc;
    if (dL_dR) {
        for (int64_t k = 0; k < (int64_t)natom * 3; ++k) dL_dR[k] = 0.0;
        for (int g = 0; g < basis->ng; ++g)
            for (int k = 0; k < 3; ++k) dL_dR[gatom[g] + (int64_t)natom * k] += gg[3 * g + k];
    }
    return THISTLE_OK;
}

const char* thistle_version(void) { return "thistle-mi355x 0.1 (gfx950)"; }

int32_t thistle_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return -THISTLE_EDEVICE;
    return n;
}

int32_t thistle_device_synchronize(int32_t device) {
    HCHECK(hipSetDevice(device));
    HCHECK(hipDeviceSynchronize());
    return THISTLE_OK;
}

int32_t thistle_plan_create(int32_t type, const ThistleBasis* basis, int64_t nmat,
                            const int32_t* mat, const double* total, int32_t norb,
                            const double* onerdm, const ThistleOpts* opts, ThistlePlan** plan) {
    if (!plan) return THISTLE_EINVAL;
    *plan = nullptr;
    return plan_create_impl(type, basis, nmat, mat, total, norb, onerdm,
                            opts ? opts : &kDefaultOpts, plan);
}

int32_t thistle_plan_execute(ThistlePlan* pl, double* result_out) {
    if (!pl) return THISTLE_EINVAL;
    HCHECK(hipSetDevice(pl->device));
    int rc = enqueue_execute(pl, true, nullptr);
    if (rc) return rc;
    if (result_out)
        HCHECK(hipMemcpyAsync(result_out, pl->S.p, sizeof(double) * pl->nq, hipMemcpyDeviceToHost,
                              pl->stream));
    HCHECK(hipStreamSynchronize(pl->stream));
    float t[5];
    for (int k = 0; k < 5; ++k) {
        HCHECK(hipEventElapsedTime(&t[k], pl->ev[k], pl->ev[k + 1]));
        pl->ms[k] = t[k];
    }
    float kt[2];
    HCHECK(hipEventElapsedTime(&kt[0], pl->kev[0], pl->kev[1]));
    HCHECK(hipEventElapsedTime(&kt[1], pl->kev[2], pl->kev[3]));
    pl->kms[0] = kt[0];
    pl->kms[1] = kt[1];
    if (pl->precision == THISTLE_MIXED)
        HCHECK(hipMemcpy(&pl->last_nsel, pl->nsel.p, sizeof(int32_t), hipMemcpyDeviceToHost));
    return THISTLE_OK;
}

int32_t thistle_plan_kernel_timings(ThistlePlan* pl, double* contract_ms, double* series_ms) {
    if (!pl) return THISTLE_EINVAL;
    if (contract_ms) *contract_ms = pl->kms[0];
    if (series_ms) *series_ms = pl->kms[1];
    return THISTLE_OK;
}

int32_t thistle_plan_mixed_fallback(ThistlePlan* pl, int32_t* count, int32_t* idx, int32_t cap,
                                    double* scale) {
    if (!pl || !count) return THISTLE_EINVAL;
    *count = 0;
    if (pl->precision != THISTLE_MIXED) return THISTLE_OK;
    HCHECK(hipSetDevice(pl->device));
    HCHECK(hipStreamSynchronize(pl->stream));
    int32_t n = 0;
    HCHECK(hipMemcpy(&n, pl->nsel.p, sizeof(int32_t), hipMemcpyDeviceToHost));
    *count = n;
    if (idx && cap >= n && n > 0)
        HCHECK(hipMemcpy(idx, pl->qsel.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    if (scale)
        HCHECK(hipMemcpy(scale, pl->Aq.p, sizeof(double) * pl->nq, hipMemcpyDeviceToHost));
    return THISTLE_OK;
}

int32_t thistle_plan_result(ThistlePlan* pl, double* result_out) {
    if (!pl || !result_out) return THISTLE_EINVAL;
    HCHECK(hipMemcpy(result_out, pl->S.p, sizeof(double) * pl->nq, hipMemcpyDeviceToHost));
    return THISTLE_OK;
}

int32_t thistle_plan_timings(ThistlePlan* pl, double* ms5, int64_t* nquart, int64_t* units) {
    if (!pl) return THISTLE_EINVAL;
    if (ms5)
        for (int k = 0; k < 5; ++k) ms5[k] = pl->ms[k];
    if (nquart) *nquart = pl->nlisted;
    if (units) *units = pl->units;
    return THISTLE_OK;
}

int32_t thistle_plan_set_quartets(ThistlePlan* pl, const int64_t* positions, int64_t n) {
    if (!pl) return THISTLE_EINVAL;
    HCHECK(hipSetDevice(pl->device));
    return set_quartets(pl, positions, n);
}

void thistle_plan_destroy(ThistlePlan* pl) { delete pl; }

int32_t thistle_total_from_twordm(const ThistleBasis* basis, int64_t nmat, const int32_t* mat,
                                  const double* total, const ThistleOpts* opts, double* result_out) {
    ThistlePlan* pl = nullptr;
    int rc = thistle_plan_create(1, basis, nmat, mat, total, 0, nullptr, opts, &pl);
    if (rc) return rc;
    rc = thistle_plan_execute(pl, result_out);
    thistle_plan_destroy(pl);
    return rc;
}

int32_t thistle_nuclei_electron(const ThistleBasis* basis, int32_t norb, const double* onerdm,
                                int32_t natom, const double* Zn, const double* geom,
                                const ThistleOpts* opts, double* result_out) {
    if (natom <= 0 || !Zn || !geom || !result_out) return THISTLE_EINVAL;
    ThistlePlan* raw = nullptr;
    int rc = thistle_plan_create(2, basis, 0, nullptr, nullptr, norb, onerdm, opts, &raw);
    if (rc) return rc;
    std::unique_ptr<ThistlePlan> pl(raw);
    // pair tables, e12 and the elastic Z; no quartet stage
    rc = run_tables_and_enum(pl.get(), false);
    if (rc) return rc;
    HCHECK(launch_zel(pl->mmod.as<double>(), pl->nmo, pl->norb, pl->onerdm.as<double>(),
                      pl->ngto, pl->Z.as<double>(), pl->stream));
    return run_nuclear(pl.get(), natom, Zn, geom, result_out);
}

int32_t thistle_elastic_from_onerdm(const ThistleBasis* basis, int32_t norb, const double* onerdm,
                                    const ThistleOpts* opts, double* result_out) {
    ThistlePlan* pl = nullptr;
    int rc = thistle_plan_create(2, basis, 0, nullptr, nullptr, norb, onerdm, opts, &pl);
    if (rc) return rc;
    rc = thistle_plan_execute(pl, result_out);
    thistle_plan_destroy(pl);
    return rc;
}

int32_t thistle_twordm_dense(const int32_t* confs, int32_t nconf, int32_t lconfs,
                             const double* civecs, int32_t nstates, int32_t state1, int32_t state2,
                             double* twordm) {
    if (!confs || !civecs || !twordm) return THISTLE_EINVAL;
    return thistle_twordm_dense_host(confs, nconf, lconfs, civecs, nstates, state1, state2, twordm);
}

int64_t thistle_twordm_select(const double* twordm, int32_t norb, double cutoff, int32_t* mat,
                              double* total, int64_t cap) {
    if (!twordm || norb <= 0) return -THISTLE_EINVAL;
    return thistle_twordm_select_host(twordm, norb, cutoff, cap > 0 ? mat : nullptr,
                                      cap > 0 ? total : nullptr, cap);
}

namespace {

#define DCHECK(x)                                                               \
    do {                                                                        \
        hipError_t e_ = (x);                                                    \
        if (e_ != hipSuccess) {                                                 \
            std::fprintf(stderr, "thistle: %s failed: %s (%s:%d)\n", #x,        \
                         hipGetErrorString(e_), __FILE__, __LINE__);            \
            return THISTLE_EDEVICE;                                             \
        }                                                                       \
    } while (0)
#define DALLOC(b, n)                                        \
    do {                                                    \
        if (dalloc(b, n)) return THISTLE_ENOMEM;            \
    } while (0)

struct StreamGuard {
    hipStream_t s = nullptr;
    ~StreamGuard() {
        if (s) (void)hipStreamDestroy(s);
    }
};

// Steps 4-5 of the ordered accumulation: the nc contributions (element
// keys, values) in the reference's global order, stably radix-sorted by
// element, then one ordered sum per element from 0 into dense (nel sums).
int sort_and_sum(DevBuf& keys, DevBuf& vals, DevBuf& keys2, DevBuf& vals2, DevBuf& temp,
                 int64_t nc, int64_t nel, DevBuf& dense, hipStream_t st) {
    int keybits = 1;
    while (((int64_t)1 << keybits) < nel) ++keybits;
    size_t tb = 0;
    DCHECK(rdm_sort(nullptr, tb, keys.as<uint32_t>(), keys2.as<uint32_t>(), vals.as<double>(),
                    vals2.as<double>(), nc, keybits, st));
    if (tb > temp.bytes) DALLOC(temp, tb);
    if (nc > 0)
        DCHECK(rdm_sort(temp.p, tb, keys.as<uint32_t>(), keys2.as<uint32_t>(), vals.as<double>(),
                        vals2.as<double>(), nc, keybits, st));
    DevBuf start;
    DALLOC(start, sizeof(int64_t) * (size_t)(nel + 1));
    DALLOC(dense, sizeof(double) * (size_t)nel);
    DCHECK(rdm_esum(keys2.as<uint32_t>(), vals2.as<double>(), nc, nel, start.as<int64_t>(),
                    dense.as<double>(), st));
    // start is freed on return: the sums are queued on st ahead of the free
    DCHECK(hipStreamSynchronize(st));
    return THISTLE_OK;
}

// The determinant builders of rdm.hip share one ordered accumulation: an
// item list in the reference's loop order ((c1, c2) pairs with ndiff <=
// maxdiff, after ndiag leading no-difference items for the bit builders),
// each item's contribution count, scanned into global positions, every
// (element, value) written at its position, a stable radix sort by element
// and one ordered sum per element from 0.  dense receives the nel sums.
enum class RdmKind { Twordm, Onerdm, Bit2, Bit1 };

int rdm_accumulate(RdmKind kind, const uint64_t* masks, int nconf, int nw, int N,
                   const double* c1v, const double* c2v, DevBuf& dense, hipStream_t st) {
    const bool bit = kind == RdmKind::Bit2 || kind == RdmKind::Bit1;
    const int maxdiff = (kind == RdmKind::Twordm || kind == RdmKind::Bit2) ? 4 : 2;
    const int64_t ndiag = bit ? nconf : 0;
    const int64_t nel = (kind == RdmKind::Twordm || kind == RdmKind::Bit2)
                            ? (int64_t)N * N * N * N
                            : (int64_t)N * N;
    DevBuf cnt, off, c1s, c2s, nds, ccnt, coff, keys, vals, keys2, vals2, temp;
    // 1. connected pairs per c1, c2 increasing
    DALLOC(cnt, sizeof(int64_t) * ((size_t)nconf + 1));
    DALLOC(off, sizeof(int64_t) * ((size_t)nconf + 1));
    DCHECK(launch_fill(reinterpret_cast<double*>(cnt.p), nconf + 1, 0.0, st));
    DCHECK(rdm_pairs(masks, nconf, nw, maxdiff, cnt.as<int64_t>(), nullptr, nullptr, nullptr,
                     nullptr, st));
    size_t tb = 0;
    DCHECK(rdm_scan_i64(nullptr, tb, cnt.as<int64_t>(), off.as<int64_t>(), nconf + 1, st));
    DALLOC(temp, tb);
    DCHECK(rdm_scan_i64(temp.p, tb, cnt.as<int64_t>(), off.as<int64_t>(), nconf + 1, st));
    int64_t np = 0;
    DCHECK(hipMemcpyAsync(&np, off.as<int64_t>() + nconf, sizeof(int64_t), hipMemcpyDeviceToHost,
                          st));
    DCHECK(hipStreamSynchronize(st));
    const int64_t ni = ndiag + np;
    DALLOC(c1s, sizeof(int32_t) * (size_t)std::max<int64_t>(ni, 1));
    DALLOC(c2s, sizeof(int32_t) * (size_t)std::max<int64_t>(ni, 1));
    DALLOC(nds, (size_t)std::max<int64_t>(ni, 1));
    if (ndiag) DCHECK(bit_diag(nconf, c1s.as<int32_t>(), c2s.as<int32_t>(), nds.as<int8_t>(), st));
    DCHECK(rdm_pairs(masks, nconf, nw, maxdiff, nullptr, off.as<int64_t>(),
                     c1s.as<int32_t>() + ndiag, c2s.as<int32_t>() + ndiag,
                     nds.as<int8_t>() + ndiag, st));
    // 2. contributions per item, scanned into positions
    DALLOC(ccnt, sizeof(int64_t) * (size_t)(ni + 1));
    DALLOC(coff, sizeof(int64_t) * (size_t)(ni + 1));
    DCHECK(launch_fill(reinterpret_cast<double*>(ccnt.p), ni + 1, 0.0, st));
    const int32_t *i1 = c1s.as<int32_t>(), *i2 = c2s.as<int32_t>();
    const int8_t* id = nds.as<int8_t>();
    switch (kind) {
        case RdmKind::Twordm:
            DCHECK(rdm_ccount(masks, nw, i1, i2, id, ni, ccnt.as<int64_t>(), st));
            break;
        case RdmKind::Onerdm:
            DCHECK(rdm1_ccount(masks, nw, i1, id, ni, ccnt.as<int64_t>(), st));
            break;
        case RdmKind::Bit2:
            DCHECK(bit2_ccount(masks, i1, i2, id, ni, (int)ndiag, ccnt.as<int64_t>(), st));
            break;
        case RdmKind::Bit1:
            DCHECK(bit1_ccount(masks, i1, i2, id, ni, (int)ndiag, ccnt.as<int64_t>(), st));
            break;
    }
    tb = 0;
    DCHECK(rdm_scan_i64(nullptr, tb, ccnt.as<int64_t>(), coff.as<int64_t>(), ni + 1, st));
    if (tb > temp.bytes) DALLOC(temp, tb);
    DCHECK(rdm_scan_i64(temp.p, tb, ccnt.as<int64_t>(), coff.as<int64_t>(), ni + 1, st));
    int64_t nc = 0;
    DCHECK(hipMemcpyAsync(&nc, coff.as<int64_t>() + ni, sizeof(int64_t), hipMemcpyDeviceToHost,
                          st));
    DCHECK(hipStreamSynchronize(st));
    // 3. every contribution at its position
    const size_t ncs = (size_t)std::max<int64_t>(nc, 1);
    DALLOC(keys, sizeof(uint32_t) * ncs);
    DALLOC(vals, sizeof(double) * ncs);
    DALLOC(keys2, sizeof(uint32_t) * ncs);
    DALLOC(vals2, sizeof(double) * ncs);
    uint32_t* k = keys.as<uint32_t>();
    double* v = vals.as<double>();
    const int64_t* co = coff.as<int64_t>();
    switch (kind) {
        case RdmKind::Twordm:
            DCHECK(rdm_cwrite(masks, nw, i1, i2, id, ni, c1v, c2v, co, N, k, v, st));
            break;
        case RdmKind::Onerdm:
            DCHECK(rdm1_cwrite(masks, nw, i1, i2, id, ni, c1v, c2v, co, N, k, v, st));
            break;
        case RdmKind::Bit2:
            DCHECK(bit2_cwrite(masks, i1, i2, id, ni, (int)ndiag, c1v, c2v, co, N, k, v, st));
            break;
        case RdmKind::Bit1:
            DCHECK(bit1_cwrite(masks, i1, i2, id, ni, (int)ndiag, c1v, co, N, k, v, st));
            break;
    }
    return sort_and_sum(keys, vals, keys2, vals2, temp, nc, nel, dense, st);
}

// 6. the selection loop (p, q, r, s) of a dense N^4 twordm: count of rows,
// written to mat (cap, 4) / total when cap >= count
int64_t twordm_select_device(const DevBuf& tw, int N, double cutoff, int32_t* mat, double* total,
                             int64_t cap, hipStream_t st) {
#define SCHECK(x)                                   \
    do {                                            \
        if ((x) != hipSuccess) return -THISTLE_EDEVICE; \
    } while (0)
#define SALLOC(b, n)                                \
    do {                                            \
        if (dalloc(b, n)) return -THISTLE_ENOMEM;   \
    } while (0)
    const int64_t n4 = (int64_t)N * N * N * N;
    const int64_t nb = (n4 + 2047) / 2048 + 1;
    DevBuf flags, pos, rows, outv;
    SALLOC(flags, (size_t)n4);
    SALLOC(pos, sizeof(int64_t) * (size_t)(n4 + nb + 1));
    SCHECK(rdm_flags(tw.as<double>(), N, cutoff, flags.as<uint8_t>(), st));
    SCHECK(launch_exclusive_scan_u8(flags.as<uint8_t>(), n4, pos.as<int64_t>(),
                                    pos.as<int64_t>() + n4 + nb, st));
    int64_t nsel = 0;
    SCHECK(hipMemcpyAsync(&nsel, pos.as<int64_t>() + n4 + nb, sizeof(int64_t),
                          hipMemcpyDeviceToHost, st));
    SCHECK(hipStreamSynchronize(st));
    if (mat && total && cap >= nsel && nsel > 0) {
        SALLOC(rows, sizeof(int32_t) * 4 * (size_t)nsel);
        SALLOC(outv, sizeof(double) * (size_t)nsel);
        SCHECK(rdm_emit(tw.as<double>(), N, flags.as<uint8_t>(), pos.as<int64_t>(),
                        rows.as<int32_t>(), outv.as<double>(), st));
        std::vector<int32_t> r4((size_t)4 * nsel);
        SCHECK(hipMemcpyAsync(r4.data(), rows.p, sizeof(int32_t) * r4.size(),
                              hipMemcpyDeviceToHost, st));
        SCHECK(hipMemcpyAsync(total, outv.p, sizeof(double) * nsel, hipMemcpyDeviceToHost, st));
        SCHECK(hipStreamSynchronize(st));
        for (int64_t i = 0; i < nsel; ++i)
            for (int k = 0; k < 4; ++k) mat[i + k * cap] = r4[(size_t)4 * i + k];
    }
    return nsel;
#undef SALLOC
#undef SCHECK
}

// device, stream, masks and the two coefficient vectors of a canonical
// punch-format determinant list
int confs_to_device(const int32_t* confs, int nconf, int lconfs, const double* c1v_h,
                    const double* c2v_h, int device, StreamGuard& sg, DevBuf& masks,
                    DevBuf& civ) {
    DCHECK(hipSetDevice(device));
    DCHECK(hipStreamCreate(&sg.s));
    const int nw = (lconfs + 63) / 64;
    DevBuf dconfs;
    DALLOC(dconfs, sizeof(int32_t) * (size_t)nconf * lconfs);
    DALLOC(masks, sizeof(uint64_t) * (size_t)nconf * nw);
    DALLOC(civ, sizeof(double) * 2 * (size_t)nconf);
    DCHECK(hipMemcpyAsync(dconfs.p, confs, dconfs.bytes, hipMemcpyHostToDevice, sg.s));
    DCHECK(hipMemcpyAsync(civ.p, c1v_h, sizeof(double) * nconf, hipMemcpyHostToDevice, sg.s));
    DCHECK(hipMemcpyAsync(civ.as<double>() + nconf, c2v_h, sizeof(double) * nconf,
                          hipMemcpyHostToDevice, sg.s));
    DCHECK(rdm_pack(dconfs.as<int32_t>(), nconf, lconfs, nw, masks.as<uint64_t>(), sg.s));
    DCHECK(hipStreamSynchronize(sg.s));
    return THISTLE_OK;
}

// createtwordm on the device (rdm.hip): count of selected rows, rows written
// to mat / total when cap >= count; -status on failure
int64_t createtwordm_device(const int32_t* confs, int nconf, int lconfs, const double* c1v_h,
                            const double* c2v_h, double cutoff, int device, int32_t* mat,
                            double* total, int64_t cap) {
    StreamGuard sg;
    DevBuf masks, civ, tw;
    int rc = confs_to_device(confs, nconf, lconfs, c1v_h, c2v_h, device, sg, masks, civ);
    if (rc) return -rc;
    const int N = lconfs / 2;
    rc = rdm_accumulate(RdmKind::Twordm, masks.as<uint64_t>(), nconf, (lconfs + 63) / 64, N,
                        civ.as<double>(), civ.as<double>() + nconf, tw, sg.s);
    if (rc) return -rc;
    return twordm_select_device(tw, N, cutoff, mat, total, cap, sg.s);
}

// createonerdm on the device (rdm.hip): the dense norb x norb onerdm
int createonerdm_device(const int32_t* confs, int nconf, int lconfs, const double* c1v_h,
                        const double* c2v_h, int device, double* onerdm) {
    StreamGuard sg;
    DevBuf masks, civ, g;
    int rc = confs_to_device(confs, nconf, lconfs, c1v_h, c2v_h, device, sg, masks, civ);
    if (rc) return rc;
    const int N = lconfs / 2;
    rc = rdm_accumulate(RdmKind::Onerdm, masks.as<uint64_t>(), nconf, (lconfs + 63) / 64, N,
                        civ.as<double>(), civ.as<double>() + nconf, g, sg.s);
    if (rc) return rc;
    DCHECK(hipMemcpyAsync(onerdm, g.p, sizeof(double) * (size_t)N * N, hipMemcpyDeviceToHost,
                          sg.s));
    DCHECK(hipStreamSynchronize(sg.s));
    return THISTLE_OK;
}

bool canonical_confs(const int32_t* confs, int nconf, int lconfs) {
    for (int j = 0; j < lconfs; ++j)
        for (int c = 0; c < nconf; ++c) {
            const int v = confs[(size_t)c + (size_t)nconf * j];
            if (v != 0 && v != ((j % 2 == 0) ? 1 : 2)) return false;
        }
    return true;
}

// createonerdm on the GPU for canonical determinant lists, on the host
// otherwise; both give the reference's doubles
int onerdm_any(const int32_t* confs, int nconf, int lconfs, const double* civecs, int nstates,
               int state1, int state2, int device, double* onerdm) {
    if (nconf <= 0 || lconfs <= 0 || (lconfs % 2) != 0 || lconfs > 256) return THISTLE_EINVAL;
    if (state1 < 1 || state2 < 1 || state1 > nstates || state2 > nstates) return THISTLE_EINVAL;
    if (!canonical_confs(confs, nconf, lconfs))
        return thistle_onerdm_host(confs, nconf, lconfs, civecs, nstates, state1, state2, onerdm);
    return createonerdm_device(confs, nconf, lconfs, civecs + (size_t)nconf * (state1 - 1),
                               civecs + (size_t)nconf * (state2 - 1), device, onerdm);
}

// ---- bit-string determinants (src/one_rdm.f90) ----------------------------

// Bit strings must be non-negative, with one alpha count and one beta count
// throughout (the excitation bookkeeping of src/one_rdm.f90:1080-1217 and
// :1468-1512 reads holes and particles of one spin as pairs); mcci_to_bit
// also needs the two counts equal (its slot lists have 2 x the first alpha
// count entries, src/one_rdm.f90:1923,1940).  nbft: the highest occupied
// orbital (src/one_rdm.f90:985).
int bit_check(const int64_t* alpha, const int64_t* beta, int64_t n, bool equal_spins,
              int* nbft) {
    if (!alpha || !beta || n <= 0 || n > (int64_t)INT32_MAX) return THISTLE_EINVAL;
    const int na = __builtin_popcountll((uint64_t)alpha[0]);
    const int nb = __builtin_popcountll((uint64_t)beta[0]);
    if (equal_spins && na != nb) return THISTLE_EINVAL;
    int64_t mx = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (alpha[i] < 0 || beta[i] < 0) return THISTLE_EINVAL;
        if (__builtin_popcountll((uint64_t)alpha[i]) != na ||
            __builtin_popcountll((uint64_t)beta[i]) != nb)
            return THISTLE_EINVAL;
        mx = std::max(mx, std::max(alpha[i], beta[i]));
    }
    int w = 0;
    while (mx) {
        ++w;
        mx >>= 1;
    }
    if (w == 0) return THISTLE_EINVAL;
    *nbft = w;
    return THISTLE_OK;
}

// mcci_to_bit writes each signed coefficient with E20.10 and
// createtwordm_bit reads it back (src/one_rdm.f90:1975, :979): the double
// nearest the value rounded to 10 significant digits.  The rounding is
// symmetric, so it is applied before the sign.
double e20_10(double x) {
    char buf[48];
    std::snprintf(buf, sizeof buf, "%.9E", x);
    return std::strtod(buf, nullptr);
}

// stream, (alpha, beta) masks and the two coefficient vectors on the device
int bits_to_device(const int64_t* alpha, const int64_t* beta, int64_t n, const double* c1,
                   const double* c2, int device, StreamGuard& sg, DevBuf& masks, DevBuf& civ) {
    DCHECK(hipSetDevice(device));
    DCHECK(hipStreamCreate(&sg.s));
    DevBuf ab;
    DALLOC(ab, sizeof(int64_t) * 2 * (size_t)n);
    DALLOC(masks, sizeof(uint64_t) * 2 * (size_t)n);
    DALLOC(civ, sizeof(double) * 2 * (size_t)n);
    DCHECK(hipMemcpyAsync(ab.p, alpha, sizeof(int64_t) * n, hipMemcpyHostToDevice, sg.s));
    DCHECK(hipMemcpyAsync(ab.as<int64_t>() + n, beta, sizeof(int64_t) * n, hipMemcpyHostToDevice,
                          sg.s));
    DCHECK(hipMemcpyAsync(civ.p, c1, sizeof(double) * n, hipMemcpyHostToDevice, sg.s));
    DCHECK(hipMemcpyAsync(civ.as<double>() + n, c2, sizeof(double) * n, hipMemcpyHostToDevice,
                          sg.s));
    DCHECK(bit_pack(ab.as<int64_t>(), ab.as<int64_t>() + n, (int)n, masks.as<uint64_t>(), sg.s));
    DCHECK(hipStreamSynchronize(sg.s));
    return THISTLE_OK;
}

// mcci_to_bit in place on the device coefficients: text rounding on the way
// in (host copies c1r / c2r), the reordering sign on the device
int mcci_in_place(const DevBuf& masks, int64_t n, const double* c1, const double* c2,
                  bool text_round, DevBuf& civ, hipStream_t st) {
    std::vector<double> r((size_t)2 * n);
    for (int64_t i = 0; i < n; ++i) {
        r[i] = text_round ? e20_10(c1[i]) : c1[i];
        r[n + i] = text_round ? e20_10(c2[i]) : c2[i];
    }
    DevBuf in;
    DALLOC(in, sizeof(double) * r.size());
    DCHECK(hipMemcpyAsync(in.p, r.data(), sizeof(double) * r.size(), hipMemcpyHostToDevice, st));
    DCHECK(mcci_phase(masks.as<uint64_t>(), (int)n, in.as<double>(), in.as<double>() + n,
                      civ.as<double>(), civ.as<double>() + n, st));
    DCHECK(hipStreamSynchronize(st));
    return THISTLE_OK;
}

// createtwordm_bit's rows and values (count first, then the rows)
int bit2_rows(const DevBuf& masks, int64_t n, int N, const DevBuf& civ, double cutoff,
              hipStream_t st, std::vector<int32_t>& mat, std::vector<double>& tot,
              int64_t& cnt) {
    DevBuf tw;
    int rc = rdm_accumulate(RdmKind::Bit2, masks.as<uint64_t>(), (int)n, 2, N, civ.as<double>(),
                            civ.as<double>() + n, tw, st);
    if (rc) return rc;
    cnt = twordm_select_device(tw, N, cutoff, nullptr, nullptr, 0, st);
    if (cnt < 0) return (int)-cnt;
    mat.assign((size_t)4 * std::max<int64_t>(cnt, 1), 0);
    tot.assign((size_t)std::max<int64_t>(cnt, 1), 0.0);
    const int64_t c2 = twordm_select_device(tw, N, cutoff, mat.data(), tot.data(), cnt, st);
    return c2 < 0 ? (int)-c2 : THISTLE_OK;
}

// one_rdm_bit's dense N x N onerdm from the device coefficients civ
int bit1_dense(const DevBuf& masks, int64_t n, int N, const double* civ, hipStream_t st,
               double* onerdm) {
    DevBuf g;
    int rc = rdm_accumulate(RdmKind::Bit1, masks.as<uint64_t>(), (int)n, 2, N, civ, civ, g, st);
    if (rc) return rc;
    DCHECK(hipMemcpyAsync(onerdm, g.p, sizeof(double) * (size_t)N * N, hipMemcpyDeviceToHost, st));
    DCHECK(hipStreamSynchronize(st));
    return THISTLE_OK;
}

// createtwordm_bit_fci on the device: every contribution of every file
// determinant in the reference's order (fci_contribs), then the ordered
// sums of rdm_accumulate's steps 4-5
int fci_accumulate(const FciArgs& a, DevBuf& dense, hipStream_t st) {
    const int64_t n4 = (int64_t)a.N * a.N * a.N * a.N;
    DevBuf cnt, coff, keys, vals, keys2, vals2, temp;
    DALLOC(cnt, sizeof(int64_t) * ((size_t)a.n + 1));
    DALLOC(coff, sizeof(int64_t) * ((size_t)a.n + 1));
    DCHECK(launch_fill(reinterpret_cast<double*>(cnt.p), a.n + 1, 0.0, st));
    DCHECK(fci_ccount(a, cnt.as<int64_t>(), st));
    size_t tb = 0;
    DCHECK(rdm_scan_i64(nullptr, tb, cnt.as<int64_t>(), coff.as<int64_t>(), a.n + 1, st));
    DALLOC(temp, tb);
    DCHECK(rdm_scan_i64(temp.p, tb, cnt.as<int64_t>(), coff.as<int64_t>(), a.n + 1, st));
    int64_t nc = 0;
    DCHECK(hipMemcpyAsync(&nc, coff.as<int64_t>() + a.n, sizeof(int64_t), hipMemcpyDeviceToHost,
                          st));
    DCHECK(hipStreamSynchronize(st));
    const size_t ncs = (size_t)std::max<int64_t>(nc, 1);
    DALLOC(keys, sizeof(uint32_t) * ncs);
    DALLOC(vals, sizeof(double) * ncs);
    DALLOC(keys2, sizeof(uint32_t) * ncs);
    DALLOC(vals2, sizeof(double) * ncs);
    DCHECK(fci_cwrite(a, coff.as<int64_t>(), keys.as<uint32_t>(), vals.as<double>(), st));
    return sort_and_sum(keys, vals, keys2, vals2, temp, nc, n4, dense, st);
}

// The inputs createtwordm_bit_fci can process: unique alpha strings
// increasing, one electron count everywhere (the tables are sized from
// the first string, :479, :565), at least norbs + 1 of them (:479 reads row
// norbs + 1), every determinant's alpha and beta string among them (:719-721
// look both up there), each string's determinants contiguous in the file
// with start / end their first / last position, and orbitals 1-31 (the
// 32-bit masks of :640-643).
int fci_check(const int64_t* alpha, const int64_t* beta, const double* civs, int64_t n,
              const int64_t* uniq, int64_t nuniq, const int64_t* irep, const int64_t* start,
              const int64_t* end, int* N) {
    if (!alpha || !beta || !civs || !uniq || !irep || !start || !end) return THISTLE_EINVAL;
    if (n <= 0 || n > (int64_t)INT32_MAX || nuniq <= 0 || nuniq > (int64_t)INT32_MAX)
        return THISTLE_EINVAL;
    const int nel = __builtin_popcountll((uint64_t)uniq[0]);
    int64_t mx = 0;
    for (int64_t u = 0; u < nuniq; ++u) {
        if (uniq[u] <= 0 || (u > 0 && uniq[u] <= uniq[u - 1])) return THISTLE_EINVAL;
        if (__builtin_popcountll((uint64_t)uniq[u]) != nel) return THISTLE_EINVAL;
        if (irep[u] < 0 || start[u] < 1 || end[u] < start[u] || end[u] > n) return THISTLE_EINVAL;
    }
    auto find = [&](int64_t s) -> int64_t {
        const int64_t* p = std::lower_bound(uniq, uniq + nuniq, s);
        return (p != uniq + nuniq && *p == s) ? (int64_t)(p - uniq) : -1;
    };
    std::vector<int64_t> first((size_t)nuniq, -1), last((size_t)nuniq, -1);
    for (int64_t i = 0; i < n; ++i) {
        mx = std::max(mx, alpha[i]);
        const int64_t ua = find(alpha[i]);
        if (ua < 0 || find(beta[i]) < 0) return THISTLE_EINVAL;
        if (first[ua] < 0) first[ua] = i;
        else if (last[ua] != i - 1) return THISTLE_EINVAL;  // not contiguous
        last[ua] = i;
    }
    for (int64_t u = 0; u < nuniq; ++u)
        if (first[u] < 0 || start[u] != first[u] + 1 || end[u] != last[u] + 1)
            return THISTLE_EINVAL;
    int w = 0;
    while (mx) {
        ++w;
        mx >>= 1;
    }
    if (w == 0 || w > 31 || nuniq < w + 1) return THISTLE_EINVAL;
    *N = w;
    return THISTLE_OK;
}

#undef DALLOC
#undef DCHECK

}  // namespace

int32_t thistle_createonerdm(const int32_t* confs, int32_t nconf, int32_t lconfs,
                             const double* civecs, int32_t nstates, int32_t state1,
                             int32_t state2, const ThistleOpts* opts, double* onerdm) {
    if (!confs || !civecs || !onerdm) return THISTLE_EINVAL;
    const ThistleOpts* o = opts ? opts : &kDefaultOpts;
    return onerdm_any(confs, nconf, lconfs, civecs, nstates, state1, state2, o->device, onerdm);
}

int64_t thistle_createtwordm(const int32_t* confs, int32_t nconf, int32_t lconfs,
                             const double* civecs, int32_t nstates, int32_t state1,
                             int32_t state2, double cutoff, const ThistleOpts* opts,
                             int32_t* mat, double* total, int64_t cap) {
    if (!confs || !civecs) return -THISTLE_EINVAL;
    if (nconf <= 0 || lconfs <= 0 || (lconfs % 2) != 0 || lconfs > 256) return -THISTLE_EINVAL;
    if (state1 < 1 || state2 < 1 || state1 > nstates || state2 > nstates) return -THISTLE_EINVAL;
    if (cap > 0 && (!mat || !total)) return -THISTLE_EINVAL;
    const ThistleOpts* o = opts ? opts : &kDefaultOpts;
    const bool canonical = canonical_confs(confs, nconf, lconfs);
    const double* c1v = civecs + (size_t)nconf * (state1 - 1);
    const double* c2v = civecs + (size_t)nconf * (state2 - 1);
    if (canonical)
        return createtwordm_device(confs, nconf, lconfs, c1v, c2v, cutoff, o->device, mat, total,
                                   cap);
    const int norb = lconfs / 2;
    std::vector<double> tw((size_t)norb * norb * norb * norb);
    int rc = thistle_twordm_dense_host(confs, nconf, lconfs, civecs, nstates, state1, state2,
                                       tw.data());
    if (rc) return -rc;
    return thistle_twordm_select_host(tw.data(), norb, cutoff, cap > 0 ? mat : nullptr,
                                      cap > 0 ? total : nullptr, cap);
}

int32_t thistle_onerdm(const int32_t* confs, int32_t nconf, int32_t lconfs, const double* civecs,
                       int32_t nstates, int32_t state1, int32_t state2, double* onerdm) {
    if (!confs || !civecs || !onerdm) return THISTLE_EINVAL;
    return thistle_onerdm_host(confs, nconf, lconfs, civecs, nstates, state1, state2, onerdm);
}

int32_t thistle_total_scattering_calculation(int32_t type, int32_t natom, const double* Zn,
                                             const double* geom, int32_t state1, int32_t state2,
                                             int32_t state3, const ThistleBasis* basis,
                                             int32_t nconf, int32_t lconfs, int32_t nstates,
                                             const int32_t* confs, const double* civecs,
                                             const ThistleOpts* opts, double* q_abs_out,
                                             double* result_out) {
    (void)state3;
    if (type != 1 && type != 2 && type != 5 && type != 6 && type != 7 && type != 8)
        return THISTLE_ETYPE;
    if (!basis || !result_out) return THISTLE_EINVAL;
    if ((type == 5 || type == 6) && (natom <= 0 || !Zn || !geom)) return THISTLE_EINVAL;
    const ThistleOpts* o = opts ? opts : &kDefaultOpts;
    if (q_abs_out)
        for (int i = 0; i < basis->nq; ++i) q_abs_out[i] = 0.0;
    const int norb = lconfs / 2;
    if (type == 5 || type == 6) {
        // electron scattering (src/main_calculation.f90:270-330):
        // Iee - 2 Ine + Inn, Iee total (5) or elastic (6) j0
        const int nq = basis->nq;
        std::vector<double> Iee(nq), Ine(nq), Inn(nq);
        std::vector<double> g((size_t)norb * norb);
        int rc = onerdm_any(confs, nconf, lconfs, civecs, nstates, state1, state2, o->device,
                            g.data());
        if (rc) return rc;
        if (type == 5) {
            rc = thistle_total_scattering_calculation(1, natom, Zn, geom, state1, state2, state3,
                                                      basis, nconf, lconfs, nstates, confs,
                                                      civecs, o, nullptr, Iee.data());
            if (rc) return rc;
            rc = thistle_nuclei_electron(basis, norb, g.data(), natom, Zn, geom, o, Ine.data());
            if (rc) return rc;
        } else {
            ThistlePlan* raw = nullptr;
            rc = thistle_plan_create(2, basis, 0, nullptr, nullptr, norb, g.data(), o, &raw);
            if (rc) return rc;
            std::unique_ptr<ThistlePlan> pl(raw);
            rc = thistle_plan_execute(pl.get(), Iee.data());
            if (rc) return rc;
            rc = run_nuclear(pl.get(), natom, Zn, geom, Ine.data());
            if (rc) return rc;
        }
        nuclear_nn(natom, Zn, geom, basis->q, nq, Inn.data());
        for (int k = 0; k < nq; ++k) result_out[k] = Iee[k] - 2 * Ine[k] + Inn[k];
        return THISTLE_OK;
    }
    if (type == 1 || type == 7) {
        // createtwordm on the device; at most norb^4 rows, so one pass when
        // that bound is small
        const double cut = o->rdm_cutoff > 0 ? o->rdm_cutoff : THISTLE_RDM_CUTOFF;
        const int64_t n4 = (int64_t)norb * norb * norb * norb;
        int64_t cap = n4 <= ((int64_t)1 << 22) ? n4 : 0;
        std::vector<int32_t> m4((size_t)4 * std::max<int64_t>(cap, 1));
        std::vector<double> tot(std::max<int64_t>(cap, 1));
        int64_t cnt = thistle_createtwordm(confs, nconf, lconfs, civecs, nstates, state1, state2,
                                           cut, o, m4.data(), tot.data(), cap);
        if (cnt < 0) return (int32_t)-cnt;
        if (cnt > cap) {
            cap = cnt;
            m4.assign((size_t)4 * cap, 0);
            tot.assign(cap, 0.0);
            int64_t c2 = thistle_createtwordm(confs, nconf, lconfs, civecs, nstates, state1,
                                              state2, cut, o, m4.data(), tot.data(), cap);
            if (c2 < 0) return (int32_t)-c2;
        }
        // (cap, 4) -> (cnt, 4) column-major
        std::vector<int32_t> mat((size_t)4 * std::max<int64_t>(cnt, 1));
        for (int k = 0; k < 4; ++k)
            for (int64_t i = 0; i < cnt; ++i) mat[(size_t)i + (size_t)k * cnt] = m4[(size_t)i + (size_t)k * cap];
        int rc = 0;
        if (type == 1)
            return thistle_total_from_twordm(basis, cnt, mat.data(), tot.data(), o, result_out);
        ThistlePlan* pl = nullptr;
        rc = thistle_plan_create(7, basis, cnt, mat.data(), tot.data(), 0, nullptr, o, &pl);
        if (rc) return rc;
        rc = thistle_plan_execute(pl, result_out);
        thistle_plan_destroy(pl);
        return rc;
    }
    std::vector<double> g((size_t)norb * norb);
    int rc = onerdm_any(confs, nconf, lconfs, civecs, nstates, state1, state2, o->device,
                        g.data());
    if (rc) return rc;
    if (type == 2) return thistle_elastic_from_onerdm(basis, norb, g.data(), o, result_out);
    // type 8: elastic j2 (src/main_calculation.f90:780-816)
    ThistlePlan* pl = nullptr;
    rc = thistle_plan_create(8, basis, 0, nullptr, nullptr, norb, g.data(), o, &pl);
    if (rc) return rc;
    rc = thistle_plan_execute(pl, result_out);
    thistle_plan_destroy(pl);
    return rc;
}

int32_t thistle_mcci_to_bit(const int64_t* alpha, const int64_t* beta, const double* c1,
                            const double* c2, int64_t n, int32_t text_round,
                            const ThistleOpts* opts, double* c1_out, double* c2_out) {
    if (!c1 || !c2 || !c1_out || !c2_out) return THISTLE_EINVAL;
    int N = 0;
    int rc = bit_check(alpha, beta, n, true, &N);
    if (rc) return rc;
    const ThistleOpts* o = opts ? opts : &kDefaultOpts;
    StreamGuard sg;
    DevBuf masks, civ;
    if ((rc = bits_to_device(alpha, beta, n, c1, c2, o->device, sg, masks, civ))) return rc;
    if ((rc = mcci_in_place(masks, n, c1, c2, text_round != 0, civ, sg.s))) return rc;
    if (hipMemcpyAsync(c1_out, civ.p, sizeof(double) * n, hipMemcpyDeviceToHost, sg.s) !=
            hipSuccess ||
        hipMemcpyAsync(c2_out, civ.as<double>() + n, sizeof(double) * n, hipMemcpyDeviceToHost,
                       sg.s) != hipSuccess ||
        hipStreamSynchronize(sg.s) != hipSuccess)
        return THISTLE_EDEVICE;
    return THISTLE_OK;
}

int64_t thistle_createtwordm_bit(const int64_t* alpha, const int64_t* beta, const double* c1,
                                 const double* c2, int64_t n, double cutoff,
                                 const ThistleOpts* opts, int32_t* mat, double* total,
                                 int64_t cap) {
    if (!c1 || !c2) return -THISTLE_EINVAL;
    if (cap > 0 && (!mat || !total)) return -THISTLE_EINVAL;
    int N = 0;
    int rc = bit_check(alpha, beta, n, false, &N);
    if (rc) return -rc;
    const ThistleOpts* o = opts ? opts : &kDefaultOpts;
    StreamGuard sg;
    DevBuf masks, civ, tw;
    if ((rc = bits_to_device(alpha, beta, n, c1, c2, o->device, sg, masks, civ))) return -rc;
    rc = rdm_accumulate(RdmKind::Bit2, masks.as<uint64_t>(), (int)n, 2, N, civ.as<double>(),
                        civ.as<double>() + n, tw, sg.s);
    if (rc) return -rc;
    return twordm_select_device(tw, N, cutoff > 0 ? cutoff : THISTLE_RDM_CUTOFF, cap > 0 ? mat : nullptr,
                                cap > 0 ? total : nullptr, cap, sg.s);
}

int32_t thistle_one_rdm_bit(const int64_t* alpha, const int64_t* beta, const double* civs,
                            int64_t n, const ThistleOpts* opts, int32_t cap_norb,
                            int32_t* norb_out, double* onerdm) {
    if (!civs || !norb_out) return THISTLE_EINVAL;
    int N = 0;
    int rc = bit_check(alpha, beta, n, false, &N);
    if (rc) return rc;
    *norb_out = N;
    if (!onerdm) return THISTLE_OK;
    if (N > cap_norb) return THISTLE_EINVAL;
    const ThistleOpts* o = opts ? opts : &kDefaultOpts;
    StreamGuard sg;
    DevBuf masks, civ;
    if ((rc = bits_to_device(alpha, beta, n, civs, civs, o->device, sg, masks, civ))) return rc;
    return bit1_dense(masks, n, N, civ.as<double>(), sg.s, onerdm);
}

int64_t thistle_createtwordm_bit_fci(const int64_t* alpha, const int64_t* beta,
                                     const double* civs, int64_t n, const int64_t* uniq,
                                     int64_t nuniq, const int64_t* irep, const int64_t* start,
                                     const int64_t* end, double cutoff, const ThistleOpts* opts,
                                     int32_t* mat, double* total, int64_t cap) {
    if (cap > 0 && (!mat || !total)) return -THISTLE_EINVAL;
    int N = 0;
    int rc = fci_check(alpha, beta, civs, n, uniq, nuniq, irep, start, end, &N);
    if (rc) return -rc;
    const ThistleOpts* o = opts ? opts : &kDefaultOpts;
    StreamGuard sg;
    DevBuf masks, civ, du, direp, keys, keys2, idx, idx2, temp, tw;
    if ((rc = bits_to_device(alpha, beta, n, civs, civs, o->device, sg, masks, civ))) return -rc;
    if (dalloc(du, sizeof(uint64_t) * nuniq) || dalloc(direp, sizeof(int64_t) * nuniq) ||
        dalloc(keys, sizeof(uint64_t) * n) || dalloc(keys2, sizeof(uint64_t) * n) ||
        dalloc(idx, sizeof(int32_t) * n) || dalloc(idx2, sizeof(int32_t) * n))
        return -THISTLE_ENOMEM;
    if (hipMemcpyAsync(du.p, uniq, sizeof(int64_t) * nuniq, hipMemcpyHostToDevice, sg.s) !=
            hipSuccess ||
        hipMemcpyAsync(direp.p, irep, sizeof(int64_t) * nuniq, hipMemcpyHostToDevice, sg.s) !=
            hipSuccess ||
        fci_keys(masks.as<uint64_t>(), (int)n, keys.as<uint64_t>(), idx.as<int32_t>(), sg.s) !=
            hipSuccess)
        return -THISTLE_EDEVICE;
    size_t tb = 0;
    if (fci_key_sort(nullptr, tb, keys.as<uint64_t>(), keys2.as<uint64_t>(), idx.as<int32_t>(),
                     idx2.as<int32_t>(), n, sg.s) != hipSuccess)
        return -THISTLE_EDEVICE;
    if (dalloc(temp, tb)) return -THISTLE_ENOMEM;
    if (fci_key_sort(temp.p, tb, keys.as<uint64_t>(), keys2.as<uint64_t>(), idx.as<int32_t>(),
                     idx2.as<int32_t>(), n, sg.s) != hipSuccess)
        return -THISTLE_EDEVICE;
    const FciArgs a{masks.as<uint64_t>(), civ.as<double>(), du.as<uint64_t>(),
                    direp.as<int64_t>(), keys2.as<uint64_t>(), idx2.as<int32_t>(), (int)n,
                    (int)nuniq, N};
    if ((rc = fci_accumulate(a, tw, sg.s))) return -rc;
    return twordm_select_device(tw, N, cutoff > 0 ? cutoff : THISTLE_RDM_CUTOFF,
                                cap > 0 ? mat : nullptr, cap > 0 ? total : nullptr, cap, sg.s);
}

int32_t thistle_total_scattering_calculation_2(int32_t type, const ThistleBasis* basis,
                                               int64_t n, const int64_t* alpha,
                                               const int64_t* beta, const double* c1,
                                               const double* c2, int32_t mcci,
                                               const ThistleOpts* opts, double* q_abs_out,
                                               double* result_out) {
    if (type != 1 && type != 2 && type != 7 && type != 8) return THISTLE_ETYPE;
    if (!basis || !result_out || !c1 || !c2) return THISTLE_EINVAL;
    const ThistleOpts* o = opts ? opts : &kDefaultOpts;
    if (q_abs_out)
        for (int i = 0; i < basis->nq; ++i) q_abs_out[i] = 0.0;
    // every type but 2 passes the file through mcci_to_bit first
    // (src/main_calculation.f90:586-604, 730-748, 786-787)
    int N = 0;
    int rc = bit_check(alpha, beta, n, type != 2, &N);
    if (rc) return rc;
    StreamGuard sg;
    DevBuf masks, civ;
    if ((rc = bits_to_device(alpha, beta, n, c1, c2, o->device, sg, masks, civ))) return rc;
    // type 1 always reads the converted file; type 7 only with mcci (its
    // default branch converts, then reads the original file,
    // src/main_calculation.f90:740-744); type 8 converts; type 2 reads the
    // original
    const bool converted = type == 1 || type == 8 || (type == 7 && mcci != 0);
    if (converted && (rc = mcci_in_place(masks, n, c1, c2, true, civ, sg.s))) return rc;
    if (type == 2 || type == 8) {
        // one_rdm_bit keeps the third column (src/one_rdm.f90:1428)
        std::vector<double> g((size_t)N * N);
        if ((rc = bit1_dense(masks, n, N, civ.as<double>() + n, sg.s, g.data()))) return rc;
        if (type == 2) return thistle_elastic_from_onerdm(basis, N, g.data(), o, result_out);
        ThistlePlan* pl = nullptr;
        rc = thistle_plan_create(8, basis, 0, nullptr, nullptr, N, g.data(), o, &pl);
        if (rc) return rc;
        rc = thistle_plan_execute(pl, result_out);
        thistle_plan_destroy(pl);
        return rc;
    }
    std::vector<int32_t> m4;
    std::vector<double> tot;
    int64_t cnt = 0;
    if ((rc = bit2_rows(masks, n, N, civ, o->rdm_cutoff > 0 ? o->rdm_cutoff : THISTLE_RDM_CUTOFF, sg.s,
                        m4, tot, cnt)))
        return rc;
    if (type == 1) return thistle_total_from_twordm(basis, cnt, m4.data(), tot.data(), o, result_out);
    ThistlePlan* pl = nullptr;
    rc = thistle_plan_create(7, basis, cnt, m4.data(), tot.data(), 0, nullptr, o, &pl);
    if (rc) return rc;
    rc = thistle_plan_execute(pl, result_out);
    thistle_plan_destroy(pl);
    return rc;
}

int64_t thistle_quartet_list(const ThistleBasis* basis, int32_t* out, int64_t cap,
                             const ThistleOpts* opts) {
    const ThistleOpts* o = opts ? opts : &kDefaultOpts;
    std::unique_ptr<ThistlePlan> pl(new ThistlePlan());
    int rc = setup_common(pl.get(), basis, o);
    if (rc) return -rc;
    rc = run_tables_and_enum(pl.get(), false);
    if (rc) return -rc;
    if (out && cap >= pl->nquart) {
        DevBuf tmp;
        if (dalloc(tmp, sizeof(int32_t) * 6 * pl->nquart)) return -THISTLE_ENOMEM;
        if (launch_unpack_quartets(pl->qlist.as<uint64_t>(), pl->nquart, tmp.as<int32_t>(),
                                   pl->stream) != hipSuccess)
            return -THISTLE_EDEVICE;
        if (hipMemcpyAsync(out, tmp.p, sizeof(int32_t) * 6 * pl->nquart, hipMemcpyDeviceToHost,
                           pl->stream) != hipSuccess)
            return -THISTLE_EDEVICE;
    }
    if (hipStreamSynchronize(pl->stream) != hipSuccess) return -THISTLE_EDEVICE;
    return pl->nquart;
}

int64_t thistle_quartet_list_screened(int32_t type, const ThistleBasis* basis, int64_t nmat,
                                      const int32_t* mat, const double* total, int32_t norb,
                                      const double* onerdm, int32_t* out, int64_t cap,
                                      const ThistleOpts* opts) {
    ThistlePlan* raw = nullptr;
    int rc = thistle_plan_create(type, basis, nmat, mat, total, norb, onerdm, opts, &raw);
    if (rc) return -rc;
    std::unique_ptr<ThistlePlan> pl(raw);
    rc = thistle_plan_execute(pl.get(), nullptr);
    if (rc) return -rc;
    // quartets with a nonzero 2-RDM block, in reference order
    if (launch_compact(pl->flags.as<uint8_t>(), pl->qlist.as<uint64_t>(), pl->nquart,
                       pl->scan.as<int64_t>(), pl->qcomp.as<uint64_t>(), pl->count.as<int64_t>(),
                       pl->stream) != hipSuccess ||
        hipStreamSynchronize(pl->stream) != hipSuccess)
        return -THISTLE_EDEVICE;
    int64_t cnt = 0;
    if (hipMemcpy(&cnt, pl->count.p, sizeof(int64_t), hipMemcpyDeviceToHost) != hipSuccess)
        return -THISTLE_EDEVICE;
    if (out && cap >= cnt && cnt > 0) {
        DevBuf tmp;
        if (dalloc(tmp, sizeof(int32_t) * 6 * cnt)) return -THISTLE_ENOMEM;
        if (launch_unpack_quartets(pl->qcomp.as<uint64_t>(), cnt, tmp.as<int32_t>(), pl->stream) !=
            hipSuccess)
            return -THISTLE_EDEVICE;
        if (hipMemcpyAsync(out, tmp.p, sizeof(int32_t) * 6 * cnt, hipMemcpyDeviceToHost,
                           pl->stream) != hipSuccess)
            return -THISTLE_EDEVICE;
        if (hipStreamSynchronize(pl->stream) != hipSuccess) return -THISTLE_EDEVICE;
    }
    return cnt;
}

int64_t thistle_unique_total(const int32_t* matrix, int64_t nrow, int32_t ncol, const double* total,
                             int32_t* rmat, double* rtot, int32_t* iuni, int32_t* irec,
                             const ThistleOpts* opts) {
    const ThistleOpts* o = opts ? opts : &kDefaultOpts;
    if (nrow < 0 || ncol < 1 || ncol > 8 || (nrow > 0 && (!matrix || !total))) return -THISTLE_EINVAL;
    int lim = ncol <= 4 ? 65535 : 255;
    for (int64_t k = 0; k < nrow * ncol; ++k)
        if (matrix[k] < 0 || matrix[k] > lim) return -THISTLE_EINVAL;
    if (hipSetDevice(o->device) != hipSuccess) return -THISTLE_EDEVICE;
    hipStream_t s = (hipStream_t)o->stream;
    int64_t hs = 16;
    while (hs < 2 * std::max<int64_t>(nrow, 1)) hs <<= 1;
    DevBuf dm, dt, keys, mn, cn, sl, sc, drm, drt, diu, dir, cnt;
    int64_t nb = (nrow + 2047) / 2048 + 1;
    if (dalloc(dm, sizeof(int32_t) * std::max<int64_t>(nrow * ncol, 1)) ||
        dalloc(dt, sizeof(double) * std::max<int64_t>(nrow, 1)) || dalloc(keys, 8 * hs) ||
        dalloc(mn, 4 * hs) || dalloc(cn, 4 * hs) || dalloc(sl, 4 * std::max<int64_t>(nrow, 1)) ||
        dalloc(sc, 8 * (nrow + nb + 1)) ||
        dalloc(drm, sizeof(int32_t) * std::max<int64_t>(nrow * ncol, 1)) ||
        dalloc(drt, sizeof(double) * std::max<int64_t>(nrow, 1)) ||
        dalloc(diu, 4 * std::max<int64_t>(nrow, 1)) || dalloc(dir, 4 * std::max<int64_t>(nrow, 1)) ||
        dalloc(cnt, 16))
        return -THISTLE_ENOMEM;
    if (nrow > 0) {
        if (hipMemcpyAsync(dm.p, matrix, sizeof(int32_t) * nrow * ncol, hipMemcpyHostToDevice, s) ||
            hipMemcpyAsync(dt.p, total, sizeof(double) * nrow, hipMemcpyHostToDevice, s))
            return -THISTLE_EDEVICE;
    }
    UniqueWork w{keys.as<uint64_t>(), mn.as<int32_t>(), cn.as<int32_t>(), sl.as<int32_t>(),
                 sc.as<int64_t>(), hs};
    if (launch_unique(dm.as<int32_t>(), nrow, ncol, dt.as<double>(), w, drm.as<int32_t>(),
                      drt.as<double>(), diu.as<int32_t>(), dir.as<int32_t>(), cnt.as<int64_t>(),
                      s) != hipSuccess)
        return -THISTLE_EDEVICE;
    int64_t u = 0;
    if (hipMemcpyAsync(&u, cnt.p, 8, hipMemcpyDeviceToHost, s) != hipSuccess) return -THISTLE_EDEVICE;
    if (hipStreamSynchronize(s) != hipSuccess) return -THISTLE_EDEVICE;
    // rmat comes back compacted to (u, ncol) column-major
    if (rmat && u > 0) {
        std::vector<int32_t> tmp((size_t)nrow * ncol);
        if (hipMemcpy(tmp.data(), drm.p, sizeof(int32_t) * nrow * ncol, hipMemcpyDeviceToHost))
            return -THISTLE_EDEVICE;
        for (int c = 0; c < ncol; ++c)
            for (int64_t k = 0; k < u; ++k) rmat[k + u * c] = tmp[k + nrow * c];
    }
    if (rtot && u > 0 && hipMemcpy(rtot, drt.p, sizeof(double) * u, hipMemcpyDeviceToHost))
        return -THISTLE_EDEVICE;
    if (iuni && u > 0 && hipMemcpy(iuni, diu.p, sizeof(int32_t) * u, hipMemcpyDeviceToHost))
        return -THISTLE_EDEVICE;
    if (irec && nrow > 0 && hipMemcpy(irec, dir.p, sizeof(int32_t) * nrow, hipMemcpyDeviceToHost))
        return -THISTLE_EDEVICE;
    return u;
}

}  // extern "C"

// ---- allocation-free unique_total --------------------------------------
// The hash table, first-occurrence flags and scan of launch_unique live in
// one file-scope device array, sized for THISTLE_NOALLOC_MAX_ROWS rows.  In
// host code `scratch` only names the symbol: its device address is taken
// once per device with hipGetSymbolAddress and that address is what the
// kernels receive.
namespace {
constexpr int64_t kNaRows = THISTLE_NOALLOC_MAX_ROWS;
constexpr int64_t kNaHash = 2 * kNaRows;                      // slots at most
constexpr int64_t kNaScan = kNaRows + (kNaRows + 2047) / 2048 + 2;
constexpr size_t kNaKeys = 0;                                  // uint64[kNaHash]
constexpr size_t kNaMin = kNaKeys + 8 * (size_t)kNaHash;       // int32[kNaHash]
constexpr size_t kNaCnt = kNaMin + 4 * (size_t)kNaHash;        // int32[kNaHash]
constexpr size_t kNaSlot = kNaCnt + 4 * (size_t)kNaHash;       // int32[kNaRows]
constexpr size_t kNaScanOff = (kNaSlot + 4 * (size_t)kNaRows + 255) & ~(size_t)255;
constexpr size_t kNaBytes = kNaScanOff + 8 * (size_t)kNaScan;  // int64[kNaScan]
static_assert(kNaBytes == THISTLE_NOALLOC_SCRATCH_BYTES, "scratch size");
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNaBytes];

namespace {
std::mutex g_na_mu;
void* g_na_ptr[64] = {};

// the device address of `scratch` on the current device (looked up once)
int noalloc_scratch(int device, unsigned char** out) {
    if (device < 0 || device >= 64) return THISTLE_EINVAL;
    std::lock_guard<std::mutex> lock(g_na_mu);
    if (!g_na_ptr[device]) {
        void* p = nullptr;
        HCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return THISTLE_EDEVICE;
        g_na_ptr[device] = p;
    }
    *out = static_cast<unsigned char*>(g_na_ptr[device]);
    return THISTLE_OK;
}
}  // namespace

extern "C" {

int32_t thistle_unique_total_noalloc(const int32_t* matrix_dev, int64_t nrow, int32_t ncol,
                                     const double* total_dev, int32_t* rmat_dev,
                                     double* rtot_dev, int32_t* iuni_dev, int32_t* irec_dev,
                                     int64_t* count_dev, const ThistleOpts* opts) {
    const ThistleOpts* o = opts ? opts : &kDefaultOpts;
    if (nrow < 0 || nrow > kNaRows || ncol < 1 || ncol > 8 || !count_dev ||
        (nrow > 0 && (!matrix_dev || !total_dev)))
        return THISTLE_EINVAL;
    HCHECK(hipSetDevice(o->device));
    unsigned char* sc = nullptr;
    int rc = noalloc_scratch(o->device, &sc);
    if (rc) return rc;
    int64_t hs = 16;
    while (hs < 2 * std::max<int64_t>(nrow, 1)) hs <<= 1;
    UniqueWork w{reinterpret_cast<uint64_t*>(sc + kNaKeys), reinterpret_cast<int32_t*>(sc + kNaMin),
                 reinterpret_cast<int32_t*>(sc + kNaCnt), reinterpret_cast<int32_t*>(sc + kNaSlot),
                 reinterpret_cast<int64_t*>(sc + kNaScanOff), hs};
    HCHECK(launch_unique(matrix_dev, nrow, ncol, total_dev, w, rmat_dev, rtot_dev, iuni_dev,
                         irec_dev, count_dev, (hipStream_t)o->stream));
    return THISTLE_OK;
}

int32_t thistle_pair_tables(const ThistleBasis* basis, double* E, double* P, double* e12,
                            const ThistleOpts* opts) {
    const ThistleOpts* o = opts ? opts : &kDefaultOpts;
    std::unique_ptr<ThistlePlan> pl(new ThistlePlan());
    int rc = setup_common(pl.get(), basis, o);
    if (rc) return rc;
    rc = run_tables_and_enum(pl.get(), false);
    if (rc) return rc;
    int L1 = pl->maxl + 1, T1 = 2 * pl->maxl + 1;
    if (E)
        HCHECK(hipMemcpyAsync(E, pl->E.p, sizeof(double) * (size_t)pl->npair * 3 * L1 * L1 * T1,
                              hipMemcpyDeviceToHost, pl->stream));
    if (e12)
        HCHECK(hipMemcpyAsync(e12, pl->e12.p, sizeof(double) * (size_t)pl->npair * pl->nq,
                              hipMemcpyDeviceToHost, pl->stream));
    std::vector<PairInfo> pi(pl->npair);
    HCHECK(hipMemcpyAsync(pi.data(), pl->P.p, sizeof(PairInfo) * pl->npair, hipMemcpyDeviceToHost,
                          pl->stream));
    HCHECK(hipStreamSynchronize(pl->stream));
    if (P)
        for (int p = 0; p < pl->npair; ++p) {
            P[3 * p + 0] = pi[p].Px;
            P[3 * p + 1] = pi[p].Py;
            P[3 * p + 2] = pi[p].Pz;
        }
    return THISTLE_OK;
}

}  // extern "C"
