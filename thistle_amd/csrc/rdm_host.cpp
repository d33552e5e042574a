// Host-side 2-RDM and 1-RDM construction from determinant lists.
//
// createtwordm (src/twordms.f90:175-556) accumulates every contribution into
// a dense twordm(p,q,r,s) in (c1, c2) loop order; the floating-point order of
// those additions decides which entries clear |Gamma| >= cutoff, so this
// builder keeps that order exactly.  Determinants are held as slot bitmasks;
// the connected (c1, c2) pairs (ndiff in {0, 2, 4}) are found in parallel and
// the accumulation then runs once, serially, in the reference order.
#include "thistle_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct Dets {
    int nconf = 0, lconfs = 0, nw = 0;
    std::vector<uint64_t> mask;   // nconf * nw words, bit j = slot j occupied
    bool canonical = true;        // value at slot j is 1 (j even) or 2 (j odd)
};

inline int conf_at(const int32_t* confs, int nconf, int c, int j) {
    return confs[(size_t)c + (size_t)nconf * (size_t)j];
}

Dets pack(const int32_t* confs, int nconf, int lconfs) {
    Dets d;
    d.nconf = nconf;
    d.lconfs = lconfs;
    d.nw = (lconfs + 63) / 64;
    d.mask.assign((size_t)nconf * d.nw, 0ull);
    for (int c = 0; c < nconf; ++c)
        for (int j = 0; j < lconfs; ++j) {
            int v = conf_at(confs, nconf, c, j);
            if (v != 0) {
                d.mask[(size_t)c * d.nw + j / 64] |= 1ull << (j % 64);
                if (v != ((j % 2 == 0) ? 1 : 2)) d.canonical = false;
            }
        }
    return d;
}

// ndiff of maxcoincidence (src/twordms.f90:63-69): columns that differ.
inline int ndiff_of(const Dets& d, const int32_t* confs, int c1, int c2) {
    if (c1 == c2) return 0;
    if (d.canonical) {
        int nd = 0;
        const uint64_t* a = &d.mask[(size_t)c1 * d.nw];
        const uint64_t* b = &d.mask[(size_t)c2 * d.nw];
        for (int w = 0; w < d.nw; ++w) nd += __builtin_popcountll(a[w] ^ b[w]);
        return nd;
    }
    int nd = 0;
    for (int j = 0; j < d.lconfs; ++j)
        nd += conf_at(confs, d.nconf, c1, j) != conf_at(confs, d.nconf, c2, j);
    return nd;
}

inline int occ_list(const Dets& d, int c, int* out) {
    int k = 0;
    for (int w = 0; w < d.nw; ++w) {
        uint64_t m = d.mask[(size_t)c * d.nw + w];
        while (m) {
            int b = __builtin_ctzll(m);
            out[k++] = w * 64 + b + 1;  // 1-based slot
            m &= m - 1;
        }
    }
    return k;
}

// nint(slot/2 + 0.1) for a 1-based slot (src/twordms.f90:158).
inline int slot_orb(int slot1) { return (slot1 + 1) / 2; }

// maxc_individual, src/twordms.f90:83-172: align c2's occupied slots onto
// c1's by swaps (one sign flip per swap) and report the unmatched slots.
int maxc_individual(const Dets& d, int c1, int c2, int* ep, int* d1, int* d2,
                    int* s1, int* s2) {
    int m1[256], m2[256];
    int cnt = occ_list(d, c1, m1);
    occ_list(d, c2, m2);
    int e = 1;
    for (int i = 0; i < cnt; ++i) {
        if (m1[i] != m2[i]) {
            for (int j = 0; j < cnt; ++j) {
                if (m1[i] == m2[j]) {
                    e = -e;
                    std::swap(m2[i], m2[j]);
                    break;
                }
            }
        }
    }
    int nd = 0;
    for (int i = 0; i < cnt; ++i)
        if (m1[i] != m2[i]) {
            d1[nd] = slot_orb(m1[i]);
            d2[nd] = slot_orb(m2[i]);
            s1[nd] = m1[i] % 2;
            s2[nd] = m2[i] % 2;
            ++nd;
        }
    *ep = e;
    return nd;
}

// ep2 of maxcoincidence for the unordered pair (no swaps), src/twordms.f90:39-61.
int ep2_of(const Dets& d, int c1, int c2) {
    if (c1 == c2) return 1;
    int a = std::min(c1, c2), b = std::max(c1, c2);
    int m1[256], m2[256];
    int k1 = occ_list(d, a, m1);
    int k2 = occ_list(d, b, m2);
    int e = 1;
    for (int i = 0; i < k1; ++i)
        if (m1[i] != m2[i])
            for (int j = 0; j < k2; ++j)
                if (m1[i] == m2[j]) {
                    e = -e;
                    break;
                }
    return e;
}

// connected pairs: for each c1 the c2 (in increasing order) with ndiff <= 4
void connected(const Dets& d, const int32_t* confs, int maxdiff,
               std::vector<std::vector<int32_t>>& nbr,
               std::vector<std::vector<int8_t>>& nds) {
    nbr.assign(d.nconf, {});
    nds.assign(d.nconf, {});
#pragma omp parallel for schedule(dynamic, 16)
    for (int c1 = 0; c1 < d.nconf; ++c1) {
        for (int c2 = 0; c2 < d.nconf; ++c2) {
            int nd = ndiff_of(d, confs, c1, c2);
            if (nd <= maxdiff) {
                nbr[c1].push_back(c2);
                nds[c1].push_back((int8_t)nd);
            }
        }
    }
}

}  // namespace

// Dense twordm(porb, rorb, sorb, qorb) accumulation of createtwordm.
int thistle_twordm_dense_host(const int32_t* confs, int nconf, int lconfs,
                              const double* civs, int nstates, int state1,
                              int state2, double* tw) {
    if (nconf <= 0 || lconfs <= 0 || (lconfs % 2) != 0) return THISTLE_EINVAL;
    if (state1 < 1 || state2 < 1 || state1 > nstates || state2 > nstates)
        return THISTLE_EINVAL;
    if (lconfs > 256) return THISTLE_EINVAL;
    const size_t N = (size_t)(lconfs / 2);
    std::memset(tw, 0, sizeof(double) * N * N * N * N);
    Dets d = pack(confs, nconf, lconfs);
    std::vector<std::vector<int32_t>> nbr;
    std::vector<std::vector<int8_t>> nds;
    connected(d, confs, 4, nbr, nds);
    auto TW = [&](int p, int r, int s, int q) -> double& {
        return tw[(size_t)(p - 1) + N * ((size_t)(r - 1) + N * ((size_t)(s - 1) + N * (size_t)(q - 1)))];
    };
    const double* c1v = civs + (size_t)nconf * (state1 - 1);
    const double* c2v = civs + (size_t)nconf * (state2 - 1);
    int occ2[256];
    for (int c1 = 0; c1 < nconf; ++c1) {
        const auto& L = nbr[c1];
        for (size_t k = 0; k < L.size(); ++k) {
            int c2 = L[k];
            int nd = nds[c1][k];
            if (nd == 4) {
                int ep, d1[128], d2[128], sp1[128], sp2[128];
                maxc_individual(d, c1, c2, &ep, d1, d2, sp1, sp2);
                double base = c1v[c1] * c2v[c2] * (double)ep;
                if (sp2[0] == sp1[0] && sp2[1] == sp1[1]) TW(d1[1], d1[0], d2[0], d2[1]) += base * 1.0;
                if (sp2[0] == sp1[1] && sp2[1] == sp1[0]) TW(d1[0], d1[1], d2[0], d2[1]) += base * -1.0;
                if (sp1[0] == sp2[1] && sp2[0] == sp1[1]) TW(d1[1], d1[0], d2[1], d2[0]) += base * -1.0;
                if (sp1[1] == sp2[1] && sp2[0] == sp1[0]) TW(d1[0], d1[1], d2[1], d2[0]) += base * 1.0;
            } else if (nd == 2) {
                int ep, d1[128], d2[128], sp1[128], sp2[128];
                maxc_individual(d, c1, c2, &ep, d1, d2, sp1, sp2);
                double base = c1v[c1] * c2v[c2] * (double)ep;
                // slots occupied in both determinants, in slot order
                int nocc = 0;
                for (int w = 0; w < d.nw; ++w) {
                    uint64_t m = d.mask[(size_t)c1 * d.nw + w] & d.mask[(size_t)c2 * d.nw + w];
                    while (m) {
                        int b = __builtin_ctzll(m);
                        occ2[nocc++] = w * 64 + b;  // 0-based slot
                        m &= m - 1;
                    }
                }
                if (!d.canonical) {
                    nocc = 0;
                    for (int i = 0; i < lconfs; ++i)
                        if (conf_at(confs, nconf, c2, i) != 0 && conf_at(confs, nconf, c1, i) != 0)
                            occ2[nocc++] = i;
                }
                if (sp2[0] == sp1[0])
                    for (int k2 = 0; k2 < nocc; ++k2) {
                        int o = slot_orb(occ2[k2] + 1);
                        TW(d1[0], o, o, d2[0]) += base * 1.0;
                    }
                for (int k2 = 0; k2 < nocc; ++k2) {
                    int sp = (occ2[k2] + 1) % 2;
                    if (sp1[0] == sp && sp2[0] == sp) {
                        int o = slot_orb(occ2[k2] + 1);
                        TW(o, d1[0], o, d2[0]) += base * -1.0;
                    }
                }
                for (int k2 = 0; k2 < nocc; ++k2) {
                    int sp = (occ2[k2] + 1) % 2;
                    if (sp1[0] == sp && sp2[0] == sp) {
                        int o = slot_orb(occ2[k2] + 1);
                        TW(d1[0], o, d2[0], o) += base * -1.0;
                    }
                }
                if (sp2[0] == sp1[0])
                    for (int k2 = 0; k2 < nocc; ++k2) {
                        int o = slot_orb(occ2[k2] + 1);
                        TW(o, d1[0], d2[0], o) += base * 1.0;
                    }
            } else if (nd == 0) {
                double base = c1v[c1] * c2v[c2] * 1.0;
                for (int i1 = 0; i1 < lconfs; ++i1) {
                    int v1 = conf_at(confs, nconf, c1, i1);
                    if (v1 == 0) continue;
                    for (int i2 = 0; i2 < lconfs; ++i2) {
                        if (i1 == i2 || conf_at(confs, nconf, c2, i2) == 0) continue;
                        int so = slot_orb(i1 + 1), qo = slot_orb(i2 + 1);
                        if (v1 == conf_at(confs, nconf, c1, i2)) TW(so, qo, so, qo) += base * -1.0;
                        TW(qo, so, so, qo) += base * 1.0;
                    }
                }
            }
        }
    }
    return THISTLE_OK;
}

// Selection loop of createtwordm (src/twordms.f90:522-550).
int64_t thistle_twordm_select_host(const double* tw, int norb, double cutoff,
                                   int32_t* mat, double* total, int64_t cap) {
    const size_t N = (size_t)norb;
    int64_t cnt = 0;
    for (int p = 1; p <= norb; ++p)
        for (int q = 1; q <= norb; ++q)
            for (int r = 1; r <= norb; ++r)
                for (int s = 1; s <= norb; ++s) {
                    double v = tw[(size_t)(p - 1) + N * ((size_t)(q - 1) + N * ((size_t)(r - 1) + N * (size_t)(s - 1)))];
                    if (std::fabs(v) >= cutoff) {
                        if (mat && cnt < cap) {
                            mat[cnt + 0 * cap] = p;
                            mat[cnt + 1 * cap] = s;
                            mat[cnt + 2 * cap] = q;
                            mat[cnt + 3 * cap] = r;
                            total[cnt] = v;
                        }
                        ++cnt;
                    }
                }
    return cnt;
}

// createonerdm, src/one_rdm.f90:109-222.
int thistle_onerdm_host(const int32_t* confs, int nconf, int lconfs,
                        const double* civs, int nstates, int state1, int state2,
                        double* onerdm) {
    if (nconf <= 0 || lconfs <= 0 || (lconfs % 2) != 0 || lconfs > 256) return THISTLE_EINVAL;
    if (state1 < 1 || state2 < 1 || state1 > nstates || state2 > nstates) return THISTLE_EINVAL;
    const int norb = lconfs / 2;
    std::memset(onerdm, 0, sizeof(double) * norb * norb);
    Dets d = pack(confs, nconf, lconfs);
    std::vector<std::vector<int32_t>> nbr;
    std::vector<std::vector<int8_t>> nds;
    connected(d, confs, 2, nbr, nds);
    const double* c1v = civs + (size_t)nconf * (state1 - 1);
    const double* c2v = civs + (size_t)nconf * (state2 - 1);
    for (int c1 = 0; c1 < nconf; ++c1) {
        const auto& L = nbr[c1];
        for (size_t k = 0; k < L.size(); ++k) {
            int c2 = L[k];
            int nd = nds[c1][k];
            if (nd == 2) {
                int ep = ep2_of(d, c1, c2);
                int d1 = 0, d2 = 0;
                for (int j = 0; j < lconfs; ++j) {
                    int a = conf_at(confs, nconf, c1, j), b = conf_at(confs, nconf, c2, j);
                    if (a != b) {
                        if (a != 0) d1 = slot_orb(j + 1);
                        else if (b != 0) d2 = slot_orb(j + 1);
                    }
                }
                onerdm[(d1 - 1) + norb * (d2 - 1)] += c1v[c1] * c2v[c2] * (double)ep;
            } else if (nd == 0) {
                for (int i1 = 0; i1 < lconfs; ++i1)
                    if (conf_at(confs, nconf, c1, i1) != 0) {
                        int so = slot_orb(i1 + 1);
                        onerdm[(so - 1) + norb * (so - 1)] += c1v[c1] * c2v[c1] * 1.0;
                    }
            }
        }
    }
    return THISTLE_OK;
}
