/*
 * CPU oracle for the THISTLE total / elastic j0 scattering path.
 *
 * TEST INFRASTRUCTURE ONLY -- the checker, never the product.  A plain-C
 * restatement of the reference Fortran (KirranderGroup/thistle).  Each
 * function names the reference lines it follows.  Built with
 * -ffp-contract=off so every floating-point operation is rounded in the
 * order written, as the reference's Fortran is.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* pi = acos(-1.0000): a single-precision acos widened to double
 * (src/Variables.f90:994,1174). */
static const double PI_SP = 3.1415927410125732421875;
/* 1.0e-30 is a default-real literal in the MD zero-skips and BesselDeriv
 * (src/Integrals.f90:4725,4800); compared after widening. */
static const float TINY_SP = 1.0e-30f;

#define DIMH 20 /* `dim` of tot_integral_k_ijkr, src/Integrals.f90:5093 */

static double ipow(double x, int n)
{
    /* x**n for integer n by binary powering (the usual Fortran lowering). */
    double r = 1.0;
    int e = n < 0 ? -n : n;
    double b = x;
    while (e) {
        if (e & 1) r *= b;
        b *= b;
        e >>= 1;
    }
    return n < 0 ? 1.0 / r : r;
}

/* ------------------------------------------------------------------ P0 */
/* P_LMN, src/Integrals.f90:3102-3135 */
static void p_lmn(double *P0, int D, int l, int m, int n, const double *q,
                  int nq)
{
    double prod = 1.0;
    int r = l + m + n;
    for (int i = 3; i <= l - 1; i += 2) prod = prod * i;
    for (int i = 3; i <= m - 1; i += 2) prod = prod * i;
    for (int i = 3; i <= n - 1; i += 2) prod = prod * i;
    for (int i = 3; i <= r + 1; i += 2) prod = prod / i;
    double sgn = ((r / 2) & 1) ? -1.0 : 1.0;
    double *dst = P0 + ((size_t)((l * D + m) * D + n)) * nq;
    for (int iq = 0; iq < nq; ++iq) dst[iq] = prod * sgn * ipow(q[iq], r);
}

/* set_P0, src/Integrals.f90:3001-3031 */
void orc_set_p0(double *P0, int lmax4, const double *q, int nq)
{
    int D = lmax4 + 1;
    memset(P0, 0, sizeof(double) * (size_t)D * D * D * nq);
    for (int k = 0; k <= lmax4; k += 2)
        for (int j = 0; j <= lmax4; j += 2)
            for (int i = 0; i <= lmax4; i += 2) p_lmn(P0, D, i, j, k, q, nq);
}

/* P_LMN_j2, src/Integrals.f90:3137-3170 */
static void p_lmn_j2(double *P0, int D, int l, int m, int n, const double *q,
                     int nq)
{
    double prod = 1.0;
    int r = l + m + n;
    for (int i = 3; i <= l - 1; i += 2) prod = prod * i;
    for (int i = 3; i <= m - 1; i += 2) prod = prod * i;
    for (int i = 3; i <= n - 1; i += 2) prod = prod * i;
    for (int i = 3; i <= r + 3; i += 2) prod = prod / i;
    double sgn = ((r / 2) & 1) ? -1.0 : 1.0;
    double *dst = P0 + ((size_t)((l * D + m) * D + n)) * nq;
    for (int iq = 0; iq < nq; ++iq)
        dst[iq] = (l + m - 2 * n) / 4.0 * prod * sgn * ipow(q[iq], r);
}

/* set_P0_j2, src/Integrals.f90:3034-3064 (times sqrt(5/pi), double pi) */
void orc_set_p0_j2(double *P0, int lmax4, const double *q, int nq)
{
    int D = lmax4 + 1;
    memset(P0, 0, sizeof(double) * (size_t)D * D * D * nq);
    for (int k = 0; k <= lmax4; k += 2)
        for (int j = 0; j <= lmax4; j += 2)
            for (int i = 0; i <= lmax4; i += 2) p_lmn_j2(P0, D, k, j, i, q, nq);
    const double f = sqrt(5.0 / acos(-1.0));
    for (size_t t = 0; t < (size_t)D * D * D * nq; ++t) P0[t] = P0[t] * f;
}

/* ------------------------------------------------------------ MD tables */
/* One top-level slice E^{l1,l2}_t of fill_md_table (src/Variables.f90:31-944):
 * first raise the "A" index with PA to l1, then the "B" index with PB to l2,
 * E^{i,j+1}_t = a E^{ij}_{t-1} + PB E^{ij}_t + (t+1) E^{ij}_{t+1}. */
static void md_slice(double PA, double PB, double a, int l1, int l2,
                     double *out /* [l1+l2+1] */)
{
    double cur[16], nxt[16];
    memset(cur, 0, sizeof cur);
    cur[0] = 1.0;
    int top = 0;
    for (int i = 1; i <= l1; ++i) {
        memset(nxt, 0, sizeof nxt);
        for (int t = 0; t <= top + 1; ++t) {
            double em1 = t > 0 ? cur[t - 1] : 0.0;
            double e0 = t <= top ? cur[t] : 0.0;
            double ep1 = t + 1 <= top ? cur[t + 1] : 0.0;
            double v;
            if (t == 0)
                v = PA * e0 + ep1;
            else if (t == top + 1)
                v = a * em1;
            else if (t == top)
                v = a * em1 + PA * e0;
            else
                v = a * em1 + PA * e0 + (double)(t + 1) * ep1;
            nxt[t] = v;
        }
        top += 1;
        memcpy(cur, nxt, sizeof cur);
    }
    for (int j = 1; j <= l2; ++j) {
        memset(nxt, 0, sizeof nxt);
        for (int t = 0; t <= top + 1; ++t) {
            double em1 = t > 0 ? cur[t - 1] : 0.0;
            double e0 = t <= top ? cur[t] : 0.0;
            double ep1 = t + 1 <= top ? cur[t + 1] : 0.0;
            double v;
            if (t == 0)
                v = PB * e0 + ep1;
            else if (t == top + 1)
                v = a * em1;
            else if (t == top)
                v = a * em1 + PB * e0;
            else
                v = a * em1 + PB * e0 + (double)(t + 1) * ep1;
            nxt[t] = v;
        }
        top += 1;
        memcpy(cur, nxt, sizeof cur);
    }
    for (int t = 0; t <= l1 + l2; ++t) out[t] = cur[t];
}

/* Hermite coefficients of one ordered GTO pair (u,v) along one axis, as
 * fill_md_table produces D(u,v,:,l(u)+1,l(v)+1): the higher-l member is the
 * "A" side of the recursion; on a tie the lower index is; Px uses the pair in
 * index order (src/Variables.f90:31-56,690-700). */
static void md_gto_pair(int u, int v, const int *l, const double *x,
                        const double *ga, double *out)
{
    int i = u < v ? u : v, j = u < v ? v : u;
    double gaP = ga[i] + ga[j];
    double Px = (ga[i] * x[i] + ga[j] * x[j]) / gaP;
    int first, second;
    if (u == v) {
        first = u;
        second = v;
    } else if (l[i] < l[j]) {
        first = j;
        second = i;
    } else {
        first = i;
        second = j;
    }
    double PA = Px - x[first];
    double PB = Px - x[second];
    double a = 0.5 / gaP;
    double tmp[16];
    md_slice(PA, PB, a, l[first], l[second], tmp);
    /* D(second,first) is a copy of D(first,second): the same numbers. */
    for (int t = 0; t <= l[u] + l[v]; ++t) out[t] = tmp[t];
}

void orc_pair_tables(int ngto, int ng, int maxl, const int *l, const int *m,
                     const int *n, const int *group, const double *ga,
                     const double *xx, const double *yy, const double *zz,
                     const double *q, int nq, double *ddx, double *ddy,
                     double *ddz, double *px, double *py, double *pz,
                     double *e12)
{
    int L1 = maxl + 1, T1 = 2 * maxl + 1;
    size_t dsz = (size_t)ng * ng * L1 * L1 * T1;
    memset(ddx, 0, sizeof(double) * dsz);
    memset(ddy, 0, sizeof(double) * dsz);
    memset(ddz, 0, sizeof(double) * dsz);
    int *gstart = (int *)malloc(sizeof(int) * ng);
    for (int g = 0; g < ng; ++g) gstart[g] = -1;
    for (int g = 0; g < ng; ++g)
        for (int j = 0; j < ngto; ++j)
            if (group[j] == g + 1) {
                gstart[g] = j;
                break;
            }
    /* pair centres and e12, src/Variables.f90:1113-1128 */
    for (int gj = 0; gj < ng; ++gj) {
        int j = gstart[gj];
        for (int gi = 0; gi < ng; ++gi) {
            int i = gstart[gi];
            double gaP = ga[i] + ga[j];
            px[gi * ng + gj] = (ga[i] * xx[i] + ga[j] * xx[j]) / gaP;
            py[gi * ng + gj] = (ga[i] * yy[i] + ga[j] * yy[j]) / gaP;
            pz[gi * ng + gj] = (ga[i] * zz[i] + ga[j] * zz[j]) / gaP;
            double dx = xx[i] - xx[j], dy = yy[i] - yy[j], dz = zz[i] - zz[j];
            double r2 = dx * dx + dy * dy + dz * dz;
            double pre = pow(PI_SP / gaP, 1.5);
            double kab = exp(-(ga[i] * ga[j]) / gaP * r2);
            for (int iq = 0; iq < nq; ++iq)
                e12[(size_t)(gi * ng + gj) * nq + iq] =
                    pre * exp(-(q[iq] * q[iq]) * 0.25 / gaP) * kab;
        }
    }
    /* Ddx repack, src/Variables.f90:1132-1150: jj outer, ii inner; the last
     * writer of a slot wins. */
    double tmp[16];
    for (int jj = 0; jj < ngto; ++jj) {
        int gj = group[jj] - 1;
        for (int ii = 0; ii < ngto; ++ii) {
            int gi = group[ii] - 1;
            size_t base = (size_t)(gi * ng + gj) * L1 * L1;
            md_gto_pair(ii, jj, l, xx, ga, tmp);
            for (int t = 0; t <= l[ii] + l[jj]; ++t)
                ddx[(base + l[ii] * L1 + l[jj]) * T1 + t] = tmp[t];
            md_gto_pair(ii, jj, m, yy, ga, tmp);
            for (int t = 0; t <= m[ii] + m[jj]; ++t)
                ddy[(base + m[ii] * L1 + m[jj]) * T1 + t] = tmp[t];
            md_gto_pair(ii, jj, n, zz, ga, tmp);
            for (int t = 0; t <= n[ii] + n[jj]; ++t)
                ddz[(base + n[ii] * L1 + n[jj]) * T1 + t] = tmp[t];
        }
    }
    free(gstart);
}

/* ------------------------------------------------------------- 2-RDM */
#define CONF(c, j) confs[(size_t)(c) + (size_t)nconf * (size_t)(j)]
#define CIV(c, s) civs[(size_t)(c) + (size_t)nconf * (size_t)(s)]

/* ndiff of maxcoincidence (src/twordms.f90:63-69): number of differing
 * columns. */
static int conf_ndiff(const int *confs, int nconf, int lconfs, int c1, int c2)
{
    if (c1 == c2) return 0;
    int nd = 0;
    for (int j = 0; j < lconfs; ++j) nd += CONF(c1, j) != CONF(c2, j);
    return nd;
}

/* orbital of a 1-based slot: nint(slot/2.0 + 0.1) (src/twordms.f90:158). */
static inline int slot_orb(int slot1) { return (int)lround(slot1 / 2.0 + 0.1); }

/* maxc_individual, src/twordms.f90:83-172 */
static int maxc_individual(const int *confs, int nconf, int lconfs, int c1,
                           int c2, int *ep, int *d1, int *d2, int *s1, int *s2)
{
    int m1[128], m2[128];
    int k1 = 0, k2 = 0;
    for (int i = 0; i < lconfs; ++i)
        if (CONF(c1, i) != 0) m1[k1++] = i + 1;
    for (int i = 0; i < lconfs; ++i)
        if (CONF(c2, i) != 0) m2[k2++] = i + 1;
    int cnt = k1;
    int e = 1;
    for (int i = 0; i < cnt; ++i) {
        if (m1[i] != m2[i]) {
            for (int j = 0; j < cnt; ++j) {
                if (m1[i] == m2[j]) {
                    e = -e;
                    int t = m2[j];
                    m2[j] = m2[i];
                    m2[i] = t;
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

/* createtwordm accumulation, src/twordms.f90:209-512.  twordm is indexed as
 * the reference indexes it: twordm(porb, rorb, sorb, qorb). */
void orc_twordm_dense(const int *confs, int nconf, int lconfs,
                      const double *civs, int nstates, int state1, int state2,
                      double *twordm)
{
    (void)nstates;
    int norb = lconfs / 2;
    size_t N = (size_t)norb;
    memset(twordm, 0, sizeof(double) * N * N * N * N);
#define TW(p, r, s, qq)                                                        \
    twordm[(size_t)((p)-1) +                                                   \
           N * ((size_t)((r)-1) + N * ((size_t)((s)-1) + N * (size_t)((qq)-1)))]
    int s1i = state1 - 1, s2i = state2 - 1;
    for (int c1 = 0; c1 < nconf; ++c1) {
        for (int c2 = 0; c2 < nconf; ++c2) {
            int nd = conf_ndiff(confs, nconf, lconfs, c1, c2);
            if (nd != 0 && nd <= 4) {
                int ep, d1[64], d2[64], sp1[64], sp2[64];
                if (nd == 4) {
                    maxc_individual(confs, nconf, lconfs, c1, c2, &ep, d1, d2,
                                    sp1, sp2);
                    double base = CIV(c1, s1i) * CIV(c2, s2i) * (double)ep;
                    int porb, rorb, sorb, qorb;
                    sorb = d2[0]; qorb = d2[1]; porb = d1[1]; rorb = d1[0];
                    if (sp2[0] == sp1[0] && sp2[1] == sp1[1])
                        TW(porb, rorb, sorb, qorb) += base * 1.0;
                    sorb = d2[0]; qorb = d2[1]; rorb = d1[1]; porb = d1[0];
                    if (sp2[0] == sp1[1] && sp2[1] == sp1[0])
                        TW(porb, rorb, sorb, qorb) += base * -1.0;
                    qorb = d2[0]; sorb = d2[1]; porb = d1[1]; rorb = d1[0];
                    if (sp1[0] == sp2[1] && sp2[0] == sp1[1])
                        TW(porb, rorb, sorb, qorb) += base * -1.0;
                    qorb = d2[0]; sorb = d2[1]; rorb = d1[1]; porb = d1[0];
                    if (sp1[1] == sp2[1] && sp2[0] == sp1[0])
                        TW(porb, rorb, sorb, qorb) += base * 1.0;
                } else if (nd == 2) {
                    maxc_individual(confs, nconf, lconfs, c1, c2, &ep, d1, d2,
                                    sp1, sp2);
                    double base = CIV(c1, s1i) * CIV(c2, s2i) * (double)ep;
                    /* four slot loops; a term fires only where the slot is
                     * occupied in both determinants (flags reset each pass) */
                    for (int i = 0; i < lconfs; ++i) {
                        if (CONF(c2, i) != 0 && CONF(c1, i) != 0 &&
                            sp2[0] == sp1[0]) {
                            int o = slot_orb(i + 1);
                            TW(d1[0], o, o, d2[0]) += base * 1.0;
                        }
                    }
                    for (int i = 0; i < lconfs; ++i) {
                        int sp = (i + 1) % 2;
                        if (CONF(c2, i) != 0 && CONF(c1, i) != 0 &&
                            sp1[0] == sp && sp2[0] == sp) {
                            int o = slot_orb(i + 1);
                            TW(o, d1[0], o, d2[0]) += base * -1.0;
                        }
                    }
                    for (int i = 0; i < lconfs; ++i) {
                        int sp = (i + 1) % 2;
                        if (CONF(c2, i) != 0 && CONF(c1, i) != 0 &&
                            sp1[0] == sp && sp2[0] == sp) {
                            int o = slot_orb(i + 1);
                            TW(d1[0], o, d2[0], o) += base * -1.0;
                        }
                    }
                    for (int i = 0; i < lconfs; ++i) {
                        if (CONF(c2, i) != 0 && CONF(c1, i) != 0 &&
                            sp2[0] == sp1[0]) {
                            int o = slot_orb(i + 1);
                            TW(o, d1[0], d2[0], o) += base * 1.0;
                        }
                    }
                }
            } else if (nd == 0) {
                double base = CIV(c1, s1i) * CIV(c2, s2i) * 1.0;
                for (int i1 = 0; i1 < lconfs; ++i1) {
                    for (int i2 = 0; i2 < lconfs; ++i2) {
                        if (i1 != i2 && CONF(c1, i1) != 0 && CONF(c2, i2) != 0) {
                            int so = slot_orb(i1 + 1), qo = slot_orb(i2 + 1);
                            if (CONF(c1, i1) == CONF(c1, i2))
                                TW(so, qo, so, qo) += base * -1.0;
                            TW(qo, so, so, qo) += base * 1.0;
                        }
                    }
                }
            }
        }
    }
#undef TW
}

int orc_twordm_select(const double *twordm, int norb, double cutoff, int *mat,
                      double *total, int cap)
{
    size_t N = (size_t)norb;
    int cnt = 0;
    for (int p = 1; p <= norb; ++p)
        for (int q = 1; q <= norb; ++q)
            for (int r = 1; r <= norb; ++r)
                for (int s = 1; s <= norb; ++s) {
                    double v = twordm[(size_t)(p - 1) +
                                      N * ((size_t)(q - 1) +
                                           N * ((size_t)(r - 1) +
                                                N * (size_t)(s - 1)))];
                    if (fabs(v) >= cutoff) {
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

/* ep2 of maxcoincidence for the unordered pair (no swapping; the sign flips
 * once per misplaced occupied slot found elsewhere), src/twordms.f90:39-61. */
static int conf_ep2(const int *confs, int nconf, int lconfs, int c1, int c2)
{
    if (c1 == c2) return 1;
    int a = c1 < c2 ? c1 : c2, b = c1 < c2 ? c2 : c1;
    int m1[128], m2[128], k1 = 0, k2 = 0;
    for (int j = 0; j < lconfs; ++j)
        if (CONF(a, j) != 0) m1[k1++] = j + 1;
    for (int j = 0; j < lconfs; ++j)
        if (CONF(b, j) != 0) m2[k2++] = j + 1;
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

void orc_pair_phase(const int *confs, int nconf, int lconfs, int c1, int c2,
                    int *ep2, int *ndiff)
{
    *ep2 = conf_ep2(confs, nconf, lconfs, c1, c2);
    *ndiff = conf_ndiff(confs, nconf, lconfs, c1, c2);
}

/* createonerdm, src/one_rdm.f90:109-222 */
void orc_onerdm(const int *confs, int nconf, int lconfs, const double *civs,
                int nstates, int state1, int state2, double *onerdm)
{
    (void)nstates;
    int norb = lconfs / 2;
    int s1i = state1 - 1, s2i = state2 - 1;
    memset(onerdm, 0, sizeof(double) * norb * norb);
    for (int c1 = 0; c1 < nconf; ++c1) {
        for (int c2 = 0; c2 < nconf; ++c2) {
            int nd = conf_ndiff(confs, nconf, lconfs, c1, c2);
            if (nd == 2) {
                int ep = conf_ep2(confs, nconf, lconfs, c1, c2);
                int d1 = 0, d2 = 0;
                for (int j = 0; j < lconfs; ++j) {
                    if (CONF(c1, j) != CONF(c2, j)) {
                        if (CONF(c1, j) != 0)
                            d1 = slot_orb(j + 1);
                        else if (CONF(c2, j) != 0)
                            d2 = slot_orb(j + 1);
                    }
                }
                onerdm[(d1 - 1) + norb * (d2 - 1)] +=
                    CIV(c1, s1i) * CIV(c2, s2i) * (double)ep;
            } else if (nd == 0) {
                for (int i1 = 0; i1 < lconfs; ++i1)
                    if (CONF(c1, i1) != 0) {
                        int so = slot_orb(i1 + 1);
                        onerdm[(so - 1) + norb * (so - 1)] +=
                            CIV(c1, s1i) * CIV(c1, s2i) * 1.0;
                    }
            }
        }
    }
}
#undef CONF
#undef CIV

/* variables_elastic Z, src/Variables.f90:1199-1216 */
void orc_elastic_z(const double *onerdm, int norb, const double *mmod, int nmo,
                   int ngto, double *Z)
{
#define MM(i, j) mmod[(size_t)(i) + (size_t)nmo * (size_t)(j)]
    for (int ii = 0; ii < ngto; ++ii)
        for (int jj = 0; jj < ngto; ++jj) {
            double temp = 0.0;
            for (int cc = 0; cc < norb; ++cc)
                for (int cc2 = 0; cc2 < norb; ++cc2) {
                    double g = onerdm[cc + norb * cc2];
                    temp = temp + g * MM(cc, ii) * MM(cc2, jj);
                    temp = temp + g * MM(cc, jj) * MM(cc2, ii);
                }
            Z[ii + (size_t)ngto * jj] = temp / 2.0;
        }
#undef MM
}

/* ------------------------------------------------------------ enumeration */
int64_t orc_quartet_list(int ng, const double *px, const double *py,
                         const double *pz, double cutoffcentre, int32_t *out)
{
    int64_t c = 0;
#define P(A, g1, g2) A[(g1) * ng + (g2)]
#define EMIT(a, b, cc, d, mult, br)                                             \
    do {                                                                       \
        if (out) {                                                             \
            int32_t *o = out + 6 * c;                                          \
            o[0] = (a) + 1; o[1] = (b) + 1; o[2] = (cc) + 1; o[3] = (d) + 1;   \
            o[4] = (mult); o[5] = (br);                                        \
        }                                                                      \
        ++c;                                                                   \
    } while (0)
    /* nest 1, src/Integrals.f90:79-82 */
    for (int i = 0; i < ng; ++i)
        for (int j = i + 1; j < ng; ++j)
            for (int k = i + 1; k < ng; ++k)
                for (int r = k + 1; r < ng; ++r) {
                    double hx = P(px, k, r) - P(px, i, j);
                    double hy = P(py, k, r) - P(py, i, j);
                    double hz = P(pz, k, r) - P(pz, i, j);
                    double h = pow(hx * hx + hy * hy + hz * hz, 0.5);
                    EMIT(i, j, k, r, 8, h < cutoffcentre);
                }
    /* nest 2, :178-180 */
    for (int i = 0; i < ng; ++i)
        for (int j = i + 1; j < ng; ++j)
            for (int r = i + 1; r < ng; ++r) {
                double hx = P(px, i, r) - P(px, i, j);
                double hy = P(py, i, r) - P(py, i, j);
                double hz = P(pz, i, r) - P(pz, i, j);
                double h = sqrt(hx * hx + hy * hy + hz * hz);
                EMIT(i, j, i, r, 4, h < cutoffcentre);
            }
    /* nest 3, :258-260 */
    for (int i = 0; i < ng; ++i)
        for (int k = 0; k < ng; ++k)
            for (int r = k + 1; r < ng; ++r) {
                double hx = P(px, k, r) - P(px, i, i);
                double hy = P(py, k, r) - P(py, i, i);
                double hz = P(pz, k, r) - P(pz, i, i);
                double h = sqrt(hx * hx + hy * hy + hz * hz);
                EMIT(i, i, k, r, 4, h < cutoffcentre);
            }
    /* nest 4, :340-341 */
    for (int i = 0; i < ng; ++i)
        for (int k = i + 1; k < ng; ++k) {
            double hx = P(px, k, k) - P(px, i, i);
            double hy = P(py, k, k) - P(py, i, i);
            double hz = P(pz, k, k) - P(pz, i, i);
            double h = sqrt(hx * hx + hy * hy + hz * hz);
            EMIT(i, i, k, k, 2, h < cutoffcentre);
        }
    /* nest 5, :416 (always the h = 0 branch) */
    for (int i = 0; i < ng; ++i) EMIT(i, i, i, i, 1, 1);
#undef EMIT
#undef P
    return c;
}

/* The six nests of elastic_integration_j2 (src/Integrals.f90:1941-2364):
 * the (i,j,i,r) family of elastic_integration splits into r > j (mult 8)
 * and r = j (mult 4); the rest is the same as orc_quartet_list. */
int64_t orc_quartet_list_elastic_j2(int ng, const double *px, const double *py,
                                    const double *pz, double cutoffcentre,
                                    int32_t *out)
{
    int64_t c = 0;
#define P(A, g1, g2) A[(g1) * ng + (g2)]
#define EMIT(a, b, cc, d, mult, br)                                             \
    do {                                                                       \
        if (out) {                                                             \
            int32_t *o = out + 6 * c;                                          \
            o[0] = (a) + 1; o[1] = (b) + 1; o[2] = (cc) + 1; o[3] = (d) + 1;   \
            o[4] = (mult); o[5] = (br);                                        \
        }                                                                      \
        ++c;                                                                   \
    } while (0)
    /* :2004-2007 */
    for (int i = 0; i < ng; ++i)
        for (int j = i + 1; j < ng; ++j)
            for (int k = i + 1; k < ng; ++k)
                for (int r = k + 1; r < ng; ++r) {
                    double hx = P(px, k, r) - P(px, i, j);
                    double hy = P(py, k, r) - P(py, i, j);
                    double hz = P(pz, k, r) - P(pz, i, j);
                    double h = pow(hx * hx + hy * hy + hz * hz, 0.5);
                    EMIT(i, j, k, r, 8, h < cutoffcentre);
                }
    /* :2083-2085 */
    for (int i = 0; i < ng; ++i)
        for (int j = i + 1; j < ng; ++j)
            for (int r = j + 1; r < ng; ++r) {
                double hx = P(px, i, r) - P(px, i, j);
                double hy = P(py, i, r) - P(py, i, j);
                double hz = P(pz, i, r) - P(pz, i, j);
                double h = sqrt(hx * hx + hy * hy + hz * hz);
                EMIT(i, j, i, r, 8, h < cutoffcentre);
            }
    /* :2145-2146 (h = 0 by construction) */
    for (int i = 0; i < ng; ++i)
        for (int j = i + 1; j < ng; ++j) {
            double hx = P(px, i, j) - P(px, i, j);
            double hy = P(py, i, j) - P(py, i, j);
            double hz = P(pz, i, j) - P(pz, i, j);
            double h = sqrt(hx * hx + hy * hy + hz * hz);
            EMIT(i, j, i, j, 4, h < cutoffcentre);
        }
    /* :2204-2206 */
    for (int i = 0; i < ng; ++i)
        for (int k = 0; k < ng; ++k)
            for (int r = k + 1; r < ng; ++r) {
                double hx = P(px, k, r) - P(px, i, i);
                double hy = P(py, k, r) - P(py, i, i);
                double hz = P(pz, k, r) - P(pz, i, i);
                double h = sqrt(hx * hx + hy * hy + hz * hz);
                EMIT(i, i, k, r, 4, h < cutoffcentre);
            }
    /* :2268-2269 */
    for (int i = 0; i < ng; ++i)
        for (int k = i + 1; k < ng; ++k) {
            double hx = P(px, k, k) - P(px, i, i);
            double hy = P(py, k, k) - P(py, i, i);
            double hz = P(pz, k, k) - P(pz, i, i);
            double h = sqrt(hx * hx + hy * hy + hz * hz);
            EMIT(i, i, k, k, 2, h < cutoffcentre);
        }
    /* :2325 (always the h = 0 branch) */
    for (int i = 0; i < ng; ++i) EMIT(i, i, i, i, 1, 1);
#undef EMIT
#undef P
    return c;
}

/* ------------------------------------------------------- Bessel series */
/* dphrec, src/Integrals.f90:7147-7184 (rbes(leps1,nq) column-major) */
void orc_dphrec(const double *x, double *rbes, int leps1, int lmax1, int nq)
{
    for (int iq = 0; iq < nq; ++iq) {
        double *R = rbes + (size_t)iq * leps1;
        double X = x[iq];
        double XX = X * X;
        for (int k = 0; k < leps1; ++k) R[k] = 0.0;
        R[leps1 - 1] = 1.0;
        R[leps1 - 2] = 1.0;
        double CX = cos(X), SX = sin(X);
        double rneu2 = CX + X * SX;
        for (int K = 3; K <= leps1; ++K) {
            long LU = leps1 - K + 2;
            /* RBES(LU-1) = RBES(LU) - XX/(4*LU*LU-1)*RBES(LU+1) */
            R[LU - 2] = R[LU - 1] - XX / (4.0 * LU * LU - 1.0) * R[LU];
        }
        double A = R[0] * rneu2 - XX / 3.0 * R[1] * CX;
        for (int K = 0; K < leps1; ++K) R[K] = R[K] / A;
        double CU = 1.0 / X;
        for (int K = 1; K <= lmax1; ++K) {
            double W = 2.0 * (K - 1);
            CU = X / (W + 1.0) * CU;
            R[K - 1] = CU * R[K - 1];
        }
    }
}

/* bessels0rr, src/Integrals.f90:6662-6739 */
void orc_bessels0rr(double *hsum, int order, const double *mu, int nq,
                    double H, const double *h_saved)
{
    double *pmu = (double *)malloc(sizeof(double) * nq);
    double *all = (double *)malloc(sizeof(double) * 100 * (size_t)nq);
    for (int i = 0; i < nq; ++i) pmu[i] = H * mu[i];
    for (int i = 0; i < nq; ++i) {
        double invpmu = mu[i] / H;
        (void)invpmu;
        hsum[i] = sin(pmu[i]) / pmu[i] * h_saved[0];
    }
    orc_dphrec(pmu, all, 100, order + 1, nq);
    for (int beta = 1; beta <= order; ++beta)
        for (int i = 0; i < nq; ++i) {
            double invpmu = mu[i] / H;
            hsum[i] = hsum[i] + all[(size_t)i * 100 + beta] *
                                    ipow(invpmu, beta) * h_saved[beta];
        }
    free(all);
    free(pmu);
}

/* rrdj0, src/Integrals.f90:7062-7091 (a[l][p] = a(l+1,p+1)) */
static void rrdj0(int lmax, double x, double a[DIMH][DIMH])
{
    memset(a, 0, sizeof(double) * DIMH * DIMH);
    a[0][0] = 1.0;
    a[1][1] = -x;
    for (int l = 2; l <= lmax; ++l) a[l][0] = (double)(-(l - 1)) * a[l - 2][0];
    for (int l = 2; l <= lmax; ++l)
        for (int p = 1; p <= l; ++p)
            a[l][p] = (double)(-(l - 1)) * a[l - 2][p] - x * a[l - 1][p - 1];
}

/* Hermite_like_coeffs, src/Integrals.f90:7037-7059 (elastic path) */
static void hermite_like(int LLmax, double Hx, double a[DIMH][DIMH])
{
    memset(a, 0, sizeof(double) * DIMH * DIMH);
    a[0][0] = 1.0;
    if (LLmax > 0) a[1][1] = -Hx;
    for (int L = 2; L <= LLmax; ++L) {
        a[L][0] = (double)(-(L - 1)) * a[L - 2][0];
        for (int ka = 1; ka <= LLmax && ka < DIMH; ++ka)
            a[L][ka] = -Hx * a[L - 1][ka - 1] - (double)(L - 1) * a[L - 2][ka];
    }
}

/* BesselDeriv, src/Integrals.f90:4708-4746 */
static void bessel_deriv(double *BD, int LL, int MM, int NN,
                         double a[DIMH][DIMH], double b[DIMH][DIMH],
                         double c[DIMH][DIMH], int LLmax)
{
    for (int i = 0; i <= LLmax; ++i) BD[i] = 0.0;
    for (int ii = 0; ii <= LL; ++ii) {
        double C1 = a[LL][ii];
        if (fabs(C1) < (double)TINY_SP) continue;
        for (int jj = 0; jj <= MM; ++jj) {
            double Ct2 = b[MM][jj];
            if (fabs(Ct2) < (double)TINY_SP) continue;
            double C2 = C1 * Ct2;
            for (int kk = 0; kk <= NN; ++kk) {
                double Ct3 = c[NN][kk];
                if (fabs(Ct3) < (double)TINY_SP) continue;
                double C3 = C2 * Ct3;
                int temp = LL + MM + NN - ii - jj - kk;
                int ceil_ = temp / 2 + temp % 2;
                int hOrder = ceil_ + ii + jj + kk;
                BD[hOrder] = BD[hOrder] + C3;
            }
        }
    }
}

/* rrdj2a, src/Integrals.f90:6902-6927 (a[l][c] = a(l+1,c+1); the power p
 * of x sits in column c = p + 2) */
static void rrdj2a(int lmax, double x, double a[DIMH][DIMH])
{
    memset(a, 0, sizeof(double) * DIMH * DIMH);
    a[0][2] = 1.0;
    a[1][3] = -x;
    for (int l = 2; l <= lmax; ++l) a[l][2] = (double)(-(l - 1)) * a[l - 2][2];
    for (int l = 2; l <= lmax; ++l)
        for (int p = 1; p <= l; ++p)
            a[l][p + 2] = (double)(-(l - 1)) * a[l - 2][p + 2] - x * a[l - 1][p + 1];
}

/* rrdj2b, src/Integrals.f90:6929-6952 */
static void rrdj2b(int lmax, double a[DIMH][DIMH], double b[DIMH][DIMH])
{
    memset(b, 0, sizeof(double) * DIMH * DIMH);
    for (int l = 1; l <= lmax; ++l)
        for (int p = 0; p <= l; ++p) {
            double carg = (l + p) / 2.0;
            b[l][p] = 2.0 * ceil(carg) * a[l][p + 2];
        }
}

/* BesselDeriv2j, src/Integrals.f90:6370-6435 */
static void bessel_deriv2j(double *BD, double hz, double h, int LL, int MM, int NN,
                           double a[DIMH][DIMH], double b[DIMH][DIMH],
                           double c[DIMH][DIMH], double ap[DIMH][DIMH],
                           double bp[DIMH][DIMH], double cp[DIMH][DIMH])
{
    const double pi = acos(-1.0);
    const double eta = 3.0 * hz * hz - h * h;
    for (int i = 0; i < DIMH; ++i) BD[i] = 0.0;
    for (int ii = 0; ii <= LL + 2; ++ii) {
        double C1 = a[LL][ii], C1p = ap[LL][ii];
        for (int jj = 0; jj <= MM + 2; ++jj) {
            double Ct2 = b[MM][jj], Ct2p = bp[MM][jj];
            for (int kk = 0; kk <= NN + 2; ++kk) {
                double Ct3 = c[NN][kk], Ct3p = cp[NN][kk];
                double C3 = C1 * Ct2 * Ct3 * eta + C1p * Ct2 * Ct3 + C1 * Ct2p * Ct3 +
                            C1 * Ct2 * Ct3p;
                int hOrder = (int)ceil((LL + MM + NN + ii + jj + kk) / 2.0 - 1.0);
                /* hOrder = -1 only when every index is 0, where C3 = 0 */
                if (hOrder >= 0) BD[hOrder] = BD[hOrder] + C3;
            }
        }
    }
    const double f = 0.25 * pow(5.0 / pi, 0.5);
    for (int i = 0; i < DIMH; ++i) BD[i] = f * BD[i];
}

/* Shared per-quartet state. */
typedef struct {
    const int *l, *m, *n, *gs, *gc;
    int maxl, ng;
    const double *ddx, *ddy, *ddz;
} quartet_ctx;

static inline const double *dd_ptr(const quartet_ctx *C, const double *dd,
                                   int gi, int gj, int la, int lb)
{
    int L1 = C->maxl + 1, T1 = 2 * C->maxl + 1;
    return dd + ((((size_t)(gi * C->ng + gj) * L1 + la) * L1 + lb) * T1);
}

/* tot_integral_k_ijkr (mode 0, src/Integrals.f90:5068-5226),
 * integral_k_ijkr (mode 1, :5730-5901), tot_integral_k_ijkr_j2 (mode 2,
 * :5230-5420) and integral_k_ijkr_j2 (mode 3, :5903-6082: mode 2's
 * coefficients with mode 1's MD-coefficient skips).  ztot is supplied as a
 * block.  Returns 0, or 2 when the coefficient arrays would overflow dim = 20
 * (the reference's `stop`). */
static int k_ijkr(const quartet_ctx *C, int gi, int gj, int gk, int gr,
                  double hx, double hy, double hz, double h,
                  const double *ztot, double cutoffz, double cutoffmd,
                  const double *q, int nq, int mode, double *h_saved_out,
                  double *F)
{
    const int elastic = mode == 1 || mode == 3, j2 = mode == 2 || mode == 3;
    const int *l = C->l, *m = C->m, *n = C->n;
    int si = C->gs[gi], sj = C->gs[gj], sk = C->gs[gk], sr = C->gs[gr];
    int LLmax = l[si] + m[si] + n[si] + l[sj] + m[sj] + n[sj] + l[sk] + m[sk] +
                n[sk] + l[sr] + m[sr] + n[sr];
    if (LLmax + 1 + (j2 ? 2 : 0) > DIMH) return 2;
    double a[DIMH][DIMH], b[DIMH][DIMH], c[DIMH][DIMH];
    double ap[DIMH][DIMH], bp[DIMH][DIMH], cp[DIMH][DIMH];
    if (j2) {
        rrdj2a(LLmax, hx, a);
        rrdj2a(LLmax, hy, b);
        rrdj2a(LLmax, hz, c);
        rrdj2b(LLmax, a, ap);
        rrdj2b(LLmax, b, bp);
        rrdj2b(LLmax, c, cp);
        for (int i = 0; i < DIMH; ++i)
            for (int j = 0; j < DIMH; ++j) cp[i][j] = -2.0 * cp[i][j];
    } else if (elastic) {
        hermite_like(LLmax, hx, a);
        hermite_like(LLmax, hy, b);
        hermite_like(LLmax, hz, c);
    } else {
        rrdj0(LLmax, hx, a);
        rrdj0(LLmax, hy, b);
        rrdj0(LLmax, hz, c);
    }
    /* h_pre2 over the simplex L+M+N <= LLmax, stored [L][M][N][beta] */
    int D = LLmax + 1;
    double *hp = (double *)calloc((size_t)D * D * D * DIMH, sizeof(double));
    double BD[DIMH];
    for (int k = 0; k <= LLmax; ++k)
        for (int j = 0; j <= LLmax - k; ++j)
            for (int i = 0; i <= LLmax - k - j; ++i) {
                for (int t = 0; t < DIMH; ++t) BD[t] = 0.0;
                if (j2)
                    bessel_deriv2j(BD, hz, h, k, j, i, a, b, c, ap, bp, cp);
                else
                    bessel_deriv(BD, k, j, i, a, b, c, LLmax);
                memcpy(hp + (((size_t)k * D + j) * D + i) * DIMH, BD,
                       sizeof(double) * DIMH);
            }
    double hs[DIMH];
    for (int t = 0; t < DIMH; ++t) hs[t] = 0.0;
    int cj = C->gc[gj], ck = C->gc[gk], cr = C->gc[gr];
    int posi = 0;
    for (int i = si; i < si + C->gc[gi]; ++i, ++posi) {
        const double *dx1row_unused = 0;
        (void)dx1row_unused;
        int posj = 0;
        for (int j = sj; j < sj + cj; ++j, ++posj) {
            int posk = 0;
            for (int k = sk; k < sk + ck; ++k, ++posk) {
                int posr = 0;
                for (int r = sr; r < sr + cr; ++r, ++posr) {
                    double zt =
                        ztot[(((size_t)posi * cj + posj) * ck + posk) * cr + posr];
                    if (fabs(zt) < cutoffz) continue;
                    const double *Dx1 = dd_ptr(C, C->ddx, gi, gj, l[i], l[j]);
                    const double *Dy1 = dd_ptr(C, C->ddy, gi, gj, m[i], m[j]);
                    const double *Dz1 = dd_ptr(C, C->ddz, gi, gj, n[i], n[j]);
                    const double *Dx2 = dd_ptr(C, C->ddx, gk, gr, l[k], l[r]);
                    const double *Dy2 = dd_ptr(C, C->ddy, gk, gr, m[k], m[r]);
                    const double *Dz2 = dd_ptr(C, C->ddz, gk, gr, n[k], n[r]);
                    for (int ll = 0; ll <= l[i] + l[j]; ++ll) {
                        double MDL = Dx1[ll];
                        if ((elastic || j2) && fabs(MDL) < (double)TINY_SP) continue;
                        double prod1 = MDL * zt;
                        for (int mm = 0; mm <= m[i] + m[j]; ++mm) {
                            double MDM = Dy1[mm];
                            if ((elastic || j2) && fabs(MDM) < (double)TINY_SP) continue;
                            double prod2 = MDM * prod1;
                            for (int nn = 0; nn <= n[i] + n[j]; ++nn) {
                                double H1 = ((ll + mm + nn) & 1) ? -1.0 : 1.0;
                                double MDN = Dz1[nn];
                                if ((elastic || j2) && fabs(MDN) < (double)TINY_SP)
                                    continue;
                                double prod3 = MDN * H1 * prod2;
                                for (int llp = 0; llp <= l[k] + l[r]; ++llp) {
                                    double MDLp = Dx2[llp];
                                    if ((elastic || j2) && fabs(MDLp) < (double)TINY_SP)
                                        continue;
                                    double prod4 = MDLp * prod3;
                                    for (int mmp = 0; mmp <= m[k] + m[r]; ++mmp) {
                                        double MDMp = Dy2[mmp];
                                        if (elastic &&
                                            fabs(MDMp) < (double)TINY_SP)
                                            continue;
                                        double prod5 = MDMp * prod4;
                                        for (int nnp = 0; nnp <= n[k] + n[r];
                                             ++nnp) {
                                            double MDNp = Dz2[nnp];
                                            double prod6 = MDNp * prod5;
                                            if (fabs(prod6) > cutoffmd) {
                                                const double *v =
                                                    hp + (((size_t)(ll + llp) * D +
                                                           (mm + mmp)) *
                                                              D +
                                                          (nn + nnp)) *
                                                             DIMH;
                                                for (int t = 0; t <= LLmax + 2 * j2; ++t)
                                                    hs[t] = hs[t] + v[t] * prod6;
                                            }
                                        }
                                    }
                                }
                            }
                        }
                    }
                }
            }
        }
    }
    free(hp);
    if (h_saved_out) memcpy(h_saved_out, hs, sizeof hs);
    if (elastic && !j2) {
        double s = 0.0;
        for (int t = 0; t < DIMH; ++t) s = s + hs[t];
        if (s == 0.0) {
            for (int iq = 0; iq < nq; ++iq) F[iq] = 0.0;
            return 0;
        }
    }
    if (j2) {
        /* bessels0rr(LLmax+2) then F / mu**2 */
        orc_bessels0rr(F, LLmax + 2, q, nq, h, hs);
        for (int iq = 0; iq < nq; ++iq) F[iq] = F[iq] / pow(q[iq], 2.0);
        return 0;
    }
    orc_bessels0rr(F, LLmax, q, nq, h, hs);
    return 0;
}

/* tot_integral_ijkr_pzero (src/Integrals.f90:4955-5067) and
 * integral_ijkr_pzero (:4751-4863): the h = 0 formula through P0. */
static void ijkr_pzero(const quartet_ctx *C, int gi, int gj, int gk, int gr,
                       const double *ztot, double cutoffz, double cutoffmd,
                       const double *P0, int D0, int nq, double *F)
{
    const int *l = C->l, *m = C->m, *n = C->n;
    int si = C->gs[gi], sj = C->gs[gj], sk = C->gs[gk], sr = C->gs[gr];
    int cj = C->gc[gj], ck = C->gc[gk], cr = C->gc[gr];
    for (int iq = 0; iq < nq; ++iq) F[iq] = 0.0;
    int posi = 0;
    for (int i = si; i < si + C->gc[gi]; ++i, ++posi) {
        int posj = 0;
        for (int j = sj; j < sj + cj; ++j, ++posj) {
            int posk = 0;
            for (int k = sk; k < sk + ck; ++k, ++posk) {
                int posr = 0;
                for (int r = sr; r < sr + cr; ++r, ++posr) {
                    double zt =
                        ztot[(((size_t)posi * cj + posj) * ck + posk) * cr + posr];
                    if (fabs(zt) < cutoffz) continue;
                    const double *Dx1 = dd_ptr(C, C->ddx, gi, gj, l[i], l[j]);
                    const double *Dy1 = dd_ptr(C, C->ddy, gi, gj, m[i], m[j]);
                    const double *Dz1 = dd_ptr(C, C->ddz, gi, gj, n[i], n[j]);
                    const double *Dx2 = dd_ptr(C, C->ddx, gk, gr, l[k], l[r]);
                    const double *Dy2 = dd_ptr(C, C->ddy, gk, gr, m[k], m[r]);
                    const double *Dz2 = dd_ptr(C, C->ddz, gk, gr, n[k], n[r]);
                    for (int ll = 0; ll <= l[i] + l[j]; ++ll) {
                        double MDL = Dx1[ll];
                        if (fabs(MDL) < (double)TINY_SP) continue;
                        double prod1 = MDL * zt;
                        for (int mm = 0; mm <= m[i] + m[j]; ++mm) {
                            double MDM = Dy1[mm];
                            if (fabs(MDM) < (double)TINY_SP) continue;
                            double prod2 = MDM * prod1;
                            for (int nn = 0; nn <= n[i] + n[j]; ++nn) {
                                double H1 = ((ll + mm + nn) & 1) ? -1.0 : 1.0;
                                double MDN = Dz1[nn];
                                if (fabs(MDN) < (double)TINY_SP) continue;
                                double prod3 = MDN * H1 * prod2;
                                for (int llp = 0; llp <= l[k] + l[r]; ++llp) {
                                    double MDLp = Dx2[llp];
                                    if (fabs(MDLp) < (double)TINY_SP) continue;
                                    double prod4 = MDLp * prod3;
                                    for (int mmp = 0; mmp <= m[k] + m[r]; ++mmp) {
                                        double MDMp = Dy2[mmp];
                                        if (fabs(MDMp) < (double)TINY_SP) continue;
                                        double prod5 = MDMp * prod4;
                                        for (int nnp = 0; nnp <= n[k] + n[r];
                                             ++nnp) {
                                            double MDNp = Dz2[nnp];
                                            double prod6 = MDNp * prod5;
                                            if (fabs(prod6) > cutoffmd) {
                                                const double *P =
                                                    P0 +
                                                    ((size_t)(((ll + llp) * D0 +
                                                               (mm + mmp)) *
                                                                  D0 +
                                                              (nn + nnp))) *
                                                        nq;
                                                for (int iq = 0; iq < nq; ++iq)
                                                    F[iq] = F[iq] + prod6 * P[iq];
                                            }
                                        }
                                    }
                                }
                            }
                        }
                    }
                }
            }
        }
    }
}

/* ---------------------------------------------------------------- drivers */
typedef struct {
    int ngto, ng, maxl;
    const int *l, *m, *n;
    int *gs, *gc;
    double *ddx, *ddy, *ddz, *px, *py, *pz, *e12, *P0;
    int D0;
} setup_t;

static int setup_build(setup_t *S, int ngto, int ng, int maxl, const int *l,
                       const int *m, const int *n, const int *group,
                       const double *ga, const double *xx, const double *yy,
                       const double *zz, const double *q, int nq)
{
    S->ngto = ngto;
    S->ng = ng;
    S->maxl = maxl;
    S->l = l;
    S->m = m;
    S->n = n;
    S->gs = (int *)malloc(sizeof(int) * ng);
    S->gc = (int *)calloc(ng, sizeof(int));
    for (int g = 0; g < ng; ++g) S->gs[g] = -1;
    /* group metadata, src/main_calculation.f90:85-93 */
    for (int g = 0; g < ng; ++g)
        for (int j = 0; j < ngto; ++j)
            if (group[j] == g + 1) {
                S->gs[g] = j;
                break;
            }
    for (int j = 0; j < ngto; ++j) S->gc[group[j] - 1]++;
    for (int g = 0; g < ng; ++g) {
        if (S->gs[g] < 0) return 3;
        for (int j = S->gs[g]; j < S->gs[g] + S->gc[g]; ++j)
            if (group[j] != g + 1) return 3; /* groups must be contiguous */
    }
    int maxlv = 0;
    for (int j = 0; j < ngto; ++j) maxlv = l[j] > maxlv ? l[j] : maxlv;
    for (int j = 0; j < ngto; ++j) {
        if (m[j] > maxl || n[j] > maxl || l[j] > maxl) return 4;
        if (l[j] > 4 || m[j] > 4 || n[j] > 4) return 1;
    }
    int L1 = maxl + 1, T1 = 2 * maxl + 1;
    size_t dsz = (size_t)ng * ng * L1 * L1 * T1;
    S->ddx = (double *)malloc(sizeof(double) * dsz);
    S->ddy = (double *)malloc(sizeof(double) * dsz);
    S->ddz = (double *)malloc(sizeof(double) * dsz);
    S->px = (double *)malloc(sizeof(double) * ng * ng);
    S->py = (double *)malloc(sizeof(double) * ng * ng);
    S->pz = (double *)malloc(sizeof(double) * ng * ng);
    S->e12 = (double *)malloc(sizeof(double) * (size_t)ng * ng * nq);
    orc_pair_tables(ngto, ng, maxl, l, m, n, group, ga, xx, yy, zz, q, nq,
                    S->ddx, S->ddy, S->ddz, S->px, S->py, S->pz, S->e12);
    /* set_P0(P0matrix, 4*maxval(l), q), src/main_calculation.f90:104-106 */
    S->D0 = 4 * maxlv + 1;
    S->P0 = (double *)malloc(sizeof(double) * (size_t)S->D0 * S->D0 * S->D0 * nq);
    orc_set_p0(S->P0, 4 * maxlv, q, nq);
    return 0;
}

static void setup_free(setup_t *S)
{
    free(S->gs);
    free(S->gc);
    free(S->ddx);
    free(S->ddy);
    free(S->ddz);
    free(S->px);
    free(S->py);
    free(S->pz);
    free(S->e12);
    free(S->P0);
}

/* z1(n, a, b), z2(n, a, b) of variables_total, src/Variables.f90:1041-1061:
 * z1 = total*(C(m1,a)C(m2,b) + C(m1,b)C(m2,a)), z2 = C(m3,a)C(m4,b) +
 * C(m3,b)C(m4,a).  Stored with n fastest: z[n + nmat*(a + ngto*b)]. */
static void build_z12(const double *mmod, int nmo, int ngto, const int *mat,
                      const double *total, int nmat, double *z1, double *z2)
{
#define MM(i, j) mmod[(size_t)(i) + (size_t)nmo * (size_t)(j)]
#pragma omp parallel for schedule(static)
    for (int b = 0; b < ngto; ++b)
        for (int a = 0; a < ngto; ++a) {
            double *o1 = z1 + (size_t)nmat * (a + (size_t)ngto * b);
            double *o2 = z2 + (size_t)nmat * (a + (size_t)ngto * b);
            for (int k = 0; k < nmat; ++k) {
                int p = mat[k] - 1, s = mat[k + nmat] - 1;
                int qq = mat[k + 2 * nmat] - 1, r = mat[k + 3 * nmat] - 1;
                int m11 = p < s ? p : s, m22 = p < s ? s : p;
                int m33 = qq < r ? qq : r, m44 = qq < r ? r : qq;
                o1[k] = total[k] * (MM(m11, a) * MM(m22, b) + MM(m11, b) * MM(m22, a));
                o2[k] = MM(m33, a) * MM(m44, b) + MM(m33, b) * MM(m44, a);
            }
        }
#undef MM
}

void orc_build_z12(const double *mmod, int nmo, int ngto, const int *mat,
                   const double *total, int nmat, double *z1, double *z2)
{
    build_z12(mmod, nmo, ngto, mat, total, nmat, z1, z2);
}

static int total_j0_impl(int ngto, int ng, int maxl, const int *l, const int *m,
                         const int *n, const int *group, const double *ga,
                         const double *xx, const double *yy, const double *zz,
                         const double *mmod, int nmo, const int *mat,
                         const double *total, int nmat, const double *q, int nq,
                         double cutoffz, double cutoffmd, double cutoffcentre,
                         int zmode, int nthreads, double *tsi, int64_t stride,
                         double *kernel_seconds, int64_t *nsampled, int j2)
{
    setup_t S;
    int rc = setup_build(&S, ngto, ng, maxl, l, m, n, group, ga, xx, yy, zz,
                         q, nq);
    if (rc) return rc;
    /* tot_integration_j2 (src/Integrals.f90:954-1422): the same nests with
     * the j2 kernel and the P0_j2 table */
    if (j2) orc_set_p0_j2(S.P0, S.D0 - 1, q, nq);
    int64_t Q = orc_quartet_list(ng, S.px, S.py, S.pz, cutoffcentre, NULL);
    int32_t *ql = (int32_t *)malloc(sizeof(int32_t) * 6 * (size_t)Q);
    orc_quartet_list(ng, S.px, S.py, S.pz, cutoffcentre, ql);
    /* stride > 1: an LL-stratified sample (every stride-th quartet of the
     * list ordered stably by LLmax) */
    int64_t *sel = NULL, nsel = Q;
    if (stride > 1) {
        int64_t cnt[21] = {0}, off[22] = {0};
        int *LLv = (int *)malloc(sizeof(int) * (size_t)Q);
        for (int64_t e = 0; e < Q; ++e) {
            const int32_t *o = ql + 6 * e;
            int s4 = 0;
            for (int k = 0; k < 4; ++k) {
                int g0 = S.gs[o[k] - 1];
                s4 += l[g0] + m[g0] + n[g0];
            }
            LLv[e] = s4 > 20 ? 20 : s4;
            cnt[LLv[e]]++;
        }
        for (int k = 0; k < 21; ++k) off[k + 1] = off[k] + cnt[k];
        int64_t *ord = (int64_t *)malloc(sizeof(int64_t) * (size_t)Q);
        for (int64_t e = 0; e < Q; ++e) ord[off[LLv[e]]++] = e;
        nsel = (Q + stride - 1) / stride;
        sel = (int64_t *)malloc(sizeof(int64_t) * (size_t)nsel);
        for (int64_t k = 0; k < nsel; ++k) sel[k] = ord[k * stride];
        free(ord);
        free(LLv);
    }
    if (nsampled) *nsampled = nsel;
    double *tker = NULL;

    double *z1 = NULL, *z2 = NULL, *W = NULL;
    int norb = 0;
    for (int k = 0; k < 4 * nmat; ++k) norb = mat[k] > norb ? mat[k] : norb;
    if (zmode == 0) {
        z1 = (double *)malloc(sizeof(double) * (size_t)nmat * ngto * ngto);
        z2 = (double *)malloc(sizeof(double) * (size_t)nmat * ngto * ngto);
        build_z12(mmod, nmo, ngto, mat, total, nmat, z1, z2);
    } else {
        /* MO-pair 2-RDM G[(p,s),(q,r)] and W[(p,s)][c,d] =
         * sum_{qr} G[(p,s),(q,r)] S^{qr}(c,d).  Then
         * sum_n z1(n,b,a) z2(n,c,d) = sum_{ps} S^{ps}(a,b) W[ps][cd]. */
        size_t NP = (size_t)norb * norb;
        double *G = (double *)calloc(NP * NP, sizeof(double));
        for (int k = 0; k < nmat; ++k) {
            int p = mat[k] - 1, s = mat[k + nmat] - 1;
            int qq = mat[k + 2 * nmat] - 1, r = mat[k + 3 * nmat] - 1;
            G[(size_t)(p * norb + s) * NP + (qq * norb + r)] += total[k];
        }
        W = (double *)calloc(NP * (size_t)ngto * ngto, sizeof(double));
#define MM(i, j) mmod[(size_t)(i) + (size_t)nmo * (size_t)(j)]
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
        for (long ps = 0; ps < (long)NP; ++ps) {
            double *Wr = W + (size_t)ps * ngto * ngto;
            for (int qq = 0; qq < norb; ++qq)
                for (int r = 0; r < norb; ++r) {
                    double g = G[(size_t)ps * NP + qq * norb + r];
                    if (g == 0.0) continue;
                    for (int c = 0; c < ngto; ++c)
                        for (int d = 0; d < ngto; ++d)
                            Wr[c * ngto + d] +=
                                g * (MM(qq, c) * MM(r, d) + MM(qq, d) * MM(r, c));
                }
        }
        free(G);
    }

    quartet_ctx C = {l, m, n, S.gs, S.gc, maxl, ng, S.ddx, S.ddy, S.ddz};
    int nth = 1;
#ifdef _OPENMP
    nth = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
    double *part = (double *)calloc((size_t)nth * nq, sizeof(double));
    tker = (double *)calloc((size_t)nth, sizeof(double));
    int err = 0;
#pragma omp parallel num_threads(nth)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *acc = part + (size_t)tid * nq;
        double *F = (double *)malloc(sizeof(double) * nq);
        double *zb = (double *)malloc(sizeof(double) * 15 * 15 * 15 * 15);
#pragma omp for schedule(static)
        for (int64_t e0 = 0; e0 < nsel; ++e0) {
            int64_t e = sel ? sel[e0] : e0;
            const int32_t *o = ql + 6 * e;
            int gi = o[0] - 1, gj = o[1] - 1, gk = o[2] - 1, gr = o[3] - 1;
            int mult = o[4], br = o[5];
            int ci = S.gc[gi], cj = S.gc[gj], ck = S.gc[gk], cr = S.gc[gr];
            /* zmode 0 follows the reference's per-quartet contraction of the
             * z11/z22 arrays, part of each quartet's work in tot_integration,
             * so the kernel timer covers it */
            double t0 = 0.0;
#ifdef _OPENMP
            if (zmode == 0) t0 = omp_get_wtime();
#endif
            /* quartet 2-RDM block: zcontrred + zcontrred2 (the dgemm calls of
             * src/Integrals.f90:109-127 and the four analogues) */
            for (int a = 0; a < ci; ++a)
                for (int b = 0; b < cj; ++b)
                    for (int c = 0; c < ck; ++c)
                        for (int d = 0; d < cr; ++d) {
                            int A = S.gs[gi] + a, B = S.gs[gj] + b;
                            int Cc = S.gs[gk] + c, Dd = S.gs[gr] + d;
                            double v1 = 0.0, v2 = 0.0;
                            if (zmode == 0) {
                                const double *z1ba = z1 + (size_t)nmat * (B + (size_t)ngto * A);
                                const double *z2cd = z2 + (size_t)nmat * (Cc + (size_t)ngto * Dd);
                                const double *z2ba = z2 + (size_t)nmat * (B + (size_t)ngto * A);
                                const double *z1cd = z1 + (size_t)nmat * (Cc + (size_t)ngto * Dd);
                                double s1 = 0.0, s2 = 0.0;
                                for (int k = 0; k < nmat; ++k) s1 += z1ba[k] * z2cd[k];
                                for (int k = 0; k < nmat; ++k) s2 += z2ba[k] * z1cd[k];
                                v1 = 0.125 * s1;
                                v2 = 0.125 * s2;
                            } else {
                                size_t NP = (size_t)norb * norb;
                                double s1 = 0.0, s2 = 0.0;
                                for (int p = 0; p < norb; ++p)
                                    for (int s = 0; s < norb; ++s) {
                                        double Sab = MM(p, A) * MM(s, B) + MM(p, B) * MM(s, A);
                                        double Scd = MM(p, Cc) * MM(s, Dd) + MM(p, Dd) * MM(s, Cc);
                                        const double *Wr = W + (size_t)(p * norb + s) * ngto * ngto;
                                        s1 += Sab * Wr[Cc * ngto + Dd];
                                        s2 += Scd * Wr[A * ngto + B];
                                    }
                                (void)NP;
                                v1 = 0.125 * s1;
                                v2 = 0.125 * s2;
                            }
                            zb[((a * cj + b) * ck + c) * cr + d] = v1 + v2;
                        }
#ifdef _OPENMP
            if (zmode != 0) t0 = omp_get_wtime();
#endif
            if (br) {
                ijkr_pzero(&C, gi, gj, gk, gr, zb, cutoffz, cutoffmd, S.P0, S.D0,
                           nq, F);
            } else {
                double hx = S.px[gk * ng + gr] - S.px[gi * ng + gj];
                double hy = S.py[gk * ng + gr] - S.py[gi * ng + gj];
                double hz = S.pz[gk * ng + gr] - S.pz[gi * ng + gj];
                /* nest 1 (the only mult-8 nest) uses **0.5, the others sqrt */
                double h = mult == 8 ? pow(hx * hx + hy * hy + hz * hz, 0.5)
                                     : sqrt(hx * hx + hy * hy + hz * hz);
                int r2 = k_ijkr(&C, gi, gj, gk, gr, hx, hy, hz, h, zb, cutoffz,
                                cutoffmd, q, nq, j2 ? 2 : 0, NULL, F);
                if (r2) {
#pragma omp atomic write
                    err = r2;
                }
            }
            const double *eij = S.e12 + (size_t)(gi * ng + gj) * nq;
            const double *ekr = S.e12 + (size_t)(gk * ng + gr) * nq;
            double mlt = (double)mult;
            for (int iq = 0; iq < nq; ++iq)
                acc[iq] = acc[iq] + mlt * F[iq] * eij[iq] * ekr[iq];
#ifdef _OPENMP
            tker[tid] += omp_get_wtime() - t0;
#endif
        }
        free(F);
        free(zb);
    }
#undef MM
    if (kernel_seconds) {
        double mx = 0.0;
        for (int t = 0; t < nth; ++t) mx = tker[t] > mx ? tker[t] : mx;
        *kernel_seconds = mx;
    }
    free(tker);
    free(sel);
    for (int iq = 0; iq < nq; ++iq) {
        double s = 0.0;
        for (int t = 0; t < nth; ++t) s += part[(size_t)t * nq + iq];
        tsi[iq] = s;
    }
    free(part);
    free(ql);
    free(z1);
    free(z2);
    free(W);
    setup_free(&S);
    return err;
}

int orc_total_j0(int ngto, int ng, int maxl, const int *l, const int *m,
                 const int *n, const int *group, const double *ga,
                 const double *xx, const double *yy, const double *zz,
                 const double *mmod, int nmo, const int *mat,
                 const double *total, int nmat, const double *q, int nq,
                 double cutoffz, double cutoffmd, double cutoffcentre,
                 int zmode, int nthreads, double *tsi)
{
    return total_j0_impl(ngto, ng, maxl, l, m, n, group, ga, xx, yy, zz, mmod,
                         nmo, mat, total, nmat, q, nq, cutoffz, cutoffmd,
                         cutoffcentre, zmode, nthreads, tsi, 1, NULL, NULL, 0);
}

int orc_total_j2(int ngto, int ng, int maxl, const int *l, const int *m,
                 const int *n, const int *group, const double *ga,
                 const double *xx, const double *yy, const double *zz,
                 const double *mmod, int nmo, const int *mat,
                 const double *total, int nmat, const double *q, int nq,
                 double cutoffz, double cutoffmd, double cutoffcentre,
                 int zmode, int nthreads, double *tsi)
{
    return total_j0_impl(ngto, ng, maxl, l, m, n, group, ga, xx, yy, zz, mmod,
                         nmo, mat, total, nmat, q, nq, cutoffz, cutoffmd,
                         cutoffcentre, zmode, nthreads, tsi, 1, NULL, NULL, 1);
}

int orc_total_j0_sample(int ngto, int ng, int maxl, const int *l,
                        const int *m, const int *n, const int *group,
                        const double *ga, const double *xx, const double *yy,
                        const double *zz, const double *mmod, int nmo,
                        const int *mat, const double *total, int nmat,
                        const double *q, int nq, double cutoffz,
                        double cutoffmd, double cutoffcentre, int zmode,
                        int nthreads, int64_t stride, double *tsi,
                        double *kernel_seconds, int64_t *nsampled)
{
    return total_j0_impl(ngto, ng, maxl, l, m, n, group, ga, xx, yy, zz, mmod,
                         nmo, mat, total, nmat, q, nq, cutoffz, cutoffmd,
                         cutoffcentre, zmode, nthreads, tsi, stride,
                         kernel_seconds, nsampled, 0);
}

/* elastic_integration (j2 = 0) and elastic_integration_j2 (j2 = 1,
 * src/Integrals.f90:1941-2364: its six nests, integral_k_ijkr_j2 and the
 * P0_j2 table set by main_calculation.f90:785). */
static int elastic_impl(int ngto, int ng, int maxl, const int *l, const int *m,
                        const int *n, const int *group, const double *ga,
                        const double *xx, const double *yy, const double *zz,
                        const double *Z, const double *q, int nq, double cutoffz,
                        double cutoffmd, double cutoffcentre, int nthreads,
                        double *tsi, int j2)
{
    setup_t S;
    int rc = setup_build(&S, ngto, ng, maxl, l, m, n, group, ga, xx, yy, zz,
                         q, nq);
    if (rc) return rc;
    if (j2) orc_set_p0_j2(S.P0, S.D0 - 1, q, nq);
    int64_t (*qlist)(int, const double *, const double *, const double *, double,
                     int32_t *) = j2 ? orc_quartet_list_elastic_j2 : orc_quartet_list;
    int64_t Q = qlist(ng, S.px, S.py, S.pz, cutoffcentre, NULL);
    int32_t *ql = (int32_t *)malloc(sizeof(int32_t) * 6 * (size_t)Q);
    qlist(ng, S.px, S.py, S.pz, cutoffcentre, ql);
    quartet_ctx C = {l, m, n, S.gs, S.gc, maxl, ng, S.ddx, S.ddy, S.ddz};
    int nth = 1;
#ifdef _OPENMP
    nth = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
    double *part = (double *)calloc((size_t)nth * nq, sizeof(double));
    int err = 0;
#pragma omp parallel num_threads(nth)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *acc = part + (size_t)tid * nq;
        double *F = (double *)malloc(sizeof(double) * nq);
        double *zb = (double *)malloc(sizeof(double) * 15 * 15 * 15 * 15);
#pragma omp for schedule(static)
        for (int64_t e = 0; e < Q; ++e) {
            const int32_t *o = ql + 6 * e;
            int gi = o[0] - 1, gj = o[1] - 1, gk = o[2] - 1, gr = o[3] - 1;
            int mult = o[4], br = o[5];
            int ci = S.gc[gi], cj = S.gc[gj], ck = S.gc[gk], cr = S.gc[gr];
            /* ztot = Za(posi,posj) * Zb(posk,posr), src/Integrals.f90:5842 */
            for (int a = 0; a < ci; ++a)
                for (int b = 0; b < cj; ++b) {
                    double z1 = Z[(S.gs[gi] + a) + (size_t)ngto * (S.gs[gj] + b)];
                    for (int c = 0; c < ck; ++c)
                        for (int d = 0; d < cr; ++d)
                            zb[((a * cj + b) * ck + c) * cr + d] =
                                Z[(S.gs[gk] + c) + (size_t)ngto * (S.gs[gr] + d)] * z1;
                }
            if (br) {
                ijkr_pzero(&C, gi, gj, gk, gr, zb, cutoffz, cutoffmd, S.P0, S.D0,
                           nq, F);
            } else {
                double hx = S.px[gk * ng + gr] - S.px[gi * ng + gj];
                double hy = S.py[gk * ng + gr] - S.py[gi * ng + gj];
                double hz = S.pz[gk * ng + gr] - S.pz[gi * ng + gj];
                /* the first nest (gi < gj, gk > gi) uses **0.5, the others
                 * sqrt */
                double h = (gi != gj && gk != gi) ? pow(hx * hx + hy * hy + hz * hz, 0.5)
                                                  : sqrt(hx * hx + hy * hy + hz * hz);
                int r2 = k_ijkr(&C, gi, gj, gk, gr, hx, hy, hz, h, zb, cutoffz,
                                cutoffmd, q, nq, j2 ? 3 : 1, NULL, F);
                if (r2) {
#pragma omp atomic write
                    err = r2;
                }
            }
            const double *eij = S.e12 + (size_t)(gi * ng + gj) * nq;
            const double *ekr = S.e12 + (size_t)(gk * ng + gr) * nq;
            double mlt = (double)mult;
            for (int iq = 0; iq < nq; ++iq)
                acc[iq] = acc[iq] + mlt * F[iq] * eij[iq] * ekr[iq];
        }
        free(F);
        free(zb);
    }
    for (int iq = 0; iq < nq; ++iq) {
        double s = 0.0;
        for (int t = 0; t < nth; ++t) s += part[(size_t)t * nq + iq];
        tsi[iq] = s;
    }
    free(part);
    free(ql);
    setup_free(&S);
    return err;
}

int orc_elastic_j0(int ngto, int ng, int maxl, const int *l, const int *m,
                   const int *n, const int *group, const double *ga,
                   const double *xx, const double *yy, const double *zz,
                   const double *Z, const double *q, int nq, double cutoffz,
                   double cutoffmd, double cutoffcentre, int nthreads,
                   double *tsi)
{
    return elastic_impl(ngto, ng, maxl, l, m, n, group, ga, xx, yy, zz, Z, q, nq,
                        cutoffz, cutoffmd, cutoffcentre, nthreads, tsi, 0);
}

int orc_elastic_j2(int ngto, int ng, int maxl, const int *l, const int *m,
                   const int *n, const int *group, const double *ga,
                   const double *xx, const double *yy, const double *zz,
                   const double *Z, const double *q, int nq, double cutoffz,
                   double cutoffmd, double cutoffcentre, int nthreads,
                   double *tsi)
{
    return elastic_impl(ngto, ng, maxl, l, m, n, group, ga, xx, yy, zz, Z, q, nq,
                        cutoffz, cutoffmd, cutoffcentre, nthreads, tsi, 1);
}

/* ------------------------------------------- electron scattering, types 5/6 */
/* BesselSum, src/Integrals.f90:6329-6368: the upward recurrence at every x */
static void bessel_sum(double *hsum, const double *mu, int nq, double H, int LLmax,
                       const double *hs)
{
    for (int iq = 0; iq < nq; ++iq) {
        const double Pmu = H * mu[iq];
        double h0 = sin(Pmu) / Pmu;
        double s = h0 * hs[0];
        if (LLmax == 1) {
            double h1 = (sin(Pmu) / (Pmu * Pmu) - cos(Pmu) / Pmu) * mu[iq] / H;
            s = s + hs[1] * h1;
        } else if (LLmax > 1) {
            const double muOH = mu[iq] / H;
            double h1 = (sin(Pmu) / (Pmu * Pmu) - cos(Pmu) / Pmu) * muOH;
            s = s + hs[1] * h1;
            for (int ra = 2; ra <= LLmax; ++ra) {
                double hr = ((double)(2 * ra - 1) / Pmu * h1 - h0 * muOH) * muOH;
                s = s + hr * hs[ra];
                h0 = h1;
                h1 = hr;
            }
        }
        hsum[iq] = s;
    }
}

/* integral_k_ij (src/Integrals.f90:6192-6326; branch 0) and integral_ij_pzero
 * (:4865-4951; branch 1) for the group pair (gi, gj) with the elastic Z. */
static void k_ij(const quartet_ctx *C, int gi, int gj, int branch, double hx, double hy,
                 double hz, double h, const double *Z, int ngto, double cutoffmd,
                 const double *P0, int D0, const double *q, int nq, double *F)
{
    const int *l = C->l, *m = C->m, *n = C->n;
    const int si = C->gs[gi], sj = C->gs[gj];
    const int LLmax = l[si] + m[si] + n[si] + l[sj] + m[sj] + n[sj];
    const int D = LLmax + 1;
    double a[DIMH][DIMH], b[DIMH][DIMH], c[DIMH][DIMH];
    double *hp = NULL;
    if (!branch) {
        hermite_like(LLmax, hx, a);
        hermite_like(LLmax, hy, b);
        hermite_like(LLmax, hz, c);
        hp = (double *)calloc((size_t)D * D * D * DIMH, sizeof(double));
        double BD[DIMH];
        for (int k = 0; k <= LLmax; ++k)
            for (int j = 0; j <= LLmax - k; ++j)
                for (int i = 0; i <= LLmax - k - j; ++i) {
                    for (int t = 0; t < DIMH; ++t) BD[t] = 0.0;
                    bessel_deriv(BD, i, j, k, a, b, c, LLmax);
                    memcpy(hp + (((size_t)i * D + j) * D + k) * DIMH, BD, sizeof BD);
                }
    }
    double hs[DIMH];
    for (int t = 0; t < DIMH; ++t) hs[t] = 0.0;
    if (branch)
        for (int iq = 0; iq < nq; ++iq) F[iq] = 0.0;
    for (int i = si; i < si + C->gc[gi]; ++i)
        for (int j = sj; j < sj + C->gc[gj]; ++j) {
            const double ztot = Z[i + (size_t)ngto * j];
            const double *Dx1 = dd_ptr(C, C->ddx, gi, gj, l[i], l[j]);
            const double *Dy1 = dd_ptr(C, C->ddy, gi, gj, m[i], m[j]);
            const double *Dz1 = dd_ptr(C, C->ddz, gi, gj, n[i], n[j]);
            for (int ll = 0; ll <= l[i] + l[j]; ++ll) {
                double MDL = Dx1[ll];
                if (fabs(MDL) < (double)TINY_SP) continue;
                double prod1 = MDL * ztot;
                for (int mm = 0; mm <= m[i] + m[j]; ++mm) {
                    double MDM = Dy1[mm];
                    if (fabs(MDM) < (double)TINY_SP) continue;
                    double prod2 = MDM * prod1;
                    for (int nn = 0; nn <= n[i] + n[j]; ++nn) {
                        double H1 = ((ll + mm + nn) & 1) ? -1.0 : 1.0;
                        double MDN = Dz1[nn];
                        if (fabs(MDN) < (double)TINY_SP) continue;
                        double prod3 = MDN * H1 * prod2;
                        if (fabs(prod3) > cutoffmd) {
                            if (branch) {
                                const double *P = P0 + ((size_t)((ll * D0 + mm) * D0 + nn)) * nq;
                                for (int iq = 0; iq < nq; ++iq) F[iq] = F[iq] + prod3 * P[iq];
                            } else {
                                const double *v = hp + (((size_t)ll * D + mm) * D + nn) * DIMH;
                                for (int t = 0; t < DIMH; ++t) hs[t] = hs[t] + v[t] * prod3;
                            }
                        }
                    }
                }
            }
        }
    if (!branch) {
        bessel_sum(F, q, nq, h, LLmax, hs);
        free(hp);
    }
}

int orc_nuclei_electron(int ngto, int ng, int maxl, const int *l, const int *m,
                        const int *n, const int *group, const double *ga,
                        const double *xx, const double *yy, const double *zz,
                        const double *Z, int natom, const double *Zn, const double *geom,
                        const double *q, int nq, double cutoffz, double cutoffmd,
                        double cutoffcentre, double *tsi)
{
    (void)cutoffz;
    setup_t S;
    int rc = setup_build(&S, ngto, ng, maxl, l, m, n, group, ga, xx, yy, zz, q, nq);
    if (rc) return rc;
    quartet_ctx C = {l, m, n, S.gs, S.gc, maxl, ng, S.ddx, S.ddy, S.ddz};
    double *F = (double *)malloc(sizeof(double) * nq);
    for (int iq = 0; iq < nq; ++iq) tsi[iq] = 0.0;
    for (int na = 0; na < natom; ++na)
        for (int i = 0; i < ng; ++i)
            for (int k = 0; k < ng; ++k) {
                double hx = geom[na + (size_t)natom * 0] - S.px[i * ng + k];
                double hy = geom[na + (size_t)natom * 1] - S.py[i * ng + k];
                double hz = geom[na + (size_t)natom * 2] - S.pz[i * ng + k];
                double h = sqrt(hx * hx + hy * hy + hz * hz);
                k_ij(&C, i, k, h < cutoffcentre, hx, hy, hz, h, Z, ngto, cutoffmd, S.P0, S.D0, q,
                     nq, F);
                const double *e = S.e12 + (size_t)(i * ng + k) * nq;
                for (int iq = 0; iq < nq; ++iq) tsi[iq] = tsi[iq] + Zn[na] * F[iq] * e[iq];
            }
    free(F);
    setup_free(&S);
    return 0;
}

/* The nuclear term of types 5/6, src/main_calculation.f90:303-312 (sinc of
 * src/main_calculation.f90:853-862). */
void orc_nuclear_nn(int natom, const double *Zn, const double *geom, const double *q, int nq,
                    double *Inn)
{
    for (int iq = 0; iq < nq; ++iq) Inn[iq] = 0.0;
    for (int i = 0; i < natom; ++i)
        for (int j = 0; j < natom; ++j)
            for (int k = 0; k < nq; ++k) {
                double s2 = 0.0;
                for (int c = 0; c < 3; ++c) {
                    double d = geom[i + (size_t)natom * c] - geom[j + (size_t)natom * c];
                    s2 = s2 + d * d;
                }
                double a = q[k] * fabs(sqrt(s2));
                double sinc = fabs(a) < 1.0e-10 ? 1.0 : sin(a) / a;
                Inn[k] = Inn[k] + Zn[i] * Zn[j] * sinc;
            }
}

/* ------------------------------------------------------------ unique_total */
int orc_unique_total(const int *matrix, int nrow, int ncol, const double *total,
                     int *rmat, double *rtot, int *iuni, int *irec)
{
    int *rec = (int *)calloc(nrow, sizeof(int));
    int *first = (int *)malloc(sizeof(int) * (nrow > 0 ? nrow : 1));
    double *tot = (double *)malloc(sizeof(double) * (nrow > 0 ? nrow : 1));
    int cnt = 0;
    for (int i = 0; i < nrow; ++i) {
        if (rec[i] != 0) continue;
        first[cnt] = i;
        ++cnt;
        double s = 0.0;
        for (int r = 0; r < nrow; ++r) {
            int same = 1;
            for (int c = 0; c < ncol; ++c)
                if (matrix[r + (size_t)nrow * c] != matrix[i + (size_t)nrow * c]) {
                    same = 0;
                    break;
                }
            if (same) {
                s = s + total[r];
                rec[r] += cnt;
            }
        }
        tot[cnt - 1] = s;
    }
    for (int u = 0; u < cnt; ++u) {
        if (iuni) iuni[u] = first[u] + 1;
        if (rtot) rtot[u] = tot[u];
        if (rmat)
            for (int c = 0; c < ncol; ++c)
                rmat[u + (size_t)cnt * c] = matrix[first[u] + (size_t)nrow * c];
    }
    if (irec)
        for (int r = 0; r < nrow; ++r) irec[r] = rec[r];
    free(rec);
    free(first);
    free(tot);
    return cnt;
}

/* ---- bit-string determinant builders (src/one_rdm.f90) ----------------- */

/* Highest orbital of the largest bit string: integer2binary_orbs of
 * maxval(icij) (src/twordms.f90:1036-1061, used at src/one_rdm.f90:985,
 * :1432, :1921). */
static int bit_nbft(const int64_t *a, const int64_t *b, int64_t n)
{
    int64_t mx = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (a[i] > mx) mx = a[i];
        if (b[i] > mx) mx = b[i];
    }
    int nb = 0;
    while (mx >= 1) {
        ++nb;
        mx /= 2;
    }
    return nb;
}

/* Orbitals strictly between lo and hi (1-based), the reference's
 * IAND(ibset(0,hi-1)-1, ibclr(-1,lo)+1) as 64-bit words. */
static uint64_t bit_between(int lo, int hi)
{
    uint64_t below_hi = (((uint64_t)1) << (hi - 1)) - 1;
    uint64_t from_lo = (~(((uint64_t)1) << lo)) + 1;
    return below_hi & from_lo;
}

static int bit_popc(uint64_t x) { return __builtin_popcountll(x); }
static int bit_ctz(uint64_t x) { return __builtin_ctzll(x); }

/* The same number of alpha and of beta electrons in every determinant. */
static int bit_counts_ok(const int64_t *a, const int64_t *b, int64_t n)
{
    for (int64_t i = 1; i < n; ++i)
        if (bit_popc((uint64_t)a[i]) != bit_popc((uint64_t)a[0]) ||
            bit_popc((uint64_t)b[i]) != bit_popc((uint64_t)b[0]))
            return 0;
    return 1;
}

/* The E20.10 write of mcci_to_bit followed by the list-directed read of
 * createtwordm_bit: the nearest double to x rounded to 10 significant
 * decimal digits. */
static double bit_text_round(double x)
{
    char buf[64];
    snprintf(buf, sizeof buf, "%.9E", x);
    return strtod(buf, NULL);
}

int orc_mcci_to_bit(const int64_t *alpha, const int64_t *beta, const double *c1,
                    const double *c2, int64_t n, int text_round, double *c1o,
                    double *c2o)
{
    /* src/one_rdm.f90:1901-1984: the slot list in orbital order (alpha
     * before beta on each orbital; beta slot j+1+nbft) and in alpha-then-beta
     * order; the number of exchanges that turn the second into the first
     * sets the sign of both coefficients. */
    if (n <= 0) return -1;
    const int nbft = bit_nbft(alpha, beta, n);
    const int ne = bit_popc((uint64_t)alpha[0]);
    const int ntotal = 2 * ne;
    int64_t *l1 = malloc(sizeof(int64_t) * (ntotal + 1));
    int64_t *l2 = malloc(sizeof(int64_t) * (ntotal + 1));
    int rc = 0;
    for (int64_t ici = 0; ici < n && rc == 0; ++ici) {
        if (bit_popc((uint64_t)alpha[ici]) != ne || bit_popc((uint64_t)beta[ici]) != ne) {
            rc = -1;
            break;
        }
        int k = 0;
        for (int j = 0; j < nbft; ++j) {
            if ((alpha[ici] >> j) & 1) l1[k++] = j + 1;
            if ((beta[ici] >> j) & 1) l1[k++] = j + 1 + nbft;
        }
        k = 0;
        for (int j = 0; j < nbft; ++j)
            if ((alpha[ici] >> j) & 1) l2[k++] = j + 1;
        for (int j = 0; j < nbft; ++j)
            if ((beta[ici] >> j) & 1) l2[k++] = j + 1 + nbft;
        int iperm = 0;
        for (int j = 0; j < ntotal - 1; ++j) {
            if (l1[j] == l2[j]) continue;
            for (int kk = j + 1; kk < ntotal; ++kk)
                if (l1[j] == l2[kk]) {
                    ++iperm;
                    int64_t t = l2[j];
                    l2[j] = l2[kk];
                    l2[kk] = t;
                    break;
                }
        }
        const double phase = (iperm & 1) ? -1.0 : 1.0;
        double x1 = phase * c1[ici], x2 = phase * c2[ici];
        if (text_round) {
            x1 = bit_text_round(x1);
            x2 = bit_text_round(x2);
        }
        c1o[ici] = x1;
        c2o[ici] = x2;
    }
    free(l1);
    free(l2);
    return rc;
}

int orc_createtwordm_bit(const int64_t *alpha, const int64_t *beta, const double *c1,
                         const double *c2, int64_t n, double *tw, int cap_norb)
{
    /* src/one_rdm.f90:931-1300.  Returns nbft; writes the dense
     * SpinFree2RDM (nbft^4, column-major) when tw != NULL and
     * nbft <= cap_norb.  -1: inconsistent electron counts. */
    if (n <= 0 || !bit_counts_ok(alpha, beta, n)) return -1;
    const int N = bit_nbft(alpha, beta, n);
    if (!tw || N > cap_norb) return N;
    const size_t NN = (size_t)N;
    memset(tw, 0, sizeof(double) * NN * NN * NN * NN);
#define TWB(p, q, r, s) \
    tw[(size_t)((p)-1) + NN * ((size_t)((q)-1) + NN * ((size_t)((r)-1) + NN * (size_t)((s)-1)))]
    const int64_t *ic[2] = {alpha, beta};
    /* no differences (src/one_rdm.f90:1009-1046) */
    for (int64_t i = 0; i < n; ++i) {
        const double c = c1[i] * c2[i];
        for (int sp = 0; sp < 2; ++sp) {
            uint64_t b1 = (uint64_t)ic[sp][i];
            while (b1) {
                const int so = bit_ctz(b1) + 1;
                b1 &= b1 - 1;
                for (int sp2 = 0; sp2 < 2; ++sp2) {
                    uint64_t b2 = (uint64_t)ic[sp2][i];
                    while (b2) {
                        const int qo = bit_ctz(b2) + 1;
                        b2 &= b2 - 1;
                        if (qo == so && sp == sp2) continue;
                        if (sp == sp2) TWB(so, qo, so, qo) = TWB(so, qo, so, qo) - c;
                        TWB(qo, so, so, qo) = TWB(qo, so, so, qo) + c;
                    }
                }
            }
        }
    }
    /* every ordered pair (src/one_rdm.f90:1054-1300) */
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < n; ++j) {
            const int nd = bit_popc((uint64_t)(alpha[i] ^ alpha[j])) +
                           bit_popc((uint64_t)(beta[i] ^ beta[j]));
            if (nd > 4) continue;
            if (nd == 4) {
                int hol[2][2] = {{0}}, par[2][2] = {{0}}, nh[2] = {0, 0};
                int nperm = 0, same = -1;
                for (int sp = 0; sp < 2; ++sp) {
                    const uint64_t x = (uint64_t)(ic[sp][i] ^ ic[sp][j]);
                    uint64_t h = x & (uint64_t)ic[sp][i], pt = x & (uint64_t)ic[sp][j];
                    int np_ = 0;
                    while (pt) {
                        par[sp][np_++] = bit_ctz(pt) + 1;
                        pt &= pt - 1;
                    }
                    while (h) {
                        hol[sp][nh[sp]++] = bit_ctz(h) + 1;
                        h &= h - 1;
                    }
                    for (int k = 0; k < nh[sp]; ++k) {
                        const int lo = hol[sp][k] < par[sp][k] ? hol[sp][k] : par[sp][k];
                        const int hi = hol[sp][k] < par[sp][k] ? par[sp][k] : hol[sp][k];
                        nperm += bit_popc((uint64_t)ic[sp][i] & bit_between(lo, hi));
                    }
                    if (nh[sp] == 2) {
                        same = sp;
                        const int a = hol[sp][0] < par[sp][0] ? hol[sp][0] : par[sp][0];
                        const int b = hol[sp][0] < par[sp][0] ? par[sp][0] : hol[sp][0];
                        const int cc = hol[sp][1] < par[sp][1] ? hol[sp][1] : par[sp][1];
                        const int d = hol[sp][1] < par[sp][1] ? par[sp][1] : hol[sp][1];
                        if (cc > a && cc < b && d > b) ++nperm;
                        break;
                    }
                }
                const double c = c1[i] * c2[j] * ((nperm & 1) ? -1.0 : 1.0);
                if (same < 0) {
                    const int so = par[0][0], qo = par[1][0], po = hol[1][0], ro = hol[0][0];
                    TWB(po, ro, so, qo) = TWB(po, ro, so, qo) + c;
                    TWB(ro, po, qo, so) = TWB(ro, po, qo, so) + c;
                } else {
                    const int so = par[same][0], qo = par[same][1], po = hol[same][1],
                              ro = hol[same][0];
                    TWB(po, ro, so, qo) = TWB(po, ro, so, qo) + c;
                    TWB(ro, po, so, qo) = TWB(ro, po, so, qo) - c;
                    TWB(po, ro, qo, so) = TWB(po, ro, qo, so) - c;
                    TWB(ro, po, qo, so) = TWB(ro, po, qo, so) + c;
                }
                continue;
            }
            if (nd == 2) {
                int sp = 0, tz = 0, tz2 = 0;
                for (sp = 0; sp < 2; ++sp) {
                    const uint64_t x = (uint64_t)(ic[sp][i] ^ ic[sp][j]);
                    const uint64_t h = x & (uint64_t)ic[sp][i], pt = x & (uint64_t)ic[sp][j];
                    if (h) {
                        tz = 1 + bit_ctz(h);
                        tz2 = 1 + bit_ctz(pt);
                        break;
                    }
                }
                const int lo = tz < tz2 ? tz : tz2, hi = tz < tz2 ? tz2 : tz;
                const int nperm = bit_popc((uint64_t)ic[sp][i] & bit_between(lo, hi));
                const double c = c1[i] * c2[j] * ((nperm & 1) ? -1.0 : 1.0);
                const int po = tz, qo = tz2;
                uint64_t b = (uint64_t)ic[sp][i] & ~(((uint64_t)1) << (po - 1));
                while (b) {
                    const int ro = bit_ctz(b) + 1;
                    b &= b - 1;
                    TWB(po, ro, ro, qo) = TWB(po, ro, ro, qo) + c;
                    TWB(ro, po, qo, ro) = TWB(ro, po, qo, ro) + c;
                    TWB(ro, po, ro, qo) = TWB(ro, po, ro, qo) - c;
                    TWB(po, ro, qo, ro) = TWB(po, ro, qo, ro) - c;
                }
                b = (uint64_t)ic[1 - sp][i];
                while (b) {
                    const int ro = bit_ctz(b) + 1;
                    b &= b - 1;
                    TWB(po, ro, ro, qo) = TWB(po, ro, ro, qo) + c;
                    TWB(ro, po, qo, ro) = TWB(ro, po, qo, ro) + c;
                }
            }
        }
#undef TWB
    return N;
}

int orc_one_rdm_bit(const int64_t *alpha, const int64_t *beta, const double *civs,
                    int64_t n, double *g, int cap_norb)
{
    /* src/one_rdm.f90:1388-1516.  Returns maxnmo; writes onerdm (maxnmo^2,
     * column-major) when g != NULL and maxnmo <= cap_norb.  civs is the
     * file's third column (the read at :1428 stores both coefficient
     * columns into civs(i), the second one last). */
    if (n <= 0 || !bit_counts_ok(alpha, beta, n)) return -1;
    const int N = bit_nbft(alpha, beta, n);
    if (!g || N > cap_norb) return N;
    const size_t NN = (size_t)N;
    memset(g, 0, sizeof(double) * NN * NN);
#define G1(p, q) g[(size_t)((p)-1) + NN * (size_t)((q)-1)]
    const int64_t *ic[2] = {alpha, beta};
    for (int64_t i = 0; i < n; ++i)
        for (int sp = 0; sp < 2; ++sp) {
            uint64_t b = (uint64_t)ic[sp][i];
            while (b) {
                const int j = bit_ctz(b);
                G1(j + 1, j + 1) = G1(j + 1, j + 1) + civs[i] * civs[i];
                b &= b - 1;
            }
        }
    for (int64_t i = 0; i + 1 < n; ++i)
        for (int64_t j = i + 1; j < n; ++j) {
            const int d1 = bit_popc((uint64_t)(alpha[i] ^ alpha[j]));
            const int d2 = bit_popc((uint64_t)(beta[i] ^ beta[j]));
            if (d1 + d2 != 2) continue;
            for (int sp = 0; sp < 2; ++sp) {
                if ((sp == 0 ? d1 : d2) != 2) continue;
                const uint64_t x = (uint64_t)(ic[sp][i] ^ ic[sp][j]);
                const int po = bit_ctz((uint64_t)ic[sp][i] & x) + 1;
                const int qo = bit_ctz((uint64_t)ic[sp][j] & x) + 1;
                const int lo = po < qo ? po : qo, hi = po < qo ? qo : po;
                const int ep = bit_popc((uint64_t)ic[sp][i] & bit_between(lo, hi));
                const double ph = (ep & 1) ? -1.0 : 1.0;
                G1(po, qo) = G1(po, qo) + civs[i] * civs[j] * ph;
                G1(qo, po) = G1(qo, po) + civs[i] * civs[j] * ph;
            }
        }
#undef G1
    return N;
}

/* createtwordm_bit_fci, src/one_rdm.f90:390-927, with the tables of single
 * and double excitations of every unique alpha string (:474-657) and the
 * main loop over the file's determinants (:668-875).  uniq holds
 * newdat(:,1) (nuniq strings, increasing), irep/start/end the driver's
 * per-string count and 1-based first / last file position
 * (src/scattering_main_routine.py:269-289).  Returns norbs, or -1 on input
 * the reference cannot process (see the header). */
int orc_createtwordm_bit_fci(const int64_t *alpha, const int64_t *beta, const double *civs,
                             int64_t n, const int64_t *uniq, int64_t nuniq,
                             const int64_t *irep, const int64_t *start, const int64_t *end,
                             double *tw, int cap_norb)
{
    if (n <= 0 || nuniq <= 0) return -1;
    int64_t mx = 0;
    for (int64_t i = 0; i < n; ++i)
        if (alpha[i] > mx) mx = alpha[i];
    int N = 0;
    while (mx >= 1) {
        ++N;
        mx /= 2;
    }
    if (N == 0) return -1;
    if (!tw || N > cap_norb) return N;
    const size_t NN = (size_t)N;
    memset(tw, 0, sizeof(double) * NN * NN * NN * NN);
#define TWF(p, q, r, s) \
    tw[(size_t)((p)-1) + NN * ((size_t)((q)-1) + NN * ((size_t)((r)-1) + NN * (size_t)((s)-1)))]
    const uint64_t comp = (N >= 64) ? ~(uint64_t)0 : ((((uint64_t)1) << N) - 1);
    const int nel = bit_popc((uint64_t)uniq[0]);
    const int nsg = nel * (N - nel) + 1, ndb = nel * (nel - 1) / 2 * (N - nel) * (N - nel - 1) / 2 + 1;
    /* single excitations (:498-559): added orbital j (unoccupied, increasing)
     * outer, removed orbital k (occupied, increasing) inner */
    int64_t *sx = calloc((size_t)nuniq * nsg, sizeof(int64_t));
    int *sh = calloc((size_t)nuniq * nsg, sizeof(int)), *sp = calloc((size_t)nuniq * nsg, sizeof(int));
    int *sph = calloc((size_t)nuniq * nsg, sizeof(int)), *sn = calloc((size_t)nuniq, sizeof(int));
    int64_t *sidx = calloc((size_t)nuniq * nsg, sizeof(int64_t));
    int64_t *dx = calloc((size_t)nuniq * ndb, sizeof(int64_t));
    int *dh1 = calloc((size_t)nuniq * ndb, sizeof(int)), *dh2 = calloc((size_t)nuniq * ndb, sizeof(int));
    int *dp1 = calloc((size_t)nuniq * ndb, sizeof(int)), *dp2 = calloc((size_t)nuniq * ndb, sizeof(int));
    int *dph = calloc((size_t)nuniq * ndb, sizeof(int)), *dn = calloc((size_t)nuniq, sizeof(int));
    int64_t *didx = calloc((size_t)nuniq * ndb, sizeof(int64_t));
    for (int64_t i = 0; i < nuniq; ++i) {
        const uint64_t s0 = (uint64_t)uniq[i];
        int c2 = 0;
        uint64_t b = s0 ^ comp;
        while (b) {
            const int j = bit_ctz(b);
            uint64_t b2 = s0;
            while (b2) {
                const int k = bit_ctz(b2);
                b2 &= b2 - 1;
                const uint64_t t = (s0 | ((uint64_t)1 << j)) & ~((uint64_t)1 << k);
                int64_t idx = -1;
                for (int64_t l = 0; l < nuniq; ++l)
                    if ((uint64_t)uniq[l] == t) {
                        idx = l;
                        break;
                    }
                if (idx < 0 || irep[idx] == 0) continue;
                const int lo = (j + 1) < (k + 1) ? (j + 1) : (k + 1);
                const int hi = (j + 1) < (k + 1) ? (k + 1) : (j + 1);
                sx[i * nsg + c2] = (int64_t)t;
                sh[i * nsg + c2] = k + 1;
                sp[i * nsg + c2] = j + 1;
                sph[i * nsg + c2] = bit_popc(s0 & bit_between(lo, hi));
                sidx[i * nsg + c2] = idx;
                ++c2;
            }
            b &= b - 1;
        }
        sn[i] = c2;
        /* double excitations (:586-657): occ i<j, unocc k<l */
        int occ[64], unocc[64], no = 0, nu = 0;
        for (int k = 0; k < N; ++k) {
            if ((s0 >> k) & 1) occ[no++] = k;
            else unocc[nu++] = k;
        }
        int c3 = 0;
        for (int a = 0; a < no - 1; ++a)
            for (int bb = a + 1; bb < no; ++bb)
                for (int k = 0; k < nu - 1; ++k)
                    for (int l = k + 1; l < nu; ++l) {
                        const uint64_t t = (((s0 & ~((uint64_t)1 << occ[a])) & ~((uint64_t)1 << occ[bb])) |
                                            ((uint64_t)1 << unocc[k])) | ((uint64_t)1 << unocc[l]);
                        int64_t idx = -1;
                        for (int64_t m = 0; m < nuniq; ++m)
                            if ((uint64_t)uniq[m] == t) {
                                idx = m;
                                break;
                            }
                        if (idx < 0 || irep[idx] == 0) continue;
                        const int l1 = occ[a] + 1 < unocc[k] + 1 ? occ[a] + 1 : unocc[k] + 1;
                        const int h1 = occ[a] + 1 < unocc[k] + 1 ? unocc[k] + 1 : occ[a] + 1;
                        const int l2 = occ[bb] + 1 < unocc[l] + 1 ? occ[bb] + 1 : unocc[l] + 1;
                        const int h2 = occ[bb] + 1 < unocc[l] + 1 ? unocc[l] + 1 : occ[bb] + 1;
                        int ep = bit_popc(s0 & bit_between(l1, h1)) + bit_popc(s0 & bit_between(l2, h2));
                        if (l2 > l1 && l2 < h1 && h2 > h1) ++ep;
                        dx[i * ndb + c3] = (int64_t)t;
                        dh1[i * ndb + c3] = occ[a] + 1;
                        dh2[i * ndb + c3] = occ[bb] + 1;
                        dp1[i * ndb + c3] = unocc[k] + 1;
                        dp2[i * ndb + c3] = unocc[l] + 1;
                        dph[i * ndb + c3] = ep;
                        didx[i * ndb + c3] = idx;
                        ++c3;
                    }
        dn[i] = c3;
    }
    /* first file position (0-based) of beta string bs in the block of
     * unique alpha string u, or -1 (findloc over Nalphbet(1,2,ss:ee)) */
#define FINDB(u, bs, out)                                              \
    do {                                                               \
        (out) = -1;                                                    \
        for (int64_t z = start[(u)] - 1; z <= end[(u)] - 1; ++z)       \
            if (beta[z] == (bs)) {                                     \
                (out) = z;                                             \
                break;                                                 \
            }                                                          \
    } while (0)
    for (int64_t c1 = 0; c1 < n; ++c1) {
        const double civ1 = civs[c1];
        const double c = civ1 * civ1;
        const int64_t *ab[2] = {alpha, beta};
        for (int s1 = 0; s1 < 2; ++s1) {
            uint64_t b1 = (uint64_t)ab[s1][c1];
            while (b1) {
                const int so = bit_ctz(b1) + 1;
                b1 &= b1 - 1;
                for (int s2 = 0; s2 < 2; ++s2) {
                    uint64_t b2 = (uint64_t)ab[s2][c1];
                    while (b2) {
                        const int qo = bit_ctz(b2) + 1;
                        b2 &= b2 - 1;
                        if (qo == so && s1 == s2) continue;
                        if (s1 == s2) TWF(so, qo, so, qo) = TWF(so, qo, so, qo) - c;
                        TWF(qo, so, so, qo) = TWF(qo, so, so, qo) + c;
                    }
                }
            }
        }
        int64_t i1 = -1, i2 = -1;
        for (int64_t m = 0; m < nuniq; ++m)
            if (uniq[m] == alpha[c1]) {
                i1 = m;
                break;
            }
        for (int64_t m = 0; m < nuniq; ++m)
            if (uniq[m] == beta[c1]) {
                i2 = m;
                break;
            }
        if (i1 < 0 || i2 < 0) {
            free(sx); free(sh); free(sp); free(sph); free(sn); free(sidx);
            free(dx); free(dh1); free(dh2); free(dp1); free(dp2); free(dph); free(dn); free(didx);
            return -1;
        }
        for (int i = 0; i < sn[i1]; ++i) {
            const int64_t u = sidx[i1 * nsg + i];
            int64_t idx;
            FINDB(u, beta[c1], idx);
            if (idx >= 0) {
                const double civ2 = civs[idx];
                const int po = sh[i1 * nsg + i], qo = sp[i1 * nsg + i], ep = sph[i1 * nsg + i];
                const double ct = (((ep & 1) ? -1.0 : 1.0) * 2.0 * civ1) * civ2;
                uint64_t b = (uint64_t)alpha[c1] & ~((uint64_t)1 << (po - 1));
                while (b) {
                    const int ro = bit_ctz(b) + 1;
                    b &= b - 1;
                    TWF(po, ro, ro, qo) = TWF(po, ro, ro, qo) + ct;
                    TWF(ro, po, qo, ro) = TWF(ro, po, qo, ro) + ct;
                    TWF(ro, po, ro, qo) = TWF(ro, po, ro, qo) - ct;
                    TWF(po, ro, qo, ro) = TWF(po, ro, qo, ro) - ct;
                }
                b = (uint64_t)beta[c1];
                while (b) {
                    const int ro = bit_ctz(b) + 1;
                    b &= b - 1;
                    TWF(po, ro, ro, qo) = TWF(po, ro, ro, qo) + ct;
                    TWF(ro, po, qo, ro) = TWF(ro, po, qo, ro) + ct;
                }
            }
            /* alpha-beta doubles: the single excitations of the beta string
             * (taken from the alpha-string table) in the block of this alpha
             * single (:783-812) */
            for (int j = 0; j < sn[i2]; ++j) {
                int64_t id2;
                FINDB(u, sx[i2 * nsg + j], id2);
                if (id2 < 0) continue;
                const double civ2 = civs[id2];
                const int ro = sh[i1 * nsg + i], so = sp[i1 * nsg + i];
                const int po = sh[i2 * nsg + j], qo = sp[i2 * nsg + j];
                const int ep = sph[i1 * nsg + i] + sph[i2 * nsg + j];
                const double ct = ((1.0 * ((ep & 1) ? -1.0 : 1.0)) * civ1) * civ2;
                TWF(po, ro, so, qo) = TWF(po, ro, so, qo) + ct;
                TWF(ro, po, qo, so) = TWF(ro, po, qo, so) + ct;
            }
        }
        for (int i = 0; i < dn[i1]; ++i) {
            int64_t idx;
            FINDB(didx[i1 * ndb + i], beta[c1], idx);
            if (idx < 0) continue;
            const double civ2 = civs[idx];
            const int ep = dph[i1 * ndb + i];
            const int so = dp1[i1 * ndb + i], qo = dp2[i1 * ndb + i];
            const int po = dh2[i1 * ndb + i], ro = dh1[i1 * ndb + i];
            const double ct = (((ep & 1) ? -1.0 : 1.0) * 2.0 * civ1) * civ2;
            TWF(po, ro, so, qo) = TWF(po, ro, so, qo) + ct;
            TWF(ro, po, so, qo) = TWF(ro, po, so, qo) - ct;
            TWF(po, ro, qo, so) = TWF(po, ro, qo, so) - ct;
            TWF(ro, po, qo, so) = TWF(ro, po, qo, so) + ct;
        }
    }
#undef FINDB
#undef TWF
    free(sx); free(sh); free(sp); free(sph); free(sn); free(sidx);
    free(dx); free(dh1); free(dh2); free(dp1); free(dp2); free(dph); free(dn); free(didx);
    return N;
}
