/*
 * CPU oracle for the THISTLE total / elastic j0 scattering path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * Fortran algorithm (KirranderGroup/thistle, src/Integrals.f90,
 * src/Variables.f90, src/twordms.f90, src/one_rdm.f90,
 * src/unique_module.f90).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker.  The
 * product path (thistle_amd/) never calls it.
 *
 * Array conventions (all match the Fortran dummies of the reference):
 *   - 2-D inputs are column-major: mmod(nmo,ngto) at mmod[i + nmo*j],
 *     confs(nconf,lconfs) at confs[c + nconf*j], civs(nconf,nstates).
 *   - group[] is 1-based and each group is contiguous.
 *   - MD tables dd[((((gi*ng+gj)*(L1)+la)*L1+lb)*T1 + t] with L1 = maxl+1,
 *     T1 = 2*maxl+1: Hermite coefficient E_t^{la,lb} of the pair whose first
 *     GTO lies in group gi (0-based) and second in gj (reference Ddx(t+1,
 *     la+1, lb+1, gj+1, gi+1)).
 *   - e12[(gi*ng+gj)*nq + iq], px/py/pz[gi*ng+gj].
 *   - P0[((L*D+M)*D+N)*nq + iq], D = 4*maxl+1.
 *
 * Parity pins for each function: tests/test_oracle_reference.py (against the
 * flang build of the reference) and tests/test_oracle_math.py (closed forms,
 * quadrature, brute force, the reference's own test vectors and data files).
 */
#ifndef THISTLE_ORACLE_H
#define THISTLE_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* set_P0 / P_LMN, src/Integrals.f90:3001-3031,3102-3135 */
void orc_set_p0(double *P0, int lmax4, const double *q, int nq);

/* fill_md_table + variables_total pair data, src/Variables.f90:8-946,
 * 1113-1150.  Outputs ddx/ddy/ddz, px/py/pz, e12. */
void orc_pair_tables(int ngto, int ng, int maxl, const int *l, const int *m,
                     const int *n, const int *group, const double *ga,
                     const double *xx, const double *yy, const double *zz,
                     const double *q, int nq, double *ddx, double *ddy,
                     double *ddz, double *px, double *py, double *pz,
                     double *e12);

/* maxcoincidence + createtwordm, src/twordms.f90:8-81,83-172,175-556.
 * Writes the dense twordm(p,q,r,s) (column-major, norb^4). */
void orc_twordm_dense(const int *confs, int nconf, int lconfs,
                      const double *civs, int nstates, int state1, int state2,
                      double *twordm);
/* Selection loop of createtwordm (src/twordms.f90:522-550): rows
 * (p,s,q,r) with value twordm(p,q,r,s), |value| >= cutoff, in p,q,r,s loop
 * order.  mat is (count,4) column-major (1-based orbital labels).  Returns
 * the count; pass mat=NULL/total=NULL to count only. */
int orc_twordm_select(const double *twordm, int norb, double cutoff,
                      int *mat, double *total, int cap);

/* maxcoincidence ep2 / ndiff for one ordered pair (src/twordms.f90:8-81). */
void orc_pair_phase(const int *confs, int nconf, int lconfs, int c1, int c2,
                    int *ep2, int *ndiff);

/* onerdm_creat / createonerdm, src/one_rdm.f90:27-46,109-222. */
void orc_onerdm(const int *confs, int nconf, int lconfs, const double *civs,
                int nstates, int state1, int state2, double *onerdm);

/* variables_elastic Z, src/Variables.f90:1199-1216.  Z(ngto,ngto) col-major. */
void orc_elastic_z(const double *onerdm, int norb, const double *mmod, int nmo,
                   int ngto, double *Z);

/* Reference quartet enumeration (tot_integration nests,
 * src/Integrals.f90:79-82,178-180,258-260,340-341,416).  Returns Q_ref; when
 * out != NULL writes 6 int32 per quartet: gi,gj,gk,gr (1-based), mult, branch
 * (1 = h < cutoffcentre -> pzero). */
int64_t orc_quartet_list(int ng, const double *px, const double *py,
                         const double *pz, double cutoffcentre, int32_t *out);
/* The six nests of elastic_integration_j2 (src/Integrals.f90:2004-2325):
 * (i,j,k,r) 8, (i,j,i,r) r>j 8, (i,j,i,j) 4, (i,i,k,r) 4, (i,i,k,k) 2,
 * (i,i,i,i) 1; same output format. */
int64_t orc_quartet_list_elastic_j2(int ng, const double *px, const double *py,
                                    const double *pz, double cutoffcentre,
                                    int32_t *out);

/* Full drivers.  type 1 = total j0 (tot_integration), type 2 = elastic j0
 * (elastic_integration).
 *   zmode 0: per-quartet contraction restating the dgemm calls (needs the
 *            nmat x ngto^2 z1/z2 arrays; small systems only);
 *   zmode 1: primitive-pair 2-RDM assembled once from the MO-pair matrix.
 * For type 1 pass mat(nmat,4) col-major + total; for type 2 pass Z.
 * tsi receives S(q).  nthreads <= 0 uses the OpenMP default. */
int orc_total_j0(int ngto, int ng, int maxl, const int *l, const int *m,
                 const int *n, const int *group, const double *ga,
                 const double *xx, const double *yy, const double *zz,
                 const double *mmod, int nmo, const int *mat,
                 const double *total, int nmat, const double *q, int nq,
                 double cutoffz, double cutoffmd, double cutoffcentre,
                 int zmode, int nthreads, double *tsi);
/* LL-stratified sample of the total-j0 quartets (every stride-th quartet
 * of the list ordered stably by LLmax); kernel_seconds = per-thread maximum
 * of the time spent in the per-quartet kernels (tot_integral_k_ijkr /
 * _pzero, bessels0rr and the e12 accumulation) with the 2-RDM blocks
 * supplied; tsi holds the sample's partial S(q).  Used for the CPU
 * baseline. */
/* tot_integration_j2 (src/Integrals.f90:954-1422), type 7 of
 * total_scattering_calculation; same arguments as orc_total_j0. */
int orc_total_j2(int ngto, int ng, int maxl, const int *l, const int *m,
                 const int *n, const int *group, const double *ga,
                 const double *xx, const double *yy, const double *zz,
                 const double *mmod, int nmo, const int *mat,
                 const double *total, int nmat, const double *q, int nq,
                 double cutoffz, double cutoffmd, double cutoffcentre,
                 int zmode, int nthreads, double *tsi);
/* set_P0_j2 (src/Integrals.f90:3034-3064), same layout as orc_set_p0 */
void orc_set_p0_j2(double *P0, int lmax4, const double *q, int nq);

int orc_total_j0_sample(int ngto, int ng, int maxl, const int *l,
                        const int *m, const int *n, const int *group,
                        const double *ga, const double *xx, const double *yy,
                        const double *zz, const double *mmod, int nmo,
                        const int *mat, const double *total, int nmat,
                        const double *q, int nq, double cutoffz,
                        double cutoffmd, double cutoffcentre, int zmode,
                        int nthreads, int64_t stride, double *tsi,
                        double *kernel_seconds, int64_t *nsampled);
int orc_elastic_j0(int ngto, int ng, int maxl, const int *l, const int *m,
                   const int *n, const int *group, const double *ga,
                   const double *xx, const double *yy, const double *zz,
                   const double *Z, const double *q, int nq, double cutoffz,
                   double cutoffmd, double cutoffcentre, int nthreads,
                   double *tsi);
/* nuclei_electron_integration (src/Integrals.f90:2493-2596): the
 * electron-nucleus term of types 5/6 with the elastic Z; geom is natom x 3
 * column-major (bohr). */
int orc_nuclei_electron(int ngto, int ng, int maxl, const int *l, const int *m,
                        const int *n, const int *group, const double *ga,
                        const double *xx, const double *yy, const double *zz,
                        const double *Z, int natom, const double *Zn, const double *geom,
                        const double *q, int nq, double cutoffz, double cutoffmd,
                        double cutoffcentre, double *tsi);
/* Sum_ij Zn_i Zn_j sinc(q R_ij) (src/main_calculation.f90:303-312). */
void orc_nuclear_nn(int natom, const double *Zn, const double *geom, const double *q, int nq,
                    double *Inn);
/* elastic_integration_j2 (type 8, src/Integrals.f90:1941-2364) with the
 * P0_j2 table; arguments as orc_elastic_j0. */
int orc_elastic_j2(int ngto, int ng, int maxl, const int *l, const int *m,
                   const int *n, const int *group, const double *ga,
                   const double *xx, const double *yy, const double *zz,
                   const double *Z, const double *q, int nq, double cutoffz,
                   double cutoffmd, double cutoffcentre, int nthreads,
                   double *tsi);

/* z11/z22 of variables_total (src/Variables.f90:1041-1061) for 2-RDM rows
 * mat(nmat,4) with values total: z1(k,a,b) = total_k (C_{p a} C_{s b} +
 * C_{p b} C_{s a}), z2(k,a,b) = C_{q a} C_{r b} + C_{q b} C_{r a}, (p,s) and
 * (q,r) each in increasing order; both (nmat, ngto, ngto) column-major. */
void orc_build_z12(const double *mmod, int nmo, int ngto, const int *mat,
                   const double *total, int nmat, double *z1, double *z2);

/* dphrec, src/Integrals.f90:7147-7184: rbes(leps1,nq) col-major. */
void orc_dphrec(const double *x, double *rbes, int leps1, int lmax1, int nq);
/* bessels0rr, src/Integrals.f90:6662-6739. */
void orc_bessels0rr(double *hsum, int order, const double *mu, int nq,
                    double H, const double *h_saved);

/* unique_total, src/unique_module.f90:90-145 (+ iuni/irec of
 * src/UniqueModule.jl:156-168).  matrix(nrow,ncol) col-major.  Returns the
 * unique count; rmat(cnt,ncol) col-major, rtot(cnt), iuni(cnt) and
 * irec(nrow) 1-based. */
int orc_unique_total(const int *matrix, int nrow, int ncol, const double *total,
                     int *rmat, double *rtot, int *iuni, int *irec);

/* mcci_to_bit, src/one_rdm.f90:1901-1984: each coefficient times the sign
 * of the alpha-then-beta reordering of its slots; text_round != 0 also
 * applies the E20.10 write and list-directed read that hand the result to
 * createtwordm_bit.  Returns -1 when a determinant's alpha or beta count
 * differs from the first alpha count (the reference's slot lists assume
 * it). */
int orc_mcci_to_bit(const int64_t *alpha, const int64_t *beta, const double *c1,
                    const double *c2, int64_t n, int text_round, double *c1o,
                    double *c2o);
/* createtwordm_bit, src/one_rdm.f90:931-1300: dense SpinFree2RDM(p,q,r,s)
 * (nbft^4, column-major) from alpha/beta bit strings (orbital k on bit
 * k-1).  Returns nbft (writes only when tw != NULL and nbft <= cap_norb),
 * or -1 when the alpha or beta counts vary.  Select rows with
 * orc_twordm_select (the loop of src/one_rdm.f90:1312-1348 is the same). */
int orc_createtwordm_bit(const int64_t *alpha, const int64_t *beta, const double *c1,
                         const double *c2, int64_t n, double *tw, int cap_norb);
/* one_rdm_bit, src/one_rdm.f90:1388-1516: onerdm (maxnmo^2, column-major);
 * civs is the third column of the bit file.  Same return convention. */
int orc_one_rdm_bit(const int64_t *alpha, const int64_t *beta, const double *civs,
                    int64_t n, double *g, int cap_norb);
/* createtwordm_bit_fci, src/one_rdm.f90:390-927: dense twordm (norbs^4,
 * norbs the highest alpha orbital) from one coefficient column civs and the
 * driver's unique-alpha tables (src/scattering_main_routine.py:269-289):
 * uniq = newdat(:,1) (nuniq strings), irep (determinants per string), start
 * / end (1-based first and last file position).  The lookups follow the
 * reference: excited strings are searched in uniq, determinants in the
 * string's [start, end] block by beta string.  Returns norbs (writes when
 * tw != NULL and norbs <= cap_norb), or -1 when a determinant's alpha or
 * beta string is not in uniq. */
int orc_createtwordm_bit_fci(const int64_t *alpha, const int64_t *beta, const double *civs,
                             int64_t n, const int64_t *uniq, int64_t nuniq,
                             const int64_t *irep, const int64_t *start, const int64_t *end,
                             double *tw, int cap_norb);

#ifdef __cplusplus
}
#endif
#endif
