/*
 * thistle_scatter.h -- C ABI of the MI355X (gfx950) total / elastic j0
 * X-ray scattering path.
 *
 * The library replaces the Fortran module procedure
 *   main_calculation_mod::total_scattering_calculation
 * (reference src/main_calculation.f90:13-152, types 1 and 2) and the
 * routines it calls: maxcoincidence/createtwordm (src/twordms.f90:8-556),
 * onerdm_creat (src/one_rdm.f90:27-222), read_twordm's consumer
 * total_scattering_calculation_2 readrdm route (src/main_calculation.f90:
 * 582-658), variables_total / variables_elastic (src/Variables.f90:8-1281),
 * tot_integration / elastic_integration and their kernels
 * (src/Integrals.f90:8-472,1575-1939,4708-5226,5730-5901,6662-6739,
 * 7037-7184), and unique_total (src/unique_module.f90:90-145).
 *
 * Conventions (those of the Fortran dummies):
 *   - 2-D arrays are column-major: mmod(nmo,ngto) at mmod[i + nmo*j],
 *     confs(nconf,lconfs), civecs(nconf,nstates), geom(natom,3),
 *     mat(nmat,4).
 *   - group[] and every orbital label in mat are 1-based; each group is a
 *     contiguous run of GTOs.
 *   - All pointers are host pointers unless a name says _dev.
 *   - Return value: THISTLE_OK (0) or one of the status codes below; the
 *     library never exits the process (the reference's `stop`s become
 *     THISTLE_ELMAX / THISTLE_ELLMAX).
 *   - Calls are thread-safe for distinct plans; one plan must not be
 *     executed from two threads at once.
 */
#ifndef THISTLE_SCATTER_H
#define THISTLE_SCATTER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    THISTLE_OK = 0,
    THISTLE_EINVAL = 1,      /* bad sizes / arguments                      */
    THISTLE_ELMAX = 2,       /* a GTO with l, m or n > 4 (fill_md_table stop,
                                src/Variables.f90:681-682,939-940)          */
    THISTLE_ELLMAX = 3,      /* quartet LLmax + 1 > 20
                                (src/Integrals.f90:5117-5120)               */
    THISTLE_EGROUP = 4,      /* a group is not one contiguous run of GTOs of
                                one shell (same l+m+n, at most 15 GTOs)     */
    THISTLE_EDEVICE = 5,     /* HIP or rocBLAS failure                       */
    THISTLE_ENOMEM = 6,      /* device allocation failed                     */
    THISTLE_ETYPE = 7,       /* calculation type other than 1 (total j0),
                                2 (elastic j0), 5 / 6 (total / elastic
                                electron), 7 (total j2) or 8 (elastic j2)    */
    THISTLE_ECUTOFF = 8      /* cutoffmd above THISTLE_CUTOFFMD_MAX         */
};

/* The reference drops every McMurchie-Davidson product with
 * |prod6| <= cutoffmd (src/Integrals.f90:5198-5203).  The device contraction
 * sums the MD products factorised per axis and keeps them all, so it equals
 * the reference only where the dropped products are negligible.  Measured
 * with the oracle (the reference's loop, screen included) on an s/p/d/f
 * three-atom case over q up to 14: cutoffmd 1e-20 -> 3.6e-14 per q,
 * 1e-16 -> 3.6e-13, 1e-12 -> 7.3e-4 (the screen now removes real terms).
 * Calls with cutoffmd above this bound return THISTLE_ECUTOFF instead of a
 * silently different S(q); the reference's default and suggested value is
 * 1e-20. */
#define THISTLE_CUTOFFMD_MAX 1e-18

enum { THISTLE_FP64 = 0, THISTLE_MIXED = 1 };

/* Basis, orbitals, q grid and cutoffs: the non-CI arguments of
 * total_scattering_calculation (src/main_calculation.f90:13-62). */
typedef struct ThistleBasis {
    int32_t ngto, ng, maxl;       /* maxl = maxval(l), sizes the MD tables */
    const double *ga;             /* [ngto] exponents                      */
    const int32_t *l, *m, *n;     /* [ngto] cartesian exponents            */
    const double *xx, *yy, *zz;   /* [ngto] centres, bohr                  */
    const int32_t *group;         /* [ngto] 1-based primitive-shell label  */
    int32_t nmo;                  /* rows of mmod                          */
    const double *mmod;           /* (nmo, ngto) MO coefficients           */
    int32_t nq;
    const double *q;              /* [nq] |q|, a.u.: finite and >= 0, else
                                     THISTLE_EINVAL                        */
    double cutoffz, cutoffmd, cutoffcentre;  /* finite, >= 0; cutoffmd <=
                                     THISTLE_CUTOFFMD_MAX                  */
} ThistleBasis;

/* The 2-RDM builders select |Gamma| >= cutoff with cutoff = 1E-30, a
 * default-real literal (src/twordms.f90:514, src/one_rdm.f90:1312): the
 * double nearest the float nearest 1e-30. */
#define THISTLE_RDM_CUTOFF 1.0000000031710769e-30

typedef struct ThistleOpts {
    int32_t precision;            /* THISTLE_FP64 or THISTLE_MIXED        */
    double rdm_cutoff;            /* |Gamma| selection threshold of
                                     createtwordm; <= 0 selects
                                     THISTLE_RDM_CUTOFF */
    void *stream;                 /* hipStream_t, or NULL for the library's */
    int32_t device;               /* HIP device ordinal                    */
    int32_t waves;                /* persistent waves for the quartet
                                     kernel; 0 = library default           */
} ThistleOpts;

/* Drop-in for main_calculation_mod::total_scattering_calculation, types 1
 * (total j0), 2 (elastic j0), 5 (total electron), 6 (elastic electron), 7
 * (total j2) and 8 (elastic j2): src/main_calculation.f90:13-15,97-152,
 * 270-330,386-404,780-816.  Zn (natom) and geom (natom x 3 column-major,
 * bohr) are used by types 5 and 6 (Iee - 2 Ine + Inn); state3 is accepted
 * for signature parity and unused.  q_abs_out receives zeros (:97-98).
 * Types 7 and 8 support shells up to f (their series runs to
 * LLmax + 2 <= 16), else THISTLE_ELLMAX. */
int32_t thistle_total_scattering_calculation(
    int32_t type, int32_t natom, const double *Zn, const double *geom,
    int32_t state1, int32_t state2, int32_t state3, const ThistleBasis *basis,
    int32_t nconf, int32_t lconfs, int32_t nstates, const int32_t *confs,
    const double *civecs, const ThistleOpts *opts, double *q_abs_out,
    double *result_out);

/* readrdm route of total_scattering_calculation_2 (src/main_calculation.f90:
 * 582-658 after read_twordm, src/one_rdm.f90:2112-2146): total j0 from a
 * 2-RDM given as rows (p,s,q,r) with values. */
int32_t thistle_total_from_twordm(const ThistleBasis *basis, int64_t nmat,
                                  const int32_t *mat, const double *total,
                                  const ThistleOpts *opts, double *result_out);

/* Adjoint of thistle_total_from_twordm.  No counterpart in the reference,
 * which has no derivatives.  Given dL_dS[nq] = dL/dS(q), returns S(q)
 * (S_out, optional), dL/d(nuclear coordinates) (dL_dR, (natom,3)
 * column-major like geom, optional) and dL/d(2-RDM values) (dL_dtotal[nmat],
 * optional).  Every shell centre must coincide with a row of geom; the MO
 * coefficients are held fixed.  The forward is recomputed; Hermite tables
 * are rebuilt with their derivatives rather than stored per quartet. */
int32_t thistle_total_scattering_adjoint(const ThistleBasis *basis,
                                         int64_t nmat, const int32_t *mat,
                                         const double *total, int32_t natom,
                                         const double *geom,
                                         const double *dL_dS,
                                         const ThistleOpts *opts,
                                         double *S_out, double *dL_dR,
                                         double *dL_dtotal);

/* nuclei_electron_integration (src/Integrals.f90:2493-2596): the
 * electron-nucleus term Ine of electron scattering (types 5/6) from a 1-RDM,
 * with nuclear charges Zn (natom) at geom (natom x 3 column-major, bohr).
 * BesselSum's upward recurrence is followed as the reference does; at q = 0
 * the q -> 0 limit is returned where the reference divides 0 by 0. */
int32_t thistle_nuclei_electron(const ThistleBasis *basis, int32_t norb,
                                const double *onerdm, int32_t natom,
                                const double *Zn, const double *geom,
                                const ThistleOpts *opts, double *result_out);

/* Elastic j0 from a given 1-RDM (norb x norb, column-major): the body of
 * case 2 after onerdm_creat (src/main_calculation.f90:137-152). */
int32_t thistle_elastic_from_onerdm(const ThistleBasis *basis, int32_t norb,
                                    const double *onerdm,
                                    const ThistleOpts *opts, double *result_out);

/* ---- 2-RDM / 1-RDM construction (host, reference accumulation order) --- */

/* maxcoincidence + createtwordm accumulation (src/twordms.f90:8-512):
 * dense twordm(p,q,r,s), norb = lconfs/2, column-major norb^4. */
int32_t thistle_twordm_dense(const int32_t *confs, int32_t nconf,
                             int32_t lconfs, const double *civecs,
                             int32_t nstates, int32_t state1, int32_t state2,
                             double *twordm);
/* Selection loop of createtwordm (src/twordms.f90:515-550): rows (p,s,q,r)
 * with value twordm(p,q,r,s), |value| >= cutoff, p,q,r,s loop order.
 * mat is (cap,4) column-major.  Returns the count (pass cap=0 to count). */
int64_t thistle_twordm_select(const double *twordm, int32_t norb,
                              double cutoff, int32_t *mat, double *total,
                              int64_t cap);
/* maxcoincidence + createtwordm + its selection loop on the GPU
 * (src/twordms.f90:8-556): rows (p,s,q,r) and values twordm(p,q,r,s) with
 * |value| >= cutoff, in the reference's p,q,r,s order and with its
 * floating-point values (each entry sums its contributions in the
 * reference's (c1, c2, statement) order).  mat is (cap,4) column-major.
 * Returns the row count, or -status; rows are written only when
 * cap >= count (pass cap = 0 to count).  Determinant lists that are not in
 * the punch reader's canonical form (1 on even slots, 2 on odd) are built by
 * the host accumulation of thistle_twordm_dense. */
int64_t thistle_createtwordm(const int32_t *confs, int32_t nconf, int32_t lconfs,
                             const double *civecs, int32_t nstates, int32_t state1,
                             int32_t state2, double cutoff, const ThistleOpts *opts,
                             int32_t *mat, double *total, int64_t cap);
/* onerdm_creat / createonerdm (src/one_rdm.f90:27-222), host accumulation:
 * onerdm is norb x norb column-major, norb = lconfs/2. */
int32_t thistle_onerdm(const int32_t *confs, int32_t nconf, int32_t lconfs,
                       const double *civecs, int32_t nstates, int32_t state1,
                       int32_t state2, double *onerdm);
/* The same 1-RDM built on the GPU (maxcoincidence + createonerdm,
 * src/one_rdm.f90:27-222): each element sums its contributions in the
 * reference's (c1, c2, slot) order, so the doubles equal thistle_onerdm's.
 * Determinant lists not in the punch reader's canonical form go through the
 * host accumulation.  The drop-in's elastic types use this entry's path. */
int32_t thistle_createonerdm(const int32_t *confs, int32_t nconf, int32_t lconfs,
                             const double *civecs, int32_t nstates, int32_t state1,
                             int32_t state2, const ThistleOpts *opts, double *onerdm);

/* ---- bit-string determinants (the file route) --------------------------
 * A determinant is two non-negative 64-bit bit strings, alpha and beta,
 * orbital k on bit k-1 (the reference's integers of the "i c1 c2 alpha
 * beta" files; its 32-bit reads hold orbitals 1-31).  Every determinant must
 * carry the same alpha count and the same beta count (THISTLE_EINVAL
 * otherwise: the reference's excitation bookkeeping assumes it).  These
 * entries build on the GPU with the reference's floating-point results. */

/* mcci_to_bit (src/one_rdm.f90:1901-1984): c*_out = sign * c*, the sign of
 * reordering the determinant's slots from orbital order (alpha before beta
 * on each orbital) to alpha-then-beta order.  text_round != 0 also applies
 * the E20.10 write and list-directed read that hand the coefficients to
 * createtwordm_bit (10 significant digits).  Needs equal alpha and beta
 * counts. */
int32_t thistle_mcci_to_bit(const int64_t *alpha, const int64_t *beta,
                            const double *c1, const double *c2, int64_t n,
                            int32_t text_round, const ThistleOpts *opts,
                            double *c1_out, double *c2_out);
/* createtwordm_bit (src/one_rdm.f90:931-1353): rows (p,s,q,r) and values
 * SpinFree2RDM(p,q,r,s) with |value| >= cutoff (<= 0: THISTLE_RDM_CUTOFF,
 * the reference's), p,q,r,s loop order, nbft = highest occupied orbital;
 * each element sums its contributions in the reference's order (the
 * no-difference pass over all determinants, then ordered pairs).  mat is
 * (cap,4) column-major.  Returns the row count or -status; rows are written
 * only when cap >= count (cap = 0 counts). */
int64_t thistle_createtwordm_bit(const int64_t *alpha, const int64_t *beta,
                                 const double *c1, const double *c2, int64_t n,
                                 double cutoff, const ThistleOpts *opts,
                                 int32_t *mat, double *total, int64_t cap);
/* one_rdm_bit (src/one_rdm.f90:1388-1537): onerdm (maxnmo x maxnmo,
 * column-major), maxnmo = highest occupied orbital, written to *norb_out.
 * civs is the file's third column (the reference reads both coefficient
 * columns into one array, the third last).  onerdm = NULL only reports
 * maxnmo; maxnmo > cap_norb returns THISTLE_EINVAL. */
int32_t thistle_one_rdm_bit(const int64_t *alpha, const int64_t *beta,
                            const double *civs, int64_t n, const ThistleOpts *opts,
                            int32_t cap_norb, int32_t *norb_out, double *onerdm);
/* createtwordm_bit_fci (src/one_rdm.f90:390-927): the FCI builder, from one
 * coefficient column civs (the file's second column) and the driver's
 * tables of unique alpha strings (src/scattering_main_routine.py:269-289):
 * uniq = newdat(:,1) (nuniq strings, increasing; newdat(:,2) is never read),
 * irep (determinants per string) and start / end (1-based first and last
 * file position).  Rows and values as thistle_createtwordm_bit's.  The
 * inputs must be ones the reference processes: one electron count in every
 * string, nuniq > norbs (:479 reads row norbs + 1), every determinant's
 * alpha and beta string among uniq, each string's determinants contiguous
 * in the file, orbitals 1-31; THISTLE_EINVAL otherwise.  Both reference
 * callers pass fci = .False. (src/scattering_main_routine.py:251,
 * src/Main.f90:142,179); this entry serves callers of the routine itself. */
int64_t thistle_createtwordm_bit_fci(const int64_t *alpha, const int64_t *beta,
                                     const double *civs, int64_t n, const int64_t *uniq,
                                     int64_t nuniq, const int64_t *irep, const int64_t *start,
                                     const int64_t *end, double cutoff, const ThistleOpts *opts,
                                     int32_t *mat, double *total, int64_t cap);
/* total_scattering_calculation_2 (src/main_calculation.f90:501-849) on the
 * bit-string route (readrdm false, fci false): type 1 total j0 and 7 total
 * j2 from createtwordm_bit, 2 elastic j0 and 8 elastic j2 from one_rdm_bit.
 * As in the reference, types 1 and 8 read the mcci_to_bit conversion, type
 * 7 reads it only when mcci != 0 and type 2 reads the original
 * coefficients; types 1, 7 and 8 need equal alpha and beta counts.
 * q_abs_out (may be NULL) is zeroed as the reference leaves it. */
int32_t thistle_total_scattering_calculation_2(int32_t type, const ThistleBasis *basis,
                                               int64_t n, const int64_t *alpha,
                                               const int64_t *beta, const double *c1,
                                               const double *c2, int32_t mcci,
                                               const ThistleOpts *opts,
                                               double *q_abs_out, double *result_out);

/* ---- device stages, exposed for parity checks ------------------------- */

/* Reference quartet enumeration (tot_integration's five nests,
 * src/Integrals.f90:79-82,178-180,258-260,340-341,416) computed on the GPU.
 * Writes 6 int32 per quartet (gi,gj,gk,gr 1-based, multiplicity, branch
 * 1 = h < cutoffcentre) in reference order.  Returns Q_ref or -status. */
int64_t thistle_quartet_list(const ThistleBasis *basis, int32_t *out,
                             int64_t cap, const ThistleOpts *opts);

/* Screened quartet list: the entries above whose 2-RDM block holds some
 * |ztot| >= cutoffz, in reference order (stable prefix-scan compaction).
 * type 1 takes the 2-RDM rows, type 2 the 1-RDM. */
int64_t thistle_quartet_list_screened(int32_t type, const ThistleBasis *basis,
                                      int64_t nmat, const int32_t *mat,
                                      const double *total, int32_t norb,
                                      const double *onerdm, int32_t *out,
                                      int64_t cap, const ThistleOpts *opts);

/* unique_total (src/unique_module.f90:90-145) on the GPU: first-occurrence
 * unique rows of matrix(nrow,ncol) (column-major, ncol <= 8) with summed
 * totals in row order; iuni/irec 1-based as src/UniqueModule.jl:156-168.
 * rmat is (nrow,ncol) column-major capacity; returns the unique count. */
int64_t thistle_unique_total(const int32_t *matrix, int64_t nrow,
                             int32_t ncol, const double *total, int32_t *rmat,
                             double *rtot, int32_t *iuni, int32_t *irec,
                             const ThistleOpts *opts);

/* Allocation-free unique_total.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops).  Same result as thistle_unique_total, on device pointers:
 * matrix_dev (nrow, ncol) column-major, total_dev [nrow]; out rmat_dev
 * (nrow, ncol) column-major with leading dimension nrow (the first *count
 * rows hold the unique rows), rtot_dev / iuni_dev [nrow] (first *count
 * valid), irec_dev [nrow], count_dev one int64 (any of rmat/rtot/iuni/irec
 * may be NULL).  Values must lie in 0..65535 (ncol <= 4) or 0..255
 * (ncol <= 8); device data is not checked.  The work is enqueued on
 * opts->stream with no allocation, copy to the host or synchronisation;
 * nrow above THISTLE_NOALLOC_MAX_ROWS returns THISTLE_EINVAL.  The hash
 * table and scan live in one static device buffer of
 * THISTLE_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends and
 * cleared by the call itself: calls that overlap on one GPU (other streams
 * or threads) race on it and must be serialised by the caller. */
#define THISTLE_NOALLOC_MAX_ROWS (1 << 18)
#define THISTLE_NOALLOC_SCRATCH_BYTES 11535376
int32_t thistle_unique_total_noalloc(const int32_t *matrix_dev, int64_t nrow,
                                     int32_t ncol, const double *total_dev,
                                     int32_t *rmat_dev, double *rtot_dev,
                                     int32_t *iuni_dev, int32_t *irec_dev,
                                     int64_t *count_dev,
                                     const ThistleOpts *opts);

/* Pair tables of variables_total (src/Variables.f90:1113-1150): for each
 * group pair gi <= gj (pair index p in row-major upper-triangle order)
 * the Hermite tables E[p][axis][la][lb][t] (strides (maxl+1)^2(2maxl+1)),
 * centres P[p][3] and e12[p][nq]. */
int32_t thistle_pair_tables(const ThistleBasis *basis, double *E, double *P,
                            double *e12, const ThistleOpts *opts);

/* ---- plan API: inputs resident in HBM across repeated executions ------ */

typedef struct ThistlePlan ThistlePlan;

/* type 1 (total j0) and 7 (total j2): mat/total rows; type 2 (elastic
 * j0) and 8 (elastic j2): onerdm (norb x norb). */
int32_t thistle_plan_create(int32_t type, const ThistleBasis *basis,
                            int64_t nmat, const int32_t *mat,
                            const double *total, int32_t norb,
                            const double *onerdm, const ThistleOpts *opts,
                            ThistlePlan **plan);
/* Runs every device stage (pair tables, enumeration, 2-RDM transform,
 * contraction + Bessel series, reduction); result_out may be NULL (the
 * result then stays on the device, retrievable by thistle_plan_result). */
int32_t thistle_plan_execute(ThistlePlan *plan, double *result_out);
int32_t thistle_plan_result(ThistlePlan *plan, double *result_out);
/* Last execution's device time per stage in ms: [pair tables, enumeration,
 * transform, quartet kernel, reduction], measured with hipEvents on the
 * plan's stream.  Also the number of quartets, waves and the units
 * (quartets x q points) of the last execution. */
int32_t thistle_plan_timings(ThistlePlan *plan, double *ms5, int64_t *nquart,
                             int64_t *units);
/* Last execution's kernel spans in ms, by hipEvents on the streams the
 * kernels run on: the contraction kernels (every LL class, plan stream,
 * first launch to last) and the series kernels (series stream, first series
 * launch to the last reduction, waits on the contractions included). */
int32_t thistle_plan_kernel_timings(ThistlePlan *plan, double *contract_ms,
                                    double *series_ms);
void thistle_plan_destroy(ThistlePlan *plan);
/* Mixed precision (THISTLE_MIXED) evaluates the series in FP32 where
 * q h > LLmax + 6 (the upward Bessel recurrence), FP64 elsewhere, and sums
 * S(q) in FP64 with compensation.  It also sums A2(q), the squares of the
 * FP32-evaluated terms' error scales |weight| sum_b (b + 2) |h_b u_b|, and
 * then re-evaluates in FP64, from the same q-independent h(beta) records,
 * every q point with |S(q)| < 4 * 2^-24 * 1e6 * sqrt(A2(q)): near the zeros
 * of S(q) the FP32 error could otherwise exceed 1e-6 |S(q)|.  After an
 * execution this returns the number of re-evaluated points, their indices
 * (idx, when cap >= count) and A2(q) (scale, nq doubles, optional).  count
 * is 0 for an FP64 plan.  No counterpart in the reference, which runs in
 * FP64. */
int32_t thistle_plan_mixed_fallback(ThistlePlan *plan, int32_t *count,
                                    int32_t *idx, int32_t cap, double *scale);

/* Restricts the plan's quartet stage to the given positions of the
 * reference enumeration (0-based, strictly increasing, < Q_ref), or
 * restores every quartet when positions is NULL.  Later executions return
 * the partial S(q) of those quartets: sums over a sample, or one shard of
 * the quartet list.  No counterpart in the reference, whose five nests
 * (src/Integrals.f90:79-416) always run whole. */
int32_t thistle_plan_set_quartets(ThistlePlan *plan, const int64_t *positions,
                                  int64_t n);

/* Batched frames (total j0).  No counterpart in the reference, which runs
 * one geometry per call.  nframes independent problems, frame f given by
 * bases[f] and its 2-RDM rows mats[f] (nmat[f] x 4, column-major) with
 * values totals[f].  The first execution tunes the series kernel's waves
 * per compute unit on eager runs and captures every frame's stages into
 * one hipGraph, which later executions replay.  S_out receives the frames'
 * S(q) concatenated in frame order (sum of bases[f].nq doubles). */
typedef struct ThistleBatch ThistleBatch;
int32_t thistle_batch_create(int32_t nframes, const ThistleBasis *bases,
                             const int64_t *nmat, const int32_t *const *mats,
                             const double *const *totals,
                             const ThistleOpts *opts, ThistleBatch **batch);
int32_t thistle_batch_execute(ThistleBatch *batch, double *S_out);
/* Device time of the last execution (ms), quartet x q-point units of the
 * batch, the series kernel's wave count per launch and whether a graph is
 * replayed (1) or the stages are enqueued eagerly (0). */
int32_t thistle_batch_info(ThistleBatch *batch, double *ms_last,
                           int64_t *units, int32_t *waves, int32_t *graph);
/* The batch's tuned configuration: partial rows (series chunks) per frame
 * and frame groups of the cross-frame quartet stage (0, 0 when the batch
 * runs per-frame launches).  No counterpart in the reference. */
int32_t thistle_batch_tuning(ThistleBatch *b, int32_t *rows, int32_t *groups);
void thistle_batch_destroy(ThistleBatch *batch);

/* Number of HIP devices visible to the library (negative on error), and a
 * device-wide synchronize of the given ordinal. */
int32_t thistle_device_count(void);
int32_t thistle_device_synchronize(int32_t device);

/* Library version string. */
const char *thistle_version(void);

#ifdef __cplusplus
}
#endif
#endif
