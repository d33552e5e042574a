// Internal declarations shared by the host and device translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/thistle_scatter.h"

// Primitive shell ("group"): all cartesian components of one primitive share
// centre and exponent (src/molden_reader_nikola.py:195-251).
struct GroupInfo {
    double x, y, z, ga;
    int32_t start, count, L, pad;  // pad: shell-signature id (see setup_common)
};

// Group pair gi <= gj, the bra or ket of a quartet.
struct PairInfo {
    double Px, Py, Pz, gaP;
    double pre, kab;           // (pi_sp/gaP)^1.5 and exp(-ga_i ga_j/gaP |AB|^2)
    int32_t gi, gj, off, nab;  // off: start of the pair block in pair space
};

// Quartet entry packing: 10 bits per group index (ng <= 1024).
#define QE_GI(e) ((int)((e)&1023u))
#define QE_GJ(e) ((int)(((e) >> 10) & 1023u))
#define QE_GK(e) ((int)(((e) >> 20) & 1023u))
#define QE_GR(e) ((int)(((e) >> 30) & 1023u))
#define QE_MULT(e) ((int)(((e) >> 40) & 15u))
#define QE_BR(e) ((int)(((e) >> 44) & 1u))

// Enumeration row: one inner loop of one of the five nests.
struct EnumRow {
    int32_t nest, i, j, k;  // fixed outer indices (unused ones = -1)
    int32_t r0, count;      // inner index runs r0 .. r0+count-1
    int64_t out;            // first output slot
};

// Per-quartet records written by the contraction and read by the series
// kernel, in class-list order: the DO series coefficients (DO = LL + 1, j2
// LL + 3) times the quartet's multiplicity (1, 2, 4 or 8: an exact scaling,
// so the series needs no multiplication by it), then h = |P_ket - P_bra|,
// 1/h (0 on the h = 0 branch) and the packed pair indices and flags of
// HS_META.  The classes' records follow one
// another (largest LL last); each class starts at the plan's offset for it.
#define HS_ROW(DO) ((DO) + 3)
// record word DO + 2: bra pair (bits 0-19), ket pair (20-39), multiplicity
// (40-43), h = 0 branch (44)
#define HS_META(pb, pk, mult, br)                                                \
    ((uint64_t)(pb) | ((uint64_t)(pk) << 20) | ((uint64_t)(mult) << 40) |        \
     ((uint64_t)(br) << 44))

// The quartet of element x of an enumeration row, and its multiplicity.
// Nests 1-5: tot_integration / elastic_integration / tot_integration_j2
// (src/Integrals.f90:79-82,178-180,258-260,340-341,416); elastic_integration_j2
// (:2004-2325) replaces nest 2 by nest 6, (i,j,i,r) with r > j and mult 8,
// and nest 7, (i,j,i,j) with mult 4.
static inline __host__ __device__ int row_quartet(const EnumRow& R, int x, int g[4]) {
    switch (R.nest) {
        case 1: g[0] = R.i; g[1] = R.j; g[2] = R.k; g[3] = x; return 8;
        case 2: g[0] = R.i; g[1] = R.j; g[2] = R.i; g[3] = x; return 4;
        case 3: g[0] = R.i; g[1] = R.i; g[2] = R.k; g[3] = x; return 4;
        case 4: g[0] = R.i; g[1] = R.i; g[2] = x; g[3] = x; return 2;
        case 6: g[0] = R.i; g[1] = R.j; g[2] = R.i; g[3] = x; return 8;
        case 7: g[0] = R.i; g[1] = x; g[2] = R.i; g[3] = x; return 4;
        default: g[0] = g[1] = g[2] = g[3] = x; return 1;
    }
}

static inline __host__ __device__ int pair_index(int gi, int gj, int ng) {
    // gi <= gj, row-major upper triangle
    return gi * ng - (gi * (gi - 1)) / 2 + (gj - gi);
}

// Electron-nucleus term of electron scattering (nuclei_electron_integration):
// per group pair (gi <= gj), every nucleus.  part: [npair_listed][nq].
struct NucArgs {
    const PairInfo* P;
    const GroupInfo* G;
    const double* E;          // pair MD tables, as SeriesArgs::E
    int maxl, ngto, nq;
    const char4* lxyz;
    const double* Z;          // elastic Z, ngto x ngto
    const double* e12;        // [npair][nq]
    const double* q;
    const double* Zn;         // [natom]
    const double* geom;       // [natom][3] column-major (natom x 3)
    int natom;
    double cutoffcentre;
    double* part;             // one row of nq per listed pair
};
// pairs: pair indices whose L_i + L_j = LP; rows written at row0 + k
hipError_t launch_nuc(const NucArgs& a, int LP, const int32_t* pairs, int npairs, int row0,
                      hipStream_t s);
// out[q] = sum over rows r of part[r][q], rows in order
hipError_t launch_rowsum(const double* part, int rows, int nq, double* out, hipStream_t s);

// createtwordm on the device (rdm.hip); see the header comment there
hipError_t rdm_pack(const int32_t* confs, int nconf, int lconfs, int nw, uint64_t* masks,
                    hipStream_t s);
hipError_t rdm_pairs(const uint64_t* masks, int nconf, int nw, int maxdiff, int64_t* cnt,
                     const int64_t* off, int32_t* c1s, int32_t* c2s, int8_t* nds, hipStream_t s);
hipError_t rdm_ccount(const uint64_t* masks, int nw, const int32_t* c1s, const int32_t* c2s,
                      const int8_t* nds, int64_t np, int64_t* cnt, hipStream_t s);
hipError_t rdm_cwrite(const uint64_t* masks, int nw, const int32_t* c1s, const int32_t* c2s,
                      const int8_t* nds, int64_t np, const double* c1v, const double* c2v,
                      const int64_t* coff, int N, uint32_t* keys, double* vals, hipStream_t s);
// createonerdm's contributions on the ndiff <= 2 pair list (counts, then
// keys (d1-1) + N (d2-1) and values at the scanned positions)
hipError_t rdm1_ccount(const uint64_t* masks, int nw, const int32_t* c1s, const int8_t* nds,
                       int64_t np, int64_t* cnt, hipStream_t s);
hipError_t rdm1_cwrite(const uint64_t* masks, int nw, const int32_t* c1s, const int32_t* c2s,
                       const int8_t* nds, int64_t np, const double* c1v, const double* c2v,
                       const int64_t* coff, int N, uint32_t* keys, double* vals, hipStream_t s);
hipError_t rdm_scan_i64(void* temp, size_t& bytes, const int64_t* in, int64_t* out, int64_t n,
                        hipStream_t s);
hipError_t rdm_sort(void* temp, size_t& bytes, const uint32_t* kin, uint32_t* kout,
                    const double* vin, double* vout, int64_t n, int keybits, hipStream_t s);
hipError_t rdm_esum(const uint32_t* keys, const double* vals, int64_t n, int64_t nel,
                    int64_t* start, double* tw, hipStream_t s);
hipError_t rdm_flags(const double* tw, int N, double cutoff, uint8_t* flags, hipStream_t s);
hipError_t rdm_emit(const double* tw, int N, const uint8_t* flags, const int64_t* pos,
                    int32_t* rows, double* out, hipStream_t s);

// Bit-string builders (createtwordm_bit, one_rdm_bit, mcci_to_bit): masks
// hold (alpha, beta) per determinant; items [0, ndiag) are the determinants'
// no-difference passes, the rest the connected pairs of rdm_pairs (nw = 2).
hipError_t bit_pack(const int64_t* alpha, const int64_t* beta, int n, uint64_t* masks,
                    hipStream_t s);
hipError_t bit_diag(int n, int32_t* c1s, int32_t* c2s, int8_t* nds, hipStream_t s);
hipError_t bit2_ccount(const uint64_t* masks, const int32_t* c1s, const int32_t* c2s,
                       const int8_t* nds, int64_t nitem, int ndiag, int64_t* cnt, hipStream_t s);
hipError_t bit2_cwrite(const uint64_t* masks, const int32_t* c1s, const int32_t* c2s,
                       const int8_t* nds, int64_t nitem, int ndiag, const double* c1v,
                       const double* c2v, const int64_t* coff, int N, uint32_t* keys, double* vals,
                       hipStream_t s);
hipError_t bit1_ccount(const uint64_t* masks, const int32_t* c1s, const int32_t* c2s,
                       const int8_t* nds, int64_t nitem, int ndiag, int64_t* cnt, hipStream_t s);
hipError_t bit1_cwrite(const uint64_t* masks, const int32_t* c1s, const int32_t* c2s,
                       const int8_t* nds, int64_t nitem, int ndiag, const double* civ,
                       const int64_t* coff, int N, uint32_t* keys, double* vals, hipStream_t s);
hipError_t mcci_phase(const uint64_t* masks, int n, const double* c1in, const double* c2in,
                      double* c1out, double* c2out, hipStream_t s);

// createtwordm_bit_fci (rdm.hip): device tables of one call
struct FciArgs {
    const uint64_t* masks;   // (alpha, beta) per determinant
    const double* civ;
    const uint64_t* uniq;    // unique alpha strings, increasing
    const int64_t* irep;
    const uint64_t* keys;    // alpha * 2^32 + beta, sorted stably
    const int32_t* kidx;
    int n, nuniq, N;
};
hipError_t fci_keys(const uint64_t* masks, int n, uint64_t* keys, int32_t* idx, hipStream_t s);
hipError_t fci_key_sort(void* temp, size_t& bytes, const uint64_t* kin, uint64_t* kout,
                        const int32_t* vin, int32_t* vout, int64_t n, hipStream_t s);
hipError_t fci_ccount(const FciArgs& a, int64_t* cnt, hipStream_t s);
hipError_t fci_cwrite(const FciArgs& a, const int64_t* coff, uint32_t* keys, double* vals,
                      hipStream_t s);

// Host helpers (rdm_host.cpp)
int thistle_twordm_dense_host(const int32_t* confs, int nconf, int lconfs,
                              const double* civs, int nstates, int state1,
                              int state2, double* tw);
int64_t thistle_twordm_select_host(const double* tw, int norb, double cutoff,
                                   int32_t* mat, double* total, int64_t cap);
int thistle_onerdm_host(const int32_t* confs, int nconf, int lconfs,
                        const double* civs, int nstates, int state1, int state2,
                        double* onerdm);

// Device launchers (kernels.hip)
struct SeriesArgs {
    const uint64_t* qlist;
    int64_t nquart;
    const GroupInfo* G;
    const PairInfo* P;
    const double* E;
    int maxl, ng, ngto;
    const char4* lxyz;
    const double* q;
    int nq, q0, nqc;          // full grid size, chunk start, chunk length
    int qs0, qs1;             // the launch's points [qs0, qs1) of the chunk
    int qlin;                 // q[k] = qa + k dq for k >= 1 (dq > 0)
    int q0off;                // ... and q[0] is not qa (e.g. 1e-10 set by hand)
    double qa, dq;
    const double* e12;        // [npair][nq] fp64
    const float* e12f;        // [npair][nq] fp32 (mixed mode)
    const double* Z;          // total: NPP x NPP lower triangle;
                              // elastic: ngto x ngto
    const float* Zf;          // fp32 copy (mixed mode), may be null
    int64_t ldz;
    int elastic, mixed;
    int j2;                   // total j2 (type 7): h2 coefficients, series / q^2
    double cutoffz;
    double* part;             // [waves][nqc] sums
    double* partc;            // [waves][nqc] Kahan compensations
    double* parta;            // mixed: [2][waves][nqc] error scale of the FP32 terms
                              // (sum of magnitudes, then of their squares)
    int64_t pstride;          // waves x nqc of the plan's partial rows
    // FP64 re-evaluation of flagged q points (mixed mode): their indices in
    // increasing order and their count, both on the device
    const int32_t* qsel;
    const int32_t* nsel;
    int pbase;                // first flagged point of this launch
    uint8_t* flags;           // [nquart] block passes the ztot screen
    double* hs;               // [nquart][17] h(beta) per quartet
    const uint64_t* quadtab;  // GTO-quadruple tables (packed, see setup_common)
    const int64_t* comboff;   // [nsig^4] table start per signature combination
    int nsig;
};

// Cross-frame work tables (batched frames): one entry per frame that has
// quartets in the class, in increasing first block blk0.
// A quartet of a class list, as the contraction reads it: three 16-byte
// loads instead of the chain class list -> quartet entry -> pairs and groups
// -> table start.
struct alignas(16) QHdr {
    uint64_t w;       // quartet entry (bits 0-44) | GTO quadruples (bits 45-60)
    int32_t e;        // position in the quartet list
    uint32_t qoff;    // start of its GTO-quadruple table in quadtab
    double rx, ry, rz;  // ket pair centre - bra pair centre
    uint32_t ob;      // bra pair block offset (bits 0-23) | L_i << 24 | L_j << 28
    uint32_t ok;      // ket pair block offset (bits 0-23) | L_k << 24 | L_r << 28
};
#define QH_NQ4(w) ((uint32_t)((w) >> 45))
#define QH_ENT(w) ((w) & ((1ull << 45) - 1ull))
#define QH_OFF(o) ((int)((o) & 0xFFFFFFu))
#define QH_LA(o) ((int)(((o) >> 24) & 15u))
#define QH_LB(o) ((int)((o) >> 28))

struct ClassWork {           // contraction: one block per quartet group
    int64_t blk0;
    int64_t n;               // quartets of the class in this frame
    const QHdr* hdr;         // the frame's class list, as headers
    double* hs;              // the frame's class records
    int64_t frame;           // index into the SeriesArgs array
};
struct SeriesWork {          // series: nchunk x ny blocks
    int64_t blk0;
    int64_t n;               // records of the class in this frame
    const double* hs;
    int32_t frame, nchunk;   // chunks = partial rows written (<= the frame's rows)
    int32_t ny, qs0, qs1;    // q slice groups over the points [qs0, qs1)
    int32_t pad;
};

struct RedWork {             // reduction: ceil(nqc / RED_Q) blocks per frame
    int64_t blk0;
    const double* part;
    const double* partc;
    double* out;
    int32_t waves, nqc;
};
struct FillWork {            // zeroing: ceil(n / 256) blocks per buffer
    int64_t blk0;
    double* p;
    int64_t n;
};
hipError_t launch_reduce_multi(const RedWork* work, int nwork, int64_t nblocks, hipStream_t s);
int reduce_q_per_block();
hipError_t launch_fill_multi(const FillWork* work, int nwork, int64_t nblocks, hipStream_t s);

// Cross-frame launches of the quartet stage (series.hip, kernels.hip)
hipError_t launch_contract_multi(const SeriesArgs* args, const ClassWork* work, int nwork,
                                 int64_t nblocks, int LL, int tmax, bool j2, hipStream_t s);
int contract_groups_per_block(int LL);
// the q sub-ranges [qs0, qs1) of a chunk of nqc points and their q slices
// per wave (1, 2 or 4); returns their number (<= 2)
int series_q_split(int nqc, bool only4, int* qs0, int* qs1, int* sq);
hipError_t launch_series_multi(const SeriesArgs* args, const SeriesWork* work, int nwork,
                               int64_t nblocks, int LL, int sq, bool j2, bool far,
                               hipStream_t s);

// Adjoint of the total j0 path (adjoint.hip)
struct AdjointArgs {
    SeriesArgs s;             // the forward's tables (Z, hs, flags, ...)
    const double* dE;         // dE/dX_AB, same layout as SeriesArgs::E
    const double* w;          // [nq] dL/dS(q)
    double* gadj;             // [nquart][18] G_beta, beta <= LL+1
    double* Wz;               // [npp][npp] dL/dZ (column-major, all reads)
    int64_t ldw;
    double* qgrad;            // [nquart][12] dL/d(centre) of the four shells
    double* scratch;          // global px tables when they exceed LDS
};
hipError_t launch_pair_dE(const GroupInfo* G, const PairInfo* P, int npair, int maxl,
                          double* dE, hipStream_t s);
hipError_t launch_series_adj(const AdjointArgs& j, int LL, const int32_t* idx,
                             int64_t nidx, int waves, hipStream_t s);
size_t contract_adj_table_doubles(int LL, int Tmax);
int contract_adj_groups(int LL);
hipError_t launch_contract_adj(const AdjointArgs& j, int LL, int Tmax, const int32_t* idx,
                               int64_t nidx, int blocks, bool glob, size_t lds_limit,
                               hipStream_t s);
hipError_t launch_group_grad(const double* qgrad, const int64_t* goff,
                             const int64_t* gent, int ng, double* out, hipStream_t s);
hipError_t launch_dtotal(const int32_t* mat, int64_t nmat, int norb, const double* W,
                         int npmo, double* dtot, hipStream_t s);

hipError_t launch_pair_info(const GroupInfo* G, PairInfo* P, int npair,
                            hipStream_t s);
hipError_t launch_pair_E(const GroupInfo* G, const PairInfo* P, int npair,
                         int maxl, double* E, hipStream_t s);
hipError_t launch_e12(const PairInfo* P, int npair, const double* q, int nq,
                      double* e12, float* e12f, hipStream_t s);
hipError_t launch_enumerate(const EnumRow* rows, int nrows, const PairInfo* P,
                            int ng, double cutoffcentre, uint64_t* out,
                            hipStream_t s);
hipError_t launch_unpack_quartets(const uint64_t* in, int64_t n, int32_t* out,
                                  hipStream_t s);
// LL class key (and position) of every quartet, the class counts, the largest
// exponent-tuple count per class and the signature combinations in use
hipError_t launch_classify(const EnumRow* rows, int nrows, const int8_t* gL, const int8_t* gsig,
                           int S, uint32_t* key, int32_t* pos, unsigned long long* cnt, int* tmax,
                           uint8_t* used, hipStream_t s);
// stable radix sort of (class key, position) pairs on the 5 key bits
hipError_t class_sort(void* temp, size_t& bytes, const uint32_t* kin, uint32_t* kout,
                      const int32_t* vin, int32_t* vout, int64_t n, hipStream_t s);
// k_contract<LL>: h(beta) of one LL class (series.hip)
size_t contract_lds_bytes(int LL, int Tmax);
hipError_t set_contract_lds_limit(size_t bytes);
hipError_t launch_contract(const SeriesArgs& a, int LL, int Tmax, const QHdr* hdr,
                           int64_t nidx, int blocks, hipStream_t s);
// the headers of a class list cls (n quartets, every class in turn)
hipError_t launch_qhdr(const int32_t* cls, int64_t n, const uint64_t* qlist, const GroupInfo* G,
                       const PairInfo* P, int ng, const int64_t* comboff, int nsig, QHdr* out,
                       hipStream_t s);
// k_series_far<LL>: the q h >= XDPH points of the class (same launch shape
// and wave partition as launch_series, added to its partial rows)
hipError_t launch_series_far(const SeriesArgs& a, int LL, int64_t nidx, int waves,
                             hipStream_t s);
// q h from which the series follows dphrec's start-100 recurrence
#define SERIES_XDPH 74.0
// k_series<LL>: S(q) of one LL class from its nidx quartet records (a.hs at
// the class start); first = 1 overwrites the partial rows
hipError_t launch_series(const SeriesArgs& a, int LL, int64_t nidx, int waves, int first,
                         hipStream_t s);
hipError_t launch_reduce(const double* part, const double* partc, int waves,
                         int nqc, double* out, hipStream_t s);
// out[q] = sum over waves of part[w][q], waves in order (no compensation)
hipError_t launch_reduce_plain(const double* part, int waves, int nqc, double* out,
                               hipStream_t s);
// Mixed mode: the q points whose FP32 error scale may exceed the per-point
// bound, |S(q)| < kappa * A(q); qsel receives them in increasing order and
// *nsel their count
hipError_t launch_qflag(const double* S, const double* Aq, int nq, double kappa,
                        int32_t* qsel, int32_t* nsel, hipStream_t s);
// FP64 S(q) of one class at the flagged points pbase .. pbase + 511 (a.qsel,
// a.nsel), one row of nqc = 512 partial sums per wave; first = 1 overwrites
hipError_t launch_series_pts(const SeriesArgs& a, int LL, int64_t nidx, int waves, int first,
                             hipStream_t s);
// S[qsel[pbase + k]] = Sfix[k] for the flagged points of one launch
hipError_t launch_qscatter(const double* Sfix, const int32_t* qsel, const int32_t* nsel,
                           int pbase, int nqc, double* S, hipStream_t s);
hipError_t launch_compact(const uint8_t* flags, const uint64_t* in, int64_t n,
                          int64_t* scratch, uint64_t* out, int64_t* count_dev,
                          hipStream_t s);
hipError_t launch_tmat2(const double* mmod, int nmo, const int2* mopairs,
                        int npmo, const int32_t* xa, const int32_t* xb,
                        int64_t npp, double* T, hipStream_t s);
hipError_t launch_gs_matrix(const double* V, int norb, const int2* mopairs,
                            int npmo, double* Gs, hipStream_t s);
hipError_t launch_zel(const double* mmod, int nmo, int norb,
                      const double* onerdm, int ngto, double* Z, hipStream_t s);
// Lower-triangle Z = T^T Gs T through the MO factorisation of T (norb <= 64):
// Mt = (Gs T)^T (npp x npmo), Cp scratch ngto x zfact_width(norb), W3 scratch
// norb x ngto x npp (unused when fused), jtab[b][a] the primitive pair of GTOs
// (a, b) for a < aend[b].  zgrid > 0: both contractions in one matrix-core
// kernel of zgrid persistent blocks (zfact_fused(norb, ngto) says where it
// applies, THISTLE_ZMFMA=0 turns it off; zfact_fused_prepare once before).
int zfact_width(int norb);
bool zfact_fused(int norb, int ngto);
hipError_t zfact_fused_prepare(int norb, int ngto);
hipError_t launch_zfact(const double* Mt, const double* C, int nmo, int norb, int ngto,
                        int64_t npp, const int32_t* jtab, const int32_t* aend_of, double* Cp,
                        double* W3, double* Z, int zgrid, hipStream_t s);
hipError_t launch_fill(double* p, int64_t n, double v, hipStream_t s);
hipError_t launch_copy(const double* in, double* out, int64_t n, hipStream_t s);
hipError_t launch_to_float(const double* in, float* out, int64_t n,
                           hipStream_t s);

// unique rows + 2-RDM densification
struct UniqueWork {
    uint64_t* keys;   // hash keys
    int32_t* minrow;  // first row per slot
    int32_t* cnt;     // rows per slot
    int32_t* slot;    // slot of each row
    int64_t* scan;    // exclusive scan of first-occurrence flags
    int64_t hsize;
};
hipError_t launch_unique(const int32_t* mat, int64_t nrow, int ncol,
                         const double* total, UniqueWork& w, int32_t* rmat,
                         double* rtot, int32_t* iuni, int32_t* irec,
                         int64_t* count_dev, hipStream_t s);
hipError_t launch_gs(const int32_t* rmat, const double* rtot, int64_t nu,
                     const int64_t* nu_dev, int64_t ldr, int norb, double* V,
                     hipStream_t s);
hipError_t launch_exclusive_scan_u8(const uint8_t* in, int64_t n,
                                    int64_t* out, int64_t* total,
                                    hipStream_t s);
