// Host AddressSanitizer self-test (test infrastructure).  Runs the host code
// of the C ABI library (argument checks, basis checks, pair and enumeration
// tables, the host 2-RDM / 1-RDM builders) and the C oracle on one input
// record, with every host allocation instrumented.  Device calls fail
// without a GPU and are expected to return THISTLE_EDEVICE there.
//
//   host_selftest record.bin
//
// record: int32 ngto, ng, nmo, nconf, lconfs, nstates, nq; int32 l, m, n,
// group [ngto]; f64 ga, xx, yy, zz [ngto]; f64 mmod (nmo, ngto) column-major;
// int32 confs (nconf, lconfs) column-major; f64 civs (nconf, nstates)
// column-major; f64 q [nq].
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/thistle_scatter.h"
#include "../../oracle/oracle.h"

template <class T>
static std::vector<T> rd(FILE* f, size_t n) {
    std::vector<T> v(n);
    if (n && fread(v.data(), sizeof(T), n, f) != n) {
        std::fprintf(stderr, "short record\n");
        std::exit(2);
    }
    return v;
}

static int fails = 0;
#define CHECK(c)                                                         \
    do {                                                                 \
        if (!(c)) {                                                      \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            ++fails;                                                     \
        }                                                                \
    } while (0)

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    auto hd = rd<int32_t>(f, 7);
    const int ngto = hd[0], ng = hd[1], nmo = hd[2], nconf = hd[3], lconfs = hd[4],
              nstates = hd[5], nq = hd[6];
    auto l = rd<int32_t>(f, ngto), m = rd<int32_t>(f, ngto), n = rd<int32_t>(f, ngto),
         group = rd<int32_t>(f, ngto);
    auto ga = rd<double>(f, ngto), xx = rd<double>(f, ngto), yy = rd<double>(f, ngto),
         zz = rd<double>(f, ngto);
    auto mmod = rd<double>(f, (size_t)nmo * ngto);
    auto confs = rd<int32_t>(f, (size_t)nconf * lconfs);
    auto civs = rd<double>(f, (size_t)nconf * nstates);
    auto q = rd<double>(f, nq);
    std::fclose(f);
    const int norb = lconfs / 2;
    int maxl = 0;
    for (int j = 0; j < ngto; ++j) maxl = std::max(maxl, std::max(l[j], std::max(m[j], n[j])));

    // ---- oracle (C restatement)
    std::vector<double> tw((size_t)norb * norb * norb * norb), tw2(tw.size());
    orc_twordm_dense(confs.data(), nconf, lconfs, civs.data(), nstates, 1, 1, tw.data());
    const int nmat = orc_twordm_select(tw.data(), norb, 1e-30, nullptr, nullptr, 0);
    std::vector<int> mat((size_t)4 * nmat);
    std::vector<double> tot(nmat);
    orc_twordm_select(tw.data(), norb, 1e-30, mat.data(), tot.data(), nmat);
    std::vector<double> g1((size_t)norb * norb), Z((size_t)ngto * ngto);
    orc_onerdm(confs.data(), nconf, lconfs, civs.data(), nstates, 1, 1, g1.data());
    orc_elastic_z(g1.data(), norb, mmod.data(), nmo, ngto, Z.data());
    const int T1 = 2 * maxl + 1, L1 = maxl + 1;
    std::vector<double> ddx((size_t)ng * ng * L1 * L1 * T1), ddy(ddx.size()), ddz(ddx.size());
    std::vector<double> px((size_t)ng * ng), py(px.size()), pz(px.size()),
        e12((size_t)ng * ng * nq);
    orc_pair_tables(ngto, ng, maxl, l.data(), m.data(), n.data(), group.data(), ga.data(),
                    xx.data(), yy.data(), zz.data(), q.data(), nq, ddx.data(), ddy.data(),
                    ddz.data(), px.data(), py.data(), pz.data(), e12.data());
    const int64_t Q = orc_quartet_list(ng, px.data(), py.data(), pz.data(), 1e-3, nullptr);
    std::vector<int32_t> ql((size_t)6 * Q);
    orc_quartet_list(ng, px.data(), py.data(), pz.data(), 1e-3, ql.data());
    CHECK(orc_quartet_list_elastic_j2(ng, px.data(), py.data(), pz.data(), 1e-3, nullptr) > 0);
    std::vector<double> S0(nq), S1(nq), S2(nq), Se(nq), Se2(nq);
    for (int zm = 0; zm < 2; ++zm)
        CHECK(orc_total_j0(ngto, ng, maxl, l.data(), m.data(), n.data(), group.data(), ga.data(),
                           xx.data(), yy.data(), zz.data(), mmod.data(), nmo, mat.data(),
                           tot.data(), nmat, q.data(), nq, 1e-15, 1e-20, 1e-3, zm, 2,
                           zm ? S1.data() : S0.data()) == 0);
    double worst = 0.0;
    for (int k = 0; k < nq; ++k) worst = std::max(worst, std::fabs(S0[k] - S1[k]) / std::fabs(S0[k]));
    CHECK(worst < 1e-12);
    CHECK(orc_total_j2(ngto, ng, maxl, l.data(), m.data(), n.data(), group.data(), ga.data(),
                       xx.data(), yy.data(), zz.data(), mmod.data(), nmo, mat.data(), tot.data(),
                       nmat, q.data(), nq, 1e-15, 1e-20, 1e-3, 1, 2, S2.data()) == 0);
    CHECK(orc_elastic_j0(ngto, ng, maxl, l.data(), m.data(), n.data(), group.data(), ga.data(),
                         xx.data(), yy.data(), zz.data(), Z.data(), q.data(), nq, 1e-15, 1e-20,
                         1e-3, 2, Se.data()) == 0);
    CHECK(orc_elastic_j2(ngto, ng, maxl, l.data(), m.data(), n.data(), group.data(), ga.data(),
                         xx.data(), yy.data(), zz.data(), Z.data(), q.data(), nq, 1e-15, 1e-20,
                         1e-3, 2, Se2.data()) == 0);
    std::vector<int> rm((size_t)4 * nmat), iu(nmat), ir(nmat);
    std::vector<double> rt(nmat);
    CHECK(orc_unique_total(mat.data(), nmat, 4, tot.data(), rm.data(), rt.data(), iu.data(),
                           ir.data()) == nmat);
    std::vector<double> xs(nq), rb((size_t)100 * nq);
    for (int k = 0; k < nq; ++k) xs[k] = 0.5 + 11.0 * k;
    orc_dphrec(xs.data(), rb.data(), 100, 8, nq);

    // ---- host code of the C ABI library
    std::vector<double> htw(tw.size()), hg1(g1.size());
    CHECK(thistle_twordm_dense(confs.data(), nconf, lconfs, civs.data(), nstates, 1, 1,
                               htw.data()) == THISTLE_OK);
    CHECK(htw == tw);
    const int64_t hn = thistle_twordm_select(htw.data(), norb, 1e-30, nullptr, nullptr, 0);
    CHECK(hn == nmat);
    std::vector<int32_t> hmat((size_t)4 * hn);
    std::vector<double> htot(hn);
    thistle_twordm_select(htw.data(), norb, 1e-30, hmat.data(), htot.data(), hn);
    CHECK(std::vector<int>(hmat.begin(), hmat.end()) == mat && htot == tot);
    CHECK(thistle_onerdm(confs.data(), nconf, lconfs, civs.data(), nstates, 1, 1, hg1.data()) ==
          THISTLE_OK);
    CHECK(hg1 == g1);
    ThistleBasis b{ngto, ng, maxl, ga.data(), l.data(), m.data(), n.data(), xx.data(), yy.data(),
                   zz.data(), group.data(), nmo, mmod.data(), nq, q.data(), 1e-15, 1e-20, 1e-3};
    ThistleOpts o{THISTLE_FP64, 1e-30, nullptr, 0, 0};
    const bool gpu = thistle_device_count() > 0;
    const int want = gpu ? THISTLE_OK : THISTLE_EDEVICE;
    for (int type : {1, 2, 7, 8}) {
        ThistlePlan* pl = nullptr;
        const int rc = (type == 1 || type == 7)
                           ? thistle_plan_create(type, &b, hn, hmat.data(), htot.data(), 0,
                                                 nullptr, &o, &pl)
                           : thistle_plan_create(type, &b, 0, nullptr, nullptr, norb, g1.data(),
                                                 &o, &pl);
        CHECK(rc == want);
        if (pl) thistle_plan_destroy(pl);
    }
    std::vector<double> qa(nq), res(nq);
    std::vector<double> Zn(1, 1.0), geom(3, 0.0);
    const int rc1 = thistle_total_scattering_calculation(
        1, 1, Zn.data(), geom.data(), 1, 1, 1, &b, nconf, lconfs, nstates, confs.data(),
        civs.data(), &o, qa.data(), res.data());
    CHECK(rc1 == want);
    std::vector<int32_t> hq((size_t)6 * Q);
    const int64_t qn = thistle_quartet_list(&b, hq.data(), Q, &o);
    CHECK(gpu ? qn == Q : qn == -THISTLE_EDEVICE);
    // argument errors come back before any device call
    ThistleBasis bad = b;
    std::vector<double> qneg(q);
    qneg[0] = -1.0;
    bad.q = qneg.data();
    ThistlePlan* pl = nullptr;
    CHECK(thistle_plan_create(1, &bad, hn, hmat.data(), htot.data(), 0, nullptr, &o, &pl) ==
          THISTLE_EINVAL);
    bad = b;
    bad.cutoffmd = 1e-9;
    CHECK(thistle_plan_create(1, &bad, hn, hmat.data(), htot.data(), 0, nullptr, &o, &pl) ==
          THISTLE_ECUTOFF);
    std::printf("host self-test: ngto=%d ng=%d nmat=%d Q=%lld gpu=%d fails=%d\n", ngto, ng, nmat,
                (long long)Q, gpu ? 1 : 0, fails);
    return fails ? 1 : 0;
}
