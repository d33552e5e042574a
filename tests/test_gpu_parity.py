"""GPU parity: the HIP path against the reference golden vectors and the oracle.

Golden vectors come from the reference's own compiled Fortran kernels
(tests/golden/make_golden.py).  Tolerances: bit-exact for integer and index
outputs; 1e-10 relative for S(q) in FP64 mode.
"""
import numpy as np
import pytest

import oracle
from thistle_amd import scattering as sc

pytestmark = pytest.mark.gpu

CASES = ["h2", "sp3", "spd2", "f2", "xyx_cas44", "xyx", "far"]
RTOL_FP64 = 1e-10


def _args(g):
    return (g["ga"], g["l"], g["m"], g["n"], g["xx"], g["yy"], g["zz"])


@pytest.mark.parametrize("name", CASES)
def test_total_j0_matches_reference(golden, name):
    g = golden(name)
    print(f"case {name}", flush=True)
    cz, cm, cc = g["cutoffs"]
    res = sc.total_from_twordm(*_args(g), g["mmod"], g["q"], g["group"],
                               g["mat"].astype(np.int32), g["total"], cz, cm, cc)
    ref = g["S_total"]
    rel = np.max(np.abs(res - ref) / np.abs(ref))
    print(f"{name}: max rel err total {rel:.3e}", flush=True)
    assert rel < RTOL_FP64


@pytest.mark.parametrize("name", CASES)
def test_elastic_j0_matches_reference(golden, name):
    g = golden(name)
    cz, cm, cc = g["cutoffs"]
    res = sc.elastic_from_onerdm(*_args(g), g["mmod"], g["q"], g["group"], g["onerdm"],
                                 cz, cm, cc)
    ref = g["S_elastic"]
    rel = np.max(np.abs(res - ref) / np.abs(ref))
    print(f"{name}: max rel err elastic {rel:.3e}", flush=True)
    assert rel < RTOL_FP64


@pytest.mark.parametrize("name", ["h2", "sp3", "spd2"])
def test_drop_in_call_matches_reference(golden, name):
    g = golden(name)
    cz, cm, cc = g["cutoffs"]
    l = g["l"]
    for typ, key in ((1, "S_total"), (2, "S_elastic")):
        q_abs, res = sc.total_scattering_calculation(
            typ, np.ones(2), np.zeros((2, 3)), 1, 1, 1, int(l.max()), l.size,
            int(g["group"].max()), g["ga"], l, g["m"], g["n"], g["xx"], g["yy"], g["zz"],
            g["mmod"], g["q"], g["q"].size, g["group"], cz, cm, cc,
            g["confs"].astype(np.int32), g["civs"])
        assert np.all(q_abs == 0.0)
        rel = np.max(np.abs(res - g[key]) / np.abs(g[key]))
        print(f"{name} type {typ}: {rel:.3e}", flush=True)
        assert rel < RTOL_FP64


@pytest.mark.parametrize("name", CASES)
def test_quartet_list_bit_exact(golden, name):
    g = golden(name)
    cz, cm, cc = g["cutoffs"]
    dd, P, e12 = oracle.pair_tables(g["l"], g["m"], g["n"], g["group"], g["ga"], g["xx"],
                                    g["yy"], g["zz"], g["q"])
    ref = oracle.quartet_list(P[0], P[1], P[2], cc)
    got = sc.quartet_list(g["ga"], g["l"], g["m"], g["n"], g["xx"], g["yy"], g["zz"],
                          g["group"], g["q"], cc)
    ng = int(g["group"].max())
    P_ = ng * (ng + 1) // 2
    assert got.shape[0] == P_ * (P_ + 1) // 2 + ng * (ng - 1) * (ng - 2) // 6
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("name", ["h2", "spd2", "xyx_cas44"])
def test_pair_tables_match_oracle(golden, name):
    g = golden(name)
    E, P, e12 = sc.pair_tables(g["ga"], g["l"], g["m"], g["n"], g["xx"], g["yy"], g["zz"],
                               g["group"], g["q"])
    dd, Po, e12o = oracle.pair_tables(g["l"], g["m"], g["n"], g["group"], g["ga"], g["xx"],
                                      g["yy"], g["zz"], g["q"])
    ng = int(g["group"].max())
    p = 0
    for gi in range(ng):
        for gj in range(gi, ng):
            for ax in range(3):
                assert np.allclose(E[p, ax], dd[ax][gi, gj], rtol=1e-14, atol=1e-300)
                assert P[p, ax] == Po[ax][gi, gj]
            assert np.allclose(e12[p], e12o[gi, gj], rtol=1e-14, atol=0)
            p += 1


@pytest.mark.parametrize("name", ["h2", "sp3", "xyx_cas44"])
def test_unique_total_matches_oracle(golden, name):
    g = golden(name)
    mat = g["mat"].astype(np.int32)
    # duplicate some rows so the sums and back-references are exercised
    rng = np.random.default_rng(7)
    pick = rng.integers(0, mat.shape[0], size=max(3, mat.shape[0] // 3))
    m2 = np.concatenate([mat, mat[pick]])
    t2 = np.concatenate([g["total"], rng.standard_normal(pick.size)])
    r_ref = oracle.unique_total(m2, t2)
    r_gpu = sc.unique_total(m2, t2)
    for a, b in zip(r_gpu, r_ref):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("name", ["sp3", "xyx_cas44"])
def test_screened_list_is_stable_compaction(golden, name):
    g = golden(name)
    cz, cm, cc = g["cutoffs"]
    full = sc.quartet_list(g["ga"], g["l"], g["m"], g["n"], g["xx"], g["yy"], g["zz"],
                           g["group"], g["q"], cc)
    scr = sc.quartet_list_screened(1, g["ga"], g["l"], g["m"], g["n"], g["xx"], g["yy"],
                                   g["zz"], g["mmod"], g["q"], g["group"],
                                   mat=g["mat"].astype(np.int32), total=g["total"],
                                   cutoffz=cz, cutoffcentre=cc)
    # every kept entry appears in the full list, in the same relative order
    keyf = {tuple(r): i for i, r in enumerate(full)}
    pos = np.array([keyf[tuple(r)] for r in scr])
    assert np.all(np.diff(pos) > 0)
    # random MOs leave no all-zero block: nothing is dropped
    assert scr.shape[0] == full.shape[0]


@pytest.mark.parametrize("name", CASES)
def test_total_j2_matches_reference(golden, name):
    """Type 7 (tot_integration_j2) against the reference; at q[0] = 1e-10 the
    j2 intensity is ~1e-20 and is held to 1e-10 of max |S|."""
    g = golden(name)
    cz, cm, cc = g["cutoffs"]
    res = sc.total_j2_from_twordm(*_args(g), g["mmod"], g["q"], g["group"],
                                  g["mat"].astype(np.int32), g["total"], cz, cm, cc)
    ref = g["S_total_j2"]
    rel = np.max(np.abs(res - ref)[1:] / np.abs(ref)[1:])
    print(f"{name}: max rel err total j2 {rel:.3e}", flush=True)
    assert rel < RTOL_FP64
    assert abs(res[0] - ref[0]) <= RTOL_FP64 * np.max(np.abs(ref))


@pytest.mark.parametrize("name", ["sp3", "spd2"])
def test_drop_in_type7(golden, name):
    g = golden(name)
    cz, cm, cc = g["cutoffs"]
    l = g["l"]
    _, res = sc.total_scattering_calculation(
        7, np.ones(2), np.zeros((2, 3)), 1, 1, 1, int(l.max()), l.size, int(g["group"].max()),
        g["ga"], l, g["m"], g["n"], g["xx"], g["yy"], g["zz"], g["mmod"], g["q"], g["q"].size,
        g["group"], cz, cm, cc, g["confs"].astype(np.int32), g["civs"])
    ref = g["S_total_j2"]
    assert np.max(np.abs(res - ref)[1:] / np.abs(ref)[1:]) < RTOL_FP64



@pytest.mark.parametrize("name", CASES)
def test_elastic_j2_matches_reference(golden, name):
    """Type 8 (elastic_integration_j2) against the reference; q[0] as type 7."""
    g = golden(name)
    cz, cm, cc = g["cutoffs"]
    res = sc.elastic_j2_from_onerdm(*_args(g), g["mmod"], g["q"], g["group"], g["onerdm"],
                                    cz, cm, cc)
    ref = g["S_elastic_j2"]
    rel = np.max(np.abs(res - ref)[1:] / np.abs(ref)[1:])
    print(f"{name}: max rel err elastic j2 {rel:.3e}", flush=True)
    assert rel < RTOL_FP64
    assert abs(res[0] - ref[0]) <= RTOL_FP64 * np.max(np.abs(ref))


@pytest.mark.parametrize("name", ["sp3", "spd2"])
def test_drop_in_type8(golden, name):
    g = golden(name)
    cz, cm, cc = g["cutoffs"]
    l = g["l"]
    _, res = sc.total_scattering_calculation(
        8, np.ones(2), np.zeros((2, 3)), 1, 1, 1, int(l.max()), l.size, int(g["group"].max()),
        g["ga"], l, g["m"], g["n"], g["xx"], g["yy"], g["zz"], g["mmod"], g["q"], g["q"].size,
        g["group"], cz, cm, cc, g["confs"].astype(np.int32), g["civs"])
    ref = g["S_elastic_j2"]
    assert np.max(np.abs(res - ref)[1:] / np.abs(ref)[1:]) < RTOL_FP64


@pytest.mark.parametrize("name", ["sp3", "xyx_cas44"])
def test_elastic_j2_list_bit_exact(golden, name):
    """The six nests of elastic_integration_j2 enumerated on the GPU: with
    random MOs no block is all zero, so the screened list is the whole list,
    equal to the oracle's row for row (indices, multiplicity, branch)."""
    g = golden(name)
    cz, cm, cc = g["cutoffs"]
    dd, P, e12 = oracle.pair_tables(g["l"], g["m"], g["n"], g["group"], g["ga"], g["xx"],
                                    g["yy"], g["zz"], g["q"][:2])
    ref = oracle.quartet_list_elastic_j2(P[0], P[1], P[2], cc)
    scr = sc.quartet_list_screened(8, g["ga"], g["l"], g["m"], g["n"], g["xx"], g["yy"],
                                   g["zz"], g["mmod"], g["q"], g["group"], onerdm=g["onerdm"],
                                   cutoffz=cz, cutoffcentre=cc)
    assert scr.shape == ref.shape and np.array_equal(scr, ref)


@pytest.mark.parametrize("name", CASES)
def test_nuclei_electron_matches_reference(golden, name):
    """Ine of electron scattering (nuclei_electron_integration) against the
    reference, within 1e-10 of max |Ine| (BesselSum's upward recurrence is
    followed as the reference does, so small q h carries its rounding)."""
    g = golden(name)
    cz, cm, cc = g["cutoffs"]
    res = sc.nuclei_electron(*_args(g), g["mmod"], g["q"], g["group"], g["onerdm"], g["Zn"],
                             g["geom"], cz, cm, cc)
    ref = g["S_ne"]
    err = np.max(np.abs(res - ref)) / np.max(np.abs(ref))
    rel = np.max(np.abs(res - ref)[1:] / np.abs(ref)[1:])
    print(f"{name}: Ine normwise err {err:.3e}, per-q (q>0) {rel:.3e}", flush=True)
    assert err < RTOL_FP64


@pytest.mark.parametrize("name", ["h2", "sp3", "spd2"])
def test_drop_in_types_5_6(golden, name):
    """Types 5 and 6: Iee - 2 Ine + Inn with the reference's pieces."""
    g = golden(name)
    cz, cm, cc = g["cutoffs"]
    l = g["l"]
    for typ, iee in ((5, "S_total"), (6, "S_elastic")):
        _, res = sc.total_scattering_calculation(
            typ, g["Zn"], g["geom"], 1, 1, 1, int(l.max()), l.size, int(g["group"].max()),
            g["ga"], l, g["m"], g["n"], g["xx"], g["yy"], g["zz"], g["mmod"], g["q"],
            g["q"].size, g["group"], cz, cm, cc, g["confs"].astype(np.int32), g["civs"])
        ref = g[iee] - 2 * g["S_ne"] + g["S_nn"]
        err = np.max(np.abs(res - ref)) / np.max(np.abs(ref))
        print(f"{name} type {typ}: normwise err {err:.3e}", flush=True)
        assert err < RTOL_FP64


@pytest.mark.parametrize("name", ["sp3", "xyx_cas44"])
def test_device_class_lists_match_host_bucketing(golden, name):
    """The LL-class lists a plan builds on the device (class keys + stable
    radix sort) against the host bucketing of thistle_plan_set_quartets over
    every quartet: the same lists give bitwise the same S(q)."""
    g = golden(name)
    cz, cm, cc = g["cutoffs"]
    plan = sc.Plan(1, *_args(g), g["mmod"], g["q"], g["group"], mat=g["mat"].astype(np.int32),
                   total=g["total"], cutoffz=cz, cutoffmd=cm, cutoffcentre=cc)
    try:
        dev = plan.execute()
        plan.set_quartets(None)
        host = plan.execute()
        assert np.array_equal(dev, host)
    finally:
        plan.close()
