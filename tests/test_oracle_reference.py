"""The CPU oracle against golden vectors written by the reference's own
Fortran kernels (flang build of src/twordms.f90, src/one_rdm.f90,
src/Integrals.f90, src/unique_module.f90, src/calculate_form_factors.f90;
fixtures made by tests/golden/make_golden.py).

These tests pin the oracle before it is used as the checker of the HIP path.
"""
import numpy as np
import pytest

import oracle

CASES = ["h2", "sp3", "spd2", "f2", "xyx_cas44", "xyx", "far"]
PI_SP = 3.1415927410125732421875   # acos(-1.0000) in single precision (src/Variables.f90:994)


def _close(S, ref, rtol=1e-12, ntol=1e-14):
    """Per q within rtol, except near a zero of S(q), where the absolute error
    is held to ntol of max |S| (both sums round differently there)."""
    err = np.abs(S - ref)
    ok = err <= rtol * np.abs(ref) + ntol * np.max(np.abs(ref))
    return bool(np.all(ok)), float(np.max(err / np.abs(ref)))


@pytest.fixture(scope="module", params=CASES)
def case(request, golden):
    return request.param, golden(request.param)


def _cut(g):
    cz, cmd, cc = (float(v) for v in g["cutoffs"])
    return dict(cutoffz=cz, cutoffmd=cmd, cutoffcentre=cc)


def test_twordm_selection_bit_exact(case):
    """L1: nonzero 2-RDM rows and values of createtwordm (src/twordms.f90:175-556)."""
    name, g = case
    mat, tot = oracle.createtwordm(g["confs"], g["civs"])
    assert np.array_equal(mat, g["mat"].astype(np.int32)), name
    assert np.array_equal(tot, g["total"]), name


def test_onerdm(case):
    """createonerdm (src/one_rdm.f90:109-222)."""
    name, g = case
    gam = oracle.onerdm(g["confs"], g["civs"])
    ref = g["onerdm"]
    n = gam.shape[0]
    np.testing.assert_allclose(gam, ref[:n, :n], rtol=0, atol=1e-14)


@pytest.mark.parametrize("zmode", [0, 1])
def test_total_j0(case, zmode):
    """tot_integration (src/Integrals.f90:8-472) with the per-quartet dgemm
    (zmode 0) and with the MO-pair factorised transform (zmode 1)."""
    name, g = case
    mat, tot = g["mat"].astype(np.int32), g["total"]
    S = oracle.total_j0(g["l"], g["m"], g["n"], g["group"], g["ga"], g["xx"], g["yy"],
                        g["zz"], g["mmod"], mat, tot, g["q"], zmode=zmode, **_cut(g))
    ok, rel = _close(S, g["S_total"])
    assert ok, (name, rel)


def test_elastic_j0(case):
    """elastic_integration (src/Integrals.f90:1575-1939) with variables_elastic's Z."""
    name, g = case
    gam = oracle.onerdm(g["confs"], g["civs"])
    Z = oracle.elastic_z(gam, g["mmod"])
    S = oracle.elastic_j0(g["l"], g["m"], g["n"], g["group"], g["ga"], g["xx"], g["yy"],
                          g["zz"], Z, g["q"], **_cut(g))
    rel = np.max(np.abs(S - g["S_elastic"]) / np.abs(g["S_elastic"]))
    assert rel < 1e-12, (name, rel)


def test_total_j2(case):
    """tot_integration_j2 (src/Integrals.f90:954-1422) with set_P0_j2: type 7.
    At q[0] = 1e-10 the j2 intensity is ~1e-20 and only its size is checked."""
    name, g = case
    mat, tot = g["mat"].astype(np.int32), g["total"]
    S = oracle.total_j2(g["l"], g["m"], g["n"], g["group"], g["ga"], g["xx"], g["yy"],
                        g["zz"], g["mmod"], mat, tot, g["q"], **_cut(g))
    ref = g["S_total_j2"]
    rel = np.max(np.abs(S - ref)[1:] / np.abs(ref)[1:])
    assert rel < 1e-12, (name, rel)
    assert abs(S[0] - ref[0]) <= 1e-14 * np.max(np.abs(ref))


def test_elastic_j2(case):
    """elastic_integration_j2 (src/Integrals.f90:1941-2364) with the P0_j2
    table: type 8.  At q[0] = 1e-10 only the size of the value is checked."""
    name, g = case
    gam = oracle.onerdm(g["confs"], g["civs"])
    Z = oracle.elastic_z(gam, g["mmod"])
    S = oracle.elastic_j2(g["l"], g["m"], g["n"], g["group"], g["ga"], g["xx"], g["yy"],
                          g["zz"], Z, g["q"], **_cut(g))
    ref = g["S_elastic_j2"]
    rel = np.max(np.abs(S - ref)[1:] / np.abs(ref)[1:])
    assert rel < 1e-12, (name, rel)
    assert abs(S[0] - ref[0]) <= 1e-14 * np.max(np.abs(ref))


def test_nuclei_electron(case):
    """nuclei_electron_integration (src/Integrals.f90:2493-2596) with
    integral_k_ij / integral_ij_pzero / BesselSum, and the nuclear sinc sum
    of types 5/6 (src/main_calculation.f90:303-312)."""
    name, g = case
    gam = oracle.onerdm(g["confs"], g["civs"])
    Z = oracle.elastic_z(gam, g["mmod"])
    Ine = oracle.nuclei_electron(g["l"], g["m"], g["n"], g["group"], g["ga"], g["xx"], g["yy"],
                                 g["zz"], Z, g["Zn"], g["geom"], g["q"], **_cut(g))
    ref = g["S_ne"]
    assert np.max(np.abs(Ine - ref)) <= 1e-13 * np.max(np.abs(ref)), name
    Inn = oracle.nuclear_nn(g["Zn"], g["geom"], g["q"])
    assert np.array_equal(Inn, g["S_nn"]), name


def test_unique_total(case):
    """L3: unique_total (src/unique_module.f90:90-145) on the 2-RDM rows."""
    name, g = case
    rmat, rtot, iuni, irec = oracle.unique_total(g["mat"], g["total"])
    assert np.array_equal(rmat, g["rmat"].astype(np.int32)), name
    assert np.array_equal(rtot, g["rtot"]), name


def test_p0_table(case):
    """set_P0 (src/Integrals.f90:3001-3031): sum and absolute sum of the table."""
    name, g = case
    P0 = oracle.set_p0(4 * int(g["l"].max()), g["q"])
    s, a = P0.sum(), np.abs(P0).sum()
    np.testing.assert_allclose([s, a], g["P0_sum"], rtol=1e-13)


def test_forward_limits(case):
    """S(q -> 0): total = N(N-1), elastic = N^2, both times (pi_sp/pi)^3 from
    e12's single-precision pi, for Loewdin-orthonormal MOs."""
    name, g = case
    N = int(np.count_nonzero(g["confs"][0]))
    f = (PI_SP / np.pi) ** 3
    assert g["q"][0] <= 1e-10
    np.testing.assert_allclose(g["S_total"][0], N * (N - 1) * f, rtol=1e-9)
    np.testing.assert_allclose(g["S_elastic"][0], N * N * f, rtol=1e-9)


@pytest.mark.parametrize("name", ["spd2", "f2", "h2"])
def test_variables_total_tables(golden, name):
    """fill_md_table's repacked Hermite tables, the pair centres, e12 (with
    the single-precision pi) and the z11/z22 2-RDM factors of the reference's
    own variables_total (src/Variables.f90:948-1153, run through
    oracle/ref_vars; fixtures vars_*.npz) against the oracle's pair tables and
    z arrays."""
    g = golden(name)
    v = golden(f"vars_{name}")
    l, m, n, group = g["l"], g["m"], g["n"], g["group"]
    dd, P, e12 = oracle.pair_tables(l, m, n, group, g["ga"], g["xx"], g["yy"], g["zz"],
                                    g["q"])
    for k, a in enumerate(("px", "py", "pz")):
        np.testing.assert_allclose(P[k], v[a], rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(e12, np.transpose(v["e12"], (1, 2, 0)), rtol=1e-14, atol=0)
    # Ddx(t, l_ii+1, l_jj+1, group(jj), group(ii)) is set for the exponents
    # that occur (src/Variables.f90:1132-1150); compare exactly those
    for ax, (key, ex) in enumerate((("ddx", l), ("ddy", m), ("ddz", n))):
        ref = np.transpose(v[key], (4, 3, 1, 2, 0))          # [gi, gj, la, lb, t]
        seen = np.zeros(ref.shape, dtype=bool)
        for ii in range(l.size):
            for jj in range(l.size):
                gi, gj = group[ii] - 1, group[jj] - 1
                seen[gi, gj, ex[ii], ex[jj], :ex[ii] + ex[jj] + 1] = True
        np.testing.assert_allclose(dd[ax][seen], ref[seen], rtol=1e-13, atol=1e-15)
    mat, tot = oracle.createtwordm(g["confs"], g["civs"])
    z1, z2 = oracle.build_z12(g["mmod"], mat, tot, l.size)
    np.testing.assert_array_equal(z1, v["z11"])
    np.testing.assert_array_equal(z2, v["z22"])


BIT_CASES = ["small", "mid", "wide", "mrci"]


@pytest.mark.parametrize("name", BIT_CASES)
def test_bit_builders(golden, name):
    """mcci_to_bit (E20.10 round trip included), createtwordm_bit on the raw
    and on the converted coefficients, and one_rdm_bit, against the
    reference's own routines (flang build of src/one_rdm.f90): identical
    doubles and rows."""
    g = golden(f"bit_{name}")
    a, b, c1, c2 = g["alpha"], g["beta"], g["c1"], g["c2"]
    o1, o2 = oracle.mcci_to_bit(a, b, c1, c2)
    assert np.array_equal(o1, g["conv_c1"]) and np.array_equal(o2, g["conv_c2"])
    m, t = oracle.createtwordm_bit(a, b, c1, c2)
    assert np.array_equal(m, g["two_mat"].astype(np.int32))
    assert np.array_equal(t, g["two_total"])
    m, t = oracle.createtwordm_bit(a, b, o1, o2)
    assert np.array_equal(m, g["route_mat"].astype(np.int32))
    assert np.array_equal(t, g["route_total"])
    assert np.array_equal(oracle.one_rdm_bit(a, b, c2), g["onerdm"])


@pytest.mark.parametrize("name", ["small", "mid", "drop"])
def test_fci_builder(golden, name):
    """createtwordm_bit_fci against the reference's own routine (flang build
    with its combo memo table zero-initialised): identical rows and doubles."""
    g = golden(f"fci_{name}")
    m, t = oracle.createtwordm_bit_fci(g["alpha"], g["beta"], g["civs"])
    assert np.array_equal(m, g["mat"].astype(np.int32))
    assert np.array_equal(t, g["total"])
