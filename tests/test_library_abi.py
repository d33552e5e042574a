"""The C-ABI library on the CPU: it loads, exports every function that
include/thistle_scatter.h declares, its host-side 2-RDM / 1-RDM builders are
bit-exact against the reference's (golden vectors), and argument errors
come back as status codes before any device call."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from thistle_amd import lib as L
from thistle_amd import scattering as sc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "thistle_scatter.h")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(L.LIBPATH):
        subprocess.run(["make", "-C", os.path.join(ROOT, "thistle_amd", "csrc")], check=True)
    return L.load()


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(thistle_\w+)\s*\(", text)))


def test_exports_every_declared_function(lib):
    names = header_functions()
    assert len(names) >= 18
    for name in names:
        assert isinstance(getattr(lib, name), C._CFuncPtr), name
    assert sorted(L.EXPORTS) == names
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIBPATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (thistle_\w+)", out))
    assert set(names) <= exported
    # nothing else leaks out with the thistle_ prefix
    assert exported == set(names)


def test_version(lib):
    assert b"gfx950" in lib.thistle_version()


CASES = ["h2", "sp3", "spd2", "f2", "xyx_cas44", "xyx", "far"]


@pytest.mark.parametrize("name", CASES)
def test_host_twordm_bit_exact(golden, name, lib):
    """L1 through the C ABI: createtwordm rows, order and values."""
    g = golden(name)
    mat, tot = sc.createtwordm(g["confs"], g["civs"])
    assert np.array_equal(mat, g["mat"].astype(np.int32))
    assert np.array_equal(tot, g["total"])


@pytest.mark.parametrize("name", CASES)
def test_host_onerdm(golden, name, lib):
    g = golden(name)
    gam = sc.onerdm(g["confs"], g["civs"])
    n = gam.shape[0]
    np.testing.assert_allclose(gam, g["onerdm"][:n, :n], rtol=0, atol=1e-14)


def test_host_rdm_kat(lib):
    """test/runtests.jl:165-179."""
    confs = np.array([[1, 0, 1, 0], [0, 2, 1, 0]], dtype=np.int32)
    civs = np.array([[0.8, 0.1], [0.6, 0.2]])
    g = sc.onerdm(confs, civs)
    assert g[0, 0] == pytest.approx(1.96, abs=1e-12) and g[1, 1] == pytest.approx(1.0, abs=1e-12)
    assert g[0, 1] == 0.0
    mat, tot = sc.createtwordm(confs, civs)
    G = np.zeros((2,) * 4)
    for (p, s, q, r), v in zip(mat - 1, tot):
        G[p, q, r, s] = v
    assert G[1, 0, 0, 1] == pytest.approx(1.0, abs=1e-12) and G[0, 0, 0, 0] == 0.0


def test_host_hf_file(golden, lib):
    """src/HF_NonZero2RM_MO_Coeffs from a closed-shell determinant."""
    d = golden("ref_rdm_files")
    mat, tot = sc.createtwordm(np.array([[1, 2] * 7]), np.ones((1, 1)))
    label = np.array([1, 2, 3, 4, 5, 11, 15])
    assert np.array_equal(label[mat - 1], d["hf_mat"].astype(np.int64))
    assert np.array_equal(tot, d["hf_total"])


def test_twordm_cutoff_screen(golden, lib):
    """A larger screen keeps the rows with |Gamma| >= cutoff, in order."""
    g = golden("xyx_cas44")
    mat, tot = sc.createtwordm(g["confs"], g["civs"])
    m2, t2 = sc.createtwordm(g["confs"], g["civs"], cutoff=1e-3)
    keep = np.abs(tot) >= 1e-3
    assert np.array_equal(m2, mat[keep]) and np.array_equal(t2, tot[keep])


def _small_basis():
    ga = np.array([1.0, 1.0, 1.0, 0.5])
    l = np.array([1, 0, 0, 0])
    m = np.array([0, 1, 0, 0])
    n = np.array([0, 0, 1, 0])
    xyz = np.zeros(4)
    group = np.array([1, 1, 1, 2])
    mmod = np.eye(2, 4)
    q = np.linspace(0.1, 1.0, 4)
    return ga, l, m, n, xyz, group, mmod, q


def _status(type_, ga, l, m, n, xyz, group, mmod, q, mat, tot, norb=0, onerdm=None,
            cutoffs=(1e-15, 1e-20, 1e-3)):
    lib = L.load()
    b = L.make_basis(ga, l, m, n, xyz, xyz, xyz, group, mmod, q, *cutoffs)
    o = L.make_opts(device=0)
    plan = C.c_void_p()
    mat = None if mat is None else np.require(mat, np.int32, ["F"])
    tot = None if tot is None else np.require(tot, np.float64)
    g = None if onerdm is None else np.require(onerdm, np.float64, ["F"])
    rc = lib.thistle_plan_create(type_, C.byref(b), 0 if mat is None else mat.shape[0],
                                 L._ptr(mat), L._ptr(tot), norb, L._ptr(g), C.byref(o),
                                 C.byref(plan))
    assert not plan.value
    return rc


def test_status_codes_before_device(lib):
    """The reference stops the process on these inputs (src/Integrals.f90:5117-5120,
    src/Variables.f90:681-682,939-940); the C ABI returns a status instead."""
    ga, l, m, n, xyz, group, mmod, q = _small_basis()
    mat = np.array([[1, 1, 1, 1]])
    tot = np.array([1.0])
    assert _status(3, ga, l, m, n, xyz, group, mmod, q, mat, tot) == 7         # ETYPE
    bad = l.copy()
    bad[0] = 5
    assert _status(1, ga, bad, m, n, xyz, group, mmod, q, mat, tot) == 2       # ELMAX
    hl, hm = l.copy(), m.copy()
    hl[0], hm[0] = 3, 2                                                        # L = 5 shell
    assert _status(1, ga, hl, hm, n, xyz, group, mmod, q, mat, tot) == 3       # ELLMAX
    assert _status(1, ga, l, m, n, xyz, np.array([1, 2, 1, 2]), mmod, q, mat, tot) == 4
    assert _status(1, ga, l, m, n, xyz, np.array([1, 1, 1, 3]), mmod, q, mat, tot) == 4
    mixed_l = n.copy()
    mixed_l[2] = 2                                     # a d component in a p shell
    assert _status(1, ga, l, m, mixed_l, xyz, group, mmod, q, mat, tot) == 4
    assert _status(1, ga, l, m, n, xyz, group, mmod, q, np.array([[1, 1, 3, 1]]), tot) == 1
    gl = l.copy()
    gl[3] = 4                                          # a g shell: j2 order 18 > 16
    assert _status(7, ga, gl, m, n, xyz, group, mmod, q, mat, tot) == 3
    assert _status(2, ga, l, m, n, xyz, group, mmod, q, None, None, 3, np.eye(3)) == 1
    assert _status(2, ga, l, m, n, xyz, group, mmod, q, None, None, 0, None) == 1


def test_q_grid_and_cutoffs_are_checked(lib):
    """|q| must be finite and >= 0 (x = q h < 0 would break the series, whose
    reference form sin(pmu)/pmu is even in q); cutoffs finite and >= 0."""
    ga, l, m, n, xyz, group, mmod, q = _small_basis()
    mat = np.array([[1, 1, 1, 1]])
    tot = np.array([1.0])
    for bad in (-0.5, np.nan, np.inf):
        qb = q.copy()
        qb[1] = bad
        assert _status(1, ga, l, m, n, xyz, group, mmod, qb, mat, tot) == 1
    for cut in ((-1.0, 1e-20, 1e-3), (1e-15, np.nan, 1e-3), (1e-15, 1e-20, -1e-3)):
        assert _status(1, ga, l, m, n, xyz, group, mmod, q, mat, tot, cutoffs=cut) == 1


def test_cutoffmd_rule(lib):
    """The device contraction keeps every MD product; the reference drops
    |prod6| <= cutoffmd (src/Integrals.f90:5198-5203).  Above
    THISTLE_CUTOFFMD_MAX the call returns THISTLE_ECUTOFF (8) instead of a
    different S(q); at and below it the arguments pass (the next status is
    the invalid 2-RDM row, checked after the basis)."""
    ga, l, m, n, xyz, group, mmod, q = _small_basis()
    tot = np.array([1.0])
    badrow = np.array([[1, 1, 3, 1]])
    assert _status(1, ga, l, m, n, xyz, group, mmod, q, badrow, tot,
                   cutoffs=(1e-15, 1e-9, 1e-3)) == 1          # row check comes first
    good = np.array([[1, 1, 1, 1]])
    for md, want in ((1e-9, 8), (1e-17, 8), (2 * L.CUTOFFMD_MAX, 8)):
        assert _status(2, ga, l, m, n, xyz, group, mmod, q, None, None, 2, np.eye(2),
                       cutoffs=(1e-15, md, 1e-3)) == want
    with pytest.raises(L.ThistleError) as e:
        sc.total_from_twordm(ga, l, m, n, xyz, xyz, xyz, mmod, q, group, good, tot,
                             cutoffmd=1e-12)
    assert e.value.code == 8
    header = open(HEADER).read()
    assert re.search(r"#define THISTLE_CUTOFFMD_MAX 1e-18", header)


def test_python_errors_are_raised(lib):
    ga, l, m, n, xyz, group, mmod, q = _small_basis()
    with pytest.raises(L.ThistleError) as e:
        sc.total_from_twordm(ga, l * 5, m, n, xyz, xyz, xyz, mmod, q, group,
                             np.array([[1, 1, 1, 1]]), np.array([1.0]))
    assert e.value.code == 2
