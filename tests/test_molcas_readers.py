"""Program-specific molden rules (src/molden_reader_nikola_{molcas,pyscf,bagel}.py)
and the Molcas CI route (src/molcas_ci_reader.py: CSF -> determinant
conversion), pinned by physics on the reference's own data files:

* the orbitals read from the Molcas and PySCF molden files are orthonormal
  in the overlap of the primitives read with them (units, normalisation,
  component order and MO assignment all enter);
* every CSF expansion is a normalised eigenfunction of S^2 at M = S, and
  the CSFs of one occupation are orthonormal;
* the Molcas CI vector converted to determinants keeps its norm, and its
  active 1-RDM is diagonal with the natural occupation numbers that the
  same Molcas log prints (the log's orbitals are the natural orbitals);
* the BAGEL CI reader (src/bagel_ci_reader.f90) on logs written to the
  columns it reads, with and without the XMS-CASPT2 rotation (the
  reference holds no BAGEL output file).
"""
import gzip
import itertools
import os
import re

import numpy as np
import pytest

import oracle
from thistle_amd import driver, molcas, readers

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _unzip(tmp_path, name):
    p = tmp_path / name
    with gzip.open(os.path.join(GOLD, name + ".gz"), "rb") as f:
        p.write_bytes(f.read())
    return str(p)


def overlap(arr):
    """Primitive overlap matrix, Gauss-Hermite per axis (exact for the
    polynomial degrees here)."""
    t, w = np.polynomial.hermite.hermgauss(10)
    a, b = arr.ga[:, None], arr.ga[None, :]
    p = a + b
    S = np.ones((arr.ga.size, arr.ga.size))
    for c, e in ((arr.x, arr.l), (arr.y, arr.m), (arr.z, arr.n)):
        A, B = c[:, None], c[None, :]
        P = (a * A + b * B) / p
        x = P[..., None] + t / np.sqrt(p)[..., None]
        f = (x - A[..., None]) ** e[:, None, None] * (x - B[..., None]) ** e[None, :, None]
        S *= np.exp(-a * b / p * (A - B) ** 2) * (f @ w) / np.sqrt(p)
    return S


# ------------------------------------------------------------- molden rules
def test_molcas_molden_orthonormal(tmp_path):
    path = _unzip(tmp_path, "molcas.rasscf.molden")
    arr, mol = readers.read_orbitals(path, decontract=True, flavour="molcas")
    assert len(mol.atoms) == 15 and arr.mo.shape[1] == 43
    # [Atoms] (AU): coordinates taken as bohr
    assert mol.atoms[1].x == pytest.approx(4.63489171)
    assert arr.occ.sum() == pytest.approx(50.0, abs=1e-4)
    C = arr.mo
    err = np.max(np.abs(C.T @ overlap(arr) @ C - np.eye(C.shape[1])))
    assert err < 5e-6, err
    # the Molpro rules stop the [Atoms] block at [GTO] only and trip on the
    # [5D] and [Charge] lines of a Molcas file
    with pytest.raises((ValueError, IndexError)):
        readers.read_orbitals(path, decontract=True, flavour="molpro")


def test_pyscf_molden_orthonormal(tmp_path):
    path = _unzip(tmp_path, "UCCSD_CO_4Bohr_Molden_631star.txt")
    arr, mol = readers.read_orbitals(path, decontract=True, flavour="pyscf")
    assert len(mol.atoms) == 2 and mol.atoms[0].z == pytest.approx(-2.284779939)
    assert arr.mo.shape[1] == 36  # 18 alpha then 18 beta orbitals, file order
    S = overlap(arr)
    for C in (arr.mo[:, :18], arr.mo[:, 18:]):
        err = np.max(np.abs(C.T @ S @ C - np.eye(18)))
        assert err < 1e-9, err
    # the occupied subset (7 alpha + 7 beta) orders energies and occupations
    assert arr.occ.shape == (14,)
    # planted mistake: the same file with angstrom coordinates (Molpro rules)
    bad, _ = readers.read_orbitals(path, decontract=True, flavour="molpro")
    Cb = bad.mo[:, :7]
    assert np.max(np.abs(Cb.T @ overlap(bad) @ Cb - np.eye(7))) > 1e-3


SMALL = """[Molden Format]
[Atoms] {unit}
O     1    8     0.0000000000     0.0000000000     0.1000000000
H     2    1     0.0000000000     0.7500000000    -0.5000000000
{tags}[GTO]
  1 0
 s    1 1.00
      5.0000000000  1.0000000000
 p    1 1.00
      1.2000000000  1.0000000000

  2 0
 s    1 1.00
      0.8000000000  1.0000000000
{gap}[MO]
{orbs}"""


def _orb(header, coeffs):
    return header + "".join(f"   {k + 1}   {v:.10f}\n" for k, v in enumerate(coeffs))


def test_molcas_rules_on_small_file(tmp_path):
    orbs = (_orb(" Sym=     2a\n Ene=   -0.5\n Spin= Alpha\n Occup=    2.0\n", [1, 2, 3, 4, 5])
            + _orb(" Sym=     1a\n Ene=  -20.0\n Spin= Alpha\n Occup=    1.5\n", [6, 7, 8, 9, 10])
            + _orb(" Sym=     3a\n Ene=    0.7\n Spin= Alpha\n Occup=    0.0\n", [9, 9, 9, 9, 9]))
    p = tmp_path / "m.molden"
    p.write_text(SMALL.format(unit="(AU)", tags="[5D]\n[Charge] (Mulliken)\n  0.1\n  0.2\n",
                              gap=" \n", orbs=orbs))
    arr, mol = readers.read_orbitals(str(p), flavour="molcas")
    assert mol.atoms[1].y == 0.75  # not rescaled
    # every orbital kept in file order; keys are the first label character
    np.testing.assert_array_equal(arr.syms, [2.0, 1.0, 3.0])
    assert arr.mo.shape == (5, 3)
    np.testing.assert_array_equal(arr.occ, [1.5, 2.0, 0.0])
    np.testing.assert_array_equal(arr.energy, [20.0, 0.5, 0.7])


def test_bagel_rules_on_small_file(tmp_path):
    orbs = (_orb(" Ene=  -20.0\n Spin= Alpha\n Occup=    2.0\n", [1, 2, 3, 4, 5])
            + _orb(" Ene=   -0.5\n Spin= Alpha\n Occup=    2.0\n", [6, 7, 8, 9, 10]))
    p = tmp_path / "b.molden"
    p.write_text(SMALL.format(unit="Angs", tags="", gap="\n\n", orbs=orbs))
    arr, mol = readers.read_orbitals(str(p), flavour="bagel")
    assert mol.atoms[1].y == pytest.approx(0.75 / readers.AU2ANG)
    np.testing.assert_array_equal(arr.energy, [20.0, 0.5])
    np.testing.assert_array_equal(arr.syms, np.arange(5))
    assert arr.mo.shape == (5, 2)


def test_unknown_flavour_and_unsupported_shell(tmp_path):
    p = tmp_path / "g.molden"
    p.write_text(SMALL.format(unit="AU", tags="", gap="\n\n", orbs=_orb(
        " Sym= 1.1\n Ene= -1.0\n Spin= Alpha\n Occup= 2.0\n", [1, 2, 3, 4, 5])).replace(
        " p    1 1.00", " g    1 1.00"))
    with pytest.raises(ValueError):
        readers.read_orbitals(str(p), flavour="molcas")  # no g shells in that reader
    with pytest.raises(ValueError):
        readers.read_orbitals(str(p), flavour="orca")


# ---------------------------------------------------------------- CSF -> SD
def _csfs(norb):
    """Every step vector over norb orbitals with 2S >= 0 throughout."""
    out = []
    for v in itertools.product("02ud", repeat=norb):
        b, ok = 0, True
        for s in v:
            b += {"u": 1, "d": -1}.get(s, 0)
            ok &= b >= 0
        if ok:
            out.append("".join(v))
    return out


def _raise_spin(vec):
    """S+ on a determinant expansion; alpha and beta slots of one orbital
    are adjacent, so moving a beta electron to its alpha slot has no sign."""
    out = {}
    for sd, c in vec.items():
        for i, s in enumerate(sd):
            if s == "b":
                t = sd[:i] + "a" + sd[i + 1:]
                out[t] = out.get(t, 0.0) + c
    return out


def test_csf2sd_spin_eigenfunctions():
    by_occ = {}
    for csf in _csfs(5):
        twoS = csf.count("u") - csf.count("d")
        det = molcas.csf2sd(csf, twoS / 2)
        vec = {sd: c for c, sd in det}
        assert sum(c * c for c in vec.values()) == pytest.approx(1.0, abs=1e-12), csf
        # highest projection: S+ annihilates an S eigenfunction with M = S
        assert all(abs(c) < 1e-12 for c in _raise_spin(vec).values()), csf
        key = (csf.replace("u", "1").replace("d", "1"), twoS)
        by_occ.setdefault(key, []).append(vec)
    # CSFs of one occupation and spin are orthonormal
    for vecs in by_occ.values():
        sds = sorted({sd for v in vecs for sd in v})
        M = np.array([[v.get(sd, 0.0) for sd in sds] for v in vecs])
        np.testing.assert_allclose(M @ M.T, np.eye(len(vecs)), atol=1e-12)


def test_csf2sd_closed_shell():
    assert molcas.csf2sd("2200", 0) == [[1.0, "2200"]]


def _natural_occupations(log):
    with open(log) as f:
        text = f.read()
    m = re.search(r"Natural orbitals and occupation numbers for root\s+1\s*\n\s*sym 1:(.*)\n",
                  text)
    return np.array([float(v) for v in m.group(1).split()])


def test_molcas_log_to_determinants(tmp_path):
    log = _unzip(tmp_path, "molcas.log")
    nroots, S, nclosed, csfs, civs = molcas.get_civecs_in_csfs(log, False)
    assert (nroots, S, nclosed, len(csfs)) == (1, 0, 21, 1764)
    civs_sd, confs = molcas.get_civs_and_confs(log, False)
    assert len(confs) == 4900 and all(len(c) == 2 * 29 for c in confs)
    v = readers.civs_to_array(civs_sd)
    # the conversion is orthogonal: the printed coefficients keep their norm
    assert np.sum(v[:, 0] ** 2) == pytest.approx(sum(float(c) ** 2 for c in civs[0]), abs=1e-12)
    g = oracle.onerdm(readers.confs_to_int(confs), v)
    act = g[21:, 21:]
    occ = _natural_occupations(log)
    # coefficients are printed to 5 decimals
    np.testing.assert_allclose(np.diag(act), occ, atol=5e-5)
    assert np.max(np.abs(act - np.diag(np.diag(act)))) < 5e-5
    np.testing.assert_allclose(np.diag(g)[:21], 2.0 * np.sum(v[:, 0] ** 2), atol=1e-12)


def test_prepare_inputs_molcas(tmp_path):
    inp = driver.prepare_inputs_molcas(_unzip(tmp_path, "molcas.rasscf.molden"),
                                       _unzip(tmp_path, "molcas.log"))
    assert inp.norb == 29 and inp.mmod.shape == (29, inp.ngto) and inp.Zn.size == 15
    arr = readers.ArrayGTO(mo=None, x=inp.xx, y=inp.yy, z=inp.zz, ga=inp.ga, l=inp.l, m=inp.m,
                           n=inp.n, group=inp.group)
    err = np.max(np.abs(inp.mmod @ overlap(arr) @ inp.mmod.T - np.eye(29)))
    assert err < 5e-6, err
    assert inp.civs.shape == (4900, 1)


# -------------------------------------------------------------- BAGEL CI log
def _bagel_log(path, nclosed, dets, vecs, rot=None):
    """A log with the lines and columns src/bagel_ci_reader.f90 reads."""
    nact = len(dets[0])
    out = ["  Cartesian basis functions are used",
           f"{'  Number of electrons':<29}{2 * nclosed + 4:>6}",
           f"{'    nstate':<18}{len(vecs):>7}",
           f"{'    nclosed':<18}{nclosed:>7}",
           f"{'    nact':<18}{nact:>7}"]
    for r, v in enumerate(vecs):
        out.append(f"     * ci vector, state {r:>3}")
        # each root prints its determinants in its own order
        order = np.argsort(-np.abs(v), kind="stable")
        out += [f"       {dets[k]}  {v[k]: .12f}" for k in order]
    if rot is not None:
        out.append("  * XMS-CASPT2 rotation matrix")
        out += ["   " + "  ".join(f"{x: .12f}" for x in row) for row in rot]
    path.write_text("\n".join(out) + "\n")


def _cas44_dets():
    dets = []
    for a in itertools.combinations(range(4), 2):
        for b in itertools.combinations(range(4), 2):
            dets.append("".join("2" if (i in a and i in b) else "a" if i in a else
                                "b" if i in b else "." for i in range(4)))
    return dets


@pytest.mark.parametrize("caspt2", [False, True])
def test_bagel_log_route(tmp_path, caspt2):
    rng = np.random.default_rng(5)
    dets = _cas44_dets()
    vecs, _ = np.linalg.qr(rng.standard_normal((36, 2)))
    vecs = vecs.T
    th = 0.3
    rot = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    p = tmp_path / "bagel.out"
    _bagel_log(p, 3, dets, vecs, rot if caspt2 else None)
    from thistle_amd import bagel
    sds, civs, ncl = bagel.get_civecs_in_sds(str(p), caspt2)
    assert ncl == 3 and len(sds) == 36 and civs.shape == (2, 36)
    # determinants in order of first appearance (root 0's print order)
    order = np.argsort(-np.abs(vecs[0]), kind="stable")
    assert sds == [dets[k] for k in order]
    expect = vecs[:, order]
    if caspt2:
        expect = rot @ expect
    np.testing.assert_allclose(civs, expect, atol=1e-12)
    confs, cv = bagel.get_civs_and_confs(str(p), caspt2)
    assert confs.shape == (36, 14) and np.all(confs[:, :6] == [1, 2, 1, 2, 1, 2])
    # every row holds the 10 electrons, 5 alpha and 5 beta
    assert np.all(np.sum(confs == 1, 1) == 5) and np.all(np.sum(confs == 2, 1) == 5)
    for r in range(2):
        g = oracle.onerdm(confs, cv, r + 1, r + 1)
        assert np.trace(g) == pytest.approx(10.0, abs=1e-10)
