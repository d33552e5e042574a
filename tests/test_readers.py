"""Readers and text formats on the input side of the path: molden basis/MOs
(src/molden_reader_nikola.py), CI punch strings (src/twordm_red.py,
src/scattering_main_routine.py:232-238), 2-RDM text (src/one_rdm.f90:
2112-2146, src/twordms.f90:521-539) and the Main.f90 text route."""
import itertools
import math
import os

import numpy as np
import pytest

from thistle_amd import driver, readers, synth

AU2ANG = 0.52917721092

MOLDEN = """[Molden Format]
[Atoms] AU
O     1    8     0.0000000000     0.0000000000     0.1000000000
H     2    1     0.0000000000     0.7500000000    -0.5000000000
[GTO]
  1 0
 s    2 1.00
      5.0000000000  0.3000000000
      1.2000000000  0.7000000000
 p    1 1.00
      1.2000000000  1.0000000000
 s    1 1.00
      5.0000000000  1.0000000000

  2 0
 s    1 1.00
      0.8000000000  1.0000000000


[MO]
 Sym=      2.1
 Ene=     -0.5000
 Spin= Alpha
 Occup=    2.000000
   1       0.1000000000
   2       0.2000000000
   3       0.3000000000
   4       0.4000000000
   5       0.5000000000
   6       0.6000000000
 Sym=      1.1
 Ene=    -20.0000
 Spin= Alpha
 Occup=    2.000000
   1       1.0000000000
   2      -1.0000000000
   3       2.0000000000
   4      -2.0000000000
   5       3.0000000000
   6      -3.0000000000
 Sym=      3.1
 Ene=      0.7000
 Spin= Alpha
 Occup=    0.000000
   1       9.0000000000
   2       9.0000000000
   3       9.0000000000
   4       9.0000000000
   5       9.0000000000
   6       9.0000000000
"""


@pytest.fixture
def molden_file(tmp_path):
    p = tmp_path / "mol.molden"
    p.write_text(MOLDEN)
    return str(p)


def test_gto_norm_is_unit_self_overlap():
    """norm * x^l y^m z^n exp(-a r^2) has unit norm (1-D quadrature per axis)."""
    x = np.linspace(-15, 15, 60001)
    for a in (0.3, 1.7):
        for l, m, n in [(0, 0, 0), (1, 0, 0), (2, 0, 0), (1, 1, 0), (3, 0, 0), (1, 1, 1),
                        (2, 1, 1), (4, 0, 0), (2, 2, 0)]:
            ov = 1.0
            for e in (l, m, n):
                ov *= np.trapezoid(x ** (2 * e) * np.exp(-2 * a * x * x), x)
            assert readers.gto_norm(a, l, m, n) ** 2 * ov == pytest.approx(1.0, rel=1e-10)


def test_molden_contracted(molden_file):
    gtos, mol = readers.read_orbitals(molden_file, decontract=False)
    # coordinates are always divided by the bohr length (src/molden_reader_nikola.py:16,109-111)
    np.testing.assert_allclose(mol.geometry()[1], np.array([0.0, 0.75, -0.5]) / AU2ANG,
                               rtol=1e-15)
    assert mol.atomic_numbers().tolist() == [8.0, 1.0]
    # O: s(2 prims) + p(1 prim, 3 comps) + s(1 prim); H: s(1 prim)
    assert gtos.ga.tolist() == [5.0, 1.2, 1.2, 1.2, 1.2, 5.0, 0.8]
    assert list(zip(gtos.l, gtos.m, gtos.n)) == [(0, 0, 0), (0, 0, 0), (1, 0, 0), (0, 1, 0),
                                                 (0, 0, 1), (0, 0, 0), (0, 0, 0)]
    # one group per primitive shared by its cartesian components
    assert gtos.group.tolist() == [1, 2, 3, 3, 3, 4, 5]
    assert gtos.x[-1] == 0.0 and gtos.y[-1] == pytest.approx(0.75 / AU2ANG)
    # occupied MOs ordered by frac(sym)*1000 + int(sym): 1.1 before 2.1
    assert gtos.mo.shape == (7, 2)
    rows = [1, 1, 2, 3, 4, 5, 6]              # contraction row of each primitive
    coef = [0.3, 0.7, 1.0, 1.0, 1.0, 1.0, 1.0]
    mo11 = np.array([1.0, -1.0, 2.0, -2.0, 3.0, -3.0])
    mo21 = np.array([0.1, 0.2, 0.3, 0.4, 0.5, 0.6])
    for k in range(7):
        nrm = readers.gto_norm(gtos.ga[k], gtos.l[k], gtos.m[k], gtos.n[k])
        assert gtos.mo[k, 0] == pytest.approx(coef[k] * nrm * mo11[rows[k] - 1], rel=1e-14)
        assert gtos.mo[k, 1] == pytest.approx(coef[k] * nrm * mo21[rows[k] - 1], rel=1e-14)


def test_molden_decontracted_merges_equal_primitives(molden_file):
    """Same atom, exponent and l,m,n: MO vectors summed, groups renumbered."""
    full, _ = readers.read_orbitals(molden_file, decontract=False)
    gtos, _ = readers.read_orbitals(molden_file, decontract=True)
    assert gtos.ga.tolist() == [5.0, 1.2, 1.2, 1.2, 1.2, 0.8]
    assert gtos.group.tolist() == [1, 2, 3, 3, 3, 4]
    np.testing.assert_allclose(gtos.mo[0], full.mo[0] + full.mo[5], rtol=1e-15)
    np.testing.assert_allclose(gtos.mo[1:5], full.mo[1:5], rtol=0)
    np.testing.assert_allclose(gtos.mo[5], full.mo[6], rtol=0)


def test_molden_component_order(tmp_path):
    """Cartesian component order of the molden format for s..g."""
    text = MOLDEN.split("[GTO]")[0] + "[GTO]\n  1 0\n" + "".join(
        f" {k}    1 1.00\n      1.0  1.0\n" for k in "spdfg") + "\n  2 0\n s    1 1.00\n" \
        "      0.8  1.0\n\n\n[MO]\n Sym= 1.1\n Ene= -1.0\n Spin= Alpha\n Occup= 2.0\n" + \
        "".join(f"  {i} 0.1\n" for i in range(1, 37))
    p = tmp_path / "order.molden"
    p.write_text(text)
    gtos, _ = readers.read_orbitals(str(p))
    got = list(zip(gtos.l.tolist(), gtos.m.tolist(), gtos.n.tolist()))
    want = ([(0, 0, 0)] + [(1, 0, 0), (0, 1, 0), (0, 0, 1)]
            + [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)]
            + [(3, 0, 0), (0, 3, 0), (0, 0, 3), (1, 2, 0), (2, 1, 0), (2, 0, 1), (1, 0, 2),
               (0, 1, 2), (0, 2, 1), (1, 1, 1)]
            + [(4, 0, 0), (0, 4, 0), (0, 0, 4), (3, 1, 0), (3, 0, 1), (1, 3, 0), (0, 3, 1),
               (1, 0, 3), (0, 1, 3), (2, 2, 0), (2, 0, 2), (0, 2, 2), (2, 1, 1), (1, 2, 1),
               (1, 1, 2)]
            + [(0, 0, 0)])
    assert got == want
    assert gtos.group.tolist() == [1, 2, 2, 2] + [3] * 6 + [4] * 10 + [5] * 15 + [6]


def test_punch_string_rewrite(tmp_path):
    """closed orbitals -> 'ab'; 0 -> 00, a -> a0, b -> 0b, 2 -> ab; a -> 1, b -> 2;
    coefficient continuation lines extend the previous determinant."""
    p = tmp_path / "ci.pun"
    p.write_text("PROGRAM * CI\n 2a0b   0.8  0.1\n          0.05\n a2b0  -0.6  0.2\n"
                 "          0.07\n")
    civs, confs = readers.twordmconst(1, str(p))
    assert confs == ["ababa0000b", "aba0ab0b00"]
    ints = readers.confs_to_int(confs)
    assert ints.tolist() == [[1, 2, 1, 2, 1, 0, 0, 0, 0, 2], [1, 2, 1, 0, 1, 2, 0, 2, 0, 0]]
    np.testing.assert_array_equal(readers.civs_to_array(civs),
                                  [[0.8, 0.1, 0.05], [-0.6, 0.2, 0.07]])
    with pytest.raises(ValueError):
        readers.twordmconst([1, 0], str(p))


def test_twordm_text_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    mat = rng.integers(1, 30, size=(50, 4)).astype(np.int32)
    tot = rng.normal(size=50) * 10.0 ** rng.integers(-30, 2, size=50)
    p = str(tmp_path / "tw.dat")
    readers.write_twordm(p, mat, tot)
    m2, t2 = readers.read_twordm(p)
    assert np.array_equal(m2, mat) and np.array_equal(t2, tot)
    with open(p, "a") as f:
        f.write("           1           2           3           4  -1.25D-04\n")
    m3, t3 = readers.read_twordm(p)
    assert m3[-1].tolist() == [1, 2, 3, 4] and t3[-1] == -1.25e-4


def test_reference_twordm_file_through_reader(golden, tmp_path):
    """The reference's 4549-row file survives the writer/reader pair."""
    d = golden("ref_rdm_files")
    p = str(tmp_path / "NonZero2RDM_MO_Coeffs")
    readers.write_twordm(p, d["cas_mat"], d["cas_total"])
    m, t = readers.read_twordm(p)
    assert len(t) == 4549
    assert np.array_equal(m, d["cas_mat"].astype(np.int32))
    assert np.array_equal(t, d["cas_total"])


def test_main_text_route(tmp_path):
    rng = np.random.default_rng(1)
    n = 7
    xx, yy, zz, ga = (rng.normal(size=n) for _ in range(4))
    l, m, nn = (rng.integers(0, 3, size=n) for _ in range(3))
    group = np.array([1, 1, 2, 3, 3, 3, 4])
    readers.write_basis_dat(str(tmp_path / "basis.dat"), xx, yy, zz, ga, l, m, nn, group)
    out = readers.read_basis_dat(str(tmp_path / "basis.dat"))
    for a, b in zip(out, (xx, yy, zz, ga, l, m, nn, group)):
        assert np.array_equal(a, b)
    mm = rng.normal(size=(3, n))
    readers.write_mos_dat(str(tmp_path / "MOs.dat"), mm)
    assert np.array_equal(readers.read_mos_dat(str(tmp_path / "MOs.dat")), mm)
    (tmp_path / "options.dat").write_text(
        "2\n8 1\n0.0 0.0 0.0\n0.0 0.0 1.8\n0.001\n1.0D-15\n1.0D-20\n.false.\n.false.\n"
        ".false.\n0.0 10.0 11\n1\n1 1 1\n'result.dat'\n.true.\n.false.\n.false.\n1\n4\n1\n"
        "1 2 1 2 1.0\n")
    o = readers.read_options_dat(str(tmp_path / "options.dat"))
    assert o.atoms.tolist() == [8.0, 1.0] and o.geom[1].tolist() == [0.0, 0.0, 1.8]
    assert (o.cutoffcentre, o.cutoffz, o.cutoffmd) == (0.001, 1e-15, 1e-20)
    assert (o.qmin, o.qmax, o.npoints, o.typec) == (0.0, 10.0, 11, 1)
    assert (o.state1, o.state2, o.state3, o.file_out) == (1, 1, 1, "result.dat")
    assert o.molpro and not o.molcas and not o.bagel
    assert o.confs.tolist() == [[1, 2, 1, 2]] and o.civs.tolist() == [[1.0]]
    q = readers.linspace_1(o.qmin, o.qmax, o.npoints)
    assert q.tolist() == [float(i) for i in range(11)]
    readers.write_result(str(tmp_path / "r.dat"), q, q * 2)
    back = np.loadtxt(str(tmp_path / "r.dat"))
    assert np.array_equal(back[:, 1], q * 2)


def test_synthetic_inputs_round_trip(tmp_path):
    """The seeded generator's molden + punch text reads back to its own basis
    (bohr centres, exponents, component order, MOs on normalised primitives)."""
    w = synth.make_small(seed=5, natom=2, spec={"s": 1, "p": 1, "d": 1}, nact=2, nelec=2,
                         ncore=1, nq=8)
    mld, pun = synth.write_inputs(w, str(tmp_path))
    inp = driver.prepare_inputs(mld, pun, w.ncore)
    assert inp.ngto == 2 * (1 + 3 + 6)
    assert inp.ng == 6
    assert inp.mmod.shape == (w.ncore + 2, inp.ngto)
    assert inp.confs.shape == (4, 2 * (w.ncore + 2))
    # every determinant has the core doubly occupied and 2 active electrons
    assert np.all(inp.confs[:, :2] == [1, 2]) and np.all(np.count_nonzero(inp.confs, 1) == 4)
    np.testing.assert_allclose(np.linalg.norm(inp.civs[:, 0]), 1.0, rtol=1e-12)
    # MOs orthonormal in the primitive overlap
    S = synth_overlap(inp)
    np.testing.assert_allclose(inp.mmod @ S @ inp.mmod.T, np.eye(inp.mmod.shape[0]),
                               atol=1e-10)


def synth_overlap(inp):
    """Primitive overlap by 1-D quadrature (independent of the generator)."""
    x = np.linspace(-14, 14, 20001)
    n = inp.ngto
    S = np.zeros((n, n))
    for i, j in itertools.product(range(n), repeat=2):
        v = 1.0
        for e1, e2, c1, c2 in ((inp.l[i], inp.l[j], inp.xx[i], inp.xx[j]),
                               (inp.m[i], inp.m[j], inp.yy[i], inp.yy[j]),
                               (inp.n[i], inp.n[j], inp.zz[i], inp.zz[j])):
            v *= np.trapezoid((x - c1) ** e1 * (x - c2) ** e2
                              * np.exp(-inp.ga[i] * (x - c1) ** 2 - inp.ga[j] * (x - c2) ** 2), x)
        S[i, j] = v
    return S


def test_default_q_clamp():
    q = driver.default_q(8.0, 5)
    assert q[0] == 1e-10 and q[-1] == 8.0
    assert math.isclose(q[1], 2.0 + 1e-11 * 0.75, rel_tol=1e-12)
    assert os.path.basename(readers.__file__) == "readers.py"
