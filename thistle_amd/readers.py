"""Input readers of the scattering path, behaviour-compatible with the reference.

* Molden: ``read_orbitals`` restates ``src/molden_reader_nikola.py:496-508``
  (``_read_MO_GTOs`` :470-493) with the same parsing rules, MO ordering,
  primitive normalisation and decontraction.  ``flavour`` selects the rules
  of the program-specific copies of that reader: "molcas"
  (``src/molden_reader_nikola_molcas.py``), "pyscf"
  (``src/molden_reader_nikola_pyscf.py``) and "bagel"
  (``src/molden_reader_nikola_bagel.py``); the differences are listed in
  ``FLAVOURS``.
* CI punch file: ``twordmconst`` restates ``src/twordm_red.py:12-139`` and
  ``confs_to_int`` the string -> integer conversion of
  ``src/scattering_main_routine.py:232-238``.
* 2-RDM text file: ``read_twordm`` restates ``src/one_rdm.f90:2112-2146``;
  ``write_twordm`` the writer format of ``src/twordms.f90:521-539``.
* ``Main.f90`` text route: ``read_basis_dat``/``read_mos_dat``/
  ``read_options_dat`` follow ``src/Main.f90:28-112`` and ``write_result``
  ``src/Main.f90:192-198``.
"""
from __future__ import annotations

import copy
import math
import re
from dataclasses import dataclass, field

import numpy as np

AU2ANG = 0.52917721092  # src/molden_reader_nikola.py:16

# cartesian component order per shell letter, src/molden_reader_nikola.py:199-238
_ANG = {
    "s": [(0, 0, 0)],
    "p": [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
    "d": [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)],
    "f": [(3, 0, 0), (0, 3, 0), (0, 0, 3), (1, 2, 0), (2, 1, 0), (2, 0, 1),
          (1, 0, 2), (0, 1, 2), (0, 2, 1), (1, 1, 1)],
    "g": [(4, 0, 0), (0, 4, 0), (0, 0, 4), (3, 1, 0), (3, 0, 1), (1, 3, 0),
          (0, 3, 1), (1, 0, 3), (0, 1, 3), (2, 2, 0), (2, 0, 2), (0, 2, 2),
          (2, 1, 1), (1, 2, 1), (1, 1, 2)],
}


def factd(n: int) -> int:
    """Double factorial with factd(n<=0) = 1 (src/molden_reader_nikola.py:347-351)."""
    r = 1
    while n > 0:
        r *= n
        n -= 2
    return r


def gto_norm(ga: float, l: int, m: int, n: int) -> float:
    """Primitive cartesian normalisation (src/molden_reader_nikola.py:53-61)."""
    normA = (2 / math.pi) ** 0.75
    normB = 2 ** (l + m + n)
    normC = ga ** ((2 * l + 2 * m + 2 * n + 3) / 4.0)
    normD = factd(2 * l - 1) * factd(2 * m - 1) * factd(2 * n - 1)
    return normA * normB * normC / normD ** 0.5


@dataclass
class Atom:
    atomic_number: int
    x: float
    y: float
    z: float


@dataclass
class Molecule:
    """Atoms in bohr (src/molecule.py:89-104 keeps the same fields)."""
    atoms: list = field(default_factory=list)

    def atomic_numbers(self) -> np.ndarray:
        return np.array([a.atomic_number for a in self.atoms], dtype=np.float64)

    def geometry(self) -> np.ndarray:
        return np.array([[a.x, a.y, a.z] for a in self.atoms], dtype=np.float64)


@dataclass
class _Prim:
    atom_idx: int
    contraction: int | None
    group: int
    ga: float
    c: float | None
    l: int
    m: int
    n: int
    mo: object = None
    x: float | None = None
    y: float | None = None
    z: float | None = None


@dataclass
class ArrayGTO:
    """Primitive cartesian GTOs and their MO coefficients (mo: ngto x nMO)."""
    mo: np.ndarray
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    ga: np.ndarray
    l: np.ndarray
    m: np.ndarray
    n: np.ndarray
    group: np.ndarray
    occ: np.ndarray | None = None
    energy: np.ndarray | None = None
    syms: np.ndarray | None = None


# Per-program rules of the molden readers.  atoms: how the [Atoms] block
# starts and ends, and whether its coordinates are divided by AU2ANG; gto_end:
# what ends the [GTO] block after a blank line; shells: the shell letters the
# program's reader expands; mo: the [MO] block rules (see _read_mo).
FLAVOURS = {
    # src/molden_reader_nikola.py:91-115,163-177,258-344
    "molpro": dict(atoms_start=lambda ln: "[Atoms]" in ln, atoms_end=lambda ln: "[GTO]" in ln,
                   to_bohr=True, gto_end="blank", shells="spdfg", mo="molpro"),
    # src/molden_reader_nikola_molcas.py:92-113,166-183,274-323
    "molcas": dict(atoms_start=lambda ln: "[ATOMS]" in ln.upper(), atoms_end=lambda ln: "[" in ln,
                   to_bohr=False, gto_end="[MO]", shells="spdf", mo="molcas"),
    # src/molden_reader_nikola_pyscf.py:109-111,168-181,296-344
    "pyscf": dict(atoms_start=lambda ln: "[Atoms]" in ln, atoms_end=lambda ln: "[GTO]" in ln,
                  to_bohr=False, gto_end="[6d]", shells="spdfg", mo="pyscf"),
    # src/molden_reader_nikola_bagel.py:222,279-324
    "bagel": dict(atoms_start=lambda ln: "[Atoms]" in ln, atoms_end=lambda ln: "[GTO]" in ln,
                  to_bohr=True, gto_end="blank", shells="spdf", mo="bagel"),
}


def _flavour(flavour):
    try:
        return FLAVOURS[flavour]
    except KeyError:
        raise ValueError(f"molden flavour must be one of {sorted(FLAVOURS)}, not {flavour!r}")


def _read_atoms(lines, flavour="molpro"):
    """[Atoms] block (src/molden_reader_nikola.py:80-115 and the copies'
    variants): the Molpro and BAGEL readers divide every coordinate by AU2ANG
    whatever the block's unit tag, the Molcas and PySCF readers take it as
    bohr."""
    fl = _flavour(flavour)
    it = iter(lines)
    for line in it:
        if fl["atoms_start"](line):
            break
    else:
        raise ValueError("molden: no [Atoms] section")
    scale = 1.0 / AU2ANG if fl["to_bohr"] else 1.0
    mol = Molecule()
    for line in it:
        if fl["atoms_end"](line):
            break
        cols = line.split()
        if fl["to_bohr"]:
            mol.atoms.append(Atom(int(cols[2]), float(cols[3]) / AU2ANG,
                                  float(cols[4]) / AU2ANG, float(cols[5]) / AU2ANG))
        else:
            mol.atoms.append(Atom(int(cols[2]), float(cols[3]) * scale,
                                  float(cols[4]) * scale, float(cols[5]) * scale))
    else:
        raise ValueError("molden: no [GTO] section")
    return mol


def _read_contractions(lines, flavour="molpro"):
    """[GTO] block -> list of primitives (src/molden_reader_nikola.py:133-255).

    One group per primitive shared by its cartesian components; the
    contraction index of each component is its MO-table row.  A blank line
    followed by an atom line starts a new atom.  What else may follow a blank
    line: another blank line (Molpro, BAGEL: skipped), the [MO] tag (Molcas:
    the block ends) or a [6d] tag (PySCF: the block ends).
    """
    fl = _flavour(flavour)
    i = 0
    while i < len(lines) and "[GTO]" not in lines[i]:
        i += 1
    if i == len(lines):
        raise ValueError("molden: no [GTO] section")
    i += 1
    atm = [int(s) for s in lines[i].split() if s.isdigit()][0]
    i += 1
    gtos = []
    contraction_counter = 1
    group_counter = 0
    while i < len(lines):
        line = lines[i]
        i += 1
        if "[MO]" in line:
            break
        if not line or line.isspace():
            nxt = lines[i] if i < len(lines) else ""
            i += 1
            if fl["gto_end"] == "[MO]" and "[MO]" in nxt:
                break
            if fl["gto_end"] == "[6d]" and "[6d]" in nxt:
                break
            if fl["gto_end"] != "[MO]" and (not nxt or nxt.isspace()):
                continue
            atm = [int(s) for s in nxt.split() if s.isdigit()][0]
            line = lines[i]
            i += 1
        spec = line.split()
        kind = spec[0].lower()
        nprim = int(spec[1])
        if kind not in _ANG or kind not in fl["shells"]:
            # the reference copies skip the components of a letter they do
            # not expand, which misaligns the MO rows: refuse instead
            raise ValueError(f"molden ({flavour}): unsupported shell '{spec[0]}'")
        ang = _ANG[kind]
        for _ in range(nprim):
            pl = lines[i].replace("D", "E")
            i += 1
            g = float(pl.split()[0])
            c = float(pl.split()[1])
            group_counter += 1
            for k, (l, m, n) in enumerate(ang):
                gtos.append(_Prim(atm, contraction_counter + k, group_counter,
                                  g, c, l, m, n))
        contraction_counter += len(ang)
    return gtos


_FLOAT_RE = re.compile(r"\d+\.\d+")


def _read_mo(lines, mo_cutoff, flavour="molpro"):
    """[MO] block (src/molden_reader_nikola.py:258-344 and the copies).

    Molpro: reads up to ``mo_cutoff`` orbitals, keeps the occupied subset and
    orders it by the key frac(sym)*1000 + int(sym) (numpy argsort); the
    permutation of the occupied subset is applied to the leading rows of the
    MO table, energies and occupations, as the reference does.
    PySCF (:296-344): the same keys and permutation, applied to the energies
    and occupations only; every MO is kept in file order.
    Molcas (:274-323): the key of an orbital is the first character of its
    symmetry label; every MO is kept in file order, energies and occupations
    are permuted, and the keys are returned as the labels.
    BAGEL (:279-324): an orbital starts at its Ene line; nothing is
    reordered and the labels are 0..len(last MO)-1.
    """
    mode = _flavour(flavour)["mo"]
    i = 0
    while i < len(lines) and "[MO]" not in lines[i]:
        i += 1
    if i == len(lines):
        raise ValueError("molden: no [MO] section")
    i += 1
    start = "Ene" if mode == "bagel" else "Sym"
    syms, energy, occ, table = [], [], [], []
    mo_counter = 0
    this = []
    while i < len(lines) and mo_counter <= mo_cutoff and "[FREQ]" not in lines[i]:
        line = lines[i]
        if start in line:
            mo_counter += 1
            if mo_counter <= mo_cutoff:
                if mode == "molcas":
                    syms.append(line.split()[1][0])
                else:
                    s = _FLOAT_RE.findall(line)
                    syms.append(s[0] if s else 1.1)
            if this:
                table.append(this)
            this = []
            if mode == "bagel":
                energy.append(float(_FLOAT_RE.findall(line)[0]))
        elif "Ene" in line:
            energy.append(float(_FLOAT_RE.findall(line)[0]))
        elif "Spin" in line:
            pass
        elif "Occup" in line:
            occ.append(float(_FLOAT_RE.findall(line)[0]))
        elif line.strip():
            this.append(float(line.split()[1]))
        i += 1
    if this:
        table.append(this)
    mo = np.array(table)
    if mode == "bagel":
        return mo.T, np.array(occ), np.array(energy), np.arange(len(this))
    syms_array = np.array([float(s) for s in syms])
    if mode == "molcas":
        keys = (syms_array - syms_array.astype(int)) * 1000 + syms_array.astype(int)
        idx = np.argsort(keys)
        return mo.T, np.array(occ)[idx], np.array(energy)[idx], keys
    occ_array = np.array([float(o) for o in occ])
    finalind = np.where(occ_array > 0.0)
    keys = syms_array[finalind]
    keys = (keys - keys.astype(int)) * 1000 + keys.astype(int)
    idx = np.argsort(keys)
    if mode == "molpro":
        mo = mo[idx, :]
    energy_array = np.array(energy)[idx]
    occ_array = np.array(occ)[idx]
    return mo.T, occ_array, energy_array, syms_array


def _reverse_contraction(gtos):
    """Merge primitives on the same atom with the same exponent and l,m,n by
    summing MO vectors, then renumber groups 1,2,3,... by changes between
    consecutive entries (src/molden_reader_nikola.py:382-422)."""
    first = copy.copy(gtos[0])
    first.mo = 0
    contracted = [first]
    for g in gtos:
        for j in contracted:
            if (g.atom_idx == j.atom_idx and g.ga == j.ga and g.l == j.l
                    and g.m == j.m and g.n == j.n):
                j.mo = j.mo + g.mo
                break
        else:
            t = copy.copy(g)
            t.contraction = None
            t.c = None
            t.mo = g.mo
            contracted.append(t)
    counter = 1
    prev = contracted[0].group
    for k in range(1, len(contracted)):
        if contracted[k].group != prev:
            counter += 1
            prev = contracted[k].group
        contracted[k].group = counter
    return contracted


def read_orbitals(file, N=10000, decontract=False, flavour="molpro"):
    """Primitive GTOs + MOs from a molden file.

    Mirrors ``read_orbitals`` (src/molden_reader_nikola.py:496-508 and the
    same function of the Molcas, PySCF and BAGEL copies, selected by
    ``flavour``): returns ``(ArrayGTO, Molecule)`` when given a file name,
    ``ArrayGTO`` when given an open file.
    """
    if isinstance(file, str):
        with open(file, "r") as f:
            text = f.read()
        lines = text.splitlines(keepends=True)
        return _read_mo_gtos(lines, N, decontract, flavour), _read_atoms(lines, flavour)
    text = file.read()
    return _read_mo_gtos(text.splitlines(keepends=True), N, decontract, flavour)


def read_occupied_orbitals(file, decontract=False, flavour="molpro"):
    """``read_orbitals`` without the orbitals of zero occupation
    (read_occupied_orbitals, src/molden_reader_nikola_molcas.py:504-518)."""
    if isinstance(file, str):
        with open(file, "r") as f:
            arr = _read_mo_gtos(f.read().splitlines(keepends=True), 10000, decontract, flavour)
    else:
        arr = _read_mo_gtos(file.read().splitlines(keepends=True), 10000, decontract, flavour)
    drop = np.where(arr.occ == 0)[0]
    arr.occ = np.delete(arr.occ, drop)
    arr.energy = np.delete(arr.energy, drop)
    arr.mo = np.delete(arr.mo, drop, 1)
    arr.syms = np.delete(arr.syms, drop)
    return arr


def _read_mo_gtos(lines, N, decontract, flavour="molpro"):
    mol = _read_atoms(lines, flavour)
    gtos = _read_contractions(lines, flavour)
    mo_table, occ, energy, syms = _read_mo(lines, N, flavour)
    for g in gtos:
        g.mo = mo_table[g.contraction - 1, :]
    for g in gtos:
        g.mo = g.c * gto_norm(g.ga, g.l, g.m, g.n) * g.mo
        g.c = None
    if decontract:
        gtos = _reverse_contraction(gtos)
    for g in gtos:
        a = mol.atoms[g.atom_idx - 1]
        g.x, g.y, g.z = a.x, a.y, a.z
    ng = len(gtos)
    arr = ArrayGTO(
        mo=np.array([np.asarray(g.mo, dtype=np.float64) for g in gtos]).reshape(ng, -1),
        x=np.array([g.x for g in gtos], dtype=np.float64),
        y=np.array([g.y for g in gtos], dtype=np.float64),
        z=np.array([g.z for g in gtos], dtype=np.float64),
        ga=np.array([g.ga for g in gtos], dtype=np.float64),
        l=np.array([g.l for g in gtos], dtype=np.int32),
        m=np.array([g.m for g in gtos], dtype=np.int32),
        n=np.array([g.n for g in gtos], dtype=np.int32),
        group=np.array([g.group for g in gtos], dtype=np.int32),
    )
    arr.occ, arr.energy, arr.syms = occ, energy, syms
    return arr


# ----------------------------------------------------------------- CI punch
def twordmconst(closed, fileread):
    """CI vectors and determinant strings from a Molpro punch-style file.

    Restates the no-symmetry branch of ``twordmconst`` (src/twordm_red.py:
    12-86); the multi-irrep branches are not part of this path and raise.
    Returns ``(civs, confs)``: civs as lists of coefficient strings per
    determinant and confs as spin-orbital strings ('a' = alpha, 'b' = beta,
    '0' = empty, two characters per orbital).
    """
    if isinstance(closed, list):
        ncirr = len(closed)
        if ncirr == 1:
            closed = closed[0]
    else:
        ncirr = 1
    firstline = None
    with open(fileread, "r") as fh:
        for line in fh:
            if line[0].isupper():
                continue
            if line[0] not in "*?-":
                firstline = line.strip().split()
                break
    if firstline is None:
        raise ValueError("punch file holds no determinant lines")
    nirr = -1
    for element in firstline:
        nirr += 1
        if "." in element:
            break
    if ncirr != nirr:
        raise ValueError(
            f"mismatch between the closed orbital entry and the number of irreps "
            f"in the CI vector: {nirr} entries are required")
    if nirr != 1:
        raise NotImplementedError("symmetry-blocked CI vectors are not supported")
    civs, confs = [], []
    with open(fileread, "r") as fh:
        for line in fh:
            if line[0].isupper() or line[0] in "*?-":
                continue
            NN = line.strip().split()
            if not NN:
                continue
            if "." not in NN[0]:
                sdet = "c" * closed + NN[0]
                sdet = (sdet.replace("0", "00").replace("a", "a0")
                        .replace("b", "0b").replace("2", "ab"))
                sdet = sdet.replace("c", "ab")
                confs.append(sdet)
                civs.append(NN[1:])
            else:
                civs[-1].extend(NN)
    return civs, confs


def confs_to_int(confs) -> np.ndarray:
    """'a' -> 1, 'b' -> 2 per character (src/scattering_main_routine.py:232-238)."""
    out = np.zeros((len(confs), len(confs[0])), dtype=np.int32)
    for i, s in enumerate(confs):
        out[i, :] = np.asarray(list(s.replace("a", "1").replace("b", "2")),
                               dtype=np.int32)
    return out


def civs_to_array(civs) -> np.ndarray:
    return np.array([[float(v) for v in row] for row in civs], dtype=np.float64)


# ----------------------------------------------------------- 2-RDM files
def read_twordm(filename):
    """``p s q r value`` rows (src/one_rdm.f90:2112-2146) -> (mat int32[n,4], total f64[n])."""
    mats, tots = [], []
    with open(filename, "r") as fh:
        for line in fh:
            tok = line.split()
            if not tok:
                continue
            mats.append([int(t) for t in tok[:4]])
            tots.append(float(tok[4].replace("D", "E")))
    return np.array(mats, dtype=np.int32).reshape(-1, 4), np.array(tots, dtype=np.float64)


def read_onerdm(filename):
    """``i j value`` rows (src/one_rdm.f90:2072-2108): maxnmo is the first
    index of the last row; unlisted elements are 0."""
    rows = []
    with open(filename, "r") as fh:
        for line in fh:
            tok = line.split()
            if tok:
                rows.append((int(tok[0]), int(tok[1]), float(tok[2].replace("D", "E"))))
    nmo = rows[-1][0] if rows else 0
    g = np.zeros((nmo, nmo), order="F")
    for i, j, v in rows:
        g[i - 1, j - 1] = v
    return g


# -------------------------------------------------- bit-string CI files
def read_bit_file(filename, numberlines=None):
    """``i c1 c2 alpha beta`` lines, the layout mcci_to_bit and
    createtwordm_bit read (src/one_rdm.f90:979, :1914): int64 alpha and beta
    bit strings (orbital k on bit k-1) and float64 coefficient columns.
    numberlines, when given, reads only that many lines, as the reference
    does."""
    al, be, c1, c2 = [], [], [], []
    with open(filename, "r") as fh:
        for line in fh:
            if numberlines is not None and len(al) >= numberlines:
                break
            tok = line.replace(",", " ").split()
            if not tok:
                continue
            c1.append(float(tok[1].replace("D", "E")))
            c2.append(float(tok[2].replace("D", "E")))
            al.append(int(tok[3]))
            be.append(int(tok[4]))
    if numberlines is not None and len(al) < numberlines:
        raise ValueError(f"{filename}: {len(al)} lines, {numberlines} expected")
    return (np.array(al, dtype=np.int64), np.array(be, dtype=np.int64),
            np.array(c1, dtype=np.float64), np.array(c2, dtype=np.float64))


def read_fci_file(filename, numberlines=None):
    """``i civ alpha beta`` lines, the FCI file createtwordm_bit_fci reads
    (src/one_rdm.f90:443-446; written at src/scattering_main_routine.py:302-308)."""
    al, be, c = [], [], []
    with open(filename, "r") as fh:
        for line in fh:
            if numberlines is not None and len(al) >= numberlines:
                break
            tok = line.split()
            if not tok:
                continue
            c.append(float(tok[1].replace("D", "E")))
            al.append(int(tok[2]))
            be.append(int(tok[3]))
    if numberlines is not None and len(al) < numberlines:
        raise ValueError(f"{filename}: {len(al)} lines, {numberlines} expected")
    return (np.array(al, dtype=np.int64), np.array(be, dtype=np.int64),
            np.array(c, dtype=np.float64))


def fci_tables(alpha):
    """newdat(:,1), irep, start and end as the driver builds them
    (src/scattering_main_routine.py:269-289): the unique alpha strings in
    increasing order, their determinant counts and their 1-based first and
    last positions."""
    alpha = np.asarray(alpha, dtype=np.int64)
    uniq, first, counts = np.unique(alpha, return_index=True, return_counts=True)
    last = np.array([np.nonzero(alpha == u)[0][-1] for u in uniq], dtype=np.int64)
    return uniq, counts.astype(np.int64), first.astype(np.int64) + 1, last + 1


def write_bit_file(filename, alpha, beta, c1, c2):
    """The same layout, coefficients written to round-trip exactly."""
    with open(filename, "w") as fh:
        for i, (a, b, x, y) in enumerate(zip(alpha, beta, c1, c2)):
            fh.write(f"{i + 1} {float(x)!r} {float(y)!r} {int(a)} {int(b)}\n")


def write_twordm(filename, mat, total):
    """Same column layout as twordm_fortran.dat (src/twordms.f90:521-539)."""
    with open(filename, "w") as fh:
        for row, v in zip(np.asarray(mat), np.asarray(total)):
            fh.write(f"{int(row[0]):12d}{int(row[1]):12d}{int(row[2]):12d}{int(row[3]):12d}   {float(v)!r}\n")


# ------------------------------------------------------- Main.f90 text route
def write_basis_dat(path, xx, yy, zz, ga, l, m, n, group):
    with open(path, "w") as f:
        f.write(f"{len(l)}\n")
        for i in range(len(l)):
            f.write(f"{float(xx[i])!r} {float(yy[i])!r} {float(zz[i])!r} {float(ga[i])!r} "
                    f"{int(l[i])} {int(m[i])} {int(n[i])} {int(group[i])}\n")


def read_basis_dat(path):
    """basis.dat: ngtos, then x y z ga l m n group per line (src/Main.f90:28-36)."""
    with open(path) as f:
        tok = f.read().split()
    ng = int(tok[0])
    rows = np.array(tok[1:1 + 8 * ng], dtype=object).reshape(ng, 8)
    xx, yy, zz, ga = (rows[:, k].astype(np.float64) for k in range(4))
    l, m, n, group = (rows[:, k].astype(np.int32) for k in range(4, 8))
    return xx, yy, zz, ga, l, m, n, group


def write_mos_dat(path, mmod):
    mmod = np.asarray(mmod)
    with open(path, "w") as f:
        f.write(f"{mmod.shape[0]} {mmod.shape[1]}\n")
        for i in range(mmod.shape[0]):
            f.write(" ".join(repr(float(v)) for v in mmod[i]) + "\n")


def read_mos_dat(path):
    """MOs.dat: norbs ngtos, then one MO per row (src/Main.f90:40-48)."""
    with open(path) as f:
        tok = f.read().split()
    norbs, ngtos = int(tok[0]), int(tok[1])
    return np.array(tok[2:2 + norbs * ngtos], dtype=np.float64).reshape(norbs, ngtos)


def linspace_1(qmin, qmax, npoints):
    """src/linspace.f90:5-22: from + range*(i-1)/(n-1)."""
    rng = qmax - qmin
    return np.array([qmin + rng * (i - 1) / (npoints - 1) for i in range(1, npoints + 1)])


@dataclass
class Options:
    atoms: np.ndarray
    geom: np.ndarray
    cutoffcentre: float
    cutoffz: float
    cutoffmd: float
    jeremyR: bool
    mcci: bool
    hf: bool
    qmin: float
    qmax: float
    npoints: int
    typec: int
    state1: int
    state2: int
    state3: int
    file_out: str
    molpro: bool
    molcas: bool
    bagel: bool
    confs: np.ndarray | None = None
    civs: np.ndarray | None = None


def _flog(s: str) -> bool:
    s = s.strip().strip(".").upper()
    return s.startswith("T")


def read_options_dat(path) -> Options:
    """options.dat positional records (src/Main.f90:47-112), list-directed."""
    with open(path) as f:
        recs = [ln.strip() for ln in f if ln.strip()]
    k = 0

    def nxt():
        nonlocal k
        k += 1
        return recs[k - 1]

    natoms = int(nxt().split()[0])
    atoms = np.array(nxt().split()[:natoms], dtype=np.float64)
    geom = np.array([nxt().split()[:3] for _ in range(natoms)], dtype=np.float64)
    cc = float(nxt().split()[0].replace("D", "E"))
    cz = float(nxt().split()[0].replace("D", "E"))
    cm = float(nxt().split()[0].replace("D", "E"))
    jr = _flog(nxt())
    mc = _flog(nxt())
    hf = _flog(nxt())
    qt = nxt().split()
    typec = int(nxt().split()[0])
    st = [int(v) for v in nxt().split()[:3]]
    fout = nxt().split()[0].strip("'\"")
    molpro, molcas, bagel = _flog(nxt()), _flog(nxt()), _flog(nxt())
    opt = Options(atoms, geom, cc, cz, cm, jr, mc, hf, float(qt[0]), float(qt[1]),
                  int(qt[2]), typec, st[0], st[1], st[2], fout, molpro, molcas, bagel)
    if molpro or molcas or bagel:
        nconfs = int(nxt().split()[0])
        lconfs = int(nxt().split()[0])
        ncivs = int(nxt().split()[0])
        if molpro and nconfs >= 2:
            raise NotImplementedError("bitwise determinant files are not part of this path")
        vals = " ".join(recs[k:]).split()
        need = nconfs * (lconfs + ncivs)
        arr = np.array(vals[:need], dtype=np.float64).reshape(nconfs, lconfs + ncivs)
        opt.confs = arr[:, :lconfs].astype(np.int32)
        opt.civs = arr[:, lconfs:]
    return opt


def write_result(path, q, result):
    """'q I' per line (src/Main.f90:192-198)."""
    with open(path, "w") as f:
        for a, b in zip(q, result):
            f.write(f" {float(a)!r} {float(b)!r}\n")
