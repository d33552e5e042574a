"""Seeded synthetic inputs for the scattering path.

Each workload is written the way the reference consumes it: a Molpro-style
molden file (Angstrom coordinates, ``Sym= n.1`` labels, ``Occup`` lines) and a
CI punch-style text file, which ``thistle_amd.readers`` then parses exactly
as the reference readers do.  MOs are random and Loewdin-orthonormal in the
overlap of the contracted cartesian basis, so that the two-electron sum rule
S(q->0) = N(N-1) holds; CI vectors are random over all Ms = 0 determinants
of a CAS, plus doubly occupied core orbitals.

Workloads (sizes are those of the named benchmark cases):
  h2    H2 on z at +-0.7 bohr, per atom 2 s + 1 p contracted shells of 3
        primitives (10 functions), CAS(2,2), 100 q in [0, 8]
  xyx   linear X-Y-X, bonds 2.2 bohr, per atom 4s2p2d uncontracted
        (66 functions), 3 core + CAS(8,8), 500 q in [0, 10]
  xy6   octahedral XY6, X-Y 2.95 bohr, X 4s3p2d1f + Y 4s2p2d1f uncontracted
        (227 functions), 20 core + CAS(10,10), 512 q in [0, 12]
  mol10 10 atoms, min separation 2.0 bohr, 3s1p1d per atom (120 functions),
        CAS(6,6), 256 q in [0, 10]
  batch 512 frames of 2-12 atoms (2s1p1d per atom), CAS(2,2)-(6,6),
        64-512 q in [0, 10], coordinates jittered by 0.1 bohr
"""
from __future__ import annotations

import itertools
import math
from dataclasses import dataclass

import numpy as np

from .readers import AU2ANG, _ANG, gto_norm


# ------------------------------------------------------------ overlaps
def _e0(la: int, lb: int, PA: float, PB: float, a: float) -> float:
    """E_0^{la,lb} by the McMurchie-Davidson recursion (textbook)."""
    E = {(0, 0, 0): 1.0}

    def get(i, j, t):
        if t < 0 or t > i + j:
            return 0.0
        return E[(i, j, t)]

    for i in range(1, la + 1):
        for t in range(i + 1):
            E[(i, 0, t)] = a * get(i - 1, 0, t - 1) + PA * get(i - 1, 0, t) + (t + 1) * get(i - 1, 0, t + 1)
    for j in range(1, lb + 1):
        for t in range(la + j + 1):
            E[(la, j, t)] = a * get(la, j - 1, t - 1) + PB * get(la, j - 1, t) + (t + 1) * get(la, j - 1, t + 1)
    return E[(la, lb, 0)]


def prim_overlap(ga, la, A, gb, lb, B):
    """<x^l y^m z^n e^{-a r^2} | ...> for unnormalised cartesian primitives."""
    p = ga + gb
    mu = ga * gb / p
    P = (ga * np.asarray(A) + gb * np.asarray(B)) / p
    AB2 = float(np.sum((np.asarray(A) - np.asarray(B)) ** 2))
    s = (math.pi / p) ** 1.5 * math.exp(-mu * AB2)
    for ax in range(3):
        s *= _e0(la[ax], lb[ax], P[ax] - A[ax], P[ax] - B[ax], 0.5 / p)
    return s


# ------------------------------------------------------------ structures
@dataclass
class Shell:
    atom: int          # 0-based atom index
    kind: str          # 's', 'p', 'd', 'f'
    exps: np.ndarray   # primitive exponents
    coefs: np.ndarray  # contraction coefficients


@dataclass
class Workload:
    name: str
    Z: np.ndarray          # atomic numbers
    coords: np.ndarray     # bohr, (natom, 3)
    shells: list
    mo: np.ndarray         # (nbf, nmo) coefficients on contracted functions
    ncore: int
    nact: int
    nelec_act: int
    civec: np.ndarray      # (ndet,)
    dets: list             # occupation strings over the active orbitals
    q: np.ndarray
    cutoff2rdm: float = 1e-30


def basis_functions(shells):
    """Contracted cartesian functions in molden order: (shell, (l,m,n))."""
    out = []
    for sh in shells:
        for lmn in _ANG[sh.kind]:
            out.append((sh, lmn))
    return out


def overlap_matrix(shells, coords):
    bfs = basis_functions(shells)
    nb = len(bfs)
    S = np.zeros((nb, nb))
    for i, (si, li) in enumerate(bfs):
        for j in range(i, nb):
            sj, lj = bfs[j]
            v = 0.0
            for ea, ca in zip(si.exps, si.coefs):
                na = gto_norm(ea, *li)
                for eb, cb in zip(sj.exps, sj.coefs):
                    nb_ = gto_norm(eb, *lj)
                    v += ca * na * cb * nb_ * prim_overlap(
                        ea, li, coords[si.atom], eb, lj, coords[sj.atom])
            S[i, j] = S[j, i] = v
    return S


def lowdin_mos(rng, S, nmo):
    nb = S.shape[0]
    C0 = rng.standard_normal((nb, nmo))
    M = C0.T @ S @ C0
    w, V = np.linalg.eigh(M)
    return C0 @ (V @ np.diag(w ** -0.5) @ V.T)


def cas_dets(nact, nelec):
    """All Ms = 0 determinants of CAS(nelec, nact): alpha-string outer."""
    na = nelec // 2
    strings = list(itertools.combinations(range(nact), na))
    dets = []
    for a in strings:
        for b in strings:
            s = []
            for o in range(nact):
                ia, ib = o in a, o in b
                s.append("2" if ia and ib else "a" if ia else "b" if ib else "0")
            dets.append("".join(s))
    return dets


def _shells_for(atom, spec, rng, lo, hi, nprim=1):
    """spec like {'s': 4, 'p': 2}; exponents log-uniform in [lo, hi]."""
    out = []
    for kind in "spdf":
        for _ in range(spec.get(kind, 0)):
            e = np.exp(rng.uniform(np.log(lo), np.log(hi), size=nprim))
            c = np.ones(nprim) if nprim == 1 else rng.standard_normal(nprim)
            out.append(Shell(atom, kind, np.sort(e)[::-1].copy(), c))
    return out


def _finish(name, rng, Z, coords, shells, ncore, nact, nelec_act, q, cutoff=1e-30):
    S = overlap_matrix(shells, coords)
    nmo = ncore + nact
    mo = lowdin_mos(rng, S, nmo)
    dets = cas_dets(nact, nelec_act)
    civ = rng.standard_normal(len(dets))
    civ /= np.linalg.norm(civ)
    return Workload(name, np.asarray(Z), np.asarray(coords, dtype=np.float64),
                    shells, mo, ncore, nact, nelec_act, civ, dets,
                    np.asarray(q, dtype=np.float64), cutoff)


def qgrid(qmax, nq):
    q = np.linspace(0.0, qmax, nq)
    if q[0] < 1e-10:
        q[0] = 1e-10  # src/scattering_main_routine.py:185-187
    return q


def make_h2(seed=1):
    rng = np.random.default_rng(seed)
    coords = np.array([[0.0, 0.0, -0.7], [0.0, 0.0, 0.7]])
    shells = []
    for a in range(2):
        shells += _shells_for(a, {"s": 2, "p": 1}, rng, 0.1, 20.0, nprim=3)
    return _finish("h2", rng, [1, 1], coords, shells, 0, 2, 2, qgrid(8.0, 100))


def make_xyx(seed=2, nact=8, nelec=8, ncore=3, nq=500, cutoff=1e-12):
    rng = np.random.default_rng(seed)
    coords = np.array([[0.0, 0.0, -2.2], [0.0, 0.0, 0.0], [0.0, 0.0, 2.2]])
    shells = []
    for a in range(3):
        shells += _shells_for(a, {"s": 4, "p": 2, "d": 2}, rng, 0.08, 40.0)
    return _finish("xyx", rng, [8, 6, 8], coords, shells, ncore, nact, nelec,
                   qgrid(10.0, nq), cutoff)


def make_far(seed=7, sep=12.0, nq=64, qmax=10.0):
    """Two atoms `sep` bohr apart and one between them, s/p/d shells, 1 core +
    CAS(2,2): q h reaches past 74 (q h up to qmax * sep), where the reference's
    dphrec, started at order 100, departs from the exact j_beta."""
    rng = np.random.default_rng(seed)
    coords = np.array([[0.0, 0.0, 0.0], [0.0, 0.3 * sep, 0.5 * sep], [0.0, 0.0, sep]])
    shells = []
    for a in range(3):
        shells += _shells_for(a, {"s": 2, "p": 1, "d": 1}, rng, 0.1, 20.0)
    return _finish("far", rng, [1, 1, 1], coords, shells, 1, 2, 2, qgrid(qmax, nq))


def make_xy6(seed=3, nact=10, nelec=10, ncore=20, nq=512):
    rng = np.random.default_rng(seed)
    d = 2.95
    coords = np.array([[0, 0, 0], [d, 0, 0], [-d, 0, 0], [0, d, 0], [0, -d, 0],
                       [0, 0, d], [0, 0, -d]], dtype=np.float64)
    shells = _shells_for(0, {"s": 4, "p": 3, "d": 2, "f": 1}, rng, 0.08, 40.0)
    for a in range(1, 7):
        shells += _shells_for(a, {"s": 4, "p": 2, "d": 2, "f": 1}, rng, 0.08, 40.0)
    return _finish("xy6", rng, [16] + [9] * 6, coords, shells, ncore, nact,
                   nelec, qgrid(12.0, nq))


def random_molecule(rng, natom, dmin=2.0, box=None):
    box = box or 1.6 * natom ** (1 / 3) * dmin
    pts = []
    while len(pts) < natom:
        p = rng.uniform(-box / 2, box / 2, size=3)
        if all(np.linalg.norm(p - o) >= dmin for o in pts):
            pts.append(p)
    return np.array(pts)


def make_mol10(seed=4, nq=256):
    rng = np.random.default_rng(seed)
    coords = random_molecule(rng, 10)
    shells = []
    for a in range(10):
        shells += _shells_for(a, {"s": 3, "p": 1, "d": 1}, rng, 0.08, 40.0)
    Z = rng.choice([1, 6, 7, 8], size=10)
    return _finish("mol10", rng, Z, coords, shells, 0, 6, 6, qgrid(10.0, nq))


def make_batch(nframes=512, seed=6, natom_range=(2, 12), nq_range=(64, 512), qmax=10.0,
               sigma=0.1):
    """Frames of uneven size for the batched launch: per frame 2-12 atoms
    (a seeded base geometry per size, jittered by N(0, sigma^2) bohr), per atom
    2 s + 1 p + 1 d uncontracted cartesian shells, CAS(2,2), CAS(4,4) or
    CAS(6,6) with a random CI vector, and 64-512 q points in [0, qmax]."""
    rng = np.random.default_rng(seed)
    bases = {n: random_molecule(rng, n) for n in range(natom_range[0], natom_range[1] + 1)}
    frames = []
    for f in range(nframes):
        natom = int(rng.integers(natom_range[0], natom_range[1] + 1))
        coords = bases[natom] + rng.normal(scale=sigma, size=(natom, 3))
        shells = []
        for a in range(natom):
            shells += _shells_for(a, {"s": 2, "p": 1, "d": 1}, rng, 0.1, 30.0)
        nact = int(rng.choice([2, 4, 6]))
        nq = int(rng.integers(nq_range[0], nq_range[1] + 1))
        frames.append(_finish(f"frame{f}", rng, rng.choice([1, 6, 7, 8], size=natom), coords,
                              shells, 0, nact, nact, qgrid(qmax, nq)))
    return frames


def make_small(seed=5, natom=2, spec=None, nact=2, nelec=2, ncore=0, nq=16,
               qmax=6.0, bond=1.6, nprim=1):
    """Small custom case for unit tests."""
    rng = np.random.default_rng(seed)
    spec = spec or {"s": 1, "p": 1}
    coords = random_molecule(rng, natom, dmin=bond) if natom > 2 else \
        np.array([[0.0, 0.0, -bond / 2], [0.0, 0.0, bond / 2]])[:natom]
    shells = []
    for a in range(natom):
        shells += _shells_for(a, spec, rng, 0.2, 8.0, nprim=nprim)
    return _finish("small", rng, [1] * natom, coords, shells, ncore, nact, nelec,
                   qgrid(qmax, nq))


# ----------------------------------------------------------------- writers
_SYM = {1: "H", 6: "C", 7: "N", 8: "O", 9: "F", 16: "S"}


def molden_text(w: Workload) -> str:
    lines = ["[Molden Format]", "[Atoms] Angs"]
    for i, (z, c) in enumerate(zip(w.Z, w.coords)):
        ang = c * AU2ANG
        lines.append(f" {_SYM.get(int(z), 'X')} {i + 1:6d} {int(z):6d} "
                     f"{float(ang[0])!r} {float(ang[1])!r} {float(ang[2])!r}")
    lines.append("[GTO]")
    first = True
    for a in range(len(w.Z)):
        if not first:
            lines.append("")
        first = False
        lines.append(f"  {a + 1} 0")
        for sh in [s for s in w.shells if s.atom == a]:
            lines.append(f" {sh.kind}   {len(sh.exps)} 1.00")
            for e, c in zip(sh.exps, sh.coefs):
                lines.append(f"   {float(e)!r}  {float(c)!r}")
    lines += ["", "", "[MO]"]
    nmo = w.mo.shape[1]
    for k in range(nmo):
        occ = 2.0 if k < w.ncore else 1.0
        lines.append(f" Sym= {k + 1}.1")
        lines.append(f" Ene= {-1.0 + 0.01 * k:.6f}")
        lines.append(" Spin= Alpha")
        lines.append(f" Occup= {occ:.6f}")
        for mu in range(w.mo.shape[0]):
            lines.append(f" {mu + 1:4d} {float(w.mo[mu, k])!r}")
    return "\n".join(lines) + "\n"


def punch_text(w: Workload) -> str:
    lines = ["MCSCF STATE 1.1 Energy -1.0", "* synthetic CI vector"]
    for d, c in zip(w.dets, w.civec):
        lines.append(f" {d}   {float(c)!r}")
    return "\n".join(lines) + "\n"


def write_inputs(w: Workload, directory) -> tuple:
    import os
    os.makedirs(directory, exist_ok=True)
    mld = os.path.join(directory, f"{w.name}.molden")
    pun = os.path.join(directory, f"{w.name}.pun")
    with open(mld, "w") as f:
        f.write(molden_text(w))
    with open(pun, "w") as f:
        f.write(punch_text(w))
    return mld, pun


WORKLOADS = {"h2": make_h2, "xyx": make_xyx, "xy6": make_xy6, "mol10": make_mol10}
