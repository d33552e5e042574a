"""BAGEL CASSCF / XMS-CASPT2 CI vectors in the determinant basis of the path.

Restates ``src/bagel_ci_reader.f90``: the log prints each root's CI vector
as determinant strings over the active orbitals ('2' doubly occupied, 'a',
'b' singly occupied, '.' empty) with their coefficients; with XMS-CASPT2 the
roots are rotated by the printed rotation matrix.  ``format_sd`` writes the
determinants as the integer spin-orbital rows the 2-RDM builder reads (1 =
alpha, 2 = beta, 0 = empty; the closed orbitals first), which is what the
Fortran reader prints for the options.dat route (src/Main.f90:102-110).
"""
from __future__ import annotations

import math
import warnings

import numpy as np


def _int_cols(line, first, last):
    """Integer read of the 1-based columns first..last of a line."""
    return int(line[first - 1:last])


def get_civecs_in_sds(filename, caspt2):
    """(sds, civs[no_roots, no_sds], no_closed) from a BAGEL log
    (get_civecs_in_sds, src/bagel_ci_reader.f90:87-275)."""
    with open(filename) as f:
        lines = f.read().splitlines()
    if not any("Cartesian basis functions are used" in ln for ln in lines):
        warnings.warn("BAGEL log: spherical basis functions? The path needs cartesian ones")
    it = iter(lines)

    def find(tag):
        for ln in it:
            if tag in ln:
                return ln
        raise ValueError(f"BAGEL log: no '{tag}' line")

    _int_cols(find("Number of electrons"), 30, 35)
    no_roots = _int_cols(find("nstate"), 19, 25)
    no_closed = _int_cols(find("nclosed"), 19, 25)
    no_active = _int_cols(find("nact"), 19, 25)
    find("ci vector, state")
    first = next(it)[7:7 + no_active]
    no_a = sum(c in "2a" for c in first)
    no_b = sum(c in "2b" for c in first)
    no_sds = math.comb(no_active, no_a) * math.comb(no_active, no_b)

    it = iter(lines)
    where, sds = {}, []
    civs = np.zeros((no_roots, no_sds))
    for root in range(no_roots):
        for ln in it:
            if "ci vector, state  " in ln:
                for _ in range(no_sds):
                    tok = next(it).split()
                    sd = tok[0][:no_active]
                    k = where.get(sd)
                    if k is None:
                        k = where[sd] = len(sds)
                        sds.append(sd)
                    civs[root, k] = float(tok[1])
                break
    if caspt2:
        for ln in it:
            if "XMS-CASPT2 rotation matrix" in ln:
                rot = np.array([[float(v) for v in next(it).split()] for _ in range(no_roots)])
                civs = rot @ civs
                break
    return sds, civs, no_closed


_SLOTS = {"2": (1, 2), "a": (1, 0), "b": (0, 2), ".": (0, 0)}


def format_sd(sds, no_closed):
    """Integer spin-orbital rows, closed orbitals first (format_sd,
    src/bagel_ci_reader.f90:55-84)."""
    norb = no_closed + len(sds[0])
    out = np.zeros((len(sds), 2 * norb), dtype=np.int32)
    out[:, :2 * no_closed] = np.tile([1, 2], no_closed)
    for i, sd in enumerate(sds):
        for j, c in enumerate(sd):
            out[i, 2 * (no_closed + j):2 * (no_closed + j) + 2] = _SLOTS[c]
    return out


def get_civs_and_confs(logfile, caspt2=False):
    """(confs int32[no_sds, 2 norb], civs f64[no_sds, no_roots]) of a BAGEL
    log (get_civs_and_confs, src/bagel_ci_reader.f90:20-50)."""
    sds, civs, no_closed = get_civecs_in_sds(logfile, caspt2)
    return format_sd(sds, no_closed), np.ascontiguousarray(civs.T)
