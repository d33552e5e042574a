"""Molcas RASSCF / CASPT2 CI vectors in the determinant basis of the path.

Restates ``src/molcas_ci_reader.py``: the CI vector printed in a Molcas log
is a list of configuration state functions (CSFs, one step letter per active
orbital: '2' doubly occupied, 'u' / 'd' singly occupied coupled up / down,
'0' empty) with their coefficients.  Each CSF is expanded in Slater
determinants (``csf2sd``), the expansions are summed per determinant
(``csf_to_slater_basis_conversion``) and the determinants are written as the
spin-orbital strings the 2-RDM builder reads (``sd_formatter``: two
characters per orbital, 'a' / 'b' / '0', the closed orbitals first).

``get_civs_and_confs`` returns what the reference's does: ``civs`` as tuples
of coefficient strings (one tuple per determinant, one entry per root) and
``confs`` as strings; ``readers.civs_to_array`` and ``readers.confs_to_int``
turn them into the arrays ``scattering.createtwordm`` takes.
"""
from __future__ import annotations

import itertools
import re

import numpy as np


def _paldus(csf):
    """Cumulative Paldus (a, b) after each orbital of a CSF: a counts the
    doubly occupied and down-coupled orbitals, b = 2S of the coupled
    electrons so far (src/molcas_ci_reader.py:56-100)."""
    a = np.zeros(len(csf))
    b = np.zeros(len(csf))
    ca = cb = 0.0
    for i, step in enumerate(csf):
        if step == "u":
            cb += 1
        elif step == "d":
            ca += 1
            cb -= 1
        elif step == "2":
            ca += 1
        elif step != "0":
            raise ValueError(f"not a CSF step vector: {csf!r}")
        a[i], b[i] = ca, cb
    return a, b


def csf2sd(csfvec, m_s):
    """Determinant expansion of one CSF at spin projection ``m_s``
    (csf2sd, src/molcas_ci_reader.py:45-153).

    Returns ``[[coefficient, determinant], ...]`` with determinants written
    per orbital as '2', 'a', 'b' or '0', in the order of the alpha
    assignments of the open shells (itertools.combinations).  The
    coefficient is the product over the open shells of the squared
    genealogical coupling factors, each written with the Paldus numbers
    after the orbital (N = 2a + b electrons, S = b/2, M = the projection of
    the electrons so far):
      up,   alpha: (S + M) / 2S                 beta: (S - M) / 2S
      down, alpha: -(S - M + 1) / (2S + 2)      beta: (S + M + 1) / (2S + 2)
    (the down-step sign is applied when b is even for alpha, odd for beta),
    times -1 for a doubly occupied orbital at odd b; the result is
    sign * sqrt(|product|).  Determinants with a zero factor are dropped.
    """
    a, b = _paldus(csfvec)
    n_open = csfvec.count("u") + csfvec.count("d")
    n_alpha = int((n_open + 2 * m_s) / 2)
    det = []
    for alpha_open in itertools.combinations(range(n_open), n_alpha):
        alpha_open = set(alpha_open)
        coeff = 1
        k = 0       # open shells passed
        na = nb = 0  # alpha, beta electrons placed before this orbital
        sd = []
        for i, step in enumerate(csfvec):
            if step == "0":
                sd.append("0")
            elif step == "2":
                sd.append("2")
                if b[i] % 2 != 0:
                    coeff *= -1
                na += 1
                nb += 1
            elif step == "u":
                if k in alpha_open:
                    sd.append("a")
                    coeff *= (a[i] + b[i] - nb)
                    na += 1
                else:
                    sd.append("b")
                    coeff *= (a[i] + b[i] - na)
                    nb += 1
                coeff /= b[i]
                k += 1
            else:  # "d"
                if k in alpha_open:
                    sd.append("a")
                    coeff *= (nb - a[i] + 1)
                    na += 1
                    if b[i] % 2 == 0:
                        coeff *= -1
                else:
                    sd.append("b")
                    coeff *= (na - a[i] + 1)
                    nb += 1
                    if b[i] % 2 != 0:
                        coeff *= -1
                coeff /= (b[i] + 2)
                k += 1
        if coeff != 0:
            det.append([np.sign(coeff) * np.sqrt(np.abs(coeff)), "".join(sd)])
    return det


def csf_to_slater_basis_conversion(csfs, civs, S):
    """One root's CI vector from the CSF to the determinant basis
    (src/molcas_ci_reader.py:10-28): determinants in order of first
    appearance, coefficients summed in CSF order, returned as strings."""
    where = {}
    sds, civecinsd = [], []
    for csf, cv in zip(csfs, civs):
        for c, sd in csf2sd(csf, S):
            k = where.get(sd)
            if k is None:
                k = where[sd] = len(sds)
                sds.append(sd)
                civecinsd.append(0.0)
            civecinsd[k] += c * float(cv)
    return sds, [str(v) for v in civecinsd]


def create_sym_csf(string, frozen, inact, ras2, second):
    """Full-orbital CSF of a symmetry-blocked log line: per irrep the frozen
    and inactive orbitals doubly occupied, its active string, the secondary
    orbitals empty (src/molcas_ci_reader.py:30-42)."""
    csf = ""
    j = 0
    for i in range(len(frozen)):
        csf += frozen[i] * "2"
        csf += inact[i] * "2"
        if ras2[i] > 0:
            csf += string[j]
            j += 1
        csf += second[i] * "0"
    return csf


def _ints_after(line, k):
    return [int(x) for x in line.split()[k:]]


def get_civecs_in_csfs(filename, caspt2):
    """CSFs and CI vectors of every root printed in a Molcas log
    (src/molcas_ci_reader.py:170-300): RASSCF printout ("printout of
    CI-coefficients larger than") or, with ``caspt2``, the CASPT2 mixed
    states ("The CI coefficients for the MIXED state nr.").  Returns
    ``(no_roots, S, no_closed, csfs, civs)``; with symmetry the CSFs span
    every orbital and no_closed is 0."""
    with open(filename) as f:
        civs = []
        sym = False
        for line in f:
            if re.search(".*Symmetry species.*", line):
                if len(line.split()) - 2 > 1:
                    sym = True
                    irrep_numbers = _ints_after(f.readline(), 2)
                break
        for line in f:
            if re.search(".*Number of closed shell electrons.*", line):
                no_closed = int(line.split()[5]) // 2
                break
        for line in f:
            if re.search(".*Spin quantum number.*", line):
                S = int(float(line.split()[3]))
                break
        if sym:
            blocks = {}
            for key in ("Frozen orbitals", "Inactive orbitals", "RAS2 orbitals",
                        "Secondary orbitals"):
                for line in f:
                    if re.search(f".*{key}.*", line, re.IGNORECASE):
                        blocks[key] = _ints_after(line, 2)
                        break
            frozen, inact = blocks["Frozen orbitals"], blocks["Inactive orbitals"]
            ras2, second = blocks["RAS2 orbitals"], blocks["Secondary orbitals"]
            n_act_irreps = sum(v > 0 for v in ras2)
            if [sum(v) for v in zip(frozen, inact, ras2, second)] != irrep_numbers:
                raise ValueError("Molcas log: orbital blocks do not add up to the irreps")
        for line in f:
            if re.search(".*Number of CSFs.*", line, re.IGNORECASE):
                no_csfs = int(line.split()[3])
                break
        if caspt2:
            for line in f:
                if re.search(".*Number of CI roots used.*", line, re.IGNORECASE):
                    no_roots = int(line.split()[5])
                    break
            head = ".*The CI coefficients for the MIXED state nr."
        else:
            for line in f:
                if re.search(r".*Number of root\(s\) required.*", line, re.IGNORECASE):
                    no_roots = int(line.split()[4])
                    break
            head = ".*printout of CI-coefficients larger than"
        csfs = []
        for _ in range(no_roots):
            while True:
                line = f.readline()
                if not line:
                    break
                if re.search(head, line):
                    csfs, data = [], []
                    if caspt2:
                        for _ in range(6):  # rule, title and column header lines
                            f.readline()
                    else:
                        f.readline()  # energy
                        f.readline()  # column header
                    for _ in range(no_csfs):
                        row = f.readline()
                        if caspt2:
                            tok = row[30:].split()
                            if sym:
                                csfs.append(create_sym_csf(tok[:-2], frozen, inact, ras2, second))
                                data.append(tok[-2])
                            else:
                                csfs.append(tok[0])
                                data.append(tok[1])
                        else:
                            if sym:
                                tok = row.split()[1:2 + n_act_irreps]
                                csfs.append(create_sym_csf(tok[:-1], frozen, inact, ras2, second))
                                data.append(tok[-1])
                            else:
                                tok = row.split()[1:3]
                                csfs.append(tok[0])
                                data.append(tok[1])
                    civs.append(data)
    if sym:
        no_closed = 0
    return no_roots, S, no_closed, csfs, civs


def sd_formatter(sds, no_closed):
    """Determinants as spin-orbital strings, closed orbitals first
    (src/molcas_ci_reader.py:303-307)."""
    return [no_closed * "ab" + s.replace("0", "00").replace("a", "a0").replace("b", "0b")
            .replace("2", "ab") for s in sds]


def transposer(array):
    return list(zip(*array))


def get_civs_and_confs(logfile, caspt2, extra_precision=False):
    """(civs, confs) of every root of a Molcas log in the determinant basis
    (src/molcas_ci_reader.py:316-338).  ``extra_precision`` (CI vectors from
    the HDF5 job files) needs h5py, which this build does not use."""
    if extra_precision:
        raise NotImplementedError("CI vectors from Molcas HDF5 files are not read by this build")
    no_roots, S, no_closed, csfs, civs = get_civecs_in_csfs(logfile, caspt2)
    sds, newcivec = [], []
    for i in range(no_roots):
        sds, v = csf_to_slater_basis_conversion(csfs, civs[i], S)
        newcivec.append(v)
    return transposer(newcivec), sd_formatter(sds, no_closed)
