"""The ``scattering_main_routine.main()`` flow, minus Molpro and the GUI.

``prepare_inputs`` goes from a molden file plus a CI punch file to the
positional arguments of ``total_scattering_calculation``, following
src/scattering_main_routine.py:111-187,232-238; ``main`` then runs the call
and writes ``{'q','I'}`` (src/scattering_main_routine.py:474).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import readers

# defaults of src/scattering_main_routine.py:174-179
CUTOFF_CENTRE = 0.001
CUTOFF_Z = 1e-15
CUTOFF_MD = 1e-20


@dataclass
class ScatteringInputs:
    Zn: np.ndarray
    geom: np.ndarray
    ga: np.ndarray
    l: np.ndarray
    m: np.ndarray
    n: np.ndarray
    xx: np.ndarray
    yy: np.ndarray
    zz: np.ndarray
    mmod: np.ndarray      # (nmo, ngto)
    group: np.ndarray     # 1-based
    confs: np.ndarray     # (nconf, 2*norb) int32
    civs: np.ndarray      # (nconf, nstates)

    @property
    def ngto(self):
        return int(self.l.size)

    @property
    def ng(self):
        return int(self.group.max())

    @property
    def maxl(self):
        return int(self.l.max())

    @property
    def norb(self):
        return int(self.confs.shape[1] // 2)


def prepare_inputs(mldfile, punfile, closed) -> ScatteringInputs:
    civs, confs = readers.twordmconst(closed, punfile)
    nmo_max = len(confs[0]) / 2
    gtos, atoms = readers.read_orbitals(mldfile, N=nmo_max, decontract=True)
    return ScatteringInputs(
        Zn=atoms.atomic_numbers(), geom=atoms.geometry(),
        ga=np.asarray(gtos.ga, dtype=np.float64),
        l=np.asarray(gtos.l, dtype=np.int32), m=np.asarray(gtos.m, dtype=np.int32),
        n=np.asarray(gtos.n, dtype=np.int32),
        xx=np.asarray(gtos.x, dtype=np.float64), yy=np.asarray(gtos.y, dtype=np.float64),
        zz=np.asarray(gtos.z, dtype=np.float64),
        mmod=np.ascontiguousarray(np.transpose(gtos.mo), dtype=np.float64),
        group=np.asarray(gtos.group, dtype=np.int32),
        confs=readers.confs_to_int(confs), civs=readers.civs_to_array(civs))


def prepare_inputs_molcas(mldfile, logfile, caspt2=False) -> ScatteringInputs:
    """The same inputs from a Molcas molden file and the CI vector printed in
    a Molcas RASSCF (or, with ``caspt2``, CASPT2) log: the commented Molcas
    route of src/scattering_main_routine.py:117-119,136 with the Molcas
    molden rules (src/molden_reader_nikola_molcas.py) and the CSF to
    determinant conversion of src/molcas_ci_reader.py:316-338."""
    from . import molcas
    civs, confs = molcas.get_civs_and_confs(logfile, caspt2)
    nmo_max = len(confs[0]) / 2
    gtos, atoms = readers.read_orbitals(mldfile, N=nmo_max, decontract=True, flavour="molcas")
    return ScatteringInputs(
        Zn=atoms.atomic_numbers(), geom=atoms.geometry(),
        ga=np.asarray(gtos.ga, dtype=np.float64),
        l=np.asarray(gtos.l, dtype=np.int32), m=np.asarray(gtos.m, dtype=np.int32),
        n=np.asarray(gtos.n, dtype=np.int32),
        xx=np.asarray(gtos.x, dtype=np.float64), yy=np.asarray(gtos.y, dtype=np.float64),
        zz=np.asarray(gtos.z, dtype=np.float64),
        mmod=np.ascontiguousarray(np.transpose(gtos.mo), dtype=np.float64),
        group=np.asarray(gtos.group, dtype=np.int32),
        confs=readers.confs_to_int(confs), civs=readers.civs_to_array(civs))


def default_q(qmax=10.0, nq=200):
    q = np.linspace(0.00000000001, qmax, nq)
    if q[0] < 1e-10:
        q[0] = 1e-10
    return q


def main(mldfile, punfile, closed, q=None, type_=1, state1=1, state2=1,
         state3=1, out=None, **kw):
    """Run one total/elastic call the way the reference driver does."""
    from .scattering import total_scattering_calculation
    inp = prepare_inputs(mldfile, punfile, closed)
    q = default_q() if q is None else np.asarray(q, dtype=np.float64)
    q_abs, res = total_scattering_calculation(
        type_, inp.Zn, inp.geom, state1, state2, state3, inp.maxl, inp.ngto,
        inp.ng, inp.ga, inp.l, inp.m, inp.n, inp.xx, inp.yy, inp.zz, inp.mmod,
        q, q.size, inp.group, CUTOFF_Z, CUTOFF_MD, CUTOFF_CENTRE, inp.confs,
        inp.civs, **kw)
    if out is not None:
        import scipy.io as sio
        sio.savemat(out, {"q": q, "I": res})
    return res, q, q_abs


def prepare_basis(mldfile, nmo_max=14, flavour="molpro") -> ScatteringInputs:
    """Basis and orbitals alone, for the bit-string file route: the reference
    sets the orbital count by hand there (Nmo_max = 14,
    src/scattering_main_routine.py:134-136) and reads no punch file."""
    gtos, atoms = readers.read_orbitals(mldfile, N=nmo_max, decontract=True, flavour=flavour)
    return ScatteringInputs(
        Zn=atoms.atomic_numbers(), geom=atoms.geometry(),
        ga=np.asarray(gtos.ga, dtype=np.float64),
        l=np.asarray(gtos.l, dtype=np.int32), m=np.asarray(gtos.m, dtype=np.int32),
        n=np.asarray(gtos.n, dtype=np.int32),
        xx=np.asarray(gtos.x, dtype=np.float64), yy=np.asarray(gtos.y, dtype=np.float64),
        zz=np.asarray(gtos.z, dtype=np.float64),
        mmod=np.ascontiguousarray(np.transpose(gtos.mo), dtype=np.float64),
        group=np.asarray(gtos.group, dtype=np.int32),
        confs=np.zeros((0, 2), dtype=np.int32), civs=np.zeros((0, 1)))


def bits_to_confs(alpha, beta, nbits=32):
    """Bit strings as punch-style slot rows: slot 2j is alpha orbital j+1
    (value 1), slot 2j+1 beta orbital j+1 (value 2), over nbits orbitals
    (integertobinary's 32 bits, src/scattering_main_routine.py:15-25,
    345-360)."""
    alpha = np.asarray(alpha, dtype=np.int64)
    beta = np.asarray(beta, dtype=np.int64)
    confs = np.zeros((alpha.size, 2 * nbits), dtype=np.int32)
    for j in range(nbits):
        confs[:, 2 * j] = (alpha >> j) & 1
        confs[:, 2 * j + 1] = 2 * ((beta >> j) & 1)
    return confs


def main_file(mldfile, fileJ, q=None, type_=1, state1=1, state2=1, nmo_max=14, mcci=False,
              read2rdm=False, fileJ_2="", out=None, **kw):
    """The driver's external-file branch (jeremyR, src/scattering_main_routine.py:
    249-457) with fci off, as the driver has it: with read2rdm the 2-RDM /
    1-RDM files go to total_scattering_calculation_2; otherwise a file of at
    most 300 lines is read as ``i civ alpha beta``, normalised and run through
    total_scattering_calculation, and a longer one goes through
    total_scattering_calculation_2 (mcci_to_bit and the bit-string
    builders)."""
    from .scattering import total_scattering_calculation, total_scattering_calculation_2
    inp = prepare_basis(mldfile, nmo_max)
    q = default_q() if q is None else np.asarray(q, dtype=np.float64)
    base = (inp.maxl, inp.ngto, inp.ng, inp.ga, inp.l, inp.m, inp.n, inp.xx, inp.yy, inp.zz,
            inp.mmod, q, q.size, inp.group, CUTOFF_Z, CUTOFF_MD, CUTOFF_CENTRE)
    with open(fileJ) as fh:
        lines = [ln.split() for ln in fh if ln.strip()]
    count = len(lines)
    if not read2rdm and count <= 300:
        civs = np.array([float(t[1]) for t in lines])
        alpha = [int(t[2]) for t in lines]
        beta = [int(t[3]) for t in lines]
        civs = civs / np.sqrt(np.sum(civs ** 2))
        confs = bits_to_confs(alpha, beta)
        q_abs, res = total_scattering_calculation(
            type_, inp.Zn, inp.geom, state1, state2, 1, *base, confs, civs, **kw)
    else:
        q_abs, res = total_scattering_calculation_2(
            type_, inp.Zn, inp.geom, state1, state2, *base, fileJ, fileJ_2, count, None,
            None, None, None, None, False, None, None, None, None, read2rdm, mcci, **kw)
    if out is not None:
        import scipy.io as sio
        sio.savemat(out, {"q": q, "I": res})
    return res, q, q_abs
