"""ctypes wrapper of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker.  See oracle.h for the restated
reference routines and their array conventions.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = C.CDLL(path)
        _declare(_LIB)
    return _LIB


_i = C.c_int
_d = C.c_double
_p = C.c_void_p


def _declare(L):
    L.orc_set_p0.argtypes = [_p, _i, _p, _i]
    L.orc_pair_tables.argtypes = [_i, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p,
                                  _p, _i, _p, _p, _p, _p, _p, _p, _p]
    L.orc_twordm_dense.argtypes = [_p, _i, _i, _p, _i, _i, _i, _p]
    L.orc_twordm_select.argtypes = [_p, _i, _d, _p, _p, _i]
    L.orc_twordm_select.restype = _i
    L.orc_pair_phase.argtypes = [_p, _i, _i, _i, _i, _p, _p]
    L.orc_onerdm.argtypes = [_p, _i, _i, _p, _i, _i, _i, _p]
    L.orc_elastic_z.argtypes = [_p, _i, _p, _i, _i, _p]
    L.orc_quartet_list.argtypes = [_i, _p, _p, _p, _d, _p]
    L.orc_quartet_list.restype = C.c_int64
    L.orc_quartet_list_elastic_j2.argtypes = [_i, _p, _p, _p, _d, _p]
    L.orc_quartet_list_elastic_j2.restype = C.c_int64
    L.orc_total_j0.argtypes = [_i, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p,
                               _i, _p, _p, _i, _p, _i, _d, _d, _d, _i, _i, _p]
    L.orc_total_j0.restype = _i
    L.orc_total_j2.argtypes = L.orc_total_j0.argtypes
    L.orc_total_j2.restype = _i
    L.orc_set_p0_j2.argtypes = [_p, _i, _p, _i]
    L.orc_total_j0_sample.argtypes = [_i, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p,
                                      _p, _i, _p, _p, _i, _p, _i, _d, _d, _d, _i, _i,
                                      C.c_int64, _p, _p, _p]
    L.orc_total_j0_sample.restype = _i
    L.orc_elastic_j0.argtypes = [_i, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p,
                                 _p, _p, _i, _d, _d, _d, _i, _p]
    L.orc_elastic_j0.restype = _i
    L.orc_elastic_j2.argtypes = L.orc_elastic_j0.argtypes
    L.orc_elastic_j2.restype = _i
    L.orc_nuclei_electron.argtypes = [_i, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i, _p,
                                      _p, _p, _i, _d, _d, _d, _p]
    L.orc_nuclei_electron.restype = _i
    L.orc_nuclear_nn.argtypes = [_i, _p, _p, _p, _i, _p]
    L.orc_build_z12.argtypes = [_p, _i, _i, _p, _p, _i, _p, _p]
    L.orc_dphrec.argtypes = [_p, _p, _i, _i, _i]
    L.orc_bessels0rr.argtypes = [_p, _i, _p, _i, _d, _p]
    L.orc_unique_total.argtypes = [_p, _i, _i, _p, _p, _p, _p, _p]
    L.orc_unique_total.restype = _i
    _l = C.c_int64
    L.orc_mcci_to_bit.argtypes = [_p, _p, _p, _p, _l, _i, _p, _p]
    L.orc_createtwordm_bit.argtypes = [_p, _p, _p, _p, _l, _p, _i]
    L.orc_one_rdm_bit.argtypes = [_p, _p, _p, _l, _p, _i]
    L.orc_createtwordm_bit_fci.argtypes = [_p, _p, _p, _l, _p, _l, _p, _p, _p, _p, _i]


def _ptr(a):
    return a.ctypes.data_as(_p) if a is not None else None


def _I(a, order="C"):
    return np.require(np.asarray(a), dtype=np.int32, requirements=[order])


def _D(a, order="C"):
    return np.require(np.asarray(a), dtype=np.float64, requirements=[order])


def set_p0(lmax4, q):
    q = _D(q)
    D = lmax4 + 1
    P0 = np.zeros((D, D, D, len(q)))
    lib().orc_set_p0(_ptr(P0), lmax4, _ptr(q), len(q))
    return P0


def pair_tables(l, m, n, group, ga, xx, yy, zz, q, maxl=None):
    l, m, n, group = _I(l), _I(m), _I(n), _I(group)
    ga, xx, yy, zz, q = _D(ga), _D(xx), _D(yy), _D(zz), _D(q)
    ngto = len(l)
    ng = int(group.max())
    maxl = int(l.max()) if maxl is None else maxl
    L1, T1 = maxl + 1, 2 * maxl + 1
    dd = [np.zeros((ng, ng, L1, L1, T1)) for _ in range(3)]
    P = [np.zeros((ng, ng)) for _ in range(3)]
    e12 = np.zeros((ng, ng, len(q)))
    lib().orc_pair_tables(ngto, ng, maxl, _ptr(l), _ptr(m), _ptr(n), _ptr(group),
                          _ptr(ga), _ptr(xx), _ptr(yy), _ptr(zz), _ptr(q), len(q),
                          _ptr(dd[0]), _ptr(dd[1]), _ptr(dd[2]), _ptr(P[0]),
                          _ptr(P[1]), _ptr(P[2]), _ptr(e12))
    return dd, P, e12


def twordm_dense(confs, civs, state1=1, state2=1):
    confs = _I(confs, "F")
    civs = _D(civs, "F")
    nconf, lconfs = confs.shape
    norb = lconfs // 2
    tw = np.zeros((norb,) * 4, order="F")
    lib().orc_twordm_dense(_ptr(confs), nconf, lconfs, _ptr(civs), civs.shape[1],
                           state1, state2, _ptr(tw))
    return tw


def twordm_select(tw, cutoff=1e-30):
    tw = _D(tw, "F")
    norb = tw.shape[0]
    cnt = lib().orc_twordm_select(_ptr(tw), norb, cutoff, None, None, 0)
    mat = np.zeros((cnt, 4), dtype=np.int32, order="F")
    tot = np.zeros(cnt)
    lib().orc_twordm_select(_ptr(tw), norb, cutoff, _ptr(mat), _ptr(tot), cnt)
    return np.ascontiguousarray(mat), tot


def createtwordm(confs, civs, state1=1, state2=1, cutoff=1e-30):
    return twordm_select(twordm_dense(confs, civs, state1, state2), cutoff)


def pair_phase(confs, c1, c2):
    """(ep2(c1,c2), ndiff(c1,c2)) of maxcoincidence for 1-based determinant
    indices, as the reference's matrices are indexed."""
    confs = _I(confs, "F")
    nconf = confs.shape[0]
    if not (1 <= c1 <= nconf and 1 <= c2 <= nconf):
        raise IndexError(f"determinant index out of 1..{nconf}: {c1}, {c2}")
    ep = np.zeros(1, dtype=np.int32)
    nd = np.zeros(1, dtype=np.int32)
    lib().orc_pair_phase(_ptr(confs), nconf, confs.shape[1], c1 - 1, c2 - 1,
                         _ptr(ep), _ptr(nd))
    return int(ep[0]), int(nd[0])


def onerdm(confs, civs, state1=1, state2=1):
    confs = _I(confs, "F")
    civs = _D(civs, "F")
    norb = confs.shape[1] // 2
    g = np.zeros((norb, norb), order="F")
    lib().orc_onerdm(_ptr(confs), confs.shape[0], confs.shape[1], _ptr(civs),
                     civs.shape[1], state1, state2, _ptr(g))
    return g


def elastic_z(g, mmod):
    g = _D(g, "F")
    mmod = _D(mmod, "F")
    nmo, ngto = mmod.shape
    Z = np.zeros((ngto, ngto), order="F")
    lib().orc_elastic_z(_ptr(g), g.shape[0], _ptr(mmod), nmo, ngto, _ptr(Z))
    return Z


def quartet_list(px, py, pz, cutoffcentre):
    px, py, pz = _D(px), _D(py), _D(pz)
    ng = px.shape[0]
    Q = lib().orc_quartet_list(ng, _ptr(px), _ptr(py), _ptr(pz), cutoffcentre, None)
    out = np.zeros((Q, 6), dtype=np.int32)
    lib().orc_quartet_list(ng, _ptr(px), _ptr(py), _ptr(pz), cutoffcentre, _ptr(out))
    return out


def total_j0(l, m, n, group, ga, xx, yy, zz, mmod, mat, total, q,
             cutoffz=1e-15, cutoffmd=1e-20, cutoffcentre=0.001, zmode=1,
             nthreads=0, maxl=None):
    l, m, n, group = _I(l), _I(m), _I(n), _I(group)
    ga, xx, yy, zz, q = _D(ga), _D(xx), _D(yy), _D(zz), _D(q)
    mmod = _D(mmod, "F")
    mat = _I(mat, "F")
    total = _D(total)
    ngto = len(l)
    ng = int(group.max())
    maxl = int(l.max()) if maxl is None else maxl
    out = np.zeros(len(q))
    rc = lib().orc_total_j0(ngto, ng, maxl, _ptr(l), _ptr(m), _ptr(n), _ptr(group),
                            _ptr(ga), _ptr(xx), _ptr(yy), _ptr(zz), _ptr(mmod),
                            mmod.shape[0], _ptr(mat), _ptr(total), len(total),
                            _ptr(q), len(q), cutoffz, cutoffmd, cutoffcentre,
                            zmode, nthreads, _ptr(out))
    if rc:
        raise RuntimeError(f"oracle total_j0 status {rc}")
    return out


def total_j2(l, m, n, group, ga, xx, yy, zz, mmod, mat, total, q,
             cutoffz=1e-15, cutoffmd=1e-20, cutoffcentre=0.001, zmode=1,
             nthreads=0, maxl=None):
    """tot_integration_j2 (type 7): total j2 S(q)."""
    l, m, n, group = _I(l), _I(m), _I(n), _I(group)
    ga, xx, yy, zz, q = _D(ga), _D(xx), _D(yy), _D(zz), _D(q)
    mmod = _D(mmod, "F")
    mat = _I(mat, "F")
    total = _D(total)
    ngto = len(l)
    ng = int(group.max())
    maxl = int(l.max()) if maxl is None else maxl
    out = np.zeros(len(q))
    rc = lib().orc_total_j2(ngto, ng, maxl, _ptr(l), _ptr(m), _ptr(n), _ptr(group),
                            _ptr(ga), _ptr(xx), _ptr(yy), _ptr(zz), _ptr(mmod),
                            mmod.shape[0], _ptr(mat), _ptr(total), len(total),
                            _ptr(q), len(q), cutoffz, cutoffmd, cutoffcentre,
                            zmode, nthreads, _ptr(out))
    if rc:
        raise RuntimeError(f"oracle total_j2 status {rc}")
    return out


def set_p0_j2(lmax4, q):
    q = _D(q)
    D = lmax4 + 1
    P0 = np.zeros((D, D, D, len(q)))
    lib().orc_set_p0_j2(_ptr(P0), lmax4, _ptr(q), len(q))
    return P0


def total_j0_sample(l, m, n, group, ga, xx, yy, zz, mmod, mat, total, q, stride,
                    cutoffz=1e-15, cutoffmd=1e-20, cutoffcentre=0.001, nthreads=0,
                    maxl=None, zmode=1):
    """Time the per-quartet kernels on an LL-stratified 1/stride sample.
    zmode 0 contracts the reference's z11/z22 arrays per quartet (its dgemm
    step) inside the timed region; zmode 1 supplies the blocks untimed.

    Returns (partial S(q), kernel seconds, quartets sampled)."""
    l, m, n, group = _I(l), _I(m), _I(n), _I(group)
    ga, xx, yy, zz, q = _D(ga), _D(xx), _D(yy), _D(zz), _D(q)
    mmod = _D(mmod, "F")
    mat = _I(mat, "F")
    total = _D(total)
    ngto = len(l)
    ng = int(group.max())
    maxl = int(l.max()) if maxl is None else maxl
    out = np.zeros(len(q))
    secs = np.zeros(1)
    ns = np.zeros(1, dtype=np.int64)
    rc = lib().orc_total_j0_sample(ngto, ng, maxl, _ptr(l), _ptr(m), _ptr(n), _ptr(group),
                                   _ptr(ga), _ptr(xx), _ptr(yy), _ptr(zz), _ptr(mmod),
                                   mmod.shape[0], _ptr(mat), _ptr(total), len(total),
                                   _ptr(q), len(q), cutoffz, cutoffmd, cutoffcentre,
                                   int(zmode), nthreads, int(stride), _ptr(out), _ptr(secs),
                                   _ptr(ns))
    if rc:
        raise RuntimeError(f"oracle total_j0_sample status {rc}")
    return out, float(secs[0]), int(ns[0])


def elastic_j0(l, m, n, group, ga, xx, yy, zz, Z, q, cutoffz=1e-15,
               cutoffmd=1e-20, cutoffcentre=0.001, nthreads=0, maxl=None, _fn="orc_elastic_j0"):
    l, m, n, group = _I(l), _I(m), _I(n), _I(group)
    ga, xx, yy, zz, q = _D(ga), _D(xx), _D(yy), _D(zz), _D(q)
    Z = _D(Z, "F")
    ngto = len(l)
    ng = int(group.max())
    maxl = int(l.max()) if maxl is None else maxl
    out = np.zeros(len(q))
    rc = getattr(lib(), _fn)(ngto, ng, maxl, _ptr(l), _ptr(m), _ptr(n), _ptr(group),
                             _ptr(ga), _ptr(xx), _ptr(yy), _ptr(zz), _ptr(Z),
                             _ptr(q), len(q), cutoffz, cutoffmd, cutoffcentre,
                             nthreads, _ptr(out))
    if rc:
        raise RuntimeError(f"oracle {_fn} status {rc}")
    return out


def elastic_j2(l, m, n, group, ga, xx, yy, zz, Z, q, cutoffz=1e-15,
               cutoffmd=1e-20, cutoffcentre=0.001, nthreads=0, maxl=None):
    """elastic_integration_j2 (type 8) with variables_elastic's Z."""
    return elastic_j0(l, m, n, group, ga, xx, yy, zz, Z, q, cutoffz, cutoffmd, cutoffcentre,
                      nthreads, maxl, _fn="orc_elastic_j2")


def nuclei_electron(l, m, n, group, ga, xx, yy, zz, Z, Zn, geom, q, cutoffz=1e-15,
                    cutoffmd=1e-20, cutoffcentre=0.001, maxl=None):
    """nuclei_electron_integration (types 5/6): Sum_na Zn Sum_(i,k) F e12."""
    l, m, n, group = _I(l), _I(m), _I(n), _I(group)
    ga, xx, yy, zz, q = _D(ga), _D(xx), _D(yy), _D(zz), _D(q)
    Z = _D(Z, "F")
    Zn = _D(np.atleast_1d(Zn))
    geom = _D(np.asarray(geom, dtype=np.float64).reshape(-1, 3), "F")
    ngto, ng = len(l), int(group.max())
    maxl = int(l.max()) if maxl is None else maxl
    out = np.zeros(len(q))
    rc = lib().orc_nuclei_electron(ngto, ng, maxl, _ptr(l), _ptr(m), _ptr(n), _ptr(group),
                                   _ptr(ga), _ptr(xx), _ptr(yy), _ptr(zz), _ptr(Z), len(Zn),
                                   _ptr(Zn), _ptr(geom), _ptr(q), len(q), cutoffz, cutoffmd,
                                   cutoffcentre, _ptr(out))
    if rc:
        raise RuntimeError(f"oracle nuclei_electron status {rc}")
    return out


def nuclear_nn(Zn, geom, q):
    """Sum_ij Zn_i Zn_j sinc(q R_ij), the nuclear term of types 5/6."""
    Zn = _D(np.atleast_1d(Zn))
    geom = _D(np.asarray(geom, dtype=np.float64).reshape(-1, 3), "F")
    q = _D(q)
    out = np.zeros(len(q))
    lib().orc_nuclear_nn(len(Zn), _ptr(Zn), _ptr(geom), _ptr(q), len(q), _ptr(out))
    return out


def quartet_list_elastic_j2(px, py, pz, cutoffcentre):
    """The six nests of elastic_integration_j2; rows as quartet_list."""
    px, py, pz = _D(px), _D(py), _D(pz)
    ng = px.shape[0]
    Q = lib().orc_quartet_list_elastic_j2(ng, _ptr(px), _ptr(py), _ptr(pz), cutoffcentre, None)
    out = np.zeros((Q, 6), dtype=np.int32)
    lib().orc_quartet_list_elastic_j2(ng, _ptr(px), _ptr(py), _ptr(pz), cutoffcentre, _ptr(out))
    return out


def build_z12(mmod, mat, total, ngto):
    """z11, z22 of variables_total, each (nmat, ngto, ngto) (column-major
    z(k,a,b) indexing, returned as numpy arrays indexed [k, a, b])."""
    mmod = _D(mmod, "F")
    mat = _I(mat, "F")
    total = _D(total)
    nmat = len(total)
    z1 = np.zeros(nmat * ngto * ngto)
    z2 = np.zeros(nmat * ngto * ngto)
    lib().orc_build_z12(_ptr(mmod), mmod.shape[0], ngto, _ptr(mat), _ptr(total), nmat,
                        _ptr(z1), _ptr(z2))
    sh = (nmat, ngto, ngto)
    return z1.reshape(sh, order="F"), z2.reshape(sh, order="F")


def dphrec(x, lmax1, leps1=100):
    x = _D(x)
    out = np.zeros((len(x), leps1))
    lib().orc_dphrec(_ptr(x), _ptr(out), leps1, lmax1, len(x))
    return out  # out[iq, k] = RBES(k+1, iq)


def bessels0rr(order, mu, H, h_saved):
    mu = _D(mu)
    hs = np.zeros(20)
    hs[:len(h_saved)] = h_saved
    out = np.zeros(len(mu))
    lib().orc_bessels0rr(_ptr(out), order, _ptr(mu), len(mu), H, _ptr(hs))
    return out


def unique_total(matrix, total):
    matrix = _I(matrix, "F")
    total = _D(total)
    nrow, ncol = matrix.shape
    rmat = np.zeros((nrow, ncol), dtype=np.int32, order="F")
    rtot = np.zeros(nrow)
    iuni = np.zeros(nrow, dtype=np.int32)
    irec = np.zeros(nrow, dtype=np.int32)
    cnt = lib().orc_unique_total(_ptr(matrix), nrow, ncol, _ptr(total), _ptr(rmat),
                                 _ptr(rtot), _ptr(iuni), _ptr(irec))
    rm = np.zeros((cnt, ncol), dtype=np.int32, order="F")
    lib().orc_unique_total(_ptr(matrix), nrow, ncol, _ptr(total), _ptr(rm), None,
                           None, None)
    return np.ascontiguousarray(rm), rtot[:cnt].copy(), iuni[:cnt].copy(), irec


def _L(a):
    return np.require(np.asarray(a), dtype=np.int64, requirements=["C"])


def mcci_to_bit(alpha, beta, c1, c2, text_round=True):
    """mcci_to_bit (src/one_rdm.f90:1901-1984): signed (and, with
    text_round, E20.10-rounded) coefficients."""
    a, b, c1, c2 = _L(alpha), _L(beta), _D(c1), _D(c2)
    o1, o2 = np.zeros(len(a)), np.zeros(len(a))
    rc = lib().orc_mcci_to_bit(_ptr(a), _ptr(b), _ptr(c1), _ptr(c2), len(a),
                               1 if text_round else 0, _ptr(o1), _ptr(o2))
    if rc != 0:
        raise ValueError("mcci_to_bit: alpha/beta electron counts differ")
    return o1, o2


def twordm_bit_dense(alpha, beta, c1, c2):
    """createtwordm_bit's SpinFree2RDM (src/one_rdm.f90:931-1300)."""
    a, b, c1, c2 = _L(alpha), _L(beta), _D(c1), _D(c2)
    N = lib().orc_createtwordm_bit(_ptr(a), _ptr(b), _ptr(c1), _ptr(c2), len(a), None, 0)
    if N < 0:
        raise ValueError("createtwordm_bit: electron counts vary")
    tw = np.zeros((N,) * 4, order="F")
    lib().orc_createtwordm_bit(_ptr(a), _ptr(b), _ptr(c1), _ptr(c2), len(a), _ptr(tw), N)
    return tw


# cutoff = 1E-30 in createtwordm_bit is a default-real literal
BIT_CUTOFF = float(np.float32(1e-30))


def createtwordm_bit(alpha, beta, c1, c2):
    """createtwordm_bit rows (p,s,q,r) and values (src/one_rdm.f90:931-1353)."""
    return twordm_select(twordm_bit_dense(alpha, beta, c1, c2), BIT_CUTOFF)


def one_rdm_bit(alpha, beta, civs):
    """one_rdm_bit (src/one_rdm.f90:1388-1516): maxnmo x maxnmo."""
    a, b, c = _L(alpha), _L(beta), _D(civs)
    N = lib().orc_one_rdm_bit(_ptr(a), _ptr(b), _ptr(c), len(a), None, 0)
    if N < 0:
        raise ValueError("one_rdm_bit: electron counts vary")
    g = np.zeros((N, N), order="F")
    lib().orc_one_rdm_bit(_ptr(a), _ptr(b), _ptr(c), len(a), _ptr(g), N)
    return g


def fci_tables(alpha, beta):
    """newdat(:,1), irep, start, end as the driver builds them
    (src/scattering_main_routine.py:269-289): unique alpha strings in
    increasing order, their determinant counts and 1-based first / last
    file positions."""
    alpha = np.asarray(alpha, dtype=np.int64)
    uniq = np.unique(alpha)
    irep = np.array([np.count_nonzero(alpha == u) for u in uniq], dtype=np.int64)
    start = np.array([np.nonzero(alpha == u)[0][0] + 1 for u in uniq], dtype=np.int64)
    end = np.array([np.nonzero(alpha == u)[0][-1] + 1 for u in uniq], dtype=np.int64)
    return uniq, irep, start, end


def twordm_bit_fci_dense(alpha, beta, civs):
    """createtwordm_bit_fci's twordm (src/one_rdm.f90:390-875)."""
    a, b, c = _L(alpha), _L(beta), _D(civs)
    uniq, irep, start, end = (_L(x) for x in fci_tables(a, b))
    args = (_ptr(a), _ptr(b), _ptr(c), len(a), _ptr(uniq), len(uniq), _ptr(irep),
            _ptr(start), _ptr(end))
    N = lib().orc_createtwordm_bit_fci(*args, None, 0)
    if N < 0:
        raise ValueError("createtwordm_bit_fci: a string is not among the unique alpha strings")
    tw = np.zeros((N,) * 4, order="F")
    lib().orc_createtwordm_bit_fci(*args, _ptr(tw), N)
    return tw


def createtwordm_bit_fci(alpha, beta, civs):
    """createtwordm_bit_fci rows (p,s,q,r) and values (src/one_rdm.f90:390-927)."""
    return twordm_select(twordm_bit_fci_dense(alpha, beta, civs), BIT_CUTOFF)
