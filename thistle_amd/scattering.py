"""Python mirror of ``main_calculation_mod`` for the total/elastic j0 path.

``total_scattering_calculation`` takes the positional arguments of the
Fortran procedure (src/main_calculation.f90:13-15, the form with ``state3``
that ``src/Main.f90:171-173`` calls) and returns ``(q_abs, result)`` like the
f2py binding used at src/scattering_main_routine.py:240-247.  Types 1 (total
j0), 2 (elastic j0), 5 / 6 (total / elastic electron scattering), 7 (total
j2) and 8 (elastic j2) run on the GPU through libthistle_hip.so; the other
types of the reference (aligned, resolved, j1) are outside this path and
raise.

Beyond the reference arguments, keyword ``precision`` selects FP64 (default)
or the mixed-precision kernels, and ``rdm_cutoff`` the |Gamma| selection
threshold of the 2-RDM (the reference hard-codes 1e-30).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as L


def _basis(ga, l, m, n, xx, yy, zz, mmod, q, group, cutoffz, cutoffmd, cutoffcentre,
           maxl=None, ng=None):
    return L.make_basis(ga, l, m, n, xx, yy, zz, group, mmod, q, cutoffz, cutoffmd,
                        cutoffcentre, maxl=maxl, ng=ng)


def total_scattering_calculation(type, Zn, geom, state1, state2, state3, maxl, ngto, ng,
                                 ga, l, m, n, xx, yy, zz, mmod, q, nq, group, cutoffz,
                                 cutoffmd, cutoffcentre, confs, civecs, *, precision=L.FP64,
                                 rdm_cutoff=L.RDM_CUTOFF, waves=0):
    """Drop-in for main_calculation_mod::total_scattering_calculation (types 1, 2,
    5, 6, 7 and 8)."""
    lib = L.load()
    q = np.asarray(q, dtype=np.float64)
    if int(nq) != q.size or int(ngto) != np.asarray(l).size:
        raise ValueError("nq / ngto do not match the array sizes")
    if int(type) not in (1, 2, 5, 6, 7, 8):
        raise NotImplementedError(
            f"type {type}: only 1 (total j0), 2 (elastic j0), 5/6 (total/elastic electron), "
            "7 (total j2) and 8 (elastic j2) are on this path")
    b = _basis(ga, l, m, n, xx, yy, zz, mmod, q, group, cutoffz, cutoffmd, cutoffcentre,
               maxl=maxl, ng=ng)
    o = L.make_opts(precision, rdm_cutoff, waves=waves)
    confs = np.require(np.asarray(confs), dtype=np.int32, requirements=["F"])
    civ = np.require(np.asarray(civecs, dtype=np.float64), requirements=["F"])
    if civ.ndim == 1:
        civ = civ.reshape(-1, 1, order="F")
    Zn = np.require(np.atleast_1d(np.asarray(Zn, dtype=np.float64)), requirements=["C"])
    geom = np.require(np.asarray(geom, dtype=np.float64).reshape(-1, 3), requirements=["F"])
    q_abs = np.zeros(q.size)
    res = np.zeros(q.size)
    rc = lib.thistle_total_scattering_calculation(
        int(type), int(Zn.size), L._ptr(Zn), L._ptr(geom), int(state1), int(state2),
        int(state3), C.byref(b), int(confs.shape[0]), int(confs.shape[1]), int(civ.shape[1]),
        L._ptr(confs), L._ptr(civ), C.byref(o), L._ptr(q_abs), L._ptr(res))
    L.check(rc, "total_scattering_calculation")
    return q_abs, res


def total_from_twordm(ga, l, m, n, xx, yy, zz, mmod, q, group, mat, total, cutoffz=1e-15,
                      cutoffmd=1e-20, cutoffcentre=0.001, *, precision=L.FP64, waves=0,
                      maxl=None):
    """readrdm route: S(q) of the total j0 path from 2-RDM rows (p,s,q,r) + values."""
    lib = L.load()
    b = _basis(ga, l, m, n, xx, yy, zz, mmod, q, group, cutoffz, cutoffmd, cutoffcentre,
               maxl=maxl)
    o = L.make_opts(precision, waves=waves)
    mat = np.require(np.asarray(mat), dtype=np.int32, requirements=["F"])
    total = np.require(np.asarray(total, dtype=np.float64), requirements=["C"])
    res = np.zeros(np.asarray(q).size)
    rc = lib.thistle_total_from_twordm(C.byref(b), int(total.size), L._ptr(mat),
                                       L._ptr(total), C.byref(o), L._ptr(res))
    L.check(rc, "total_from_twordm")
    return res


def total_j2_from_twordm(ga, l, m, n, xx, yy, zz, mmod, q, group, mat, total, cutoffz=1e-15,
                         cutoffmd=1e-20, cutoffcentre=0.001, *, precision=L.FP64, waves=0,
                         maxl=None):
    """Total j2 S(q) (type 7, tot_integration_j2) from 2-RDM rows (p,s,q,r) + values."""
    plan = Plan(7, ga, l, m, n, xx, yy, zz, mmod, q, group, mat=mat, total=total,
                cutoffz=cutoffz, cutoffmd=cutoffmd, cutoffcentre=cutoffcentre,
                precision=precision, waves=waves, maxl=maxl)
    try:
        return plan.execute()
    finally:
        plan.close()


def total_adjoint(ga, l, m, n, xx, yy, zz, mmod, q, group, mat, total, dL_dS, geom=None,
                  cutoffz=1e-15, cutoffmd=1e-20, cutoffcentre=0.001, *, want_geom=True,
                  want_total=True, maxl=None):
    """S(q), dL/dR and dL/d(2-RDM values) for a given dL/dS(q).

    No reference counterpart (the reference has no derivatives).  ``geom``
    (natom, 3) in bohr must contain every shell centre; the MO coefficients
    are held fixed.  Returns ``(S, dL_dR or None, dL_dtotal or None)``.
    """
    lib = L.load()
    b = _basis(ga, l, m, n, xx, yy, zz, mmod, q, group, cutoffz, cutoffmd, cutoffcentre,
               maxl=maxl)
    o = L.make_opts(L.FP64)
    mat = np.require(np.asarray(mat), dtype=np.int32, requirements=["F"])
    total = np.require(np.asarray(total, dtype=np.float64), requirements=["C"])
    w = np.require(np.asarray(dL_dS, dtype=np.float64), requirements=["C"])
    if w.size != np.asarray(q).size:
        raise ValueError("dL_dS must have one entry per q point")
    S = np.zeros(w.size)
    dR = gptr = None
    natom = 0
    if want_geom:
        if geom is None:
            raise ValueError("geom is required for the coordinate gradient")
        gptr = np.require(np.asarray(geom, dtype=np.float64).reshape(-1, 3), requirements=["F"])
        natom = gptr.shape[0]
        dR = np.zeros((natom, 3), order="F")
    dT = np.zeros(total.size) if want_total else None
    rc = lib.thistle_total_scattering_adjoint(C.byref(b), int(total.size), L._ptr(mat),
                                              L._ptr(total), natom, L._ptr(gptr), L._ptr(w),
                                              C.byref(o), L._ptr(S), L._ptr(dR), L._ptr(dT))
    L.check(rc, "total_scattering_adjoint")
    return S, (np.ascontiguousarray(dR) if dR is not None else None), dT


def elastic_from_onerdm(ga, l, m, n, xx, yy, zz, mmod, q, group, onerdm, cutoffz=1e-15,
                        cutoffmd=1e-20, cutoffcentre=0.001, *, precision=L.FP64, waves=0,
                        maxl=None):
    lib = L.load()
    b = _basis(ga, l, m, n, xx, yy, zz, mmod, q, group, cutoffz, cutoffmd, cutoffcentre,
               maxl=maxl)
    o = L.make_opts(precision, waves=waves)
    g = np.require(np.asarray(onerdm, dtype=np.float64), requirements=["F"])
    res = np.zeros(np.asarray(q).size)
    rc = lib.thistle_elastic_from_onerdm(C.byref(b), int(g.shape[0]), L._ptr(g), C.byref(o),
                                         L._ptr(res))
    L.check(rc, "elastic_from_onerdm")
    return res


def nuclei_electron(ga, l, m, n, xx, yy, zz, mmod, q, group, onerdm, Zn, geom, cutoffz=1e-15,
                    cutoffmd=1e-20, cutoffcentre=0.001, maxl=None):
    """nuclei_electron_integration: the electron-nucleus term Ine of electron
    scattering (types 5/6) from a 1-RDM; geom (natom, 3) in bohr."""
    lib = L.load()
    b = _basis(ga, l, m, n, xx, yy, zz, mmod, q, group, cutoffz, cutoffmd, cutoffcentre,
               maxl=maxl)
    o = L.make_opts()
    g = np.require(np.asarray(onerdm, dtype=np.float64), requirements=["F"])
    Zn = np.require(np.atleast_1d(np.asarray(Zn, dtype=np.float64)), requirements=["C"])
    geom = np.require(np.asarray(geom, dtype=np.float64).reshape(-1, 3), requirements=["F"])
    res = np.zeros(np.asarray(q).size)
    rc = lib.thistle_nuclei_electron(C.byref(b), int(g.shape[0]), L._ptr(g), int(Zn.size),
                                     L._ptr(Zn), L._ptr(geom), C.byref(o), L._ptr(res))
    L.check(rc, "nuclei_electron")
    return res


def elastic_j2_from_onerdm(ga, l, m, n, xx, yy, zz, mmod, q, group, onerdm, cutoffz=1e-15,
                           cutoffmd=1e-20, cutoffcentre=0.001, *, precision=L.FP64, waves=0,
                           maxl=None):
    """Elastic j2 (type 8, elastic_integration_j2) from a 1-RDM."""
    plan = Plan(8, ga, l, m, n, xx, yy, zz, mmod, q, group, onerdm=onerdm, cutoffz=cutoffz,
                cutoffmd=cutoffmd, cutoffcentre=cutoffcentre, precision=precision, waves=waves,
                maxl=maxl)
    try:
        return plan.execute()
    finally:
        plan.close()


# --------------------------------------------------- 2-RDM / 1-RDM
def createtwordm(confs, civs, state1=1, state2=1, cutoff=L.RDM_CUTOFF, device=None):
    """maxcoincidence + createtwordm: rows (p,s,q,r) and values (src/twordms.f90).

    device=None builds it on the GPU when one is present (thistle_createtwordm)
    and on the host otherwise; both give the reference's rows and values."""
    lib = L.load()
    confs = np.require(np.asarray(confs), dtype=np.int32, requirements=["F"])
    civs = np.require(np.asarray(civs, dtype=np.float64), requirements=["F"])
    if civs.ndim == 1:
        civs = civs.reshape(-1, 1, order="F")
    nconf, lconfs = confs.shape
    norb = lconfs // 2
    if device is None:
        device = L.device_count() > 0
    if device:
        # at most norb^4 rows: one pass when that bound is small
        cap = norb ** 4 if norb ** 4 <= (1 << 22) else 0
        mat = np.zeros((max(cap, 1), 4), dtype=np.int32, order="F")
        tot = np.zeros(max(cap, 1))
        cnt = lib.thistle_createtwordm(L._ptr(confs), nconf, lconfs, L._ptr(civs), civs.shape[1],
                                       state1, state2, cutoff, None, L._ptr(mat), L._ptr(tot),
                                       cap)
        if cnt < 0:
            L.check(-cnt, "createtwordm")
        if cnt > cap:
            mat = np.zeros((cnt, 4), dtype=np.int32, order="F")
            tot = np.zeros(cnt)
            cnt2 = lib.thistle_createtwordm(L._ptr(confs), nconf, lconfs, L._ptr(civs),
                                            civs.shape[1], state1, state2, cutoff, None,
                                            L._ptr(mat), L._ptr(tot), cnt)
            if cnt2 < 0:
                L.check(-cnt2, "createtwordm")
            cap = cnt
        return np.ascontiguousarray(mat[:cnt]), tot[:cnt].copy()
    tw = np.zeros((norb,) * 4, order="F")
    L.check(lib.thistle_twordm_dense(L._ptr(confs), nconf, lconfs, L._ptr(civs), civs.shape[1],
                                     state1, state2, L._ptr(tw)), "twordm_dense")
    cnt = lib.thistle_twordm_select(L._ptr(tw), norb, cutoff, None, None, 0)
    mat = np.zeros((cnt, 4), dtype=np.int32, order="F")
    tot = np.zeros(cnt)
    lib.thistle_twordm_select(L._ptr(tw), norb, cutoff, L._ptr(mat), L._ptr(tot), cnt)
    return np.ascontiguousarray(mat), tot


def onerdm(confs, civs, state1=1, state2=1, device=None):
    """onerdm_creat / createonerdm: the norb x norb 1-RDM (src/one_rdm.f90:27-222).

    device=None builds it on the GPU when one is present (thistle_createonerdm)
    and on the host otherwise; both give the reference's doubles."""
    lib = L.load()
    confs = np.require(np.asarray(confs), dtype=np.int32, requirements=["F"])
    civs = np.require(np.asarray(civs, dtype=np.float64), requirements=["F"])
    if civs.ndim == 1:
        civs = civs.reshape(-1, 1, order="F")
    norb = confs.shape[1] // 2
    g = np.zeros((norb, norb), order="F")
    if device is None:
        device = L.device_count() > 0
    if device:
        L.check(lib.thistle_createonerdm(L._ptr(confs), confs.shape[0], confs.shape[1],
                                         L._ptr(civs), civs.shape[1], state1, state2, None,
                                         L._ptr(g)), "createonerdm")
    else:
        L.check(lib.thistle_onerdm(L._ptr(confs), confs.shape[0], confs.shape[1], L._ptr(civs),
                                   civs.shape[1], state1, state2, L._ptr(g)), "onerdm")
    return g


# ------------------------------------------------- bit-string determinants
def _bits(alpha, beta, *coefs):
    a = np.require(np.asarray(alpha), dtype=np.int64, requirements=["C"])
    b = np.require(np.asarray(beta), dtype=np.int64, requirements=["C"])
    cs = [np.require(np.asarray(c, dtype=np.float64), requirements=["C"]) for c in coefs]
    if a.ndim != 1 or a.shape != b.shape or any(c.shape != a.shape for c in cs):
        raise ValueError("alpha, beta and the coefficients must be 1-D of one length")
    return (a, b, *cs)


def mcci_to_bit(alpha, beta, c1, c2, text_round=True):
    """mcci_to_bit (src/one_rdm.f90:1901-1984) on the GPU: each coefficient
    times the sign of reordering its determinant's slots from orbital order
    to alpha-then-beta order; text_round applies the E20.10 write / read that
    hands them to createtwordm_bit."""
    lib = L.load()
    a, b, c1, c2 = _bits(alpha, beta, c1, c2)
    o1, o2 = np.zeros(a.size), np.zeros(a.size)
    L.check(lib.thistle_mcci_to_bit(L._ptr(a), L._ptr(b), L._ptr(c1), L._ptr(c2), a.size,
                                    1 if text_round else 0, None, L._ptr(o1), L._ptr(o2)),
            "mcci_to_bit")
    return o1, o2


def createtwordm_bit(alpha, beta, c1, c2, cutoff=L.RDM_CUTOFF):
    """createtwordm_bit (src/one_rdm.f90:931-1353) on the GPU: rows (p,s,q,r)
    and values, bit-identical to the reference's."""
    lib = L.load()
    a, b, c1, c2 = _bits(alpha, beta, c1, c2)
    args = (L._ptr(a), L._ptr(b), L._ptr(c1), L._ptr(c2), a.size, cutoff, None)
    cnt = lib.thistle_createtwordm_bit(*args, None, None, 0)
    if cnt < 0:
        L.check(-cnt, "createtwordm_bit")
    mat = np.zeros((max(cnt, 1), 4), dtype=np.int32, order="F")
    tot = np.zeros(max(cnt, 1))
    c2n = lib.thistle_createtwordm_bit(*args, L._ptr(mat), L._ptr(tot), cnt)
    if c2n < 0:
        L.check(-c2n, "createtwordm_bit")
    return np.ascontiguousarray(mat[:cnt]), tot[:cnt].copy()


def one_rdm_bit(alpha, beta, civs):
    """one_rdm_bit (src/one_rdm.f90:1388-1537) on the GPU: the maxnmo x maxnmo
    1-RDM; civs is the file's third column."""
    lib = L.load()
    a, b, c = _bits(alpha, beta, civs)
    norb = C.c_int32(0)
    L.check(lib.thistle_one_rdm_bit(L._ptr(a), L._ptr(b), L._ptr(c), a.size, None, 0,
                                    C.byref(norb), None), "one_rdm_bit")
    g = np.zeros((norb.value, norb.value), order="F")
    L.check(lib.thistle_one_rdm_bit(L._ptr(a), L._ptr(b), L._ptr(c), a.size, None, norb.value,
                                    C.byref(norb), L._ptr(g)), "one_rdm_bit")
    return g


def createtwordm_bit_fci(alpha, beta, civs, uniq=None, irep=None, start=None, end=None,
                         cutoff=L.RDM_CUTOFF):
    """createtwordm_bit_fci (src/one_rdm.f90:390-927) on the GPU: rows
    (p,s,q,r) and values, bit-identical to the reference's.  uniq (the
    driver's newdat(:,1)), irep, start and end default to the driver's
    construction from the alpha strings."""
    from . import readers
    lib = L.load()
    a, b, c = _bits(alpha, beta, civs)
    if uniq is None:
        uniq, irep, start, end = readers.fci_tables(a)
    u, r, s0, e0 = (np.require(np.asarray(x), dtype=np.int64, requirements=["C"])
                    for x in (uniq, irep, start, end))
    args = (L._ptr(a), L._ptr(b), L._ptr(c), a.size, L._ptr(u), u.size, L._ptr(r), L._ptr(s0),
            L._ptr(e0), cutoff, None)
    cnt = lib.thistle_createtwordm_bit_fci(*args, None, None, 0)
    if cnt < 0:
        L.check(-cnt, "createtwordm_bit_fci")
    mat = np.zeros((max(cnt, 1), 4), dtype=np.int32, order="F")
    tot = np.zeros(max(cnt, 1))
    c2n = lib.thistle_createtwordm_bit_fci(*args, L._ptr(mat), L._ptr(tot), cnt)
    if c2n < 0:
        L.check(-c2n, "createtwordm_bit_fci")
    return np.ascontiguousarray(mat[:cnt]), tot[:cnt].copy()


def total_scattering_calculation_2(type, Zn, geom, state1, state2, maxl, ngto, ng, ga, l, m,
                                   n, xx, yy, zz, mmod, q, nq, group, cutoffz, cutoffmd,
                                   cutoffcentre, fileJ, fileJ_2, numberlines, newdat, start,
                                   end, start_2, end_2, fci, irep1, irep2, ordering1, ordering2,
                                   readrdm, mcci, *, precision=L.FP64, waves=0):
    """Drop-in for main_calculation_mod::total_scattering_calculation_2
    (src/main_calculation.f90:501-849), types 1, 2, 7 and 8.

    fileJ is the ``i c1 c2 alpha beta`` file (numberlines lines); with readrdm
    types 1 / 7 read the 2-RDM file fileJ_2 and types 2 / 8 the 1-RDM file
    fileJ instead.  With fci (which both of the reference's callers switch
    off: src/scattering_main_routine.py:251, src/Main.f90:142,179) types 1 / 7
    build the 2-RDM with createtwordm_bit_fci from the ``i civ alpha beta``
    file and newdat(:,1), irep1, start and end; start_2, end_2, irep2 and the
    orderings are read by none of the routines.  Returns (q_abs, result)."""
    from . import readers
    lib = L.load()
    q = np.asarray(q, dtype=np.float64)
    if int(nq) != q.size or int(ngto) != np.asarray(l).size:
        raise ValueError("nq / ngto do not match the array sizes")
    if int(type) not in (1, 2, 7, 8):
        raise NotImplementedError(
            f"type {type}: only 1 (total j0), 2 (elastic j0), 7 (total j2) and 8 (elastic j2) "
            "are on this path")
    q_abs = np.zeros(q.size)
    kw = dict(cutoffz=cutoffz, cutoffmd=cutoffmd, cutoffcentre=cutoffcentre,
              precision=precision, waves=waves, maxl=maxl)
    basis = (ga, l, m, n, xx, yy, zz, mmod, q, group)
    if readrdm:
        if int(type) in (1, 7):
            mat, tot = readers.read_twordm(fileJ_2)
            f = total_from_twordm if int(type) == 1 else total_j2_from_twordm
            return q_abs, f(*basis, mat, tot, **kw)
        g = readers.read_onerdm(fileJ)
        f = elastic_from_onerdm if int(type) == 2 else elastic_j2_from_onerdm
        return q_abs, f(*basis, g, **kw)
    if fci:
        if int(type) not in (1, 7):
            raise NotImplementedError(
                f"type {type} with fci: the reference reads the FCI file only in the "
                "2-RDM types 1 and 7 (src/main_calculation.f90:587-588, 731-732)")
        a, b, c = readers.read_fci_file(fileJ, int(numberlines))
        nd = np.asarray(newdat)
        uniq = nd[:, 0] if nd.ndim == 2 else nd
        mat, tot = createtwordm_bit_fci(a, b, c, uniq, irep1, start, end)
        f = total_from_twordm if int(type) == 1 else total_j2_from_twordm
        return q_abs, f(*basis, mat, tot, **kw)
    a, b, c1, c2 = readers.read_bit_file(fileJ, int(numberlines))
    a, b, c1, c2 = _bits(a, b, c1, c2)
    bb = _basis(ga, l, m, n, xx, yy, zz, mmod, q, group, cutoffz, cutoffmd, cutoffcentre,
                maxl=maxl, ng=ng)
    o = L.make_opts(precision, waves=waves)
    res = np.zeros(q.size)
    rc = lib.thistle_total_scattering_calculation_2(
        int(type), C.byref(bb), a.size, L._ptr(a), L._ptr(b), L._ptr(c1), L._ptr(c2),
        1 if mcci else 0, C.byref(o), L._ptr(q_abs), L._ptr(res))
    L.check(rc, "total_scattering_calculation_2")
    return q_abs, res


# ----------------------------------------------------------- device stages
def quartet_list(ga, l, m, n, xx, yy, zz, group, q, cutoffcentre=0.001):
    """Reference-order quartet list computed on the GPU: int32 [Q, 6]."""
    lib = L.load()
    b = _basis(ga, l, m, n, xx, yy, zz, None, q, group, 0.0, 0.0, cutoffcentre)
    o = L.make_opts()
    cnt = lib.thistle_quartet_list(C.byref(b), None, 0, C.byref(o))
    if cnt < 0:
        raise L.ThistleError(-cnt, "quartet_list")
    out = np.zeros((cnt, 6), dtype=np.int32)
    cnt2 = lib.thistle_quartet_list(C.byref(b), L._ptr(out), cnt, C.byref(o))
    if cnt2 < 0:
        raise L.ThistleError(-cnt2, "quartet_list")
    return out


def quartet_list_screened(type_, ga, l, m, n, xx, yy, zz, mmod, q, group, mat=None,
                          total=None, onerdm=None, cutoffz=1e-15, cutoffcentre=0.001):
    lib = L.load()
    b = _basis(ga, l, m, n, xx, yy, zz, mmod, q, group, cutoffz, 0.0, cutoffcentre)
    o = L.make_opts()
    mat_ = np.require(np.asarray(mat if mat is not None else np.zeros((0, 4))), dtype=np.int32,
                      requirements=["F"])
    tot_ = np.require(np.asarray(total if total is not None else np.zeros(0),
                                 dtype=np.float64), requirements=["C"])
    g = None if onerdm is None else np.require(np.asarray(onerdm, dtype=np.float64),
                                               requirements=["F"])
    norb = 0 if g is None else g.shape[0]
    args = (int(type_), C.byref(b), int(tot_.size), L._ptr(mat_), L._ptr(tot_), norb, L._ptr(g))
    cnt = lib.thistle_quartet_list_screened(*args, None, 0, C.byref(o))
    if cnt < 0:
        raise L.ThistleError(-cnt, "quartet_list_screened")
    out = np.zeros((cnt, 6), dtype=np.int32)
    cnt2 = lib.thistle_quartet_list_screened(*args, L._ptr(out), cnt, C.byref(o))
    if cnt2 < 0:
        raise L.ThistleError(-cnt2, "quartet_list_screened")
    return out


def unique_total(matrix, total):
    """GPU unique_total: (rmat, rtot, iuni, irec) with 1-based iuni/irec."""
    lib = L.load()
    mat = np.require(np.asarray(matrix), dtype=np.int32, requirements=["F"])
    tot = np.require(np.asarray(total, dtype=np.float64), requirements=["C"])
    nrow, ncol = mat.shape
    rmat = np.zeros((max(nrow, 1), ncol), dtype=np.int32, order="F")
    rtot = np.zeros(max(nrow, 1))
    iuni = np.zeros(max(nrow, 1), dtype=np.int32)
    irec = np.zeros(max(nrow, 1), dtype=np.int32)
    o = L.make_opts()
    u = lib.thistle_unique_total(L._ptr(mat), nrow, ncol, L._ptr(tot), L._ptr(rmat),
                                 L._ptr(rtot), L._ptr(iuni), L._ptr(irec), C.byref(o))
    if u < 0:
        raise L.ThistleError(-u, "unique_total")
    rm = np.asfortranarray(rmat.reshape(-1, order="F")[: u * ncol].reshape(u, ncol, order="F"))
    return np.ascontiguousarray(rm), rtot[:u].copy(), iuni[:u].copy(), irec[:nrow].copy()


def pair_tables(ga, l, m, n, xx, yy, zz, group, q, maxl=None):
    """E[p][axis][la][lb][t], P[p][3], e12[p][nq] for group pairs gi <= gj."""
    lib = L.load()
    b = _basis(ga, l, m, n, xx, yy, zz, None, q, group, 0.0, 0.0, 0.001, maxl=maxl)
    ng, mx, nq = b.ng, b.maxl, b.nq
    mxe = max(mx, int(np.max(m)), int(np.max(n)))
    npair = ng * (ng + 1) // 2
    E = np.zeros((npair, 3, mxe + 1, mxe + 1, 2 * mxe + 1))
    P = np.zeros((npair, 3))
    e12 = np.zeros((npair, nq))
    L.check(lib.thistle_pair_tables(C.byref(b), L._ptr(E), L._ptr(P), L._ptr(e12),
                                    C.byref(L.make_opts())), "pair_tables")
    return E, P, e12


class Plan:
    """Device-resident problem: repeated executions reuse every HBM buffer."""

    def __init__(self, type_, ga, l, m, n, xx, yy, zz, mmod, q, group, mat=None, total=None,
                 onerdm=None, cutoffz=1e-15, cutoffmd=1e-20, cutoffcentre=0.001,
                 precision=L.FP64, waves=0, maxl=None):
        self.lib = L.load()
        self._b = _basis(ga, l, m, n, xx, yy, zz, mmod, q, group, cutoffz, cutoffmd,
                         cutoffcentre, maxl=maxl)
        self._o = L.make_opts(precision, waves=waves)
        self.nq = int(np.asarray(q).size)
        self._mat = np.require(np.asarray(mat if mat is not None else np.zeros((0, 4))),
                               dtype=np.int32, requirements=["F"])
        self._tot = np.require(np.asarray(total if total is not None else np.zeros(0),
                                          dtype=np.float64), requirements=["C"])
        self._g = None if onerdm is None else np.require(
            np.asarray(onerdm, dtype=np.float64), requirements=["F"])
        h = C.c_void_p()
        rc = self.lib.thistle_plan_create(
            int(type_), C.byref(self._b), int(self._tot.size), L._ptr(self._mat),
            L._ptr(self._tot), 0 if self._g is None else self._g.shape[0], L._ptr(self._g),
            C.byref(self._o), C.byref(h))
        L.check(rc, "plan_create")
        self.h = h

    def execute(self, fetch=True):
        out = np.zeros(self.nq) if fetch else None
        L.check(self.lib.thistle_plan_execute(self.h, L._ptr(out)), "plan_execute")
        return out

    def result(self):
        out = np.zeros(self.nq)
        L.check(self.lib.thistle_plan_result(self.h, L._ptr(out)), "plan_result")
        return out

    def mixed_fallback(self):
        """After a mixed-precision execution: (indices of the q points
        re-evaluated in FP64, A2(q) the sum of the squared error scales of
        the FP32-evaluated terms)."""
        n = C.c_int32()
        idx = np.zeros(self.nq, dtype=np.int32)
        scale = np.zeros(self.nq)
        L.check(self.lib.thistle_plan_mixed_fallback(self.h, C.byref(n), L._ptr(idx), self.nq,
                                                     L._ptr(scale)), "plan_mixed_fallback")
        return idx[:n.value].copy(), scale

    def set_quartets(self, positions=None):
        """Restrict later executions to these 0-based positions of the
        reference quartet enumeration (strictly increasing); None restores
        every quartet."""
        if positions is None:
            L.check(self.lib.thistle_plan_set_quartets(self.h, None, 0), "plan_set_quartets")
            return
        pos = np.require(np.asarray(positions, dtype=np.int64), requirements=["C"])
        L.check(self.lib.thistle_plan_set_quartets(self.h, L._ptr(pos), int(pos.size)),
                "plan_set_quartets")

    def timings(self):
        ms = np.zeros(5)
        nq = C.c_int64()
        units = C.c_int64()
        L.check(self.lib.thistle_plan_timings(self.h, L._ptr(ms), C.byref(nq), C.byref(units)),
                "plan_timings")
        kc, ks = C.c_double(), C.c_double()
        L.check(self.lib.thistle_plan_kernel_timings(self.h, C.byref(kc), C.byref(ks)),
                "plan_kernel_timings")
        return {"pair_tables": ms[0], "enumeration": ms[1], "transform": ms[2],
                "quartets": ms[3], "reduction": ms[4], "nquart": nq.value, "units": units.value,
                "contract": kc.value, "series": ks.value}

    def close(self):
        if getattr(self, "h", None):
            self.lib.thistle_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many total-j0 frames executed together (thistle_batch_*).

    ``frames`` is a sequence of (ga, l, m, n, xx, yy, zz, mmod, q, group, mat,
    total).  ``execute()`` returns the list of per-frame S(q); the first call
    tunes the launch and captures the batch into one hipGraph.
    """

    def __init__(self, frames, precision=L.FP64, cutoffz=1e-15, cutoffmd=1e-20,
                 cutoffcentre=0.001):
        lib = L.load()
        nf = len(frames)
        bases = (L.Basis * nf)()
        self._keep = []
        mats = (C.c_void_p * nf)()
        tots = (C.c_void_p * nf)()
        nmat = np.zeros(nf, dtype=np.int64)
        self.nq = []
        for f, fr in enumerate(frames):
            ga, l, m, n, xx, yy, zz, mmod, q, group, mat, total = fr
            b = _basis(ga, l, m, n, xx, yy, zz, mmod, q, group, cutoffz, cutoffmd, cutoffcentre)
            bases[f] = b
            self._keep.append(b._keep)
            mt = np.require(np.asarray(mat), dtype=np.int32, requirements=["F"])
            tt = np.require(np.asarray(total, dtype=np.float64), requirements=["C"])
            self._keep += [mt, tt]
            mats[f] = mt.ctypes.data
            tots[f] = tt.ctypes.data
            nmat[f] = tt.size
            self.nq.append(int(np.asarray(q).size))
        self._keep += [bases, mats, tots, nmat]
        o = L.make_opts(precision)
        self._h = C.c_void_p()
        L.check(lib.thistle_batch_create(nf, C.cast(bases, C.c_void_p), L._ptr(nmat),
                                         C.cast(mats, C.c_void_p), C.cast(tots, C.c_void_p),
                                         C.byref(o), C.byref(self._h)), "batch_create")
        self._lib = lib
        self._off = np.concatenate([[0], np.cumsum(self.nq)])

    def execute(self, fetch=True):
        out = np.zeros(int(self._off[-1])) if fetch else None
        L.check(self._lib.thistle_batch_execute(self._h, L._ptr(out)), "batch_execute")
        if not fetch:
            return None
        return [out[self._off[f]:self._off[f + 1]].copy() for f in range(len(self.nq))]

    def info(self):
        ms = C.c_double()
        units = C.c_int64()
        waves = C.c_int32()
        graph = C.c_int32()
        L.check(self._lib.thistle_batch_info(self._h, C.byref(ms), C.byref(units), C.byref(waves),
                                             C.byref(graph)), "batch_info")
        rows, groups = C.c_int32(), C.c_int32()
        L.check(self._lib.thistle_batch_tuning(self._h, C.byref(rows), C.byref(groups)),
                "batch_tuning")
        return {"ms": ms.value, "units": units.value, "waves": waves.value,
                "graph": bool(graph.value), "rows": rows.value, "groups": groups.value}

    def close(self):
        if self._h:
            self._lib.thistle_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
