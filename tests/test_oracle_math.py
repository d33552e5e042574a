"""Pins of the CPU oracle that do not come from the reference build: closed
forms, invariants, brute force on tiny inputs and the known-answer tests of
the reference's own test suite (test/runtests.jl, docs/src/kernels.md,
src/UniqueModule.jl doctests, src/pzerotable.py) and data files
(src/HF_NonZero2RM_MO_Coeffs, src/NonZero2RDM_MO_Coeffs).
"""
import itertools
import math
from collections import Counter

import numpy as np
import pytest
from numpy.polynomial.hermite import hermval
from scipy.special import spherical_jn

import oracle

PI_SP = 3.1415927410125732421875
COMPS = {0: [(0, 0, 0)], 1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
         2: [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)],
         3: [(3, 0, 0), (0, 3, 0), (0, 0, 3), (1, 2, 0), (2, 1, 0), (2, 0, 1), (1, 0, 2),
             (0, 1, 2), (0, 2, 1), (1, 1, 1)]}


def shells_basis(shells):
    """(L, exponent, centre) shells -> flat cartesian GTO arrays, one group each."""
    cols = {k: [] for k in ("ga", "l", "m", "n", "xx", "yy", "zz", "group")}
    for g, (L, a, c) in enumerate(shells):
        for (i, j, k) in COMPS[L]:
            for key, v in zip(cols, (a, i, j, k, c[0], c[1], c[2], g + 1)):
                cols[key].append(v)
    return {k: np.array(v) for k, v in cols.items()}


# ---------------------------------------------------------------- a8: MD tables
def test_md_tables_hermite_identity():
    """E_t^{la,lb}: (x-A)^la (x-B)^lb e^{-a(x-A)^2-b(x-B)^2} = K_AB sum_t E_t Lambda_t(x),
    Lambda_t = (d/dP)^t e^{-p(x-P)^2} = p^{t/2} H_t(sqrt(p)(x-P)) e^{-p(x-P)^2}
    (McMurchie-Davidson; Helgaker, Joergensen, Olsen eq. 9.5.6)."""
    shells = [(3, 0.7, (0.1, -0.4, 0.3)), (2, 1.9, (1.2, 0.5, -0.8)),
              (3, 0.35, (-0.6, 1.1, 0.9)), (1, 5.0, (0.1, -0.4, 0.3))]
    b = shells_basis(shells)
    dd, P, _ = oracle.pair_tables(b["l"], b["m"], b["n"], b["group"], b["ga"], b["xx"],
                                  b["yy"], b["zz"], np.array([0.0]))
    xs = np.linspace(-3, 3, 25)
    worst = 0.0
    for ax in range(3):
        for gi, gj in itertools.product(range(len(shells)), repeat=2):
            a, A = shells[gi][1], shells[gi][2][ax]
            bb, B = shells[gj][1], shells[gj][2][ax]
            p = a + bb
            Pc = (a * A + bb * B) / p
            assert P[ax][gi, gj] == pytest.approx(Pc, abs=1e-15)
            K = math.exp(-a * bb / p * (A - B) ** 2)
            y = math.sqrt(p) * (xs - Pc)
            for la in range(shells[gi][0] + 1):
                for lb in range(shells[gj][0] + 1):
                    lhs = (xs - A) ** la * (xs - B) ** lb * np.exp(-a * (xs - A) ** 2
                                                                   - bb * (xs - B) ** 2)
                    rhs = K * np.exp(-y * y) * sum(
                        dd[ax][gi, gj, la, lb, t] * p ** (t / 2) * hermval(y, [0] * t + [1])
                        for t in range(la + lb + 1))
                    worst = max(worst, np.max(np.abs(lhs - rhs)) / np.max(np.abs(lhs)))
                    # no coefficient beyond t = la + lb
                    assert np.all(dd[ax][gi, gj, la, lb, la + lb + 1:] == 0.0)
    assert worst < 1e-13, worst


# ------------------------------------------------------------------- a9: e12
def test_e12_overlap_and_q_decay():
    """e12(0) = (pi_sp/pi)^1.5 * overlap of the two s Gaussians (quadrature), and
    e12(q)/e12(0) = |FT of the product Gaussian| / norm = exp(-q^2/4p) (quadrature)."""
    shells = [(0, 0.8, (0.0, 0.2, -0.3)), (0, 2.5, (0.7, -0.1, 0.4)), (0, 0.3, (-1.0, 0.5, 0.2))]
    b = shells_basis(shells)
    q = np.array([0.0, 0.5, 1.7, 4.0])
    _, _, e12 = oracle.pair_tables(b["l"], b["m"], b["n"], b["group"], b["ga"], b["xx"],
                                   b["yy"], b["zz"], q)
    x = np.linspace(-12, 12, 24001)
    for gi, gj in itertools.product(range(3), repeat=2):
        a, A = shells[gi][1], np.array(shells[gi][2])
        bb, B = shells[gj][1], np.array(shells[gj][2])
        ov = 1.0
        for ax in range(3):
            ov *= np.trapezoid(np.exp(-a * (x - A[ax]) ** 2 - bb * (x - B[ax]) ** 2), x)
        np.testing.assert_allclose(e12[gi, gj, 0], ov * (PI_SP / np.pi) ** 1.5, rtol=1e-10)
        p = a + bb
        g = np.exp(-p * x * x)
        for iq, qq in enumerate(q):
            # q along one axis: the other two axes' transforms cancel in the ratio
            ratio = np.trapezoid(g * np.cos(qq * x), x) / np.trapezoid(g, x)
            np.testing.assert_allclose(e12[gi, gj, iq] / e12[gi, gj, 0], ratio, rtol=1e-9)


# -------------------------------------------------------------------- a5: P0
# src/pzerotable.py:4-184 (sorted exponent triple -> value at nu = q)
PZERO_TABLE = {(0, 0, 0): (1, 0, 1), (0, 0, 2): (-1, 2, 3), (0, 2, 2): (1, 4, 15),
               (2, 2, 2): (-1, 6, 105), (0, 0, 4): (1, 4, 5), (0, 2, 4): (-1, 6, 35),
               (2, 2, 4): (1, 8, 315), (2, 4, 4): (-1, 10, 1155), (4, 4, 4): (1, 12, 5005),
               (0, 0, 6): (-1, 6, 7), (4, 4, 6): (-1, 14, 15015)}


def test_p0_sphere_average_and_table():
    """P0(L,M,N,q) = <(i q x)^L (i q y)^M (i q z)^N> over the unit sphere:
    Gauss-Legendre in cos(theta) x uniform phi (exact for these polynomials),
    plus the hand table of src/pzerotable.py."""
    lmax4 = 12
    q = np.array([0.3, 1.1, 2.5])
    P0 = oracle.set_p0(lmax4, q)
    ct, wt = np.polynomial.legendre.leggauss(40)
    nphi = 64
    phi = 2 * np.pi * np.arange(nphi) / nphi
    st = np.sqrt(1 - ct * ct)
    X = np.outer(st, np.cos(phi))
    Y = np.outer(st, np.sin(phi))
    Zc = np.outer(ct, np.ones(nphi))
    W = np.outer(wt, np.ones(nphi)) / (2 * nphi)
    for L, M, N in itertools.product(range(lmax4 + 1), repeat=3):
        avg = np.sum(W * X ** L * Y ** M * Zc ** N)
        r = L + M + N
        ipow = [1, 1j, -1, -1j][r % 4]
        expect = (ipow * avg * q ** r).real if r % 2 == 0 else np.zeros_like(q)
        err = np.abs(P0[L, M, N] - expect)
        assert np.all(err <= 1e-11 * np.abs(expect) + 1e-13 * q ** r), (L, M, N, err)
    for (a, b, c), (sg, pw, den) in PZERO_TABLE.items():
        for perm in set(itertools.permutations((a, b, c))):
            np.testing.assert_allclose(P0[perm[0], perm[1], perm[2]], sg * q ** pw / den,
                                       rtol=1e-13)


# --------------------------------------------------------- a14: Bessel series
def test_dphrec_matches_spherical_jn():
    """dphrec's normalised backward recurrence gives j_l(x) for x below ~75."""
    x = np.concatenate([np.geomspace(1e-4, 1.0, 12), np.linspace(1.3, 70.0, 60)])
    R = oracle.dphrec(x, 21)
    for l in range(21):
        j = spherical_jn(l, x)
        scale = np.maximum(np.abs(j), 1e-3 * np.max(np.abs(j)))
        err = np.max(np.abs(R[:, l] - j) / scale)
        assert err < 1e-11, (l, err)


def test_bessels0rr_series_closed_form():
    """F(q) = h0 sin(qH)/(qH) + sum_{b>=1} h_b j_b(qH) (q/H)^b (src/Integrals.f90:6662-6739)."""
    rng = np.random.default_rng(7)
    mu = np.linspace(0.05, 9.0, 40)
    for order in (0, 3, 8, 16):
        for H in (0.37, 2.1, 7.5):
            h = rng.normal(size=order + 1)
            F = oracle.bessels0rr(order, mu, H, h)
            ref = sum(h[b] * spherical_jn(b, mu * H) * (mu / H) ** b for b in range(order + 1))
            scale = np.sum(np.abs(h)[None, :] * np.abs((mu / H)[:, None]) ** np.arange(order + 1),
                           axis=1)
            assert np.max(np.abs(F - ref) / scale) < 1e-12, (order, H)
    assert np.all(oracle.bessels0rr(5, mu, 1.3, np.zeros(6)) == 0.0)


def test_form_factor_kat():
    """test/runtests.jl:57-80 and docs/src/kernels.md:56-79: an s-s-s-s quartet
    with zcontr = zcontr2 = 0.25 at unit distance gives 0.5 j0(mu h), and with
    mu in 1/Angstrom and h in bohr: 0.49976668, 0.49790243."""
    bohr = 0.529177210903
    F = oracle.bessels0rr(0, np.array([0.1, 0.3]) * bohr, 1.0, [0.25 + 0.25])
    np.testing.assert_array_equal(np.round(F, 8), [0.49976668, 0.49790243])
    mu = np.array([0.1, 0.2])
    h = math.hypot(0.05, math.hypot(0.02, 0.01))
    F = oracle.bessels0rr(0, mu, h, [0.75])
    np.testing.assert_allclose(F, 0.75 * np.sin(mu * h) / (mu * h), rtol=1e-15)


# ---------------------------------------------------- a11: quartet enumeration
def _orbit(t):
    i, j, k, r = t
    out = set()
    for a, b in ((i, j), (j, i)):
        for c, d in ((k, r), (r, k)):
            out.add((a, b, c, d))
            out.add((c, d, a, b))
    return out


@pytest.mark.parametrize("ng", [1, 2, 3, 5, 8, 13, 30])
def test_quartet_list_invariants(ng):
    """L2: count Q_ref = P(P+1)/2 + C(ng,3), multiplicities sum to ng^4 and
    every ordered group quadruple is covered exactly once by the orbits of the
    listed quartets (nest 2 lists (i,j,i,r) and (i,r,i,j) with 4 each); the
    h = 0 branch flag is h < cutoffcentre or nest 5."""
    rng = np.random.default_rng(ng)
    px, py, pz = (rng.normal(size=(ng, ng)) for _ in range(3))
    for a in (px, py, pz):
        a[:] = a + a.T
        # a few coincident pair centres (exactly, and just inside the cutoff)
        if ng >= 3:
            a[0, 2] = a[2, 0] = a[1, 1]
            a[1, 2] = a[2, 1] = a[0, 0] + 3e-4
    cc = 0.001
    L = oracle.quartet_list(px, py, pz, cc)
    P = ng * (ng + 1) // 2
    assert len(L) == P * (P + 1) // 2 + math.comb(ng, 3)
    assert int(L[:, 4].sum()) == ng ** 4
    if ng <= 13:
        cover = Counter()
        for gi, gj, gk, gr, mult, br in L.tolist():
            orb = _orbit((gi, gj, gk, gr))
            for o in orb:
                cover[o] += mult / len(orb)
        assert len(cover) == ng ** 4
        assert set(cover) == set(itertools.product(range(1, ng + 1), repeat=4))
        assert all(abs(v - 1.0) < 1e-12 for v in cover.values())
    # 1-based group labels, as the reference's loops
    assert L[:, :4].min() >= 1 and L[:, :4].max() <= ng
    for gi, gj, gk, gr, mult, br in (L - np.array([1, 1, 1, 1, 0, 0])).tolist():
        h = math.sqrt((px[gk, gr] - px[gi, gj]) ** 2 + (py[gk, gr] - py[gi, gj]) ** 2
                      + (pz[gk, gr] - pz[gi, gj]) ** 2)
        nest5 = gi == gj == gk == gr
        assert br == int(nest5 or h < cc)
        assert mult == {True: 1}.get(nest5, mult) and mult in (1, 2, 4, 8)


@pytest.mark.parametrize("ng", [1, 2, 3, 5, 8, 13])
def test_quartet_list_elastic_j2_invariants(ng):
    """elastic_integration_j2's six nests (src/Integrals.f90:2004-2325): the
    (i,j,i,r) family is listed once per unordered {j, r} (mult 8) plus
    (i,j,i,j) (mult 4); multiplicities still sum to ng^4 and the orbits still
    cover every ordered group quadruple exactly once."""
    rng = np.random.default_rng(100 + ng)
    px, py, pz = (rng.normal(size=(ng, ng)) for _ in range(3))
    for a in (px, py, pz):
        a[:] = a + a.T
    cc = 0.001
    L = oracle.quartet_list_elastic_j2(px, py, pz, cc)
    L0 = oracle.quartet_list(px, py, pz, cc)
    m = np.arange(ng)[::-1]  # ng - 1 - i
    n_j0_nest2 = int(np.sum(m * m))
    n_j2_nest2 = int(np.sum(m * (m - 1) // 2 + m))
    assert len(L) == len(L0) - n_j0_nest2 + n_j2_nest2
    assert int(L[:, 4].sum()) == ng ** 4
    cover = Counter()
    for gi, gj, gk, gr, mult, br in L.tolist():
        orb = _orbit((gi, gj, gk, gr))
        for o in orb:
            cover[o] += mult / len(orb)
    assert set(cover) == set(itertools.product(range(1, ng + 1), repeat=4))
    assert all(abs(v - 1.0) < 1e-12 for v in cover.values())
    # (i,j,i,j) sits at h = 0: always the h = 0 branch
    for gi, gj, gk, gr, mult, br in L.tolist():
        if gi == gk and gj == gr:
            assert br == 1 and mult in (1, 4)


# ------------------------------------- a10-a15: whole path, s shells, brute force
def _s_case(seed, nat=4, nmo=3):
    rng = np.random.default_rng(seed)
    cent = rng.uniform(-1.5, 1.5, size=(nat, 3))
    shells = []
    for a in range(nat):
        for e in (0.35, 1.4):
            shells.append((0, float(e * rng.uniform(0.8, 1.25)), tuple(cent[a])))
    b = shells_basis(shells)
    ngto = len(shells)
    C = rng.normal(size=(nmo, ngto))
    return b, C


def _w_s(b, q):
    """e12(ij) e12(kl) j0(q |P_ij - P_kl|) for s Gaussians, all ordered quadruples."""
    a = b["ga"]
    R = np.stack([b["xx"], b["yy"], b["zz"]], axis=1)
    p = a[:, None] + a[None, :]
    Pc = (a[:, None, None] * R[:, None, :] + a[None, :, None] * R[None, :, :]) / p[..., None]
    d2 = np.sum((R[:, None, :] - R[None, :, :]) ** 2, axis=-1)
    K = np.exp(-a[:, None] * a[None, :] / p * d2)
    e12 = (PI_SP / p[..., None]) ** 1.5 * np.exp(-q[None, None, :] ** 2 / (4 * p[..., None])) \
        * K[..., None]
    n = len(a)
    h = np.linalg.norm(Pc.reshape(n * n, 1, 3) - Pc.reshape(1, n * n, 3), axis=-1)
    x = h[..., None] * q[None, None, :]
    j0 = np.where(x > 0, np.sin(x) / np.where(x > 0, x, 1.0), 1.0)
    E = e12.reshape(n * n, -1)
    return (E[:, None, :] * E[None, :, :] * j0).reshape(n, n, n, n, len(q))


def test_total_j0_s_shells_brute_force():
    """S(q) = sum_n Gamma_n sum_{ijkl} C_pi C_sj C_qk C_rl e12 e12 j0 for rows
    (p,s,q,r): the five nests with multiplicities 8/4/4/2/1 and the symmetrised
    transform zcontrred + zcontrred2 reproduce the unrestricted ordered sum."""
    b, C = _s_case(3)
    confs = np.array([[1, 2, 0, 0, 0, 0], [1, 0, 0, 2, 0, 0], [0, 0, 1, 2, 0, 0],
                      [0, 2, 1, 0, 0, 0], [0, 0, 0, 0, 1, 2]], dtype=np.int32)
    civs = np.random.default_rng(4).normal(size=(5, 1))
    civs /= np.linalg.norm(civs)
    mat, tot = oracle.createtwordm(confs, civs)
    q = np.array([1e-10, 0.4, 1.3, 2.9, 5.5])
    S = oracle.total_j0(b["l"], b["m"], b["n"], b["group"], b["ga"], b["xx"], b["yy"],
                        b["zz"], C, mat, tot, q)
    w = _w_s(b, q)
    ref = np.zeros_like(q)
    for (p_, s_, q_, r_), v in zip(mat - 1, tot):
        ref += v * np.einsum("i,j,k,l,ijklq->q", C[p_], C[s_], C[q_], C[r_], w)
    np.testing.assert_allclose(S, ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())


def test_elastic_j0_s_shells_brute_force():
    """S_el(q) = sum_{ijkl} Z_ij Z_kl e12 e12 j0, Z = 1/2 sum gamma (C C + C C)."""
    b, C = _s_case(5)
    confs = np.array([[1, 2, 0, 0, 0, 0], [1, 0, 0, 2, 0, 0], [0, 2, 1, 0, 0, 0]],
                     dtype=np.int32)
    civs = np.array([[0.8], [0.5], [np.sqrt(1 - 0.89)]])
    gam = oracle.onerdm(confs, civs)
    Z = oracle.elastic_z(gam, C)
    Zr = 0.5 * (C.T @ gam @ C + (C.T @ gam @ C).T)
    np.testing.assert_allclose(Z, Zr, rtol=1e-13, atol=1e-15)
    q = np.array([1e-10, 0.7, 2.2, 6.0])
    S = oracle.elastic_j0(b["l"], b["m"], b["n"], b["group"], b["ga"], b["xx"], b["yy"],
                          b["zz"], Z, q)
    ref = np.einsum("ij,kl,ijklq->q", Zr, Zr, _w_s(b, q))
    np.testing.assert_allclose(S, ref, rtol=1e-12)


# ------------------------------------------------------------ a6/a7: 2-RDM
def _apply(det, ops):
    """Apply a product of (slot, create?) operators right to left to an
    occupation tuple; returns (sign, det) or None."""
    occ = list(det)
    sign = 1
    for slot, create in reversed(ops):
        if occ[slot] == (1 if create else 0):
            return None
        if sum(occ[:slot]) % 2:
            sign = -sign
        occ[slot] = 1 if create else 0
    return sign, tuple(occ)


def _brute_rdms(confs, c):
    """Spin-summed Gamma(p,q,r,s) = sum_{st} <a+_{ps} a+_{qt} a_{rt} a_{ss}> and
    gamma(p,s) = sum_s <a+_{ps} a_{ss}> by second quantisation; spin orbital
    slot 2k is alpha of orbital k, 2k+1 beta (odd / even input columns)."""
    nslot = confs.shape[1]
    norb = nslot // 2
    dets = {tuple(int(v != 0) for v in row): ci for row, ci in zip(confs, c)}
    G = np.zeros((norb,) * 4)
    g = np.zeros((norb, norb))
    for det, cj in dets.items():
        for p, s in itertools.product(range(norb), repeat=2):
            for sg in range(2):
                res = _apply(det, [(2 * p + sg, True), (2 * s + sg, False)])
                if res and res[1] in dets:
                    g[p, s] += dets[res[1]] * cj * res[0]
        for p, q, r, s in itertools.product(range(norb), repeat=4):
            for sg, tg in itertools.product(range(2), repeat=2):
                res = _apply(det, [(2 * p + sg, True), (2 * q + tg, True),
                                   (2 * r + tg, False), (2 * s + sg, False)])
                if res and res[1] in dets:
                    G[p, q, r, s] += dets[res[1]] * cj * res[0]
    return G, g


def _ms0_dets(norb, nel):
    out = []
    na = nel // 2
    for A in itertools.combinations(range(norb), na):
        for B in itertools.combinations(range(norb), na):
            row = [0] * (2 * norb)
            for a in A:
                row[2 * a] = 1
            for bb in B:
                row[2 * bb + 1] = 2
            out.append(row)
    return np.array(out, dtype=np.int32)


@pytest.mark.parametrize("norb,nel,seed", [(2, 2, 1), (3, 2, 2), (4, 4, 3)])
def test_rdms_second_quantisation(norb, nel, seed):
    """createtwordm / createonerdm against brute-force second quantisation on
    Ms = 0 CAS spaces, plus the trace and partial-trace identities."""
    confs = _ms0_dets(norb, nel)
    c = np.random.default_rng(seed).normal(size=(len(confs), 1))
    c /= np.linalg.norm(c)
    G = oracle.twordm_dense(confs, c)
    g = oracle.onerdm(confs, c)
    Gb, gb = _brute_rdms(confs, c[:, 0])
    np.testing.assert_allclose(G, Gb, atol=1e-13)
    np.testing.assert_allclose(g, gb, atol=1e-13)
    assert np.einsum("pqqp->", G) == pytest.approx(nel * (nel - 1), abs=1e-12)
    np.testing.assert_allclose(np.einsum("pqqs->ps", G), (nel - 1) * g, atol=1e-12)


def test_ci_bookkeeping_kat():
    """test/runtests.jl:165-179: two determinants over two orbitals."""
    confs = np.array([[1, 0, 1, 0], [0, 2, 1, 0]], dtype=np.int32)
    civs = np.array([[0.8, 0.1], [0.6, 0.2]])
    assert oracle.pair_phase(confs, 1, 1) == (1, 0)
    assert oracle.pair_phase(confs, 1, 2) == (1, 2)
    assert oracle.pair_phase(confs, 2, 1) == (1, 2)
    g = oracle.onerdm(confs, civs)
    assert g[0, 0] == pytest.approx(1.96, abs=1e-12)
    assert g[1, 1] == pytest.approx(1.0, abs=1e-12)
    assert g[0, 1] == 0.0
    tw = oracle.twordm_dense(confs, civs)
    assert tw[1, 0, 0, 1] == pytest.approx(1.0, abs=1e-12)
    assert tw[0, 0, 0, 0] == 0.0


def test_hf_twordm_file(golden):
    """src/HF_NonZero2RM_MO_Coeffs is createtwordm on a closed-shell
    determinant of 7 doubly occupied MOs labelled 1-5, 11, 15: rows, order
    and values."""
    d = golden("ref_rdm_files")
    confs = np.array([[1, 2] * 7], dtype=np.int32)
    mat, tot = oracle.createtwordm(confs, np.ones((1, 1)))
    label = np.array([1, 2, 3, 4, 5, 11, 15])
    assert np.array_equal(label[mat - 1], d["hf_mat"].astype(np.int64))
    assert np.array_equal(tot, d["hf_total"])
    assert len(tot) == 2 * 7 * 7 - 7


def test_cas_twordm_file_invariants(golden):
    """src/NonZero2RDM_MO_Coeffs (13 MOs, 10 electrons): trace N(N-1), a
    symmetric 1-RDM from the partial trace, rows in first-index-fastest order
    of twordm(p,q,r,s) and no duplicates."""
    d = golden("ref_rdm_files")
    mat, tot = d["cas_mat"].astype(np.int32), d["cas_total"]
    p, s, q, r = (mat[:, i] - 1 for i in range(4))
    G = np.zeros((13,) * 4)
    G[p, q, r, s] = tot
    assert np.einsum("pqqp->", G) == pytest.approx(90.0, abs=1e-9)
    g9 = np.einsum("pqqs->ps", G)
    np.testing.assert_allclose(g9, g9.T, atol=1e-10)
    assert np.trace(g9) / 9.0 == pytest.approx(10.0, abs=1e-9)
    lin = ((s * 13 + r) * 13 + q) * 13 + p
    assert np.all(np.diff(lin) > 0)
    rmat, rtot, _, _ = oracle.unique_total(mat, tot)
    assert np.array_equal(rmat, mat) and np.array_equal(rtot, tot)


# ---------------------------------------------------------------- a17: unique
def test_unique_total_doctests():
    """src/UniqueModule.jl:45-68, 124-154."""
    rmat, rtot, iuni, irec = oracle.unique_total(np.array([[1, 2], [3, 4], [1, 2]]),
                                                 np.array([10.0, 20.0, 1.5]))
    assert rmat.tolist() == [[1, 2], [3, 4]]
    assert rtot.tolist() == [11.5, 20.0]
    assert iuni.tolist() == [1, 2] and irec.tolist() == [1, 2, 1]
    rmat, _, iuni, irec = oracle.unique_total(np.array([[1, 2], [3, 4], [1, 2], [3, 4]]),
                                              np.zeros(4))
    assert rmat.tolist() == [[1, 2], [3, 4]]
    assert iuni.tolist() == [1, 2] and irec.tolist() == [1, 2, 1, 2]


def test_unique_total_brute_force():
    rng = np.random.default_rng(9)
    m = rng.integers(1, 4, size=(400, 4))
    t = rng.normal(size=400)
    rmat, rtot, iuni, irec = oracle.unique_total(m, t)
    first, sums = {}, {}
    for i, row in enumerate(map(tuple, m.tolist())):
        first.setdefault(row, i)
        sums[row] = sums.get(row, 0.0) + t[i]
    keys = sorted(first, key=first.get)
    assert rmat.tolist() == [list(k) for k in keys]
    np.testing.assert_allclose(rtot, [sums[k] for k in keys], rtol=1e-14, atol=1e-15)
    assert iuni.tolist() == [first[k] + 1 for k in keys]
    pos = {k: i + 1 for i, k in enumerate(keys)}
    assert irec.tolist() == [pos[tuple(r)] for r in m.tolist()]
