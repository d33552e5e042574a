"""Host side of the bit-string route: the file readers, the boundary's
argument checks (which run before any device call) and properties of the
oracle's bit builders that the mathematics fixes."""
import os

import numpy as np
import pytest

import oracle
from thistle_amd import lib as L
from thistle_amd import readers


def test_bit_file_round_trip(tmp_path, golden):
    g = golden("bit_mid")
    p = str(tmp_path / "bits.dat")
    readers.write_bit_file(p, g["alpha"], g["beta"], g["c1"], g["c2"])
    a, b, c1, c2 = readers.read_bit_file(p)
    assert np.array_equal(a, g["alpha"]) and np.array_equal(b, g["beta"])
    assert np.array_equal(c1, g["c1"]) and np.array_equal(c2, g["c2"])
    a2, _, _, _ = readers.read_bit_file(p, 10)
    assert np.array_equal(a2, g["alpha"][:10])
    with pytest.raises(ValueError):
        readers.read_bit_file(p, len(a) + 1)


def test_read_onerdm(tmp_path):
    p = tmp_path / "onerdm.dat"
    p.write_text("1 1 1.5\n1 3 0.25\n2 2 0.5D0\n3 3 2.0\n")
    g = readers.read_onerdm(str(p))
    ref = np.zeros((3, 3))
    ref[0, 0], ref[0, 2], ref[1, 1], ref[2, 2] = 1.5, 0.25, 0.5, 2.0
    assert np.array_equal(g, ref)


def _inv_parity(a, b):
    """(-1)^#{(beta j, alpha k): k > j}: the sign of moving every alpha slot
    ahead of every beta slot from orbital order."""
    n = 0
    for j in range(64):
        if (b >> j) & 1:
            n += bin(a >> (j + 1)).count("1")
    return -1.0 if n & 1 else 1.0


def test_mcci_sign_is_the_reordering_parity(golden):
    """The swap procedure of mcci_to_bit (restated literally in the oracle)
    gives the parity of the alpha-then-beta reordering, the closed form the
    device kernel evaluates."""
    for name in ("small", "mid", "wide", "mrci"):
        g = golden(f"bit_{name}")
        o1, _ = oracle.mcci_to_bit(g["alpha"], g["beta"], g["c1"], g["c2"], text_round=False)
        ph = np.array([_inv_parity(int(a), int(b)) for a, b in zip(g["alpha"], g["beta"])])
        assert np.array_equal(o1, ph * g["c1"]), name


def test_bit_2rdm_invariants(golden):
    """createtwordm_bit's SpinFree2RDM for a normalised vector (c1 = c2):
    trace sum_pq G(p,q,q,p) = N(N-1) and the pair symmetry G(p,q,r,s) =
    G(q,p,s,r); a 1e-12 tolerance covers the rounding of the sums."""
    g = golden("bit_mid")
    a, b, c = g["alpha"], g["beta"], g["c1"]
    c = c / np.linalg.norm(c)
    tw = oracle.twordm_bit_dense(a, b, c, c)
    ne = bin(int(a[0])).count("1") + bin(int(b[0])).count("1")
    tr = np.einsum("pqqp->", tw)
    assert abs(tr - ne * (ne - 1)) < 1e-12 * ne * ne
    assert np.allclose(tw, tw.transpose(1, 0, 3, 2), atol=1e-13)
    # contracting one pair gives the 1-RDM: sum_q G(p,q,q,s) = (N-1) g(p,s)
    g1 = oracle.one_rdm_bit(a, b, c)
    red = np.einsum("pqqs->ps", tw)
    assert np.allclose(red, (ne - 1) * g1, atol=1e-12)


def _fock_2rdm(dets, c, norb, slot):
    """G(a,b,c,d) = sum over spins s, t of <Psi| a+_{a s} a+_{b t} a_{c t}
    a_{d s} |Psi> by second quantisation on the spin-orbital Fock space;
    slot(k, spin) numbers spin-orbitals, each determinant is the product of
    its creators in increasing slot order."""

    def occ(det):
        al, be = det
        return sorted([slot(k, 0) for k in range(norb) if (al >> k) & 1] +
                      [slot(k, 1) for k in range(norb) if (be >> k) & 1])

    def apply(ops, state):
        out = {}
        for slots, amp in state.items():
            cur, sgn, ok = list(slots), amp, True
            for kind, so in reversed(ops):
                if kind == "a":
                    if so not in cur:
                        ok = False
                        break
                    i = cur.index(so)
                    sgn *= (-1) ** i
                    cur.pop(i)
                else:
                    if so in cur:
                        ok = False
                        break
                    i = sum(1 for x in cur if x < so)
                    sgn *= (-1) ** i
                    cur.insert(i, so)
            if ok:
                out[tuple(cur)] = out.get(tuple(cur), 0.0) + sgn
        return out

    psi = {}
    for det, ci in zip(dets, c):
        psi[tuple(occ(det))] = psi.get(tuple(occ(det)), 0.0) + ci
    G = np.zeros((norb,) * 4)
    for idx in np.ndindex(*G.shape):
        pa, pb, pc, pd = idx
        v = 0.0
        for s1 in (0, 1):
            for s2 in (0, 1):
                ops = [("c", slot(pa, s1)), ("c", slot(pb, s2)), ("a", slot(pc, s2)),
                       ("a", slot(pd, s1))]
                for k, amp in apply(ops, psi).items():
                    v += psi.get(k, 0.0) * amp
        G[idx] = v
    return G


def test_bit_2rdm_brute_force():
    """createtwordm_bit equals the spin-summed 2-RDM of second quantisation
    with the alpha-then-beta slot order its phases assume; after
    mcci_to_bit, coefficients given in orbital order (alpha before beta on
    each orbital, the MCCI convention) give the same state's 2-RDM."""
    norb = 4
    dets = [(0b0011, 0b0101), (0b0101, 0b0011), (0b1001, 0b0110), (0b0110, 0b1010),
            (0b1100, 0b0011)]
    a = np.array([d[0] for d in dets], dtype=np.int64)
    b = np.array([d[1] for d in dets], dtype=np.int64)
    c = np.array([0.7, -0.5, 0.3, 0.2, -0.35])
    c /= np.linalg.norm(c)
    ref = _fock_2rdm(dets, c, norb, lambda k, s: k + norb * s)
    assert np.allclose(oracle.twordm_bit_dense(a, b, c, c), ref, atol=1e-14)
    ref_i = _fock_2rdm(dets, c, norb, lambda k, s: 2 * k + s)
    o1, o2 = oracle.mcci_to_bit(a, b, c, c, text_round=False)
    assert np.allclose(oracle.twordm_bit_dense(a, b, o1, o2), ref_i, atol=1e-14)
    # the 1-RDM: sum_s <a+_{p s} a_{q s}>, the contraction of the 2-RDM
    ne = 4
    g1 = oracle.one_rdm_bit(a, b, c)
    assert np.allclose(np.einsum("pqqs->ps", ref), (ne - 1) * g1, atol=1e-14)


def test_boundary_rejects_bad_bit_strings():
    lib = L.load()
    p = lambda x: L._ptr(x)  # noqa: E731
    c = np.ones(3)
    out = np.zeros(3)
    good_a = np.array([3, 5, 6], dtype=np.int64)
    good_b = np.array([3, 3, 5], dtype=np.int64)
    # varying alpha count
    bad_a = np.array([3, 7, 6], dtype=np.int64)
    assert lib.thistle_createtwordm_bit(p(bad_a), p(good_b), p(c), p(c), 3, 0.0, None, None,
                                        None, 0) == -1
    # negative bit string
    neg = np.array([3, -5, 6], dtype=np.int64)
    assert lib.thistle_createtwordm_bit(p(neg), p(good_b), p(c), p(c), 3, 0.0, None, None,
                                        None, 0) == -1
    # mcci_to_bit needs equal alpha and beta counts
    b1 = np.array([1, 2, 4], dtype=np.int64)
    assert lib.thistle_mcci_to_bit(p(good_a), p(b1), p(c), p(c), 3, 1, None, p(out),
                                   p(out)) == 1
    # n = 0
    assert lib.thistle_one_rdm_bit(p(good_a), p(good_b), p(c), 0, None, 0,
                                   p(np.zeros(1, np.int32)), None) == 1
    # rows requested without buffers
    assert lib.thistle_createtwordm_bit(p(good_a), p(good_b), p(c), p(c), 3, 0.0, None, None,
                                        None, 5) == -1
    assert lib.thistle_total_scattering_calculation_2(3, None, 3, p(good_a), p(good_b), p(c),
                                                      p(c), 0, None, None, p(out)) == 7
    assert L.RDM_CUTOFF == float(np.float32(1e-30))


def test_fci_boundary_rejects_what_the_reference_cannot_process(golden):
    """createtwordm_bit_fci's input checks, all before any device call."""
    lib = L.load()
    g = golden("fci_small")
    a, b, c = g["alpha"].copy(), g["beta"].copy(), g["civs"].copy()
    p = L._ptr

    def call(a, b, c, uniq, irep, start, end):
        return lib.thistle_createtwordm_bit_fci(p(a), p(b), p(c), a.size, p(uniq), uniq.size,
                                                p(irep), p(start), p(end), 0.0, None, None,
                                                None, 0)

    uniq, irep, start, end = readers.fci_tables(a)
    # a determinant moved out of its alpha string's block
    perm = np.arange(a.size)
    perm[[3, 40]] = perm[[40, 3]]
    u2, r2, s2, e2 = readers.fci_tables(a[perm])
    assert call(a[perm].copy(), b[perm].copy(), c[perm].copy(), u2, r2, s2, e2) == -1
    # a zero-padded string row (the driver's padding when there are more
    # unique beta strings than alpha strings)
    z = np.concatenate([[0], uniq]).astype(np.int64)
    assert call(a, b, c, z, np.concatenate([[0], irep]), np.concatenate([[1], start]),
                np.concatenate([[1], end])) == -1
    # a beta string that is no alpha string
    b2 = b.copy()
    b2[0] = 0b10001 | 0b00100  # three electrons
    assert call(a, b2, c, uniq, irep, start, end) == -1
    # fewer unique strings than norbs + 1
    keep = np.isin(a, uniq[:4]) & np.isin(b, uniq[:4])
    u3, r3, s3, e3 = readers.fci_tables(a[keep])
    assert call(a[keep].copy(), b[keep].copy(), c[keep].copy(), u3, r3, s3, e3) == -1


def test_fci_tables_match_the_driver():
    a = np.array([3, 3, 5, 5, 5, 6], dtype=np.int64)
    uniq, irep, start, end = readers.fci_tables(a)
    assert uniq.tolist() == [3, 5, 6] and irep.tolist() == [2, 3, 1]
    assert start.tolist() == [1, 3, 6] and end.tolist() == [2, 5, 6]
