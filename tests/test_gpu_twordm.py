"""createtwordm and createonerdm on the GPU (thistle_createtwordm,
thistle_createonerdm) against the reference's own
rows and values (flang build of src/twordms.f90, the golden fixtures) and
against the host accumulation, bit for bit: same rows in the same order and
identical doubles.  Covers ndiff 0 / 2 / 4 pairs, core orbitals, transition
2-RDMs (state1 != state2), duplicate determinants, more than 64 spin-orbital
slots (two mask words) and non-canonical determinant lists (host route)."""
import itertools

import numpy as np
import pytest

from thistle_amd import scattering as sc

pytestmark = pytest.mark.gpu


def cas_confs(ncore, nact, nalpha, nbeta, rng=None, nmax=None):
    """Canonical determinant rows (1 on even slots, 2 on odd): doubly
    occupied core, then every (alpha, beta) occupation of the active space."""
    norb = ncore + nact
    rows = []
    for ao in itertools.combinations(range(nact), nalpha):
        for bo in itertools.combinations(range(nact), nbeta):
            r = np.zeros(2 * norb, dtype=np.int32)
            for k in range(ncore):
                r[2 * k], r[2 * k + 1] = 1, 2
            for a in ao:
                r[2 * (ncore + a)] = 1
            for b in bo:
                r[2 * (ncore + b) + 1] = 2
            rows.append(r)
    confs = np.array(rows, dtype=np.int32)
    if rng is not None:
        confs = confs[rng.permutation(len(confs))]
    if nmax is not None:
        confs = confs[:nmax]
    return confs


def both(confs, civs, state1=1, state2=1, cutoff=1e-30):
    mg, tg = sc.createtwordm(confs, civs, state1, state2, cutoff, device=True)
    mh, th = sc.createtwordm(confs, civs, state1, state2, cutoff, device=False)
    return mg, tg, mh, th


@pytest.mark.parametrize("name", ["h2", "sp3", "spd2", "f2", "xyx_cas44"])
def test_reference_goldens(name, golden):
    g = golden(name)
    mat, tot = sc.createtwordm(g["confs"], g["civs"], device=True)
    assert np.array_equal(mat, g["mat"].astype(np.int32)), name
    assert np.array_equal(tot, g["total"]), name


@pytest.mark.parametrize("ncore,nact,na,nb,seed", [(0, 6, 3, 3, 1), (3, 6, 3, 2, 2),
                                                   (5, 8, 4, 4, 3), (2, 7, 2, 4, 4)])
def test_cas_vs_host(ncore, nact, na, nb, seed):
    rng = np.random.default_rng(seed)
    confs = cas_confs(ncore, nact, na, nb, rng)
    civs = rng.standard_normal((len(confs), 2))
    civs /= np.linalg.norm(civs, axis=0)
    for s1, s2 in [(1, 1), (1, 2), (2, 1)]:
        mg, tg, mh, th = both(confs, civs, s1, s2)
        print(f"cas({nact},{na}+{nb}) core {ncore} states {s1},{s2}: ndet={len(confs)} "
              f"rows={len(th)}", flush=True)
        assert len(th) > 0
        assert np.array_equal(mg, mh) and np.array_equal(tg, th)


def test_screening_cutoff():
    rng = np.random.default_rng(7)
    confs = cas_confs(2, 6, 3, 3, rng)
    civs = rng.standard_normal(len(confs))
    civs /= np.linalg.norm(civs)
    for cut in (1e-30, 1e-6, 1e-3):
        mg, tg, mh, th = both(confs, civs, cutoff=cut)
        assert np.array_equal(mg, mh) and np.array_equal(tg, th)


def test_two_mask_words():
    """34 orbitals: 68 slots, two 64-bit words per determinant."""
    rng = np.random.default_rng(11)
    confs = cas_confs(30, 4, 2, 2, rng)
    civs = rng.standard_normal(len(confs))
    civs /= np.linalg.norm(civs)
    mg, tg, mh, th = both(confs, civs)
    assert np.array_equal(mg, mh) and np.array_equal(tg, th)


def test_duplicates_and_single():
    rng = np.random.default_rng(5)
    confs = cas_confs(1, 4, 2, 2, rng)
    confs = np.vstack([confs, confs[:3]])  # repeated determinants: ndiff 0 with c1 != c2
    civs = rng.standard_normal(len(confs))
    mg, tg, mh, th = both(confs, civs)
    assert np.array_equal(mg, mh) and np.array_equal(tg, th)
    one = confs[:1]
    mg, tg, mh, th = both(one, np.array([1.0]))
    assert np.array_equal(mg, mh) and np.array_equal(tg, th)


def test_noncanonical_goes_host():
    rng = np.random.default_rng(9)
    confs = cas_confs(1, 4, 2, 2, rng)
    confs[0, confs[0] == 1] = 2  # an alpha slot carrying 2: not the punch reader's form
    civs = rng.standard_normal(len(confs))
    mg, tg, mh, th = both(confs, civs)
    assert np.array_equal(mg, mh) and np.array_equal(tg, th)


# ------------------------------------------------------------ createonerdm
def both1(confs, civs, state1=1, state2=1):
    return (sc.onerdm(confs, civs, state1, state2, device=True),
            sc.onerdm(confs, civs, state1, state2, device=False))


@pytest.mark.parametrize("name", ["h2", "sp3", "spd2", "f2", "xyx_cas44"])
def test_onerdm_reference_goldens(name, golden):
    g = golden(name)
    gd, gh = both1(g["confs"], g["civs"])
    assert np.array_equal(gd, gh), name
    n = gd.shape[0]
    np.testing.assert_allclose(gd, g["onerdm"][:n, :n], rtol=0, atol=1e-14)


@pytest.mark.parametrize("ncore,nact,na,nb,seed", [(0, 6, 3, 3, 1), (3, 6, 3, 2, 2),
                                                   (5, 8, 4, 4, 3), (30, 4, 2, 2, 4)])
def test_onerdm_cas_vs_host(ncore, nact, na, nb, seed):
    """CAS spaces with core, a transition 1-RDM, and 68 slots (two mask words)."""
    rng = np.random.default_rng(seed)
    confs = cas_confs(ncore, nact, na, nb, rng)
    civs = rng.standard_normal((len(confs), 2))
    civs /= np.linalg.norm(civs, axis=0)
    for s1, s2 in ((1, 1), (1, 2)):
        gd, gh = both1(confs, civs, s1, s2)
        assert np.array_equal(gd, gh), (ncore, nact, s1, s2)
    gd, _ = both1(confs, civs)
    assert np.trace(gd) == pytest.approx(2 * ncore + na + nb, abs=1e-12)


def test_onerdm_duplicates_and_noncanonical():
    rng = np.random.default_rng(7)
    confs = cas_confs(1, 4, 2, 2, rng)
    dup = np.vstack([confs, confs[:3]])  # ndiff 0 with c1 != c2
    civs = rng.standard_normal(len(dup))
    gd, gh = both1(dup, civs)
    assert np.array_equal(gd, gh)
    nc = confs.copy()
    nc[0, nc[0] == 1] = 2  # not the punch reader's form: host route
    civs = rng.standard_normal(len(nc))
    gd, gh = both1(nc, civs)
    assert np.array_equal(gd, gh)
