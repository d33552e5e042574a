"""The closed form that the GPU createtwordm uses in place of maxc_individual's
swap alignment (src/twordms.f90:83-172), checked against the swap procedure
itself on random slot sets.

The procedure: for each position i of c1's occupied slots m1 (ascending), if
m2[i] differs, find m1[i] in m2 and swap it into place, flipping the sign.
The closed form, with a / b the slots unique to c1 / c2 (ascending) and C the
common ones:
  single (one each): b lands at a's place, sign (-1)^#C strictly between;
  double: b goes to a's places in ascending order, except when all of a lie
  above (or below) all of b with an odd number of C between the groups,
  where the b are exchanged; sign (-1)^(#C(a0, b') + #C(a1, b'') + exchanged).
"""
import numpy as np


def swap_align(m1, m2):
    m2 = list(m2)
    e = 1
    for i in range(len(m1)):
        if m1[i] != m2[i]:
            for j in range(len(m2)):
                if m1[i] == m2[j]:
                    e = -e
                    m2[i], m2[j] = m2[j], m2[i]
                    break
    return e, m2


def closed_form(common, a, b):
    def btw(x, y):
        lo, hi = min(x, y), max(x, y)
        return sum(1 for c in common if lo < c < hi)
    if len(a) == 1:
        return (-1) ** btw(a[0], b[0]), [b[0]]
    sw = 0
    if a[0] > b[1]:
        sw = btw(b[1], a[0]) % 2
    elif a[1] < b[0]:
        sw = btw(a[1], b[0]) % 2
    bs = [b[1], b[0]] if sw else [b[0], b[1]]
    return (-1) ** (btw(a[0], bs[0]) + btw(a[1], bs[1]) + sw), bs


def test_closed_form_matches_swaps():
    rng = np.random.default_rng(2024)
    for _ in range(20000):
        k = int(rng.integers(1, 3))
        n = int(rng.integers(k, 40))
        slots = rng.choice(np.arange(1, 96), size=n + k, replace=False).tolist()
        common, a, b = slots[:n - k], sorted(slots[n - k:n]), sorted(slots[n:n + k])
        m1, m2 = sorted(common + a), sorted(common + b)
        e, fin = swap_align(m1, m2)
        ce, cb = closed_form(common, a, b)
        assert e == ce
        assert [fin[m1.index(x)] for x in a] == cb


def test_separated_groups_both_orders_occur():
    """The exchange case is real: both orders appear when a lies above b."""
    seen = set()
    for cm in range(4):
        common = list(range(10, 10 + cm))
        b, a = [1, 2], [50, 51]
        m1, m2 = sorted(common + a), sorted(common + b)
        _, fin = swap_align(m1, m2)
        seen.add(tuple(fin[m1.index(x)] for x in a))
    assert seen == {(1, 2), (2, 1)}
