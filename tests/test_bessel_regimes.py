"""The small-argument Bessel regime of the series kernel (series.hip f_miller,
adjoint.hip b_series): Miller's backward recurrence for g_b = j_b(x) / x^b
started 16 orders above LL from (0, 1), normalised by j0 or by j1 = x g_1,
whichever is larger in magnitude.  Restated here in float64 numpy, step for
step, and checked against a long-double recurrence started 80 orders up on
[0, xup), xup = 0.9 LL + 1 (above it the kernel recurs upward).

The error is measured against the envelope of g_b: |g_b| below the turning
point x = sqrt(b (b + 1)), the spherical Hankel modulus / x^b above it.  A
planted mistake (normalising by j0 alone) fails the bound near x = pi.
"""
import numpy as np
import pytest

START = 16  # orders above LL (SeriesRegime::nstart - LL)


def _reference(LL, x):
    x = np.asarray(x, dtype=np.longdouble)
    x2 = x * x
    vp1 = np.zeros_like(x)
    v = np.ones_like(x)
    out = np.zeros((LL + 1,) + x.shape, dtype=np.longdouble)
    for b in range(LL + 80, 0, -1):
        vm = (2 * b + 1) * v - x2 * vp1
        vp1, v = v, vm
        if b - 1 <= LL:
            out[b - 1] = vm
        s = np.abs(v) + np.abs(vp1)
        big = s > 1e300
        if big.any():
            v[big] /= s[big]
            vp1[big] /= s[big]
            out[:, big] /= s[big]
    j0 = np.sin(x) / x
    j1 = (j0 - np.cos(x)) / x
    pick = np.abs(j0) >= np.abs(j1)
    return out * np.where(pick, j0 / out[0], j1 / (x * out[1]))


def _hankel_over_xb(LL, x, g):
    x = np.asarray(x, dtype=np.longdouble)
    ys = [-np.cos(x) / x, -np.cos(x) / x ** 2 - np.sin(x) / x]
    for b in range(1, LL):
        ys.append((2 * b + 1) / x * ys[-1] - ys[-2])
    y = np.array([ys[b] / x ** b for b in range(LL + 1)])
    return np.sqrt(g * g + y * y)


def _device_miller(LL, x, start, by_j1=True):
    x = np.asarray(x, dtype=np.float64)
    x2 = x * x
    vp1 = np.zeros_like(x)
    v = np.ones_like(x)
    for b in range(LL + start, LL, -1):
        vm = (2 * b + 1) * v - x2 * vp1
        vp1, v = v, vm
    vals = {LL: v.copy()}
    for b in range(LL, 0, -1):
        vm = (2 * b + 1) * v - x2 * vp1
        vp1, v = v, vm
        vals[b - 1] = v.copy()
    j0 = np.sin(x) / x
    j1 = (j0 - np.cos(x)) / x
    pick = np.abs(j0) >= np.abs(j1) if by_j1 else np.ones_like(x, dtype=bool)
    nrm = np.where(pick, j0 / vals[0], j1 / (x * vals[1]))
    return np.array([vals[b] * nrm for b in range(LL + 1)])


def _max_err(LL, by_j1=True, npts=2500):
    xup = 0.9 * LL + 1.0
    x = np.linspace(1e-4, xup, npts, endpoint=False)
    g = _reference(LL, x)
    tp = np.sqrt(np.arange(LL + 1) * (np.arange(LL + 1) + 1.0))
    env = np.where(x[None, :] < tp[:, None], np.abs(g), _hankel_over_xb(LL, x, g))
    got = _device_miller(LL, x, START, by_j1)
    return float(np.max(np.abs(got - g) / env.astype(np.float64)))


@pytest.mark.parametrize("LL", list(range(1, 17)))
def test_miller_start_16_within_envelope(LL):
    assert _max_err(LL) < 4e-15


def test_j0_only_normalisation_fails_near_its_zero():
    # the planted mistake: x = pi lies in the Miller regime from LL = 3 on
    assert _max_err(6, by_j1=False) > 1e-13
