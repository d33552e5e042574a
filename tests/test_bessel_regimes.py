"""The small-argument Bessel regime of the series kernel (series.hip f_miller,
adjoint.hip b_series): Miller's backward recurrence for g_b = j_b(x) / x^b
started M(LL) orders above LL from (0, 1) (series_math.hpp miller_orders),
normalised by j0 or by j1 = x g_1, whichever is larger in magnitude, on
[0, xup); the upward recurrence from j0, j1 above xup = LL - 1 (0.5 at
LL = 1, 1.2 at LL = 2).  Both are restated here in float64 numpy, step for
step, and checked against a long-double recurrence started 80 orders up.

The error is measured against the envelope of g_b: |g_b| below the turning
point x = sqrt(b (b + 1)), the spherical Hankel modulus / x^b above it.  A
planted mistake (normalising by j0 alone) fails the bound near x = pi.
"""
import numpy as np
import pytest

# orders above LL at which Miller's recurrence starts (SeriesRegime::nstart -
# LL): the least start meeting 4e-15 on [0, xup) plus two, at most 16
MILLER = {1: 7, 2: 9, 3: 10, 4: 11, 5: 12, 6: 13, 7: 13, 8: 14, 9: 15, 10: 15,
          11: 16, 12: 16, 13: 16, 14: 16, 15: 16, 16: 16}


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


def _xup(LL):
    return 0.5 if LL == 1 else 1.2 if LL == 2 else LL - 1.0


def _max_err(LL, start, by_j1=True, xup=None, npts=2500):
    xup = _xup(LL) if xup is None else xup
    x = np.linspace(1e-4, xup, npts, endpoint=False)
    g = _reference(LL, x)
    tp = np.sqrt(np.arange(LL + 1) * (np.arange(LL + 1) + 1.0))
    env = np.where(x[None, :] < tp[:, None], np.abs(g), _hankel_over_xb(LL, x, g))
    got = _device_miller(LL, x, start, by_j1)
    return float(np.max(np.abs(got - g) / env.astype(np.float64)))


@pytest.mark.parametrize("LL", list(range(1, 17)))
def test_miller_start_within_envelope(LL):
    assert _max_err(LL, MILLER[LL]) < 4e-15
    # and three orders fewer would not hold (the start is not arbitrary)
    if LL >= 2:
        assert _max_err(LL, MILLER[LL] - 3) > 4e-15


def test_j0_only_normalisation_fails_near_its_zero():
    # the planted mistake: x = pi lies in the Miller regime from LL = 5 on
    assert _max_err(6, MILLER[6], by_j1=False) > 1e-13


def _device_upward(LL, x, h=1.3):
    # f_up64: u_b = j_b (q/h)^b = g_b q^(2b),
    # u_{b+1} = (2b+1)/h^2 u_b - (q/h)^2 u_{b-1}, with q = x / h
    x = np.asarray(x, dtype=np.float64)
    q = x / h
    invh = 1.0 / h
    ih2 = invh * invh
    j0 = np.sin(x) / x
    nr2 = -((q * invh) * (q * invh))
    us = [j0, (j0 - np.cos(x)) * ih2]
    for b in range(1, LL):
        us.append((2 * b + 1) * ih2 * us[-1] + nr2 * us[-2])
    return np.array([us[b] / q ** (2 * b) for b in range(LL + 1)])


@pytest.mark.parametrize("LL", list(range(1, 17)))
def test_upward_from_xup_within_envelope(LL):
    x = np.linspace(_xup(LL), 0.9 * LL + 8.0, 3000)
    g = _reference(LL, x)
    tp = np.sqrt(np.arange(LL + 1) * (np.arange(LL + 1) + 1.0))
    env = np.where(x[None, :] < tp[:, None], np.abs(g), _hankel_over_xb(LL, x, g))
    err = np.max(np.abs(_device_upward(LL, x) - g) / env.astype(np.float64))
    assert err < 2e-14
    # and it would not hold a unit lower (the boundary is not arbitrary)
    if LL >= 4:
        xl = np.linspace(_xup(LL) - 1.0, _xup(LL), 800)
        gl = _reference(LL, xl)
        envl = np.where(xl[None, :] < tp[:, None], np.abs(gl), _hankel_over_xb(LL, xl, gl))
        assert np.max(np.abs(_device_upward(LL, xl) - gl) / envl.astype(np.float64)) > 2e-14


# ---- the upward regime's sum: Clenshaw's recurrence (series.hip f_up64)

def _long_g(LL, x):
    # g_b = j_b / x^b in long double, Miller started far above max(x)
    x = np.asarray(x, dtype=np.longdouble)
    x2 = x * x
    vp1 = np.zeros_like(x)
    v = np.ones_like(x)
    out = np.zeros((LL + 1,) + x.shape, dtype=np.longdouble)
    for b in range(LL + int(x.max()) + 90, 0, -1):
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


def _device_clenshaw(LL, x, h, hs, flip=1.0):
    # f_up64 step for step; flip = -1 plants a sign error in (q/h)^2
    q = x / h
    invh = 1.0 / h
    ih2 = invh * invh
    j0 = np.sin(x) / x
    cx = np.cos(x)
    cb = [(2 * b + 1) * ih2 for b in range(LL + 1)]
    r = q * invh
    nr2 = -flip * (r * r)
    y1, y2 = np.full_like(x, hs[LL]), np.zeros_like(x)
    if LL >= 2:
        y2 = y1
        y1 = cb[LL - 1] * y2 + hs[LL - 1]
        for b in range(LL - 2, 0, -1):
            y0 = cb[b] * y1 + (nr2 * y2 + hs[b])
            y2, y1 = y1, y0
    u1 = (j0 - cx) * cb[0]
    return y1 * u1 + j0 * (nr2 * y2 + hs[0])


def _clenshaw_err(LL, flip=1.0):
    rng = np.random.default_rng(100 + LL)
    worst = 0.0
    for h in (0.3, 5.0):
        x = np.linspace(_xup(LL), 74.0, 1500)
        g = _long_g(LL, x)
        env_g = _hankel_over_xb(LL, x, g)
        q = x.astype(np.longdouble) / np.longdouble(h)
        u = np.array([g[b] * q ** (2 * b) for b in range(LL + 1)])
        eu = np.array([env_g[b] * q ** (2 * b) for b in range(LL + 1)])
        hs = rng.standard_normal(LL + 1) * np.exp(rng.uniform(-3.0, 3.0, LL + 1))
        ref = np.sum(hs[:, None].astype(np.longdouble) * u, axis=0)
        env = np.sum(np.abs(hs[:, None]).astype(np.longdouble) * eu, axis=0)
        got = _device_clenshaw(LL, x, h, hs, flip)
        worst = max(worst, float(np.max(np.abs(got - ref) / env)))
    return worst


@pytest.mark.parametrize("LL", list(range(1, 17)))
def test_clenshaw_sum_within_envelope(LL):
    # sum_b h(b) u_b within 4e-15 of sum_b |h(b)| times u_b's envelope
    assert _clenshaw_err(LL) < 4e-15


def test_clenshaw_sign_error_fails():
    assert _clenshaw_err(6, flip=-1.0) > 1e-3
