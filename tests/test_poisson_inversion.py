"""The Poisson inversion of the mixed-precision kernels (pst_device.hpp), on
the CPU: its method against the exact inverse CDF.

* Asymptotic inversion (10 <= lambda < 2^16): a float32 model of
  poisson_asym_f32 with the coefficient tables read from pst_device.hpp; the
  model's root x~ must lie within the kernel's exactness band POIS_EPS of the
  exact root x* of Q(x*, lambda) = u (scipy), over lambda and u including the
  far tails, and the band must be rare enough for the FP64 decisions.  The
  GPU's own outputs are checked draw by draw in tests/test_gpu_poisson.py.
* Sequential-search inversion (lambda < 10, FP32): a uniform of 1.0, or just
  below it, above the FP32 sum of the pmf must still end the search well
  before its cap (the search stops when the sum stops growing).
"""
import os
import re

import numpy as np
import pytest
from scipy import optimize, special

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
f32 = np.float32


def _tables():
    text = open(os.path.join(ROOT, "py_stoch_amd", "csrc", "pst_device.hpp")).read()
    out = {}
    for name, body in re.findall(r"constexpr float POIS_([A-Z0-9]+)\[\d+\] = \{([^}]*)\};", text):
        out[name] = [f32(float(v.strip().rstrip("f"))) for v in body.split(",")]
    assert set(out) == {"QW", "D0", "D1", "EC", "ET"}
    return out


T = _tables()


def _horner(c, z):
    acc = c[-1]
    for a in c[-2::-1]:
        acc = f32(np.float64(acc) * np.float64(z) + np.float64(a))  # fused multiply-add
    return acc


def _asym(lam, u):
    """float32 model of poisson_asym_f32: (L, y, eps, central, s)."""
    with np.errstate(all="ignore"):  # w far outside the central range only for lambda < 60
        return _asym_model(lam, u)


def _asym_model(lam, u):
    lam, u = f32(lam), f32(min(u, 1 - 2.0 ** -25))
    x = f32(2 * u - 1)
    lw = f32(-np.log(np.float64(f32(4 * f32(u * f32(1 - u))))))
    p = _horner(T["EC"], f32(lw - 2.5)) if lw < 5 else _horner(T["ET"], f32(f32(np.sqrt(lw)) - 3))
    s = f32(f32(1.41421356) * f32(p * x))
    rs = f32(1 / np.sqrt(np.float64(lam)))
    w = f32(s * rs)
    sl, L, rl = f32(lam * rs), f32(np.floor(lam)), f32(1 / np.float64(lam))
    y = f32(f32(lam - L) + f32(f32(s * sl) * _horner(T["QW"], w)))
    y = f32(y + f32(_horner(T["D0"], w) + f32(_horner(T["D1"], w) * rl)))
    eps = f32(0.025) * rl * rl + f32(20 * 2.0 ** -24) * (1 + abs(s)) * sl
    return float(L), float(y), float(eps), (-0.7 <= w <= 1.0), float(s)


def _xstar(lam, u):
    """exact root x* of Q(x*, lam) = u (Q the regularized upper incomplete gamma)."""
    lam, u = float(lam), min(float(u), 1 - 2.0 ** -25)
    fn = (lambda a: special.gammaincc(a, lam) - u) if u <= 0.5 else (lambda a: (1 - u) - special.gammainc(a, lam))
    return optimize.brentq(fn, max(1e-9, lam - 12 * np.sqrt(lam) - 5), lam + 12 * np.sqrt(lam) + 10,
                           xtol=1e-12, rtol=1e-15)


def test_asymptotic_inversion_within_its_band():
    rng = np.random.default_rng(7)
    lams = np.concatenate([[10.0, 10.5, 12.0, 31.0, 110.0, 4096.0, 65535.0],
                           np.exp(rng.uniform(np.log(10), np.log(65536), 13))])
    worst, n, flagged = 0.0, 0, 0
    for lam in lams:
        ks = np.concatenate([rng.integers(0, 2 ** 24, 150), [0, 1, 2, 2 ** 24 - 2, 2 ** 24 - 1]])
        for k in ks:
            u = (k + 1) * 2.0 ** -24
            L, y, eps, central, s = _asym(lam, u)
            if not central:
                continue  # decided in FP64 by the kernel
            err = abs((L + y) - _xstar(f32(lam), u))
            worst = max(worst, err / eps)
            fr = y - np.floor(y)
            flagged += (fr < eps) or (fr > 1 - eps)
            n += 1
    assert n > 2500
    assert worst < 0.5, worst  # the band is at least twice the model's error
    assert flagged / n < 2e-3


def test_band_is_needed_planted_coefficient_error():
    """A plausible slip -- the median correction 1/3 dropped from D0 -- puts
    the root outside the band: the check above would catch it."""
    saved = T["D0"][0]
    T["D0"][0] = f32(0.0)
    try:
        L, y, eps, central, s = _asym(110.0, 0.5)
        assert abs((L + y) - _xstar(f32(110.0), 0.5)) > 10 * eps
    finally:
        T["D0"][0] = saved


def _search_f32(lam, u, stop_when_flat=True, cap=100):
    """float32 model of the kernels' two-step sequential search (lambda < 10)."""
    lam, u = f32(lam), f32(u)
    p = f32(np.exp(-np.float64(lam)))
    F, K, it = p, 0, 0
    live = u > F
    while it < cap and live:
        p1 = f32(f32(p * lam) * f32(1.0 / (it + 1)))
        F1 = f32(F + p1)
        p2 = f32(f32(p1 * lam) * f32(1.0 / (it + 2)))
        F2 = f32(F1 + p2)
        K = it + 2 if u > F1 else it + 1
        live = (u > F2) and (F2 > F if stop_when_flat else True)
        p, F = p2, F2
        it += 2
    return K


@pytest.mark.parametrize("u", [1.0, 1 - 2.0 ** -24, 1 - 2.0 ** -23])
def test_search_ends_well_before_its_cap(u):
    lams = np.linspace(0.01, 9.99, 500)
    ks = [_search_f32(l, u) for l in lams]
    assert max(ks) < 40
    if u == 1.0:
        # without the stop rule some means run into the cap (the sum ends below 1.0)
        assert max(_search_f32(l, u, stop_when_flat=False) for l in lams) >= 99
