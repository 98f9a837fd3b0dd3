"""Known answers for the leap decision of P1_P2.py:96-173 and the
interpolation rule of SaveInterpolatedDynamics (P1_P2.py:542-584), computed
by hand from the reference's formulas."""
import math

import numpy as np
import pytest

from py_stoch_amd import models
from py_stoch_amd.mod.parser.tl_parser import tables_of


def _f32(x):
    return float(np.float32(x))


def test_lotka_volterra_equilibrium_tau(oracle_mod):
    a = models.make_args(models.lotka_volterra(), E=[0, 1])
    prop, a0, crit, tau1, use_ssa = oracle_mod.leap_decision(tables_of(a), a.rate_fn(a.c), [1000, 1000])
    assert list(prop) == [10000.0, 10000.0, 10000.0] and a0 == 30000.0
    assert list(crit) == [0, 0, 0]
    # mu = (0, 0), sigma^2 = (2e4, 2e4), g = 2 (H = 2, H_type = 1)
    # num = max(0.03 * 1000 / 2, 1) = 15, tau1 = fminf(15/|0|, 225/2e4) rounded through FP32
    assert tau1 == _f32(15.0 ** 2 / 2e4)
    assert not use_ssa  # 0.01125 > 10 / 3e4


def test_schlogl_propensities_and_switch(oracle_mod):
    a = models.make_args(models.schlogl(), E=[0])
    prop, a0, crit, tau1, use_ssa = oracle_mod.leap_decision(tables_of(a), a.rate_fn(a.c), [250, 100000, 200000])
    expect = [3e-7 / 2 * 1e5 * 250 * 249, 1e-4 / 6 * 250 * 249 * 248, 1e-3 * 2e5, 3.5 * 250]
    np.testing.assert_allclose(prop, expect, rtol=1e-14)
    assert prop[0] == pytest.approx(933.75) and prop[3] == 875.0
    # X: g = 3 + 1/249 + 2/248; mu = a1 - a2 + a3 - a4; sigma^2 = sum a
    g = 3.0 + 1.0 / 249.0 + 2.0 / 248.0
    num = _f32(max(_f32(0.03 * 250 / g), 1.0))
    mu = expect[0] - expect[1] + expect[2] - expect[3]
    s2 = sum(expect)
    ref = _f32(min(_f32(num / _f32(abs(mu))), _f32(num * num / s2)))
    assert tau1 == pytest.approx(ref, rel=1e-7)
    assert use_ssa == (tau1 < 10.0 / a0)


def test_critical_reactions_on_net_change(oracle_mod):
    a = models.make_args(models.lotka_volterra(), E=[0, 1])
    # X = 9 < 10 * 1: reaction 2 (X + Y -> 2Y, V_X = -1) is critical; Y = 5000 keeps 3 non-critical
    _, _, crit, _, _ = oracle_mod.leap_decision(tables_of(a), a.rate_fn(a.c), [9, 5000])
    assert list(crit) == [0, 1, 0]
    _, _, crit, _, _ = oracle_mod.leap_decision(tables_of(a), a.rate_fn(a.c), [10, 9])
    assert list(crit) == [0, 0, 1]  # 1 * 10 > 10 is false for X; 10 > 9 for Y (reaction 3)


@pytest.mark.parametrize("h, ht, x, expect", [
    (1, 1, 10, 1.0), (2, 1, 10, 2.0), (2, 2, 10, 2.0 + 1.0 / 9.0), (2, 2, 1, math.inf),
    (3, 1, 10, 3.0), (3, 2, 10, 1.5 * (2.0 + 1.0 / 9.0)), (3, 3, 10, 3.0 + 1.0 / 9.0 + 2.0 / 8.0),
    (3, 3, 3, 3.0 + 0.5 + 2.0), (3, 3, 2, math.inf), (4, 1, 10, 0.0), (0, 0, 5, 0.0)])
def test_calculate_g_branches(oracle_mod, h, ht, x, expect):
    assert oracle_mod.calc_g(h, ht, x) == expect


@pytest.mark.parametrize("x, xp, t_old, tau, rec, expect", [
    (10, 20, 0.0, 1.0, 0.25, 12),    # 2.5 rounds half to even
    (10, 20, 0.0, 1.0, 0.35, 14),    # 3.5 -> 4
    (10, 20, 0.0, 1.0, 0.0, 10),
    (10, 20, 0.0, 1.0, 1.0, 20),
    (50, 30, 2.0, 0.5, 2.25, 40),    # -10 exactly
    (50, 31, 2.0, 0.5, 2.125, 45),   # -4.75 -> -5
    (7, 7, 1.0, 0.1, 1.05, 7),
    (0, 2147483520, 0.0, 1.0, 1.0, 2147483520),  # largest float below 2^31 converts exactly
    (0, 1 << 24, 0.0, 1.0, 1.0, 1 << 24)])
def test_interpolation_rule(oracle_mod, x, xp, t_old, tau, rec, expect):
    assert oracle_mod.interpolate(x, xp, t_old, tau, rec) == expect
