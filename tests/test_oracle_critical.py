"""Pinning the oracle's critical-reaction branch -- tau_2 = (1/a0c) ln(1/rho_1)
and SingleCriticalReaction (P1_P2.py:175-220, 488-507) -- against exact laws.

* Draw level (oracle hook on the same function the simulation calls): with
  a non-critical propensity 1000 beside critical ones 2 and 6, tau_2 must be
  Exp(a0c = 8) and the pick must take reaction j with probability a_j / a0c;
  with a finite tau_1 there is no pick with probability exp(-a0c tau_1).
* Trajectory level: when every consuming reaction is critical each leap is
  one exact firing, and the recorded counts follow the hypoexponential
  "midpoint" law of tests/laws.py (the reference interpolates samples inside
  a leap); the split of two competing decays is k2 / (k1 + k2), exactly so
  once the state has absorbed.
* Two plausible mistakes planted in the oracle (tau_2 from a0, the pick
  normalised by a0) must fail.
"""
import numpy as np
import pytest
from scipy import stats

from py_stoch_amd import models
from py_stoch_amd.mod.parser.tl_parser import tables_of

import laws

A = np.array([1000.0, 2.0, 6.0])
CRIT = np.array([0, 1, 1])
A0C = 8.0


def _draw_gaps(tau, j, tau1=np.inf):
    """(KS of tau against its exact law, chi-square p-value of the pick)."""
    if np.isinf(tau1):
        ks = stats.kstest(tau, stats.expon(scale=1.0 / A0C).cdf).statistic
        picked = j
    else:
        ks = abs(np.mean(j < 0) - np.exp(-A0C * tau1))  # P(no pick)
        picked = j[j >= 0]
    counts = np.array([(picked == 1).sum(), (picked == 2).sum()])
    expect = picked.size * np.array([2.0, 6.0]) / A0C
    return ks, stats.chisquare(counts, expect).pvalue, counts


def test_critical_draw_laws(oracle_mod):
    n = 200000
    tau, j = oracle_mod.critical_draws(A, CRIT, np.inf, seed=41, n=n)
    ks, p, counts = _draw_gaps(tau, j)
    assert (j >= 1).all()  # only critical reactions are picked
    assert ks < 1.95 / np.sqrt(n), ks
    assert p > 1e-4, counts
    # tau_1 = 0.1: no critical firing with probability exp(-0.8); tau = min
    tau, j = oracle_mod.critical_draws(A, CRIT, 0.1, seed=42, n=n)
    gap, p, counts = _draw_gaps(tau, j, tau1=0.1)
    assert gap < 5 * np.sqrt(0.25 / n), gap
    assert p > 1e-4, counts
    assert np.all(tau[j < 0] == np.float32(0.1)) and np.all(tau[j >= 0] < 0.1)


@pytest.mark.parametrize("plant", ["tau2_a0", "pick_a0"])
def test_planted_draw_mistakes_are_caught(oracle_mod, plant):
    n = 20000
    with oracle_mod.planted(plant):
        tau, j = oracle_mod.critical_draws(A, CRIT, np.inf, seed=43, n=n)
    ks, p, _ = _draw_gaps(tau, j)
    assert ks > 1.95 / np.sqrt(n) or p < 1e-4
    tau, j = oracle_mod.critical_draws(A, CRIT, np.inf, seed=43, n=n)  # knob reset
    ks, p, _ = _draw_gaps(tau, j)
    assert ks < 1.95 / np.sqrt(n) and p > 1e-4


def _run(oracle_mod, spn, E, t_max, ita, U, seed, **kw):
    a = models.make_args(spn, E=E, t_max=t_max, ita=ita, U=U)
    O, flags, st = oracle_mod.simulate(tables_of(a), a.rate_fn(a.c), a.x_0, method=oracle_mod.METHOD_TAU_LEAPING,
                                       t_max=t_max, I=a.I, E=a.E, n_sims=U, seed=seed, **kw)
    assert flags.sum() == 0
    return a, O, st


def test_pure_death_all_critical(oracle_mod):
    n, k = 50, 1.0
    a, O, st = _run(oracle_mod, laws.pure_death(k=k, x0=n), [0], 3.0, 7, 4000, seed=31, n_c=laws.BIG_NC)
    # every step is a leap made of exactly one critical firing
    assert st["ssa_steps"] == 0 and st["rejected_leaps"] == 0
    assert st["leaps"] == st["events"] > 0
    for f in range(1, 7):
        t = float(a.I[f])
        laws.assert_law(O[0, f], laws.midpoint_death_law(n, k, t), "X(%g)" % t)


def test_competing_decays_split(oracle_mod):
    k1, k2, n = 1.0, 3.0, 30
    a, O, st = _run(oracle_mod, laws.competing_decays(k1, k2, n), [0, 1], 6.0, 7, 4000, seed=32,
                    n_c=laws.BIG_NC)
    assert st["ssa_steps"] == 0 and st["leaps"] == st["events"] > 0
    p = k2 / (k1 + k2)
    for f in (1, 2):  # t = 1, 2: still decaying
        t = float(a.I[f])
        laws.assert_law(O[0, f], laws.midpoint_death_law(n, k1 + k2, t), "X(%g)" % t)
        laws.assert_law(O[1, f], laws.midpoint_split_law(n, k1 + k2, p, t), "Y(%g)" % t)
    # t = 6: every simulation has absorbed at X = 0 (P(X > 0) < 1e-8)
    assert (O[0, -1] == 0).all()
    laws.assert_law(O[1, -1], laws.binom(n, p), "Y(6)")
