"""Pinning the oracle's simulation algorithm against closed forms and exact laws.

* birth-death 0 -> X (10), X -> 0 (0.1) from X0 = 0: X(t) is exactly
  Poisson(100 (1 - exp(-0.1 t)));
* first-order chain s1 -> s2 -> s3 (rates 1, 1) from (n, 0, 0): s1(t) is
  Binomial(n, exp(-t)) and s2(t) Binomial(n, t exp(-t));
* Schloegl (the reference's examples/schlogl_sbml.xml structure): the law of
  X(t) from the truncated chemical master equation (scipy expm_multiply);
* forced SSA through the tau-leaping driver equals pure SSA draw for draw;
* tau-leaping against SSA on the same model (the reference's own TL-vs-SSA
  comparison in examples/lotkaVolterra_sim.py:98-135, as assertions).
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy import stats
from scipy.sparse.linalg import expm_multiply

from py_stoch_amd import models
from py_stoch_amd.mod.parser.tl_parser import tables_of


def _run(oracle_mod, spn, E, t_max, ita, U, method, seed=0, **kw):
    a = models.make_args(spn, E=E, t_max=t_max, ita=ita, U=U)
    O, flags, st = oracle_mod.simulate(tables_of(a), a.rate_fn(a.c), a.x_0, method=method, t_max=t_max,
                                       I=a.I, E=a.E, n_sims=U, seed=seed, **kw)
    return a, O, flags, st


def test_birth_death_exact_poisson_law(oracle_mod):
    a, O, flags, st = _run(oracle_mod, models.birth_death(), [0], 100.0, 101, 4000, oracle_mod.METHOD_PURE_SSA)
    assert flags.sum() == 0
    for f in (5, 10, 20, 50, 100):
        t = float(a.I[f])
        mean = 100.0 * (1.0 - np.exp(-0.1 * t))
        x = O[0, f]
        assert abs(x.mean() - mean) < 5 * np.sqrt(mean / x.size)
        # KS distance to the exact CDF, compared at the integers
        ks = np.abs(np.searchsorted(np.sort(x), np.arange(0, 200), side="right") / x.size
                    - stats.poisson.cdf(np.arange(0, 200), mean)).max()
        assert ks < 1.63 / np.sqrt(x.size) + 0.01
    assert (O[0, 0] == 0).all()


def test_planted_rate_error_is_caught(oracle_mod):
    """Doubling the death rate (a plausible table mistake) must fail the law test."""
    a, O, _, _ = _run(oracle_mod, models.birth_death(death=0.2), [0], 100.0, 101, 2000, oracle_mod.METHOD_PURE_SSA)
    mean = 100.0 * (1.0 - np.exp(-0.1 * 50.0))
    assert abs(O[0, 50].mean() - mean) > 5 * np.sqrt(mean / 2000)


def test_chain_binomial_laws(oracle_mod):
    from py_stoch_amd.models import _spn
    n = 200
    spn = _spn(3, [({0: 1}, {1: 1}, "parameter_0*species_0"), ({1: 1}, {2: 1}, "parameter_1*species_1")],
               [n, 0, 0], [1.0, 1.0])
    a, O, flags, st = _run(oracle_mod, spn, [0, 1], 3.0, 7, 3000, oracle_mod.METHOD_PURE_SSA, seed=4)
    for f in (1, 2, 4, 6):
        t = float(a.I[f])
        p1, p2 = np.exp(-t), t * np.exp(-t)
        for q, p in ((0, p1), (1, p2)):
            x = O[q, f]
            assert abs(x.mean() - n * p) < 5 * np.sqrt(n * p * (1 - p) / x.size)
            assert abs(x.var() - n * p * (1 - p)) < 0.12 * n * p * (1 - p)


def _schlogl_cme(t_end, xmax=1400):
    c1, c2, c3, c4, A, B = 3e-7, 1e-4, 1e-3, 3.5, 1e5, 2e5
    x = np.arange(xmax + 1, dtype=float)
    birth = c1 / 2 * A * x * (x - 1) + c3 * B
    death = c2 / 6 * x * (x - 1) * (x - 2) + c4 * x
    birth[-1] = 0.0
    Q = sp.diags([-(birth + death), birth[:-1], death[1:]], [0, -1, 1], format="csc")
    p0 = np.zeros(xmax + 1)
    p0[250] = 1.0
    return expm_multiply(Q * t_end, p0)


def test_schlogl_ssa_matches_master_equation(oracle_mod):
    p = _schlogl_cme(5.0)
    a, O, flags, st = _run(oracle_mod, models.schlogl(), [0], 5.0, 11, 1500, oracle_mod.METHOD_PURE_SSA, seed=8)
    x = O[0, -1]
    cdf = np.cumsum(p)
    emp = np.searchsorted(np.sort(x), np.arange(p.size), side="right") / x.size
    assert np.abs(emp - cdf).max() < 1.63 / np.sqrt(x.size) + 0.01
    mean = (np.arange(p.size) * p).sum()
    sd = np.sqrt((np.arange(p.size) ** 2 * p).sum() - mean ** 2)
    assert abs(x.mean() - mean) < 5 * sd / np.sqrt(x.size)


def test_forced_ssa_equals_pure_ssa(oracle_mod):
    # t_max = 0.5 is exact in float32, so the last sample is reached in both drivers
    spn = models.lotka_volterra(x0=300, y0=200)
    _, O1, _, s1 = _run(oracle_mod, spn, [0, 1], 0.5, 11, 64, oracle_mod.METHOD_PURE_SSA, seed=3)
    _, O2, _, s2 = _run(oracle_mod, spn, [0, 1], 0.5, 11, 64, oracle_mod.METHOD_FORCED_SSA, seed=3)
    assert np.array_equal(O1, O2)
    assert s1["events"] == s2["events"]


def test_tau_leaping_agrees_with_ssa_lotka_volterra(oracle_mod):
    """Means agree at eta = 0.03; the standard deviation carries the known
    tau-leaping bias of the oscillating network at eta = 0.03 (about 20% at
    t = 1) and converges to SSA's as eta shrinks (within 5% at eta = 0.01)."""
    spn = models.lotka_volterra()
    U = 3000
    _, Os, _, ss = _run(oracle_mod, spn, [0, 1], 1.0, 11, U, oracle_mod.METHOD_PURE_SSA, seed=6)
    _, Ot, _, st = _run(oracle_mod, spn, [0, 1], 1.0, 11, U, oracle_mod.METHOD_TAU_LEAPING, seed=5)
    _, Ot2, _, st2 = _run(oracle_mod, spn, [0, 1], 1.0, 11, U, oracle_mod.METHOD_TAU_LEAPING, seed=5, eta=0.01)
    assert st["leaps"] > 0 and st2["leaps"] > st["leaps"]
    for q in (0, 1):
        for f in (2, 5, 10):
            ms, sds = Os[q, f].mean(), Os[q, f].std()
            for O in (Ot, Ot2):
                assert abs(O[q, f].mean() - ms) < 5 * np.sqrt(2 * sds ** 2 / U) + 0.005 * ms
            assert abs(Ot[q, f].std() - sds) < 0.35 * sds
            assert abs(Ot2[q, f].std() - sds) < 0.05 * sds + 4 * sds / np.sqrt(U)
            assert abs(Ot2[q, f].std() - sds) <= abs(Ot[q, f].std() - sds) + 2 * sds / np.sqrt(U)


def test_tau_leaping_agrees_with_ssa_schlogl(oracle_mod):
    # bimodal: KS between the two methods' X(5) (epsilon = 0.03 leap bias is small)
    _, Ot, _, st = _run(oracle_mod, models.schlogl(), [0], 5.0, 6, 1500, oracle_mod.METHOD_TAU_LEAPING, seed=9)
    _, Os, _, _ = _run(oracle_mod, models.schlogl(), [0], 5.0, 6, 1500, oracle_mod.METHOD_PURE_SSA, seed=10)
    assert st["leaps"] > 0 and st["ssa_steps"] > 0
    assert stats.ks_2samp(Ot[0, -1], Os[0, -1]).statistic < 0.07


def test_absorbing_state_fills_remaining_samples(oracle_mod):
    spn = models.birth_death(birth=0.0, death=1.0, x0=20)
    a, O, flags, st = _run(oracle_mod, spn, [0], 30.0, 31, 200, oracle_mod.METHOD_TAU_LEAPING, seed=1)
    assert flags.sum() == 0
    assert (O[0, -1] == 0).all()
    # once absorbed at 0 a simulation stays at 0
    for s in range(200):
        z = np.nonzero(O[0, :, s] == 0)[0]
        assert (O[0, z[0]:, s] == 0).all()


def test_tau_halving_on_negative_states(oracle_mod):
    # no critical reactions (n_c = 0) and a loose bound (eta = 1): Poisson(lambda ~ x)
    # leaps overshoot below zero and must be halved until valid
    spn = models.birth_death(birth=0.0, death=1.0, x0=50)
    a, O, flags, st = _run(oracle_mod, spn, [0], 10.0, 11, 500, oracle_mod.METHOD_TAU_LEAPING, seed=2,
                           n_c=0, eta=1.0)
    assert st["rejected_leaps"] > 0
    assert O.min() >= 0
    assert flags.sum() == 0


def test_sim_offset_consistency(oracle_mod):
    spn = models.lotka_volterra()
    _, Ofull, _, _ = _run(oracle_mod, spn, [0, 1], 0.5, 6, 128, oracle_mod.METHOD_TAU_LEAPING, seed=4)
    _, Otail, _, _ = _run(oracle_mod, spn, [0, 1], 0.5, 6, 64, oracle_mod.METHOD_TAU_LEAPING, seed=4, sim_offset=64)
    assert np.array_equal(Ofull[:, :, 64:], Otail)


def test_thread_count_does_not_change_results(oracle_mod):
    spn = models.schlogl()
    _, O1, _, s1 = _run(oracle_mod, spn, [0], 2.0, 5, 256, oracle_mod.METHOD_TAU_LEAPING, seed=4, threads=1)
    _, O8, _, s8 = _run(oracle_mod, spn, [0], 2.0, 5, 256, oracle_mod.METHOD_TAU_LEAPING, seed=4, threads=8)
    assert np.array_equal(O1, O8) and s1 == s8


def test_step_budget_ends_simulations(oracle_mod):
    """max_steps: every simulation stops after that many leap decisions + SSA
    steps with the STEPS flag; up to the stop its trajectory is the unbudgeted
    one (the budget changes nothing before it is reached)."""
    a = models.make_args(models.lotka_volterra(), E=[0, 1], t_max=10.0, ita=101, U=64)
    kw = dict(t_max=10.0, I=a.I, E=a.E, n_sims=64, seed=1)
    O, F, st = oracle_mod.simulate(tables_of(a), a.rate_fn(a.c), a.x_0, max_steps=300, **kw)
    O2, F2, st2 = oracle_mod.simulate(tables_of(a), a.rate_fn(a.c), a.x_0, **kw)
    assert (F == oracle_mod.FLAG_STEPS).all() and (F2 == 0).all()
    assert st["leaps"] + st["ssa_steps"] <= 64 * 300
    # up to the sample where the budget ran out, each trajectory is the unbudgeted
    # one; from there on the samples are zero
    for s in range(64):
        nz = np.nonzero(O[:, :, s].any(axis=0))[0]
        f = nz[-1] + 1
        assert 1 <= f < O.shape[1] and (O[:, f:, s] == 0).all()
        assert np.array_equal(O[:, :f, s], O2[:, :f, s])


def test_flagged_events_are_the_events_of_flagged_simulations(oracle_mod):
    """stats.flagged_events is the sum of the event counts of exactly the
    simulations that end flagged: rerun each simulation alone (same stream,
    through sim_offset) and add up the flagged ones."""
    a = models.make_args(models.lotka_volterra(), E=[0, 1], t_max=2.0, ita=21, U=48)
    tabs, rates = tables_of(a), a.rate_fn(a.c)
    kw = dict(t_max=2.0, I=a.I, E=a.E, seed=9, max_steps=300)
    _, F, st = oracle_mod.simulate(tabs, rates, a.x_0, n_sims=48, **kw)
    assert 0 < st["flagged"] < 48  # some simulations run out of steps, some do not
    total = flagged = 0
    for s in range(48):
        _, f1, s1 = oracle_mod.simulate(tabs, rates, a.x_0, n_sims=1, sim_offset=s, **kw)
        assert f1[0] == F[s]
        total += s1["events"]
        flagged += s1["events"] if f1[0] != 0 else 0
        assert s1["flagged_events"] == (s1["events"] if f1[0] != 0 else 0)
    assert total == st["events"] and flagged == st["flagged_events"] > 0


def test_infinite_tau_flag_matches_extinction_law(oracle_mod):
    """Immigration of X beside the decay of Y from Y0 = 5 < n_c (every decay
    critical, so tau_1 = inf throughout): a simulation is flagged
    FLAG_TAU_INF exactly when Y is extinct before t_end -- no reactant of a
    non-critical reaction and no critical propensity left -- which for a pure
    death process has probability (1 - e^(-t))^5; before the flag its
    samples hold the state (Y only decreases), after it zeros.  (An
    unflagged simulation can still show Y = 0 at t_end: the last sample is
    interpolated inside the step that crosses it, P1_P2.py:542-584.)"""
    spn = models._spn(2, [({}, {0: 1}, "parameter_0"), ({1: 1}, {}, "parameter_1*species_1")], [0, 5], [10.0, 1.0])
    U, t_end = 20000, 2.0
    a = models.make_args(spn, E=[0, 1], t_max=t_end, ita=21, U=U)
    O, flags, st = oracle_mod.simulate(tables_of(a), a.rate_fn(a.c), a.x_0, method=oracle_mod.METHOD_TAU_LEAPING,
                                       t_max=t_end, I=a.I, E=a.E, n_sims=U, seed=5)
    assert set(np.unique(flags)) <= {0, oracle_mod.FLAG_TAU_INF}
    p = (1.0 - np.exp(-t_end)) ** 5
    f = (flags != 0).mean()
    assert abs(f - p) < 5 * np.sqrt(p * (1 - p) / U), (f, p)
    Y = O[1]
    ok = flags == 0
    assert (np.diff(Y[:, ok], axis=0) <= 0).all() and (Y[0] == 5).all()
    # a flagged simulation: its samples are the state up to the flag, zeros after
    for s in np.nonzero(~ok)[0][:200]:
        nz = np.nonzero(O[:, :, s].any(axis=0))[0]
        last = nz.max() + 1 if nz.size else 0
        assert not O[:, last:, s].any()
        assert (np.diff(Y[:last, s]) <= 0).all()
