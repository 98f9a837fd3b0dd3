"""FP64 GPU path vs the CPU oracle: bit-exact int32 trajectories, flags and
counters on the same Philox stream (every test marked gpu)."""
import numpy as np
import pytest

from py_stoch_amd import models
from py_stoch_amd.mod.parser.tl_parser import tables_of

pytestmark = pytest.mark.gpu


def _compare(native, oracle_mod, spn, E, t_max, ita, U, method, seed=7, precision=0, **extra):
    a = models.make_args(spn, E=E, t_max=t_max, ita=ita, U=U)
    tabs = tables_of(a)
    rates = a.rate_fn(a.c)
    kw = dict(method=method, t_max=t_max, I=a.I, E=a.E, n_sims=U, seed=seed, **extra)
    print("case: U=%d ita=%d method=%d seed=%d" % (U, ita, method, seed), flush=True)
    O_g, f_g, s_g = native.simulate(tabs, rates, a.x_0, precision=precision, **kw)
    O_c, f_c, s_c = oracle_mod.simulate(tabs, rates, a.x_0, **kw)
    assert O_g.shape == O_c.shape
    mism = np.argwhere(O_g != O_c)
    assert mism.shape[0] == 0, "first mismatch %s gpu=%d cpu=%d" % (
        mism[0].tolist(), O_g[tuple(mism[0])], O_c[tuple(mism[0])])
    assert np.array_equal(f_g, f_c)
    for k in ("events", "leaps", "rejected_leaps", "ssa_steps", "flagged", "flagged_events"):
        assert s_g[k] == s_c[k], (k, s_g[k], s_c[k])
    return O_g, s_g


def test_birth_death_pure_ssa(native, oracle_mod):
    _compare(native, oracle_mod, models.birth_death(), [0], 100.0, 101, 1024, native.METHOD_PURE_SSA, seed=0)


def test_birth_death_forced_ssa(native, oracle_mod):
    _compare(native, oracle_mod, models.birth_death(), [0], 100.0, 101, 1024, native.METHOD_FORCED_SSA, seed=0)


def test_birth_death_tau_leaping(native, oracle_mod):
    _compare(native, oracle_mod, models.birth_death(), [0], 100.0, 101, 1000, native.METHOD_TAU_LEAPING)


def test_lotka_volterra_tau_leaping(native, oracle_mod):
    _, st = _compare(native, oracle_mod, models.lotka_volterra(), [0, 1], 10.0, 1000, 2048,
                     native.METHOD_TAU_LEAPING, seed=2)
    assert st["leaps"] > 0


def test_lotka_volterra_full_size(native, oracle_mod):
    """The throughput workload at its full size (65,536 simulations, t_end =
    10, 1000 samples, seed 2) is bit-exact too, including the explosive
    simulations stopped at a count of 2^30 and their flags."""
    _, st = _compare(native, oracle_mod, models.lotka_volterra(), [0, 1], 10.0, 1000, 65536,
                     native.METHOD_TAU_LEAPING, seed=2)
    assert st["leaps"] > 0 and st["ssa_steps"] > 0 and st["flagged"] > 0


def test_schlogl_tau_leaping(native, oracle_mod):
    _, st = _compare(native, oracle_mod, models.schlogl(), [0], 10.0, 100, 1024, native.METHOD_TAU_LEAPING, seed=3)
    assert st["leaps"] > 0 and st["ssa_steps"] > 0


def test_schlogl_pure_ssa(native, oracle_mod):
    _compare(native, oracle_mod, models.schlogl(), [0], 2.0, 21, 512, native.METHOD_PURE_SSA, seed=3)


def test_ragged_and_tiny_sim_counts(native, oracle_mod):
    for U in (1, 2, 3, 63, 65, 257):
        _compare(native, oracle_mod, models.lotka_volterra(), [1, 0], 2.0, 17, U, native.METHOD_TAU_LEAPING, seed=U)


def test_absorbing_state(native, oracle_mod):
    # pure death from 30: every simulation reaches the absorbing a0 == 0 state
    spn = models.birth_death(birth=0.0, death=1.0, x0=30)
    O, _ = _compare(native, oracle_mod, spn, [0], 50.0, 51, 300, native.METHOD_TAU_LEAPING, seed=5)
    assert (O[0, -1] == 0).all()


def test_sim_offset_reproducible(native, oracle_mod):
    a = models.make_args(models.lotka_volterra(), E=[0, 1], t_max=1.0, ita=11, U=128)
    tabs, rates = tables_of(a), a.rate_fn(a.c)
    kw = dict(method=0, t_max=1.0, I=a.I, E=a.E, seed=9)
    full, _, _ = native.simulate(tabs, rates, a.x_0, n_sims=128, **kw)
    tail, _, _ = native.simulate(tabs, rates, a.x_0, n_sims=64, sim_offset=64, **kw)
    assert np.array_equal(full[:, :, 64:], tail)


def test_batch_equals_separate_runs(native, oracle_mod):
    """Jobs of different networks and parameter rows in one persistent launch
    give, job by job, the same bits as the CPU oracle run of that job."""
    rng = np.random.default_rng(7)
    specs = []
    a_lv = models.make_args(models.lotka_volterra(), E=[0, 1], t_max=2.0, ita=21, U=64)
    rates_lv = np.array([a_lv.rate_fn(a_lv.c) * 10.0 ** rng.uniform(-0.3, 0.3, 3) for _ in range(6)])
    specs.append((a_lv, rates_lv, np.tile(a_lv.x_0, (6, 1)), 6 * 100, 100, 0))
    a_sc = models.make_args(models.schlogl(), E=[0], t_max=2.0, ita=11, U=64)
    rates_sc = np.array([a_sc.rate_fn(a_sc.c) * 10.0 ** rng.uniform(-0.2, 0.2, 4) for _ in range(4)])
    specs.append((a_sc, rates_sc, np.tile(a_sc.x_0, (4, 1)), 4 * 90, 90, 0))
    a_bd = models.make_args(models.birth_death(), E=[0], t_max=20.0, ita=21, U=64)
    specs.append((a_bd, a_bd.rate_fn(a_bd.c)[None, :], a_bd.x_0[None, :], 333, 333, 2))
    jobs, refs = [], []
    for a, rates, x0, n, per, method in specs:
        kw = dict(method=method, t_max=float(a.t_max), I=a.I, E=a.E, n_sims=n, sims_per_set=per, seed=13)
        jobs.append(native.Job(tables_of(a), rates, x0, **kw))
        refs.append(oracle_mod.simulate(tables_of(a), rates, x0, **kw))
    Os, Fs, st = native.batch_simulate(jobs)
    ev = 0
    for q, (O, F, (Oc, Fc, sc)) in enumerate(zip(Os, Fs, refs)):
        print("job", q, O.shape, flush=True)
        assert np.array_equal(O, Oc), "job %d differs" % q
        assert np.array_equal(F, Fc)
        ev += sc["events"]
    assert st["events"] == ev


def test_persistent_refill_many_sims_per_lane(native, oracle_mod):
    """More simulations than resident lanes: lanes refill from the work queue."""
    a = models.make_args(models.birth_death(), E=[0], t_max=5.0, ita=6, U=64)
    n = 300000
    kw = dict(method=2, t_max=5.0, I=a.I, E=a.E, n_sims=n, seed=21)
    O, F, st = native.simulate(tables_of(a), a.rate_fn(a.c), a.x_0, **kw)
    Oc, Fc, sc = oracle_mod.simulate(tables_of(a), a.rate_fn(a.c), a.x_0, **kw)
    assert np.array_equal(O, Oc)
    assert st["events"] == sc["events"]


@pytest.mark.parametrize("specialize", [False, True])
def test_step_budget_bit_exact(native, oracle_mod, specialize):
    """The per-simulation step budget ends simulations at the same step, with
    the same samples and flags, as the oracle (Lotka-Volterra: the budget hits
    during leaps; Schlogl: also during SSA bursts)."""
    for spn, E, t_max, ita, budget in ((models.lotka_volterra(), [0, 1], 10.0, 101, 300),
                                       (models.schlogl(), [0], 10.0, 51, 180)):
        a = models.make_args(spn, E=E, t_max=t_max, ita=ita, U=512)
        kw = dict(method=0, t_max=t_max, I=a.I, E=a.E, n_sims=512, seed=3, max_steps=budget)
        print("case: budget %d specialize=%s" % (budget, specialize), flush=True)
        O_g, f_g, s_g = native.simulate(tables_of(a), a.rate_fn(a.c), a.x_0, specialize=specialize, **kw)
        O_c, f_c, s_c = oracle_mod.simulate(tables_of(a), a.rate_fn(a.c), a.x_0, **kw)
        assert np.array_equal(f_g, f_c) and (f_c & native.FLAG_STEPS).any()
        assert np.array_equal(O_g, O_c)
        for k in ("events", "leaps", "ssa_steps", "flagged", "flagged_events"):
            assert s_g[k] == s_c[k], k


@pytest.mark.parametrize("specialize", [False, True])
def test_infinite_tau_is_flagged(native, oracle_mod, specialize):
    """Immigration of X beside a decay of Y that starts critical (Y0 = 5 <
    n_c): once Y is extinct no species is a reactant of a non-critical
    reaction (tau_1 = inf, P1_P2.py:390-394) and no critical propensity is
    left (tau_2 = inf), where the reference would draw curand_poisson(inf).
    The simulation stops with PST_FLAG_TAU_INF and zeros after its last
    sample, as in the oracle: FP64 bit-exact, mixed in proportion."""
    spn = models._spn(2, [({}, {0: 1}, "parameter_0"), ({1: 1}, {}, "parameter_1*species_1")], [0, 5], [10.0, 1.0])
    U, ita = 2048, 51
    a = models.make_args(spn, E=[0, 1], t_max=5.0, ita=ita, U=U)
    kw = dict(method=0, t_max=5.0, I=a.I, E=a.E, n_sims=U, seed=7)
    print("case: infinite tau, specialize=%s" % specialize, flush=True)
    O_g, f_g, s_g = native.simulate(tables_of(a), a.rate_fn(a.c), a.x_0, specialize=specialize, **kw)
    O_c, f_c, s_c = oracle_mod.simulate(tables_of(a), a.rate_fn(a.c), a.x_0, **kw)
    assert set(np.unique(f_c)) == {0, native.FLAG_TAU_INF}
    assert np.array_equal(f_g, f_c) and np.array_equal(O_g, O_c)
    for k in ("events", "leaps", "flagged", "flagged_events"):
        assert s_g[k] == s_c[k], k
    flagged = f_c != 0
    assert 0.5 < flagged.mean() < 1.0 and not O_c[:, -1, flagged].any()
    O_m, f_m, _ = native.simulate(tables_of(a), a.rate_fn(a.c), a.x_0, specialize=specialize, precision=1, **kw)
    assert set(np.unique(f_m)) <= {0, native.FLAG_TAU_INF}
    p, q = (f_m != 0).mean(), flagged.mean()
    assert abs(p - q) < 5 * np.sqrt(q * (1 - q) / U) + 0.01, (p, q)
    assert not O_m[:, -1, f_m != 0].any()


@pytest.mark.parametrize("specialize", [False, True])
def test_irregular_sample_times(native, oracle_mod, specialize):
    """Sample times that are not a uniform grid: clusters closer than one leap
    (a step crosses several at once, each interpolated on its own fraction,
    P1_P2.py:542-584), repeated times (the same state twice) and long gaps;
    FP64 bit-exact with the oracle."""
    rng = np.random.default_rng(21)
    I = np.sort(np.concatenate([[0.0], rng.uniform(0.0, 10.0, 40), np.full(3, 2.5), 5.0 + 1e-4 * np.arange(6)]))
    I = I.astype(np.float32)
    U = 1024
    a = models.make_args(models.lotka_volterra(), E=[0, 1], t_max=10.0, ita=I.size, U=U)
    kw = dict(method=0, t_max=10.0, I=I, E=a.E, n_sims=U, seed=4)
    print("case: irregular sample times, specialize=%s" % specialize, flush=True)
    O_g, f_g, s_g = native.simulate(tables_of(a), a.rate_fn(a.c), a.x_0, specialize=specialize, **kw)
    O_c, f_c, s_c = oracle_mod.simulate(tables_of(a), a.rate_fn(a.c), a.x_0, **kw)
    assert np.array_equal(f_g, f_c) and np.array_equal(O_g, O_c)
    assert s_g["events"] == s_c["events"] and s_g["leaps"] == s_c["leaps"]
    same = np.nonzero(np.diff(I) == 0)[0]
    assert same.size >= 2 and np.array_equal(O_c[:, same], O_c[:, same + 1])
