"""Network-specific kernels compiled at run time (specialize=True): the FP64
kernel gives the same bits as the CPU oracle on the same Philox stream, like
the table-driven kernel, for every method; jobs of several networks in one
plan are grouped per network (every test marked gpu)."""
import numpy as np
import pytest

from py_stoch_amd import models
from py_stoch_amd.mod.parser.tl_parser import tables_of

from devbuf import DeviceArray as _DeviceArray

pytestmark = pytest.mark.gpu


def _compare(native, oracle_mod, spn, E, t_max, ita, U, method, seed=7):
    a = models.make_args(spn, E=E, t_max=t_max, ita=ita, U=U)
    tabs = tables_of(a)
    rates = a.rate_fn(a.c)
    kw = dict(method=method, t_max=t_max, I=a.I, E=a.E, n_sims=U, seed=seed)
    print("case: specialised U=%d ita=%d method=%d seed=%d" % (U, ita, method, seed), flush=True)
    O_g, f_g, s_g = native.simulate(tabs, rates, a.x_0, precision=native.PRECISION_FP64, specialize=True, **kw)
    O_c, f_c, s_c = oracle_mod.simulate(tabs, rates, a.x_0, **kw)
    mism = np.argwhere(O_g != O_c)
    assert mism.shape[0] == 0, "first mismatch %s gpu=%d cpu=%d" % (
        mism[0].tolist(), O_g[tuple(mism[0])], O_c[tuple(mism[0])])
    assert np.array_equal(f_g, f_c)
    for k in ("events", "leaps", "rejected_leaps", "ssa_steps", "flagged", "flagged_events"):
        assert s_g[k] == s_c[k], (k, s_g[k], s_c[k])
    return O_g, s_g


def test_lotka_volterra_tau_leaping(native, oracle_mod):
    _, st = _compare(native, oracle_mod, models.lotka_volterra(), [0, 1], 10.0, 1000, 2048,
                     native.METHOD_TAU_LEAPING, seed=2)
    assert st["leaps"] > 0 and st["ssa_steps"] > 0


def test_schlogl_tau_leaping(native, oracle_mod):
    _compare(native, oracle_mod, models.schlogl(), [0], 10.0, 100, 1024, native.METHOD_TAU_LEAPING, seed=3)


def test_schlogl_launch_beyond_resident_wavefronts(native, oracle_mod):
    """262,144 simulations need 4096 wavefronts, more than the plain Schloegl
    network kernel holds resident, so the launch takes the build compiled for
    three wavefronts per SIMD: still bit-exact."""
    _, st = _compare(native, oracle_mod, models.schlogl(), [0], 10.0, 11, 262144, native.METHOD_TAU_LEAPING, seed=3)
    assert st["leaps"] > 0 and st["ssa_steps"] > 0


@pytest.mark.parametrize("method", [0, 2])
def test_occupancy_build_equals_plain_build_mixed(native, method):
    """Mixed precision: the last 1024 of 262,144 simulations (occupancy build)
    equal the same simulations run on their own (sim_offset; 1024 simulations
    fit the plain build) bit for bit."""
    U, tail = 262144, 1024
    a = models.make_args(models.schlogl(), E=[0], t_max=10.0, ita=11, U=U)
    tabs, rates = tables_of(a), a.rate_fn(a.c)
    kw = dict(method=method, precision=native.PRECISION_MIXED, specialize=True, t_max=10.0, I=a.I, E=a.E, seed=9)
    print("case: mixed schlogl method=%d, %d sims vs the last %d alone" % (method, U, tail), flush=True)
    O, f, _ = native.simulate(tabs, rates, a.x_0, n_sims=U, **kw)
    O_t, f_t, _ = native.simulate(tabs, rates, a.x_0, n_sims=tail, sim_offset=U - tail, **kw)
    assert np.array_equal(O[:, :, U - tail:], O_t)
    assert np.array_equal(f[U - tail:], f_t)


def test_birth_death_pure_and_forced_ssa(native, oracle_mod):
    _compare(native, oracle_mod, models.birth_death(), [0], 100.0, 101, 1024, native.METHOD_PURE_SSA, seed=0)
    _compare(native, oracle_mod, models.birth_death(), [0], 100.0, 101, 1024, native.METHOD_FORCED_SSA, seed=0)


def test_recorded_species_order_and_ragged_counts(native, oracle_mod):
    for U in (1, 65, 257):
        _compare(native, oracle_mod, models.lotka_volterra(), [1, 0], 2.0, 17, U, native.METHOD_TAU_LEAPING, seed=U)


def test_absorbing_state(native, oracle_mod):
    spn = models.birth_death(birth=0.0, death=1.0, x0=30)
    O, _ = _compare(native, oracle_mod, spn, [0], 50.0, 51, 300, native.METHOD_TAU_LEAPING, seed=5)
    assert (O[0, -1] == 0).all()


def test_batch_groups_networks(native, oracle_mod):
    """Two LV parameter sweeps and one Schlogl sweep in one plan: two compiled
    kernels, every job bit-identical to its oracle run."""
    rng = np.random.default_rng(11)
    jobs, refs = [], []
    a_lv = models.make_args(models.lotka_volterra(), E=[0, 1], t_max=2.0, ita=21, U=64)
    a_sc = models.make_args(models.schlogl(), E=[0], t_max=2.0, ita=11, U=64)
    for a, sets, per, M in ((a_lv, 5, 70, 3), (a_sc, 3, 50, 4), (a_lv, 2, 40, 3)):
        rates = np.array([a.rate_fn(a.c) * 10.0 ** rng.uniform(-0.3, 0.3, M) for _ in range(sets)])
        x0 = np.tile(a.x_0, (sets, 1))
        kw = dict(method=0, t_max=2.0, I=a.I, E=a.E, n_sims=sets * per, sims_per_set=per, seed=17)
        jobs.append(native.Job(tables_of(a), rates, x0, specialize=True, **kw))
        refs.append(oracle_mod.simulate(tables_of(a), rates, x0, **kw))
    Os, Fs, st = native.batch_simulate(jobs)
    for q, (O, F, (Oc, Fc, _)) in enumerate(zip(Os, Fs, refs)):
        print("job", q, O.shape, flush=True)
        assert np.array_equal(O, Oc), "job %d differs" % q
        assert np.array_equal(F, Fc)
    assert st["events"] == sum(r[2]["events"] for r in refs)


def test_mixed_specialised_matches_table_kernel(native):
    """The mixed-precision kernel with the network compiled in samples the
    same law as the table-driven one (constant folding may change FP32
    contractions, so trajectories are compared in distribution): two-sample
    KS distance of both species at every fifth sample below the 0.01% critical
    value for 4096 + 4096 draws."""
    a = models.make_args(models.lotka_volterra(), E=[0, 1], t_max=5.0, ita=51, U=4096)
    kw = dict(method=0, t_max=5.0, I=a.I, E=a.E, n_sims=4096, precision=native.PRECISION_MIXED)
    O_t, _, s_t = native.simulate(tables_of(a), a.rate_fn(a.c), a.x_0, seed=4, **kw)
    O_s, _, s_s = native.simulate(tables_of(a), a.rate_fn(a.c), a.x_0, seed=5, specialize=True, **kw)
    crit = 2.23 * np.sqrt(2.0 / 4096)
    for k in range(2):
        for f in range(5, 51, 5):
            x, y = np.sort(O_t[k, f]), np.sort(O_s[k, f])
            grid = np.concatenate([x, y])
            d = np.abs(np.searchsorted(x, grid, side="right") - np.searchsorted(y, grid, side="right")).max() / 4096
            assert d < crit, (k, f, d)
    assert abs(s_t["events"] / s_s["events"] - 1.0) < 0.02


def test_simulator_run_batch_equals_single_runs(native):
    """CuTauLeaping().run_batch over a parameter sweep (what sim_maker
    produces) gives, set by set, the arrays of separate run() calls."""
    from py_stoch_amd.mod.simulator.cuTauLeaping.cuTauLeaping import CuTauLeaping
    args = []
    for c2 in (0.005, 0.01, 0.02):
        spn = models.lotka_volterra(c2=c2)
        a = models.make_args(spn, E=[0, 1], t_max=2.0, ita=21, U=300, seed=8)
        args.append(a)
    sim = CuTauLeaping()
    batch = sim.run_batch(args)
    assert all(a.U == 320 for a in args)  # rounded up as run() does (block 64 below 640)
    for a, O in zip(args, batch):
        single = CuTauLeaping().run(a)
        assert np.array_equal(O, single)


def _mm_stationary_cdf(k, vmax, km, n_max=400):
    """Birth-death chain 0 -> X (k), X -> 0 (vmax X / (km + X)): detailed
    balance gives pi(n) / pi(n-1) = k / (vmax n / (km + n))."""
    logp = np.zeros(n_max + 1)
    for n in range(1, n_max + 1):
        logp[n] = logp[n - 1] + np.log(k) - np.log(vmax * n / (km + n))
    p = np.exp(logp - logp.max())
    return np.cumsum(p / p.sum())


@pytest.mark.parametrize("precision", [0, 1])
def test_michaelis_menten_stationary_law(native, precision):
    """Pure SSA with a compiled non-mass-action kinetic law reaches the exact
    stationary law: one-sample KS distance at t = 100 below 0.025 (the 0.01%
    critical value is 0.0246 for 8192 draws of a continuous law; discreteness
    only lowers the statistic).  The chain relaxes at rate d/dn[vmax n /
    (km + n)] = 0.1 at its mean n = 40, where it starts, so e^-10 of the
    start remains."""
    from py_stoch_amd.mod.parser.g_parser import GArgs, GParser
    a = GArgs()
    a.E, a.t_max, a.ita, a.U = [0], 100.0, 11, 8192
    a = GParser.parse(a, models.michaelis_menten_birth_death(k=20.0, vmax=25.0, km=10.0, x0=40))
    print("case: michaelis-menten precision=%d" % precision, flush=True)
    O, f, st = native.simulate(tables_of(a), a.rate_fn(a.c), a.x_0, method=native.METHOD_PURE_SSA,
                               precision=precision, t_max=100.0, I=a.I, E=a.E, n_sims=8192, seed=12,
                               specialize=True)
    assert (f == 0).all()
    x = O[0, -1]
    cdf = _mm_stationary_cdf(20.0, 25.0, 10.0)
    emp = np.searchsorted(np.sort(x), np.arange(cdf.shape[0]), side="right") / x.shape[0]
    assert np.abs(emp - cdf).max() < 0.025


@pytest.mark.parametrize("precision", [0, 1])
def test_immigration_birth_death_mean(native, precision):
    """Tau-leaping with a compiled polynomial hazard nu + b X: the mean follows
    m(t) = nu / (d - b) (1 - e^{-(d - b) t}) (linear propensities), within
    4 standard errors plus 2% for the leap bias."""
    a = models.make_args(models.immigration_birth_death(nu=50.0, b=0.5, d=1.0), E=[0], t_max=10.0, ita=11, U=8192)
    print("case: immigration-birth-death precision=%d" % precision, flush=True)
    O, f, st = native.simulate(tables_of(a), a.rate_fn(a.c), a.x_0, method=native.METHOD_TAU_LEAPING,
                               precision=precision, t_max=10.0, I=a.I, E=a.E, n_sims=8192, seed=13,
                               specialize=True)
    assert (f == 0).all() and st["leaps"] > 0
    t = np.asarray(a.I, np.float64)
    m = 50.0 / 0.5 * (1.0 - np.exp(-0.5 * t))
    mean = O[0].mean(axis=-1)
    se = np.sqrt(np.maximum(O[0].var(axis=-1), 1.0) / O.shape[-1])
    assert np.all(np.abs(mean - m) <= 4 * se + 0.02 * m)


@pytest.mark.parametrize("precision", [0, 1])
def test_autotuned_plan_gives_the_same_results(native, oracle_mod, precision):
    """The autotuner changes only the launch shape: after it, the FP64 plan is
    still bit-exact against the oracle and the mixed plan reproduces an
    untuned plan's trajectories."""
    a = models.make_args(models.lotka_volterra(), E=[0, 1], t_max=4.0, ita=41, U=8192)
    kw = dict(method=0, t_max=4.0, I=a.I, E=a.E, n_sims=8192, seed=12, precision=precision, specialize=True)
    tabs, rates = tables_of(a), a.rate_fn(a.c)
    print("case: autotune precision=%d" % precision, flush=True)
    O1, O2 = _DeviceArray((2, 41, 8192)), _DeviceArray((2, 41, 8192))
    try:
        p1 = native.Plan(tabs, rates, a.x_0, **kw)
        p2 = native.Plan(tabs, rates, a.x_0, **kw)
        ch = p1.autotune(O1.ptr.value, 8192, fraction=0.5)
        assert ch["lanes_shift"] in (0, 1) and ch["ssa_slice"] in (3, 7, 15) and ch["waves_trim"] in (0, 1)
        assert 0.0 < ch["kernel_ms"] <= ch["default_ms"]
        p1.execute(O1.ptr.value, 8192)
        p2.execute(O2.ptr.value, 8192)
        s1, s2 = p1.stats(), p2.stats()
        g1, g2 = O1.get(), O2.get()
    finally:
        O1.free()
        O2.free()
    assert np.array_equal(g1, g2)
    assert all(s1[k] == s2[k] for k in ("events", "leaps", "ssa_steps", "flagged", "flagged_events"))
    if precision == 0:
        Oc, _, sc = oracle_mod.simulate(tabs, rates, a.x_0, method=0, t_max=4.0, I=a.I, E=a.E, n_sims=8192, seed=12)
        assert np.array_equal(g1, Oc) and s1["events"] == sc["events"]


def test_waves_trim_gives_the_same_results(native):
    """Holding one resident wavefront per SIMD back (262,144 pure-SSA
    simulations: four resident wavefronts per SIMD become three, and the work
    queue refills lanes) changes only the launch: trajectories, flags and
    counters are identical."""
    U = 262144
    a = models.make_args(models.schlogl(), E=[0], t_max=10.0, ita=11, U=U)
    kw = dict(method=2, t_max=10.0, I=a.I, E=a.E, n_sims=U, seed=4, precision=1, specialize=True)
    tabs, rates = tables_of(a), a.rate_fn(a.c)
    print("case: waves_trim 0 vs 1, mixed schlogl pure SSA, %d sims" % U, flush=True)
    O1, O2 = _DeviceArray((1, 11, U)), _DeviceArray((1, 11, U))
    F1, F2 = _DeviceArray((U,)), _DeviceArray((U,))
    try:
        p1 = native.Plan(tabs, rates, a.x_0, **kw)
        p2 = native.Plan(tabs, rates, a.x_0, **kw)
        p2.set_launch(0, 7, 1)
        p1.execute(O1.ptr.value, U, F1.ptr.value)
        p2.execute(O2.ptr.value, U, F2.ptr.value)
        s1, s2 = p1.stats(), p2.stats()
        g1, g2, f1, f2 = O1.get(), O2.get(), F1.get(), F2.get()
    finally:
        for b in (O1, O2, F1, F2):
            b.free()
    assert np.array_equal(g1, g2) and np.array_equal(f1, f2)
    assert all(s1[k] == s2[k] for k in ("events", "ssa_steps", "flagged", "flagged_events"))
    assert s1["ssa_steps"] > 0


def test_autotuned_batch_plan_gives_the_same_results(native):
    """A batch of two networks (one launch group each): after the autotuner the
    trajectories and counters equal those of an untuned plan."""
    from py_stoch_amd import workloads
    specs = [workloads.lotka_volterra(U=4096)[0], workloads.schlogl(U=8192)[0]]
    for s in specs:
        s.kw["I"] = s.kw["I"][:40]
        s.kw["t_max"] = float(s.kw["I"][-1])
    print("case: autotune batch", flush=True)
    bufs1 = [_DeviceArray((s.kappa, s.ita, s.n_sims)) for s in specs]
    bufs2 = [_DeviceArray((s.kappa, s.ita, s.n_sims)) for s in specs]
    try:
        p1 = native.BatchPlan(workloads.jobs(specs, native.PRECISION_MIXED))
        p2 = native.BatchPlan(workloads.jobs(specs, native.PRECISION_MIXED))
        ptrs1, ptrs2 = [b.ptr.value for b in bufs1], [b.ptr.value for b in bufs2]
        strides = [s.n_sims for s in specs]
        ch = p1.autotune(ptrs1, strides, fraction=0.5)
        assert ch["lanes_shift"] in (0, 1) and ch["ssa_slice"] in (3, 7, 15) and ch["waves_trim"] in (0, 1)
        p1.execute(ptrs1, strides)
        p2.execute(ptrs2, strides)
        s1, s2 = p1.stats(), p2.stats()
        for b1, b2 in zip(bufs1, bufs2):
            assert np.array_equal(b1.get(), b2.get())
    finally:
        for b in bufs1 + bufs2:
            b.free()
    assert all(s1[k] == s2[k] for k in ("events", "leaps", "ssa_steps", "flagged", "flagged_events"))
