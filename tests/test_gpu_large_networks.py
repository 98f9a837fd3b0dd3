"""The large-network kernel (one wavefront per simulation, mixed precision):
networks beyond the register-resident size classes, checked in distribution
against the FP64 CPU oracle (every test marked gpu)."""
import numpy as np
import pytest

from py_stoch_amd import models
from py_stoch_amd.mod.parser.tl_parser import tables_of

pytestmark = pytest.mark.gpu


def _ks(x, y):
    x, y = np.sort(x), np.sort(y)
    g = np.concatenate([x, y])
    return np.abs(np.searchsorted(x, g, side="right") / x.size - np.searchsorted(y, g, side="right") / y.size).max()


def _padded_lotka_volterra(extra=20):
    """Lotka-Volterra with `extra` species that take part in no reaction:
    the same dynamics, but too many species for the register kernels."""
    reactions = [({0: 1}, {0: 2}, "parameter_0*species_0"),
                 ({0: 1, 1: 1}, {1: 2}, "parameter_1*species_0*species_1"),
                 ({1: 1}, {}, "parameter_2*species_1")]
    return models._spn(2 + extra, reactions, [1000, 1000] + [7] * extra, [10.0, 0.01, 10.0])


def test_padded_lotka_volterra_matches_oracle(native, oracle_mod):
    """KS distance between the wave kernel and the oracle (4096 simulations
    each) at every fifth sample of both species below the 0.01% critical
    value 2.23 sqrt(2/4096); the inert species stay at their initial count."""
    a = models.make_args(_padded_lotka_volterra(), E=[0, 1, 2], t_max=2.0, ita=41, U=4096)
    tabs = tables_of(a)
    kw = dict(method=0, t_max=2.0, I=a.I, E=a.E, n_sims=4096)
    print("case: padded lotka-volterra N=%d" % a.N, flush=True)
    O_w, f_w, s_w = native.simulate(tabs, a.rate_fn(a.c), a.x_0, precision=native.PRECISION_MIXED, seed=21, **kw)
    O_c, f_c, s_c = oracle_mod.simulate(tabs, a.rate_fn(a.c), a.x_0, seed=22, **kw)
    assert s_w["leaps"] > 0
    crit = 2.23 * np.sqrt(2.0 / 4096)
    for k in range(2):
        for f in range(5, 41, 5):
            assert _ks(O_w[k, f], O_c[k, f]) < crit, (k, f)
    assert (O_w[2] == 7).all()
    assert abs(s_w["events"] / s_c["events"] - 1.0) < 0.02


def test_pure_ssa_padded_birth_death(native):
    """Pure SSA on the wave kernel: birth-death padded to 20 species follows
    X(t) ~ Poisson(100 (1 - e^{-t/10})) (mean within 4 standard errors)."""
    reactions = [({}, {0: 1}, "parameter_0"), ({0: 1}, {}, "parameter_1*species_0")]
    spn = models._spn(20, reactions, [0] * 20, [10.0, 0.1])
    a = models.make_args(spn, E=[0], t_max=50.0, ita=11, U=4096)
    O, f, st = native.simulate(tables_of(a), a.rate_fn(a.c), a.x_0, method=native.METHOD_PURE_SSA,
                               precision=native.PRECISION_MIXED, t_max=50.0, I=a.I, E=a.E, n_sims=4096, seed=5)
    t = np.asarray(a.I, np.float64)
    lam = 100.0 * (1.0 - np.exp(-0.1 * t))
    m = O[0].mean(axis=-1)
    assert np.all(np.abs(m - lam)[1:] <= 4.0 * np.sqrt(lam[1:] / 4096))


def test_random_network_matches_oracle_and_is_deterministic(native, oracle_mod):
    """A random mass-action network with 30 species and 120 reactions: per
    species, the final-time mean of the wave kernel lies within 5 standard
    errors of the oracle's (flagged simulations left out on both sides), and
    two runs of the kernel give the same bits."""
    rng = np.random.default_rng(5)
    spn = models.random_network(rng, n_species=30, n_reactions=120)
    a = models.make_args(spn, E=list(range(30)), t_max=0.5, ita=6, U=2048)
    tabs = tables_of(a)
    kw = dict(method=0, t_max=0.5, I=a.I, E=a.E, n_sims=2048)
    print("case: random network 30 x 120", flush=True)
    O_w, f_w, s_w = native.simulate(tabs, a.rate_fn(a.c), a.x_0, precision=native.PRECISION_MIXED, seed=31, **kw)
    O_w2, f_w2, _ = native.simulate(tabs, a.rate_fn(a.c), a.x_0, precision=native.PRECISION_MIXED, seed=31, **kw)
    assert np.array_equal(O_w, O_w2) and np.array_equal(f_w, f_w2)
    O_c, f_c, s_c = oracle_mod.simulate(tabs, a.rate_fn(a.c), a.x_0, seed=32, **kw)
    kw_ok, kc_ok = f_w == 0, f_c == 0
    assert kw_ok.mean() > 0.9 and kc_ok.mean() > 0.9
    x, y = O_w[:, -1, kw_ok].astype(np.float64), O_c[:, -1, kc_ok].astype(np.float64)
    se = np.sqrt(x.var(axis=-1) / x.shape[-1] + y.var(axis=-1) / y.shape[-1])
    z = np.abs(x.mean(axis=-1) - y.mean(axis=-1)) / np.maximum(se, 1e-9)
    assert z.max() < 5.0, z
