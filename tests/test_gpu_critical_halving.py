"""The critical-reaction leap (tau_2 and SingleCriticalReaction,
P1_P2.py:175-220, 488-507) and tau halving (P1_P2.py:228-259) on the GPU.

FP64 table-driven and network kernels: bit-exact against the oracle, whose
critical branch is pinned by tests/test_oracle_critical.py, on configurations
that exercise each branch (every consuming reaction critical; critical X
beside a leapt reservoir, a0 >> a0c; n_c = 0 and eta = 1, where leaps
overshoot below zero and are halved).  Mixed precision (all three kernel
families): the same configurations against the exact laws of tests/laws.py,
or against the FP64 oracle in distribution where no closed form exists
(every test marked gpu)."""
import numpy as np
import pytest

from py_stoch_amd import models
from py_stoch_amd.mod.parser.tl_parser import tables_of

import laws

pytestmark = pytest.mark.gpu

CASES = {
    # name: (network, recorded species, t_max, ita, U, n_c, eta, seed)
    "death_critical": (lambda: laws.pure_death(1.0, 50), [0], 3.0, 7, 4000, laws.BIG_NC, 0.03, 31),
    "decays_critical": (lambda: laws.competing_decays(1.0, 3.0, 30), [0, 1], 6.0, 7, 4000, laws.BIG_NC, 0.03, 32),
    "reservoir": (lambda: laws.decays_beside_reservoir(), [0, 1, 2], 1.0, 5, 4000, 10, 0.03, 34),
    "halving": (lambda: models.birth_death(birth=0.0, death=1.0, x0=50), [0], 10.0, 11, 2000, 0, 1.0, 2),
}


def _args(name, padded=False):
    net, E, t_max, ita, U, n_c, eta, seed = CASES[name]
    spn = laws.padded(net()) if padded else net()
    a = models.make_args(spn, E=E, t_max=t_max, ita=ita, U=U)
    kw = dict(method=0, t_max=t_max, I=a.I, E=a.E, n_sims=U, n_c=n_c, eta=eta)
    return a, tables_of(a), a.rate_fn(a.c), kw, seed


@pytest.mark.parametrize("specialize", [False, True])
@pytest.mark.parametrize("name", list(CASES))
def test_fp64_bit_exact(native, oracle_mod, name, specialize):
    a, tabs, rates, kw, seed = _args(name)
    print("case: %s fp64 specialize=%s U=%d" % (name, specialize, kw["n_sims"]), flush=True)
    O_g, f_g, s_g = native.simulate(tabs, rates, a.x_0, precision=native.PRECISION_FP64, specialize=specialize,
                                    seed=seed, **kw)
    O_c, f_c, s_c = oracle_mod.simulate(tabs, rates, a.x_0, seed=seed, **kw)
    mism = np.argwhere(O_g != O_c)
    assert mism.shape[0] == 0, "first mismatch %s gpu=%d cpu=%d" % (
        mism[0].tolist(), O_g[tuple(mism[0])], O_c[tuple(mism[0])])
    assert np.array_equal(f_g, f_c) and f_g.sum() == 0
    for k in ("events", "leaps", "rejected_leaps", "ssa_steps", "flagged", "flagged_events"):
        assert s_g[k] == s_c[k], (k, s_g[k], s_c[k])
    if name == "halving":
        assert s_g["rejected_leaps"] > 0
    else:
        assert s_g["ssa_steps"] == 0 and s_g["leaps"] > 0


# mixed precision: register kernels (table-driven, network compiled in) and the
# large-network kernel (the same network padded with inert species)
MIXED = [("table", False, False), ("network", True, False), ("large", False, True)]


def _mixed(native, name, specialize, pad, seed_shift=100):
    a, tabs, rates, kw, seed = _args(name, padded=pad)
    print("case: %s mixed specialize=%s padded=%s" % (name, specialize, pad), flush=True)
    O, f, st = native.simulate(tabs, rates, a.x_0, precision=native.PRECISION_MIXED, specialize=specialize,
                               seed=seed + seed_shift, **kw)
    assert f.sum() == 0
    return a, O, st


@pytest.mark.parametrize("kind, specialize, pad", MIXED)
def test_mixed_critical_laws(native, kind, specialize, pad):
    a, O, st = _mixed(native, "death_critical", specialize, pad)
    assert st["ssa_steps"] == 0 and st["leaps"] == st["events"] > 0
    for f in range(1, 7):
        t = float(a.I[f])
        laws.assert_law(O[0, f], laws.midpoint_death_law(50, 1.0, t), "%s X(%g)" % (kind, t))
    a, O, st = _mixed(native, "decays_critical", specialize, pad)
    for f in (1, 2):
        t = float(a.I[f])
        laws.assert_law(O[0, f], laws.midpoint_death_law(30, 4.0, t), "%s X(%g)" % (kind, t))
        laws.assert_law(O[1, f], laws.midpoint_split_law(30, 4.0, 0.75, t), "%s Y(%g)" % (kind, t))
    assert (O[0, -1] == 0).all()
    laws.assert_law(O[1, -1], laws.binom(30, 0.75), "%s Y(6)" % kind)


@pytest.mark.parametrize("kind, specialize, pad", MIXED)
@pytest.mark.parametrize("name", ["reservoir", "halving"])
def test_mixed_matches_oracle_in_distribution(native, oracle_mod, name, kind, specialize, pad):
    """Two-sample KS between mixed precision and the FP64 oracle at every
    sample of every recorded species, below the 0.01% critical value."""
    _, O, st = _mixed(native, name, specialize, pad)
    a, tabs, rates, kw, seed = _args(name)  # the network without padding
    O_c, f_c, s_c = oracle_mod.simulate(tabs, rates, a.x_0, seed=seed, **kw)
    if name == "halving":
        assert st["rejected_leaps"] > 0 and s_c["rejected_leaps"] > 0
    crit = 2.23 * np.sqrt(2.0 / kw["n_sims"])
    for q in range(O_c.shape[0]):
        for f in range(1, O_c.shape[1]):
            assert laws.ks2(O[q, f], O_c[q, f]) < crit, (kind, name, q, f)
