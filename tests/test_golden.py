"""The committed golden fixtures (tests/golden, written by make_golden.py):
Philox4x32-10 against the published Random123 known answers, the oracle
against its own committed trajectories (any change to the oracle's
arithmetic, stream layout or sampling shows up here), and the Schlogl
master-equation fixture against the mathematics it comes from."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.linalg import expm_multiply

from py_stoch_amd import models
from py_stoch_amd.mod.parser.tl_parser import tables_of

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MODELS = {"bd": (models.birth_death, [0]), "lv": (models.lotka_volterra, [0, 1]), "sc": (models.schlogl, [0])}


def golden_case(name):
    """(args, method, seed, O, flags, stats) of one committed oracle run."""
    d = np.load(os.path.join(GOLDEN, "oracle_trajectories.npz"))
    method, t_max, ita, U, seed = d[name + "_settings"]
    make, E = MODELS[name]
    a = models.make_args(make(), E=E, t_max=float(t_max), ita=int(ita), U=int(U))
    return a, int(method), int(seed), d[name + "_O"], d[name + "_flags"], d[name + "_stats"]


def test_philox_matches_published_known_answers(oracle_mod):
    with open(os.path.join(GOLDEN, "philox4x32_10_kat.json")) as f:
        kat = json.load(f)["vectors"]
    assert len(kat) == 3
    for v in kat:
        assert oracle_mod.philox(v["counter"], v["key"]) == v["output"]


@pytest.mark.parametrize("name", ["bd", "lv", "sc"])
def test_oracle_reproduces_committed_trajectories(oracle_mod, name):
    a, method, seed, O, flags, stats = golden_case(name)
    O2, f2, st = oracle_mod.simulate(tables_of(a), a.rate_fn(a.c), a.x_0, method=method, t_max=a.t_max, I=a.I,
                                     E=a.E, n_sims=a.U, seed=seed)
    assert np.array_equal(O2, O)
    assert np.array_equal(f2, flags)
    assert [st[k] for k in ("events", "leaps", "rejected_leaps", "ssa_steps", "flagged", "flagged_events")] == stats.tolist()


def test_golden_trajectories_obey_the_models():
    # birth-death is pure SSA: counts start at 0, and the mean follows 100 (1 - exp(-0.1 t))
    a, _, _, O, _, _ = golden_case("bd")
    assert (O[0, 0] == 0).all()
    m = 100.0 * (1.0 - np.exp(-0.1 * a.I.astype(float)))
    z = (O[0].mean(axis=1) - m) / np.sqrt(np.maximum(m, 1e-9) / O.shape[2])
    assert np.abs(z[1:]).max() < 4.5
    for name in ("lv", "sc"):
        a, _, _, O, flags, stats = golden_case(name)
        assert (O[:, 0, :] == np.asarray(a.x_0)[a.E][:, None]).all()  # sample 0 is x0 in every simulation
        assert stats[1] > 0                                        # leaps were taken
        assert (O >= 0).all() and (flags == 0).all()


def test_schlogl_master_equation_fixture():
    d = np.load(os.path.join(GOLDEN, "schlogl_cme.npz"))
    p, t = d["pmf"], d["t"]
    assert (p >= -1e-15).all() and np.allclose(p.sum(axis=1), 1.0, atol=1e-10)
    x = np.arange(p.shape[1])
    assert abs((p[-1] * x).sum() - 316.59) < 0.01 and abs(p[-1, :300].sum() - 0.5135) < 1e-4
    # bimodal at t = 10: local maxima near 83 and 563
    assert 70 < np.argmax(p[-1, :300]) < 100 and 540 < 300 + np.argmax(p[-1, 300:]) < 590
    # an independent propagation (two half steps) reproduces t = 1
    c1, c2, c3, c4, A, B, x0, xmax = d["params"]
    xf = np.arange(int(xmax) + 1, dtype=float)
    birth = c1 / 2 * A * xf * (xf - 1) + c3 * B
    death = c2 / 6 * xf * (xf - 1) * (xf - 2) + c4 * xf
    birth[-1] = 0.0
    Q = sp.diags([-(birth + death), birth[:-1], death[1:]], [0, -1, 1], format="csc")
    p0 = np.zeros(xf.size)
    p0[int(x0)] = 1.0
    half = expm_multiply(Q * (t[0] / 2), expm_multiply(Q * (t[0] / 2), p0))
    assert np.abs(half - p[0]).max() < 1e-9


def test_oracle_ssa_matches_master_equation_fixture(oracle_mod):
    d = np.load(os.path.join(GOLDEN, "schlogl_cme.npz"))
    t_end = float(d["t"][2])  # 5
    a = models.make_args(models.schlogl(), E=[0], t_max=t_end, ita=11, U=2000)
    O, _, _ = oracle_mod.simulate(tables_of(a), a.rate_fn(a.c), a.x_0, method=oracle_mod.METHOD_PURE_SSA,
                                  t_max=t_end, I=a.I, E=a.E, n_sims=2000, seed=21)
    x = np.sort(O[0, -1])
    cdf = np.cumsum(d["pmf"][2])
    emp = np.searchsorted(x, np.arange(cdf.size), side="right") / x.size
    assert np.abs(emp - cdf).max() < 1.95 / np.sqrt(x.size)
