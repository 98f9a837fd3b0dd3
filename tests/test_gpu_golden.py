"""GPU kernels against the committed golden fixtures (tests/golden): the
FP64 table-driven and network kernels reproduce the committed oracle
trajectories bit for bit, and the mixed-precision kernel samples the exact
Schlogl law of the committed master-equation solution (every test marked
gpu; nothing here runs the oracle)."""
import os

import numpy as np
import pytest

from py_stoch_amd import models
from py_stoch_amd.mod.parser.tl_parser import tables_of
from test_golden import GOLDEN, golden_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("specialize", [False, True])
@pytest.mark.parametrize("name", ["bd", "lv", "sc"])
def test_fp64_kernels_reproduce_golden_trajectories(native, name, specialize):
    a, method, seed, O, flags, stats = golden_case(name)
    print("case: golden %s specialize=%d" % (name, specialize), flush=True)
    O_g, f_g, st = native.simulate(tables_of(a), a.rate_fn(a.c), a.x_0, method=method, precision=native.PRECISION_FP64,
                                   t_max=a.t_max, I=a.I, E=a.E, n_sims=a.U, seed=seed, specialize=specialize)
    mism = np.argwhere(O_g != O)
    assert mism.shape[0] == 0, "first mismatch %s gpu=%d golden=%d" % (
        mism[0].tolist(), O_g[tuple(mism[0])], O[tuple(mism[0])])
    assert np.array_equal(f_g, flags)
    assert [st[k] for k in ("events", "leaps", "rejected_leaps", "ssa_steps", "flagged", "flagged_events")] == stats.tolist()


@pytest.mark.parametrize("method", [0, 2])
def test_mixed_schlogl_matches_master_equation(native, method):
    """Mixed precision, 65,536 simulations: the KS distance of X(t) to the
    exact law at t = 1, 2, 5, 10 stays below 0.02 (the accuracy bar of the
    mixed path) for tau-leaping, and within the 99.9% sampling band for pure
    SSA."""
    d = np.load(os.path.join(GOLDEN, "schlogl_cme.npz"))
    U = 65536
    a = models.make_args(models.schlogl(), E=[0], t_max=10.0, ita=11, U=U)
    print("case: mixed schlogl method=%d" % method, flush=True)
    O, flags, st = native.simulate(tables_of(a), a.rate_fn(a.c), a.x_0, method=method, precision=native.PRECISION_MIXED,
                                   t_max=10.0, I=a.I, E=a.E, n_sims=U, seed=31)
    assert (flags == 0).all()
    band = 1.95 / np.sqrt(U)
    for k, t in enumerate(d["t"]):
        f = int(round(float(t)))               # samples at t = 0, 1, ..., 10
        assert abs(float(a.I[f]) - float(t)) < 1e-6
        x = np.sort(O[0, f])
        cdf = np.cumsum(d["pmf"][k])
        ks = np.abs(np.searchsorted(x, np.arange(cdf.size), side="right") / U - cdf).max()
        assert ks < (0.02 if method == 0 else band), (float(t), ks)
