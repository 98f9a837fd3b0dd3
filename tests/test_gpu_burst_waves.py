"""SSA bursts on paired wavefronts (launch shape burst_waves = 1): a network
kernel of a tau-leaping job runs in workgroups of four decide and four burst
wavefronts, and a simulation's SSA fallback burst moves through LDS to the
paired burst wavefront and back.  Only where a step runs changes, so the FP64
kernel stays bit-exact against the oracle and every precision reproduces the
unpaired launch's trajectories, flags and counters (every test marked gpu)."""
import numpy as np
import pytest

from py_stoch_amd import models, workloads
from py_stoch_amd.mod.parser.tl_parser import tables_of

from devbuf import DeviceArray

pytestmark = pytest.mark.gpu

_KEYS = ("events", "leaps", "rejected_leaps", "ssa_steps", "flagged", "flagged_events", "leap_decisions",
         "sample_rows", "uniforms")


def _run_pair(native, tabs, rates, x0, kw, U, kappa, ita, shapes):
    """Execute one plan per launch shape into its own buffers; returns the
    trajectories, flags and counters of each."""
    out = []
    for shape in shapes:
        O, F = DeviceArray((kappa, ita, U)), DeviceArray((U,))
        try:
            p = native.Plan(tabs, rates, x0, **kw)
            p.set_launch(*shape)
            p.execute(O.ptr.value, U, F.ptr.value)
            st = p.stats()
            out.append((O.get(), F.get(), st))
            p.close()
        finally:
            O.free()
            F.free()
    return out


@pytest.mark.parametrize("slice_", [3, 7])
def test_fp64_paired_is_bit_exact(native, oracle_mod, slice_):
    a = models.make_args(models.lotka_volterra(), E=[0, 1], t_max=10.0, ita=1000, U=2048)
    tabs, rates = tables_of(a), a.rate_fn(a.c)
    kw = dict(method=0, t_max=10.0, I=a.I, E=a.E, n_sims=2048, seed=2, precision=native.PRECISION_FP64,
              specialize=True)
    print("case: fp64 lotka-volterra 2048 sims, burst waves, slice %d" % slice_, flush=True)
    (O, F, st), = _run_pair(native, tabs, rates, a.x_0, kw, 2048, 2, 1000, [(0, slice_, 0, 1)])
    Oc, Fc, sc = oracle_mod.simulate(tabs, rates, a.x_0, method=0, t_max=10.0, I=a.I, E=a.E, n_sims=2048, seed=2)
    assert np.array_equal(O, Oc) and np.array_equal(F, Fc)
    for k in ("events", "leaps", "rejected_leaps", "ssa_steps", "flagged", "flagged_events"):
        assert st[k] == sc[k], (k, st[k], sc[k])
    assert st["ssa_steps"] > 0


def test_mixed_paired_equals_unpaired_full_workload(native):
    """The 65,536-simulation Lotka-Volterra workload, mixed precision."""
    U = 65536
    a = models.make_args(models.lotka_volterra(), E=[0, 1], t_max=10.0, ita=1000, U=U)
    tabs, rates = tables_of(a), a.rate_fn(a.c)
    kw = dict(method=0, t_max=10.0, I=a.I, E=a.E, n_sims=U, seed=2, precision=native.PRECISION_MIXED,
              specialize=True)
    print("case: mixed lotka-volterra %d sims, paired vs unpaired" % U, flush=True)
    (O1, F1, s1), (O2, F2, s2) = _run_pair(native, tabs, rates, a.x_0, kw, U, 2, 1000, [(0, 3, 0, 0), (0, 7, 0, 1)])
    assert np.array_equal(O1, O2) and np.array_equal(F1, F2)
    assert all(s1[k] == s2[k] for k in _KEYS), [(k, s1[k], s2[k]) for k in _KEYS if s1[k] != s2[k]]
    assert s2["ssa_steps"] > 0


@pytest.mark.parametrize("precision", [0, 1])
def test_paired_refill_and_step_budget(native, precision):
    """Schloegl tau-leaping (SSA bursts in the low mode) with more
    simulations than decide lanes, so lanes refill from the work queue, and a
    step budget that ends simulations inside bursts (flags and the events of
    flagged simulations cross the handoff)."""
    U = 300000
    a = models.make_args(models.schlogl(), E=[0], t_max=10.0, ita=11, U=U)
    tabs, rates = tables_of(a), a.rate_fn(a.c)
    kw = dict(method=0, t_max=10.0, I=a.I, E=a.E, n_sims=U, seed=5, precision=precision, specialize=True,
              max_steps=2500)
    print("case: schlogl %d sims, precision %d, step budget, paired vs unpaired" % (U, precision), flush=True)
    (O1, F1, s1), (O2, F2, s2) = _run_pair(native, tabs, rates, a.x_0, kw, U, 1, 11, [(0, 7, 0, 0), (0, 7, 0, 1)])
    assert np.array_equal(O1, O2) and np.array_equal(F1, F2)
    assert all(s1[k] == s2[k] for k in _KEYS), [(k, s1[k], s2[k]) for k in _KEYS if s1[k] != s2[k]]
    assert s2["flagged"] > 0 and s2["flagged_events"] > 0 and s2["ssa_steps"] > 0


def test_paired_batch_switches_jobs(native):
    """A batch with two parameter sets of one network as two jobs (one
    kernel, so each decide wavefront moves from job to job and its burst
    wavefront follows) and a second network: paired equals unpaired."""
    lv = workloads.lotka_volterra(U=4096)[0]
    lv2 = workloads.lotka_volterra(U=3000)[0]
    lv2.rates = lv2.rates * 1.3
    sc = workloads.schlogl(U=8192)[0]
    specs = [lv, lv2, sc]
    for s in specs:
        s.kw["I"] = s.kw["I"][:60]
        s.kw["t_max"] = float(s.kw["I"][-1])
    print("case: batch of three jobs, paired vs unpaired", flush=True)
    res = []
    for bw in (0, 1):
        bufs = [DeviceArray((s.kappa, s.ita, s.n_sims)) for s in specs]
        flg = [DeviceArray((s.n_sims,)) for s in specs]
        try:
            p = native.BatchPlan(workloads.jobs(specs, native.PRECISION_MIXED))
            p.set_launch(0, 7, 0, bw)
            p.execute([b.ptr.value for b in bufs], [s.n_sims for s in specs], [f.ptr.value for f in flg])
            st = p.stats()
            res.append(([b.get() for b in bufs], [f.get() for f in flg], st))
        finally:
            for b in bufs + flg:
                b.free()
    (O1, F1, s1), (O2, F2, s2) = res
    for a_, b_ in zip(O1 + F1, O2 + F2):
        assert np.array_equal(a_, b_)
    assert all(s1[k] == s2[k] for k in _KEYS), [(k, s1[k], s2[k]) for k in _KEYS if s1[k] != s2[k]]
