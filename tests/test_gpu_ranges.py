"""Sample ranges and graph-captured chunks: a run split at sample boundaries
(every simulation stopped and saved to device memory, then resumed) gives the
same trajectories, flags and counters as one execute -- FP64 against the
oracle too -- and a captured graph of the ranges, replayed several times
with other work in between, gives the same result every time (every test
marked gpu)."""
import numpy as np
import pytest

from py_stoch_amd import models
from py_stoch_amd.mod.parser.tl_parser import tables_of

from devbuf import DeviceArray

pytestmark = pytest.mark.gpu

_KEYS = ("events", "leaps", "rejected_leaps", "ssa_steps", "flagged", "flagged_events", "sample_rows")


def _lv(U, ita, t_max=10.0):
    a = models.make_args(models.lotka_volterra(), E=[0, 1], t_max=t_max, ita=ita, U=U)
    return a, tables_of(a), a.rate_fn(a.c)


def _shape(layout, kappa, ita, U):
    return (kappa, ita, U) if layout == 0 else (kappa, (ita + 7) // 8, U, 8)


def _whole_and_ranges(native, make_plan, kappa, ita, U, ends, layout=0, shape=None):
    """One execute, then the same plan as consecutive ranges into fresh
    buffers; returns ((O, F, stats), (O, F, summed stats, positions))."""
    out = []
    for split in (False, True):
        O, F = DeviceArray(_shape(layout, kappa, ita, U)), DeviceArray((U,))
        try:
            p = make_plan()
            if shape:
                p.set_launch(*shape)
            if not split:
                p.execute(O.ptr.value, U, F.ptr.value)
                st = p.stats()
                pos = None
            else:
                st, pos = {k: 0 for k in _KEYS}, []
                for f in ends:
                    p.execute_range([O.ptr.value], [U], f, [F.ptr.value])
                    s = p.stats()
                    for k in _KEYS:
                        st[k] += s[k]
                    pos.append(p.range_position())
            out.append((O.get(), F.get(), st, pos))
            p.close()
        finally:
            O.free()
            F.free()
    return out


@pytest.mark.parametrize("specialize,shape", [(False, None), (True, None), (True, (0, 7, 0, 1))])
def test_fp64_ranges_bit_exact(native, oracle_mod, specialize, shape):
    U, ita = 2048, 1000
    a, tabs, rates = _lv(U, ita)
    kw = dict(method=0, t_max=10.0, I=a.I, E=a.E, n_sims=U, seed=2, precision=0, specialize=specialize)
    print("case: fp64 ranges, specialize %s, shape %s" % (specialize, shape), flush=True)
    (O1, F1, s1, _), (O2, F2, s2, pos) = _whole_and_ranges(
        native, lambda: native.Plan(tabs, rates, a.x_0, **kw), 2, ita, U, [1, 200, 457, 999, 1000], shape=shape)
    assert pos == [1, 200, 457, 999, 0]
    assert np.array_equal(O1, O2) and np.array_equal(F1, F2)
    assert all(s1[k] == s2[k] for k in _KEYS), [(k, s1[k], s2[k]) for k in _KEYS]
    Oc, Fc, sc = oracle_mod.simulate(tabs, rates, a.x_0, method=0, t_max=10.0, I=a.I, E=a.E, n_sims=U, seed=2)
    assert np.array_equal(O2, Oc) and np.array_equal(F2, Fc) and s2["events"] == sc["events"]
    assert s2["ssa_steps"] > 0


@pytest.mark.parametrize("shape", [(0, 7, 0, 0), (0, 7, 0, 1)])
def test_mixed_ranges_blocks_step_budget(native, shape):
    """Mixed precision, blocked layout (ranges end on multiples of 8), more
    simulations than lanes (refill from the queue) and a step budget that ends
    simulations inside ranges and inside bursts."""
    U, ita = 100000, 100
    a = models.make_args(models.schlogl(), E=[0], t_max=10.0, ita=ita, U=U)
    tabs, rates = tables_of(a), a.rate_fn(a.c)
    kw = dict(method=0, t_max=10.0, I=a.I, E=a.E, n_sims=U, seed=6, precision=1, specialize=True, max_steps=3000,
              layout=native.LAYOUT_BLOCKS)
    print("case: mixed schlogl ranges, blocks, shape %s" % (shape,), flush=True)
    (O1, F1, s1, _), (O2, F2, s2, pos) = _whole_and_ranges(
        native, lambda: native.Plan(tabs, rates, a.x_0, **kw), 1, ita, U, [8, 48, 96, 100], layout=1, shape=shape)
    assert pos == [8, 48, 96, 0]
    assert np.array_equal(O1, O2) and np.array_equal(F1, F2)
    assert all(s1[k] == s2[k] for k in _KEYS), [(k, s1[k], s2[k]) for k in _KEYS]
    assert s2["flagged"] > 0 and s2["ssa_steps"] > 0


def test_graph_replays_equal_eager(native):
    """Four ranges captured in a graph; three replays, each after eager work
    on the same plan and device, each equal to the eager run."""
    U, ita = 65536, 1000
    a, tabs, rates = _lv(U, ita)
    kw = dict(method=0, t_max=10.0, I=a.I, E=a.E, n_sims=U, seed=2, precision=1, specialize=True,
              layout=native.LAYOUT_BLOCKS)
    shp = _shape(1, 2, ita, U)
    E, G, X = DeviceArray(shp), DeviceArray(shp), DeviceArray((2, 100, 4096))
    try:
        p = native.Plan(tabs, rates, a.x_0, **kw)
        p.set_launch(0, 7, 0, 1)
        print("case: graph of four ranges, eager reference", flush=True)
        p.execute(E.ptr.value, U)
        se = p.stats()
        eager = E.get()
        for f in (256, 504, 752, 1000):
            p.execute_range([G.ptr.value], [U], f)
        p.stats()
        got = G.get()
        bad = np.argwhere(got != eager)
        assert bad.shape[0] == 0, "ranges: %d entries differ, first %s" % (bad.shape[0], bad[0].tolist())
        p.graph_create([G.ptr.value], [U], [256, 504, 752, 1000])
        other_a, other_t, other_r = _lv(4096, 100, 1.0)
        other = native.Plan(other_t, other_r, other_a.x_0, method=0, t_max=1.0, I=other_a.I, E=other_a.E,
                            n_sims=4096, seed=9, precision=1)
        for rep in range(3):
            print("case: graph replay %d" % rep, flush=True)
            other.execute(X.ptr.value, 4096)
            other.stats()
            p.graph_launch()
            sg = p.stats()
            got = G.get()
            bad = np.argwhere(got != eager)
            assert bad.shape[0] == 0, "replay %d: %d entries differ, first %s (graph %d, eager %d); sims %s" % (
                rep, bad.shape[0], bad[0].tolist(), got[tuple(bad[0])], eager[tuple(bad[0])],
                np.unique(bad[:, 2])[:10].tolist())
            assert all(sg[k] == se[k] for k in _KEYS), (rep, [(k, sg[k], se[k]) for k in _KEYS])
            assert sg["kernel_ms"] > 0.0
        # a changed launch shape invalidates the graph
        p.set_launch(0, 7, 0, 0)
        with pytest.raises(native.NativeError):
            p.graph_launch()
        p.close()
        other.close()
    finally:
        for b in (E, G, X):
            b.free()


def test_range_argument_checks(native):
    """Errors raised on the host before any launch."""
    U, ita = 1024, 20
    a, tabs, rates = _lv(U, ita, 1.0)
    O = DeviceArray((2, ita, U))
    try:
        p = native.Plan(tabs, rates, a.x_0, method=0, t_max=1.0, I=a.I, E=a.E, n_sims=U, seed=1, precision=1)
        p.execute_range([O.ptr.value], [U], 5)
        with pytest.raises(native.NativeError):
            p.execute_range([O.ptr.value], [U], 5)  # must advance
        p.execute_range([O.ptr.value], [U], 20)
        assert p.range_position() == 0
        p.close()
        pb = native.Plan(tabs, rates, a.x_0, method=0, t_max=1.0, I=a.I, E=a.E, n_sims=U, seed=1, precision=1,
                         layout=native.LAYOUT_BLOCKS)
        with pytest.raises(native.NativeError):
            pb.execute_range([O.ptr.value], [U], 5)  # blocked layout: multiples of 8
        with pytest.raises(native.NativeError):
            pb.graph_create([O.ptr.value], [U], [8, 16])  # does not reach the last sample
        pb.close()
    finally:
        O.free()


def _batch_jobs(native, layout):
    """Three FP64 jobs (one precision per batch) on three kernels:
    Lotka-Volterra on the table kernel, Schlogl (fewer samples) on its network
    kernel, and a Lotka-Volterra sweep of two parameter rows with a step
    budget on the Lotka-Volterra network kernel."""
    lv = models.make_args(models.lotka_volterra(), E=[0, 1], t_max=5.0, ita=96, U=3000)
    sc = models.make_args(models.schlogl(), E=[0], t_max=5.0, ita=40, U=2000)
    lv_t, sc_t = tables_of(lv), tables_of(sc)
    rates2 = np.stack([lv.rate_fn(lv.c), lv.rate_fn(lv.c) * np.array([1.2, 0.9, 1.1])])
    x02 = np.stack([lv.x_0, lv.x_0 // 2])
    return [
        native.Job(lv_t, lv.rate_fn(lv.c), lv.x_0, method=0, t_max=5.0, I=lv.I, E=lv.E, n_sims=3000, seed=4,
                   precision=0, layout=layout),
        native.Job(sc_t, sc.rate_fn(sc.c), sc.x_0, method=0, t_max=5.0, I=sc.I, E=sc.E, n_sims=2000, seed=5,
                   precision=0, specialize=True, layout=layout),
        native.Job(lv_t, rates2, x02, method=0, t_max=5.0, I=lv.I, E=lv.E, n_sims=4000, sims_per_set=2000,
                   seed=6, precision=0, specialize=True, layout=layout, max_steps=4000),
    ]


@pytest.mark.parametrize("layout", [0, 1])
def test_batch_plan_ranges_bit_exact(native, layout):
    """A batch plan split into ranges (jobs of different sample counts: a
    range past a job's last sample ends that job) equals one execute of the
    same batch: trajectories, flags and summed counters."""
    ends = [16, 40, 64, 96]
    out = []
    for split in (False, True):
        jobs = _batch_jobs(native, layout)
        shapes = [_shape(layout, j.kappa, j.ita, j.n_sims) for j in jobs]
        bufs = [DeviceArray(s) for s in shapes]
        flags = [DeviceArray((j.n_sims,)) for j in jobs]
        try:
            p = native.BatchPlan(jobs)
            O_ptrs, strides, F_ptrs = [b.ptr.value for b in bufs], [j.n_sims for j in jobs], [f.ptr.value for f in flags]
            print("case: batch ranges, layout %d, split %s" % (layout, split), flush=True)
            if not split:
                p.execute(O_ptrs, strides, F_ptrs)
                st = p.stats()
            else:
                st = {k: 0 for k in _KEYS}
                for f in ends:
                    p.execute_range(O_ptrs, strides, f, F_ptrs)
                    s = p.stats()
                    for k in _KEYS:
                        st[k] += s[k]
                assert p.range_position() == 0
            out.append(([b.get() for b in bufs], [f.get() for f in flags], st))
            p.close()
        finally:
            for b in bufs + flags:
                b.free()
    (O1, F1, s1), (O2, F2, s2) = out
    for q in range(3):
        assert np.array_equal(O1[q], O2[q]), "job %d trajectories differ" % q
        assert np.array_equal(F1[q], F2[q]), "job %d flags differ" % q
    assert all(s1[k] == s2[k] for k in _KEYS), [(k, s1[k], s2[k]) for k in _KEYS]
    assert s2["ssa_steps"] > 0 and s2["leaps"] > 0


def test_batch_ranges_refuse_large_network_job(native):
    """Sample ranges run the lane kernels only: a batch with a job for the
    large-network kernel refuses them before any launch."""
    rng = np.random.default_rng(3)
    big = models.make_args(models.random_network(rng, 24, 70), E=[0, 1], t_max=0.2, ita=16, U=256)
    job = native.Job(tables_of(big), big.rate_fn(big.c), big.x_0, method=0, t_max=0.2, I=big.I, E=big.E,
                     n_sims=256, seed=2, precision=1, max_steps=2000)
    O, F = DeviceArray((2, 16, 256)), DeviceArray((256,))
    try:
        p = native.BatchPlan([job])
        with pytest.raises(native.NativeError):
            p.execute_range([O.ptr.value], [256], 8, [F.ptr.value])
        p.close()
    finally:
        O.free()
        F.free()
