"""A batch plan with several kernel launches spreads them over the caller's
stream and up to three library streams (forked and joined with events).  One
mixed-precision batch with a job for every kernel kind -- the table kernels
of the 4x4 and 8x8 size classes, two network kernels and the large-network
kernel, five launches over four streams -- must give, job by job, the same
bits as that job run alone (one launch, one stream).  Every kernel runs
through the C ABI."""
import numpy as np
import pytest

from py_stoch_amd import models
from py_stoch_amd.mod.parser.tl_parser import tables_of
from py_stoch_amd.workloads import mass_action_tables

pytestmark = pytest.mark.gpu


def _random_job(native, seed, N, M, n_sims):
    spn = models.random_network(np.random.default_rng(seed), n_species=N, n_reactions=M)
    tabs, rates = mass_action_tables(spn)
    I = np.linspace(0.0, 1.0, num=11).astype(np.float32)
    return native.Job(tabs, rates, np.asarray(spn.M, np.int32), method=0, precision=native.PRECISION_MIXED,
                      t_max=1.0, I=I, E=np.arange(N, dtype=np.int32), n_sims=n_sims, seed=seed,
                      max_steps=3000)


def test_every_kernel_kind_in_one_batch(native):
    a_lv = models.make_args(models.lotka_volterra(), E=[0, 1], t_max=2.0, ita=21, U=4096)
    a_sc = models.make_args(models.schlogl(), E=[0], t_max=2.0, ita=21, U=8192)
    jobs = [
        # table kernel, 4x4 class
        native.Job(tables_of(a_lv), a_lv.rate_fn(a_lv.c)[None, :], a_lv.x_0[None, :], method=0,
                   precision=native.PRECISION_MIXED, t_max=2.0, I=a_lv.I, E=a_lv.E, n_sims=4096, seed=21),
        # table kernel, 8x8 class
        _random_job(native, 31, 6, 8, 2048),
        # network kernels (one launch each)
        native.Job(tables_of(a_lv), a_lv.rate_fn(a_lv.c)[None, :], a_lv.x_0[None, :], method=0,
                   precision=native.PRECISION_MIXED, t_max=2.0, I=a_lv.I, E=a_lv.E, n_sims=4096, seed=22,
                   specialize=True),
        native.Job(tables_of(a_sc), a_sc.rate_fn(a_sc.c)[None, :], a_sc.x_0[None, :], method=0,
                   precision=native.PRECISION_MIXED, t_max=2.0, I=a_sc.I, E=a_sc.E, n_sims=8192, seed=23,
                   specialize=True),
        # large-network kernel (beyond 16 species x 32 reactions)
        _random_job(native, 41, 20, 40, 512),
    ]
    print("case: five-launch mixed batch", flush=True)
    Os, Fs, st = native.batch_simulate(jobs)
    events = 0
    for q, job in enumerate(jobs):
        print("case: job %d alone" % q, flush=True)
        O1, F1, s1 = native.batch_simulate([job])
        assert np.array_equal(Os[q], O1[0]), "job %d differs" % q
        assert np.array_equal(Fs[q], F1[0]), "job %d flags differ" % q
        events += s1["events"]
    assert st["events"] == events
