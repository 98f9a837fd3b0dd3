"""Blocked trajectory layout (LAYOUT_BLOCKS, O[kappa][ceil(ita/8)][U][8]):
every kernel kind writes the same samples as in the rows layout, the last
block is padded with zeros, and pst_blocks_to_rows gives back the rows
array bit for bit (every test marked gpu)."""
import numpy as np
import pytest

from py_stoch_amd import models, workloads
from py_stoch_amd.mod.parser.tl_parser import tables_of

from devbuf import DeviceArray

pytestmark = pytest.mark.gpu


def _plan_pair(native, make_plan, kappa, ita, U, shape=None):
    """Run the same plan in rows and in blocks layout; returns (rows, blocks
    converted on the host, blocks converted by pst_blocks_to_rows, raw blocks)."""
    ita8 = (ita + 7) // 8
    R, B, C = DeviceArray((kappa, ita, U)), DeviceArray((kappa, ita8, U, 8)), DeviceArray((kappa, ita, U))
    try:
        for layout, buf in ((native.LAYOUT_ROWS, R), (native.LAYOUT_BLOCKS, B)):
            p = make_plan(layout)
            if shape:
                p.set_launch(*shape)
            p.execute(buf.ptr.value, U)
            p.stats()
            p.close()
        native.blocks_to_rows(B.ptr.value, U, C.ptr.value, U, kappa, ita, U)
        rows, blocks, conv = R.get(), B.get(), C.get()
    finally:
        for b in (R, B, C):
            b.free()
    host = blocks.transpose(0, 1, 3, 2).reshape(kappa, ita8 * 8, U)[:, :ita]
    return rows, host, conv, blocks


def _check(rows, host, conv, blocks, ita):
    assert np.array_equal(rows, host)
    assert np.array_equal(rows, conv)
    pad = (ita + 7) // 8 * 8 - ita
    if pad:
        assert not blocks[:, -1, :, 8 - pad:].any(), "padding of the last block is not zero"


@pytest.mark.parametrize("precision,specialize,shape", [
    (0, False, None), (1, False, None), (0, True, None), (1, True, None), (1, True, (0, 7, 0, 1)),
])
def test_blocks_equal_rows(native, precision, specialize, shape):
    """Lotka-Volterra tau-leaping with SSA bursts, 37 samples (a partial last
    block) and a step budget that ends simulations early (the rest of their
    samples written at once), more simulations than lanes."""
    U, ita = 70000, 37
    a = models.make_args(models.lotka_volterra(), E=[1, 0], t_max=10.0, ita=ita, U=U)
    tabs, rates = tables_of(a), a.rate_fn(a.c)
    kw = dict(method=0, t_max=10.0, I=a.I, E=a.E, n_sims=U, seed=3, precision=precision, specialize=specialize,
              max_steps=3000)
    print("case: layout, precision %d, specialize %s, shape %s" % (precision, specialize, shape), flush=True)
    res = _plan_pair(native, lambda lay: native.Plan(tabs, rates, a.x_0, layout=lay, **kw), 2, ita, U, shape)
    _check(*res, ita)


def test_blocks_equal_rows_large_network(native):
    """The large-network kernel (one simulation per wavefront) stores the
    blocked layout sample by sample."""
    rng = np.random.default_rng(11)
    spn = models.random_network(rng, 30, 120)
    U, ita = 512, 21
    a = models.make_args(spn, E=list(range(30)), t_max=0.5, ita=ita, U=U)
    tabs, rates = tables_of(a), a.rate_fn(a.c)
    kw = dict(method=0, t_max=0.5, I=a.I, E=a.E, n_sims=U, seed=5, precision=1, max_steps=5000)
    print("case: layout, large network", flush=True)
    res = _plan_pair(native, lambda lay: native.Plan(tabs, rates, a.x_0, layout=lay, **kw), 30, ita, U)
    _check(*res, ita)


def test_noalloc_blocks_equal_rows(native):
    """pst_run_noalloc honours the layout too (table-driven kernel)."""
    import ctypes
    U, ita = 4096, 13
    a = models.make_args(models.schlogl(), E=[0], t_max=5.0, ita=ita, U=U)
    tabs, rates = tables_of(a), a.rate_fn(a.c)
    hip = ctypes.CDLL("libamdhip64.so.7")

    def dev(arr):
        p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(arr.nbytes)) == 0
        assert hip.hipMemcpy(p, arr.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(arr.nbytes), 1) == 0
        return p
    r_d = dev(np.ascontiguousarray(np.atleast_2d(rates), np.float64))
    x_d = dev(np.ascontiguousarray(np.atleast_2d(a.x_0), np.int32))
    i_d = dev(np.ascontiguousarray(a.I, np.float32))
    ita8 = (ita + 7) // 8
    R, B = DeviceArray((1, ita, U)), DeviceArray((1, ita8, U, 8))
    try:
        kw = dict(method=0, precision=1, t_max=5.0, ita=ita, E=a.E, n_sims=U, seed=8)
        print("case: layout, noalloc", flush=True)
        native.run_noalloc(tabs, r_d.value, x_d.value, 1, i_d.value, R.ptr.value, U, layout=native.LAYOUT_ROWS, **kw)
        native.run_noalloc(tabs, r_d.value, x_d.value, 1, i_d.value, B.ptr.value, U, layout=native.LAYOUT_BLOCKS, **kw)
        rows, blocks = R.get(), B.get()
    finally:
        R.free()
        B.free()
        for p in (r_d, x_d, i_d):
            hip.hipFree(p)
    host = blocks.transpose(0, 1, 3, 2).reshape(1, ita8 * 8, U)[:, :ita]
    assert np.array_equal(rows, host)
    assert not blocks[:, -1, :, ita - 8 * (ita8 - 1):].any()


def test_batch_blocks_equal_rows(native):
    """A batch plan with a network-kernel job and a large-network job."""
    lv = workloads.lotka_volterra(U=5000)[0]
    lv.kw["I"], lv.kw["t_max"] = lv.kw["I"][:45], float(lv.kw["I"][44])
    rng = np.random.default_rng(12)
    big = models.make_args(models.random_network(rng, 24, 70), E=[0, 3, 5], t_max=0.2, ita=19, U=300)
    jobs_kw = [
        (tables_of(models.make_args(models.lotka_volterra(), E=[0, 1], t_max=1.0, ita=45, U=5000)), lv.rates,
         lv.x0, dict(lv.kw), 2, lv.ita, lv.n_sims, True),
        (tables_of(big), big.rate_fn(big.c), big.x_0,
         dict(method=0, t_max=0.2, I=big.I, E=big.E, n_sims=300, seed=6, max_steps=4000), 3, 19, 300, False),
    ]
    out = {}
    for layout in (0, 1):
        jobs = [native.Job(t, r, x, precision=1, specialize=sp, layout=layout, **kw)
                for t, r, x, kw, _, _, _, sp in jobs_kw]
        p = native.BatchPlan(jobs)
        bufs = [DeviceArray((k, ita, U) if layout == 0 else (k, (ita + 7) // 8, U, 8))
                for _, _, _, _, k, ita, U, _ in jobs_kw]
        try:
            print("case: layout %d, batch" % layout, flush=True)
            p.execute([b.ptr.value for b in bufs], [j[6] for j in jobs_kw])
            p.stats()
            out[layout] = [b.get() for b in bufs]
        finally:
            for b in bufs:
                b.free()
        p.close()
    for (_, _, _, _, k, ita, U, _), rows, blocks in zip(jobs_kw, out[0], out[1]):
        ita8 = (ita + 7) // 8
        host = blocks.transpose(0, 1, 3, 2).reshape(k, ita8 * 8, U)[:, :ita]
        assert np.array_equal(rows, host)
