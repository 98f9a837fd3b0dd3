"""The C ABI library (include/pst.h) without a GPU: it loads, exports every
function the header declares, and rejects malformed input with a status code
and a message before touching the device."""
import ctypes
import os
import re

import numpy as np
import pytest

from py_stoch_amd import _native, models
from py_stoch_amd.mod.parser.tl_parser import tables_of

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "pst.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pst_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_function(native):
    names = _declared_functions()
    assert len(names) >= 9
    lib = native.load()
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_native.EXPORTED_SYMBOLS)


def test_abi_version_and_limits(native):
    assert native.load().pst_abi_version() == 7 == native.ABI_VERSION
    assert native.max_network() == (64, 256)


def test_large_networks_need_mixed_precision(native):
    """Beyond the FP64 size classes (16 x 16, 16 x 32 compiled) networks run
    on the large-network kernel, which is mixed precision only; beyond 64
    species or 256 reactions nothing runs.  Both are refused on the host."""
    rng = np.random.default_rng(1)
    a = models.make_args(models.random_network(rng, n_species=20, n_reactions=30), E=[0], t_max=1.0, ita=3, U=64)
    with pytest.raises(native.NativeError, match="larger networks run in mixed precision"):
        native.Plan(tables_of(a), a.rate_fn(a.c), a.x_0, precision=native.PRECISION_FP64, t_max=1.0, I=a.I,
                    E=a.E, n_sims=64)
    b = models.make_args(models.random_network(rng, n_species=70, n_reactions=12), E=[0], t_max=1.0, ita=3, U=64)
    with pytest.raises(native.NativeError, match="N<=64"):
        native.Plan(tables_of(b), b.rate_fn(b.c), b.x_0, precision=native.PRECISION_MIXED, t_max=1.0, I=b.I,
                    E=b.E, n_sims=64)


def _args():
    a = models.make_args(models.lotka_volterra(), E=[0, 1], t_max=1.0, ita=11, U=64)
    return a, tables_of(a), a.rate_fn(a.c)


@pytest.mark.parametrize("mutate, msg", [
    (lambda t, kw: kw.update(E=[0, 5]), "recorded species"),
    (lambda t, kw: kw.update(n_c=-1), "n_c"),
    (lambda t, kw: kw.update(eta=0.0), "eta"),
    (lambda t, kw: kw.update(ssa_burst=0), "ssa_burst"),
    (lambda t, kw: kw.update(max_steps=-1), "max_steps"),
    (lambda t, kw: kw.update(precision=7), "precision"),
    (lambda t, kw: t.update(V=np.array([[0, 0, 1, 0], [0, 0, -1, 0]])), "two entries"),
    (lambda t, kw: t.update(V=np.array([[1, 1, 1, 0], [0, 0, -1, 0]])), "species-major"),
    (lambda t, kw: t.update(V=np.array([[0, 9, 1, 0]])), "out of range"),
    (lambda t, kw: t.update(prop_step=np.full((2, 3), 2)), "prop_step"),
])
def test_invalid_input_is_rejected_with_message(native, mutate, msg):
    a, tabs, rates = _args()
    kw = dict(method=0, t_max=1.0, I=a.I, E=a.E, n_sims=64, seed=1)
    mutate(tabs, kw)
    with pytest.raises(native.NativeError, match=msg):
        native.simulate(tabs, rates, a.x0 if hasattr(a, "x0") else a.x_0, **kw)


def test_negative_rate_and_count_rejected(native):
    a, tabs, rates = _args()
    with pytest.raises(native.NativeError, match="rate"):
        native.simulate(tabs, -rates, a.x_0, t_max=1.0, I=a.I, E=a.E, n_sims=8)
    with pytest.raises(native.NativeError, match="initial count"):
        native.simulate(tabs, rates, -a.x_0, t_max=1.0, I=a.I, E=a.E, n_sims=8)


def test_pure_ssa_entry_rejected_on_tau_leaping_call(native):
    a, tabs, rates = _args()
    with pytest.raises(native.NativeError, match="pst_gillespie_run"):
        lib = native.load()
        m = _native._Marshalled(tabs, rates, a.x_0, 2, 0, 10, 0.03, 100, 1.0, a.I, a.E, 8, None, 0, 0)
        O = np.zeros((2, 11, 8), np.int32)
        rc = lib.pst_tau_leaping_run(ctypes.byref(m.net), ctypes.byref(m.cfg),
                                     m.rates.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                     m.x0.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), 1,
                                     O.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), None, None)
        native._check(rc, "pst_tau_leaping_run")


def test_autotune_argument_checks(native):
    """pst_plan_autotune rejects a null plan or result and a fraction outside
    (0, 1] before touching a device."""
    L = native.load()
    out = native.PstLaunch()
    assert L.pst_plan_autotune(None, None, None, None, None, 0.5, ctypes.byref(out)) == -1  # PST_EINVAL
    assert b"null" in L.pst_last_error()


def test_set_launch_argument_checks(native):
    L = native.load()
    assert L.pst_plan_set_launch(None, 0, 7, 0, 0) == -1  # PST_EINVAL: null plan


def test_range_and_graph_argument_checks(native):
    """The sample-range and graph entry points reject a null plan before
    touching a device."""
    import ctypes
    L = native.load()
    pos = ctypes.c_int32()
    assert L.pst_plan_execute_range(None, None, None, None, 8, None) == -1
    assert L.pst_plan_range_position(None, ctypes.byref(pos)) == -1
    assert L.pst_plan_graph_create(None, None, None, None, None, 0) == -1
    assert L.pst_plan_graph_launch(None, None) == -1
    assert L.pst_blocks_to_rows(None, 0, None, 0, 1, 8, 4, None) == -1  # strides below n_sims
    assert b"strides" in L.pst_last_error()


def test_noalloc_argument_checks(native):
    """pst_run_noalloc runs the table-driven kernels only and rejects what
    they cannot take before touching a device."""
    def err(tables, **kw):
        a = dict(ita=11, n_sims=64, t_max=1.0, E=[0])
        a.update(kw)
        with pytest.raises(native.NativeError) as e:
            native.run_noalloc(tables, 1, 1, 1, 1, 1, a["n_sims"], **a)
        return str(e.value)
    a = models.make_args(models.lotka_volterra(), E=[0, 1], t_max=1.0, ita=11, U=64)
    tabs = tables_of(a)
    big = tables_of(models.make_args(models.random_network(np.random.default_rng(1), 20, 30), E=[0]))
    assert "beyond the largest table-driven size class" in err(big)
    hz = dict(tabs)
    hz["hazard_c"], hz["n_params"] = ["C(0)*X(0)", "C(1)*X(0)*X(1)", "C(2)*X(1)"], 3
    assert "compiled hazard" in err(hz)
    L = native.load()
    assert L.pst_run_noalloc(None, None, None, None, 1, None, None, 0, None, None, None) == -1
    assert b"null" in L.pst_last_error()
