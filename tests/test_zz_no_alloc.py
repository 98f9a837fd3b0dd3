"""pst_run_noalloc, the entry point that allocates no device memory, against
the existing device-resident entry point (pst_plan_create + pst_plan_execute,
table-driven kernel, same settings): identical trajectories, flags and
counters, on consecutive calls that reuse the library's one device array --
each later call with a smaller input, across networks, methods and both
precisions.  The caller's buffers come from hipMalloc of the HIP runtime the
library itself uses (every test marked gpu)."""
import ctypes

import numpy as np
import pytest

from py_stoch_amd import models
from py_stoch_amd.mod.parser.tl_parser import tables_of

pytestmark = pytest.mark.gpu

COUNTERS = ("events", "leaps", "rejected_leaps", "ssa_steps", "flagged", "uniforms", "ptrs_iterations",
            "inversion_steps", "leap_decisions", "sample_rows", "flagged_events")


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libpst_hip.so links."""

    def __init__(self, native):
        native.load()
        self.L = ctypes.CDLL("libamdhip64.so.7")  # already loaded by the library: the same runtime
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipDeviceSynchronize.argtypes = []
        self.live = []

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), max(arr.nbytes, 4)) == 0
        self.live.append(p.value)
        assert self.L.hipMemcpy(p.value, arr.ctypes.data, arr.nbytes, 1) == 0  # hipMemcpyHostToDevice
        return p.value

    def get(self, ptr, like):
        out = np.empty_like(like)
        assert self.L.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # hipMemcpyDeviceToHost
        return out

    def free(self):
        assert self.L.hipDeviceSynchronize() == 0
        for p in self.live:
            self.L.hipFree(p)
        self.live = []


def _one(native, hip, spn, E, t_max, ita, U, method, precision, seed):
    a = models.make_args(spn, E=E, t_max=t_max, ita=ita, U=U)
    tabs, rates, x0 = tables_of(a), np.atleast_2d(a.rate_fn(a.c)), np.atleast_2d(a.x_0)
    kw = dict(method=method, precision=precision, t_max=t_max, E=a.E, n_sims=U, seed=seed)
    print("case: noalloc N=%d M=%d U=%d method=%d precision=%d" % (a.N, a.M, U, method, precision), flush=True)
    marker_O = np.full((len(E), ita, U), -7, np.int32)
    marker_F = np.full((U,), -7, np.int32)
    # the existing entry point
    O_ref_d, F_ref_d = hip.put(marker_O), hip.put(marker_F)
    plan = native.Plan(tabs, rates, x0, I=a.I, specialize=False, **kw)
    plan.execute(O_ref_d, U, F_ref_d)
    st_ref = plan.stats()
    plan.close()
    O_ref, F_ref = hip.get(O_ref_d, marker_O), hip.get(F_ref_d, marker_F)
    # every input in device memory
    r_d = hip.put(rates.astype(np.float64))
    x_d = hip.put(x0.astype(np.int32))
    I_d = hip.put(np.asarray(a.I, np.float32))
    O_d, F_d = hip.put(marker_O), hip.put(marker_F)
    st = native.run_noalloc(tabs, r_d, x_d, 1, I_d, O_d, U, F_d, None, ita=ita, **kw)
    O, F = hip.get(O_d, marker_O), hip.get(F_d, marker_F)
    assert (O_ref != -7).all()
    assert np.array_equal(O, O_ref), "trajectories differ at %s" % np.argwhere(O != O_ref)[:3].tolist()
    assert np.array_equal(F, F_ref)
    for k in COUNTERS:
        assert st[k] == st_ref[k], (k, st[k], st_ref[k])
    hip.free()
    return st


def test_noalloc_equals_plan_on_consecutive_calls(native):
    hip = _Hip(native)
    rng = np.random.default_rng(3)
    small = models.random_network(rng, 6, 7)  # an 8 x 8 size-class network
    seq = [
        (models.lotka_volterra(), [0, 1], 2.0, 101, 4096, native.METHOD_TAU_LEAPING, native.PRECISION_FP64, 2),
        (models.birth_death(), [0], 20.0, 21, 1000, native.METHOD_PURE_SSA, native.PRECISION_FP64, 0),
        (models.lotka_volterra(), [0, 1], 2.0, 51, 3000, native.METHOD_TAU_LEAPING, native.PRECISION_MIXED, 4),
        (models.schlogl(), [0], 2.0, 21, 700, native.METHOD_TAU_LEAPING, native.PRECISION_MIXED, 3),
        (small, list(range(6)), 0.5, 11, 300, native.METHOD_TAU_LEAPING, native.PRECISION_FP64, 5),
        (models.birth_death(death=1.0, birth=0.0, x0=50), [0], 5.0, 6, 100, native.METHOD_TAU_LEAPING,
         native.PRECISION_FP64, 6),
    ]
    for spn, E, t_max, ita, U, method, precision, seed in seq:
        st = _one(native, hip, spn, E, t_max, ita, U, method, precision, seed)
        assert st["events"] > 0
