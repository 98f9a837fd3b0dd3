"""ctypes binding of libpst_hip.so (include/pst.h).

The product path has no CPU fallback: if the HIP library is missing or the
device is not a gfx950, every call raises.  Build the library with
``python -c "import __graft_entry__ as g; g.build()"`` or ``make -C
py_stoch_amd/csrc``.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libpst_hip.so")
ABI_VERSION = 7  # PST_ABI_VERSION of include/pst.h

METHOD_TAU_LEAPING = 0
METHOD_FORCED_SSA = 1
METHOD_PURE_SSA = 2

PRECISION_FP64 = 0
PRECISION_MIXED = 1
LAYOUT_ROWS = 0      # O[kappa][ita][o_stride]
LAYOUT_BLOCKS = 1    # O[kappa][ceil(ita/8)][o_stride][8] (device arrays of plans and run_noalloc)

FLAG_OVERFLOW = 1
FLAG_HALVING = 2
FLAG_TAU_INF = 4
FLAG_PTRS = 8
FLAG_STEPS = 16
FLAG_STALL = 32

_P_I32 = ctypes.POINTER(ctypes.c_int32)
_P_F64 = ctypes.POINTER(ctypes.c_double)
_P_F32 = ctypes.POINTER(ctypes.c_float)


class PstNetwork(ctypes.Structure):
    _fields_ = [
        ("n_species", ctypes.c_int32),
        ("n_reactions", ctypes.c_int32),
        ("a_size", ctypes.c_int32),
        ("A", _P_I32),
        ("v_size", ctypes.c_int32),
        ("V", _P_I32),
        ("H", _P_I32),
        ("H_type", _P_I32),
        ("prop_order", _P_I32),
        ("prop_step", _P_I32),
        ("hazards", ctypes.POINTER(ctypes.c_char_p)),
        ("n_params", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class PstConfig(ctypes.Structure):
    _fields_ = [
        ("method", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("n_c", ctypes.c_int32),
        ("ssa_burst", ctypes.c_int32),
        ("eta", ctypes.c_double),
        ("t_max", ctypes.c_double),
        ("ita", ctypes.c_int32),
        ("kappa", ctypes.c_int32),
        ("I", _P_F32),
        ("E", _P_I32),
        ("n_sims", ctypes.c_int32),
        ("sims_per_set", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("sim_offset", ctypes.c_uint64),
        ("specialize", ctypes.c_int32),
        ("layout", ctypes.c_int32),
        ("max_steps", ctypes.c_int64),
    ]


class PstStats(ctypes.Structure):
    _fields_ = [
        ("events", ctypes.c_int64),
        ("leaps", ctypes.c_int64),
        ("rejected_leaps", ctypes.c_int64),
        ("ssa_steps", ctypes.c_int64),
        ("flagged", ctypes.c_int64),
        ("uniforms", ctypes.c_int64),
        ("ptrs_iterations", ctypes.c_int64),
        ("inversion_steps", ctypes.c_int64),
        ("leap_decisions", ctypes.c_int64),
        ("sample_rows", ctypes.c_int64),
        ("flagged_events", ctypes.c_int64),
        ("asym_draws", ctypes.c_int64),
        ("exact_draws", ctypes.c_int64),
        ("kernel_ms", ctypes.c_double),
    ]

    def as_dict(self):
        d = {name: int(getattr(self, name)) for name, _ in self._fields_ if name != "kernel_ms"}
        d["kernel_ms"] = float(self.kernel_ms)
        return d


class PstLaunch(ctypes.Structure):
    _fields_ = [
        ("lanes_shift", ctypes.c_int32),
        ("ssa_slice", ctypes.c_int32),
        ("waves_trim", ctypes.c_int32),
        ("burst_waves", ctypes.c_int32),
        ("kernel_ms", ctypes.c_double),
        ("default_ms", ctypes.c_double),
    ]

    def as_dict(self):
        return {"lanes_shift": int(self.lanes_shift), "ssa_slice": int(self.ssa_slice),
                "waves_trim": int(self.waves_trim), "burst_waves": int(self.burst_waves),
                "kernel_ms": float(self.kernel_ms), "default_ms": float(self.default_ms)}


class PstJob(ctypes.Structure):
    _fields_ = [
        ("net", PstNetwork),
        ("cfg", PstConfig),
        ("rates", _P_F64),
        ("x0", _P_I32),
        ("n_sets", ctypes.c_int32),
    ]


EXPORTED_SYMBOLS = (
    "pst_poisson_probe_mixed",
    "pst_run_noalloc",
    "pst_tau_leaping_run",
    "pst_gillespie_run",
    "pst_batch_run",
    "pst_batch_plan_create",
    "pst_batch_plan_execute",
    "pst_plan_create",
    "pst_plan_execute",
    "pst_plan_read_stats",
    "pst_plan_destroy",
    "pst_plan_set_stream",
    "pst_max_network",
    "pst_last_error",
    "pst_abi_version",
    "pst_plan_autotune",
    "pst_plan_set_launch",
    "pst_blocks_to_rows",
    "pst_plan_execute_range",
    "pst_plan_range_position",
    "pst_plan_graph_create",
    "pst_plan_graph_launch",
)

_lib = None


class NativeError(RuntimeError):
    """A libpst_hip call returned an error status."""


def load():
    """Load libpst_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeError(
            "libpst_hip.so is not built (%s); run __graft_entry__.build() or "
            "make -C py_stoch_amd/csrc" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    run_args = [ctypes.POINTER(PstNetwork), ctypes.POINTER(PstConfig), _P_F64, _P_I32,
                ctypes.c_int32, _P_I32, _P_I32, ctypes.POINTER(PstStats)]
    L.pst_tau_leaping_run.argtypes = run_args
    L.pst_tau_leaping_run.restype = ctypes.c_int
    L.pst_gillespie_run.argtypes = run_args
    L.pst_gillespie_run.restype = ctypes.c_int
    L.pst_batch_run.argtypes = [ctypes.POINTER(PstJob), ctypes.c_int32, ctypes.POINTER(_P_I32),
                                ctypes.POINTER(_P_I32), ctypes.POINTER(PstStats)]
    L.pst_batch_run.restype = ctypes.c_int
    L.pst_batch_plan_create.argtypes = [ctypes.POINTER(PstJob), ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]
    L.pst_batch_plan_create.restype = ctypes.c_int
    L.pst_batch_plan_execute.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                         ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_void_p),
                                         ctypes.c_void_p]
    L.pst_batch_plan_execute.restype = ctypes.c_int
    L.pst_plan_autotune.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                    ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_void_p),
                                    ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(PstLaunch)]
    L.pst_plan_autotune.restype = ctypes.c_int
    L.pst_plan_set_launch.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                      ctypes.c_int32]
    L.pst_plan_set_launch.restype = ctypes.c_int
    L.pst_blocks_to_rows.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                                     ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p]
    L.pst_blocks_to_rows.restype = ctypes.c_int
    _P_VP = ctypes.POINTER(ctypes.c_void_p)
    L.pst_plan_execute_range.argtypes = [ctypes.c_void_p, _P_VP, ctypes.POINTER(ctypes.c_int64), _P_VP,
                                         ctypes.c_int32, ctypes.c_void_p]
    L.pst_plan_execute_range.restype = ctypes.c_int
    L.pst_plan_range_position.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]
    L.pst_plan_range_position.restype = ctypes.c_int
    L.pst_plan_graph_create.argtypes = [ctypes.c_void_p, _P_VP, ctypes.POINTER(ctypes.c_int64), _P_VP,
                                        ctypes.POINTER(ctypes.c_int32), ctypes.c_int32]
    L.pst_plan_graph_create.restype = ctypes.c_int
    L.pst_plan_graph_launch.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.pst_plan_graph_launch.restype = ctypes.c_int
    L.pst_plan_create.argtypes = [ctypes.POINTER(PstNetwork), ctypes.POINTER(PstConfig), _P_F64,
                                  _P_I32, ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]
    L.pst_plan_create.restype = ctypes.c_int
    L.pst_plan_execute.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                   ctypes.c_void_p, ctypes.c_void_p]
    L.pst_plan_execute.restype = ctypes.c_int
    L.pst_plan_read_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(PstStats)]
    L.pst_plan_read_stats.restype = ctypes.c_int
    L.pst_plan_set_stream.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64]
    L.pst_plan_set_stream.restype = ctypes.c_int
    L.pst_plan_destroy.argtypes = [ctypes.c_void_p]
    L.pst_plan_destroy.restype = None
    L.pst_run_noalloc.argtypes = [ctypes.POINTER(PstNetwork), ctypes.POINTER(PstConfig), ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(PstStats)]
    L.pst_run_noalloc.restype = ctypes.c_int
    L.pst_poisson_probe_mixed.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_uint32),
                                          ctypes.c_int64, _P_I32, _P_I32]
    L.pst_poisson_probe_mixed.restype = ctypes.c_int
    L.pst_max_network.argtypes = [_P_I32, _P_I32]
    L.pst_max_network.restype = None
    L.pst_last_error.argtypes = []
    L.pst_last_error.restype = ctypes.c_char_p
    L.pst_abi_version.argtypes = []
    L.pst_abi_version.restype = ctypes.c_int
    if L.pst_abi_version() != ABI_VERSION:
        raise NativeError("%s implements C ABI %d, this binding expects %d: rebuild it" % (
            LIB_PATH, L.pst_abi_version(), ABI_VERSION))
    _lib = L
    return L


def _check(rc, what):
    if rc != 0:
        msg = load().pst_last_error().decode(errors="replace")
        raise NativeError("%s failed (status %d): %s" % (what, rc, msg))


def _ptr(arr, ctype):
    return arr.ctypes.data_as(ctypes.POINTER(ctype))


class _Marshalled:
    """Keeps the numpy arrays behind a PstNetwork/PstConfig alive."""

    def __init__(self, tables, rates, x0, method, precision, n_c, eta, ssa_burst, t_max, I, E,
                 n_sims, sims_per_set, seed, sim_offset, specialize=False, max_steps=0, layout=0):
        N, M = int(tables["N"]), int(tables["M"])
        self.A = np.ascontiguousarray(tables["A"], np.int32).reshape(-1, 4)
        self.V = np.ascontiguousarray(tables["V"], np.int32).reshape(-1, 4)
        self.H = np.ascontiguousarray(tables["H"], np.int32).reshape(N)
        self.Ht = np.ascontiguousarray(tables["H_type"], np.int32).reshape(N)
        self.po = np.ascontiguousarray(tables["prop_order"], np.int32).reshape(N, M)
        self.ps = np.ascontiguousarray(tables["prop_step"], np.int32).reshape(N, M)
        hz = tables.get("hazard_c")
        n_rates = int(tables.get("n_params", 0)) if hz else M
        self.rates = np.ascontiguousarray(np.atleast_2d(np.asarray(rates, np.float64)).reshape(-1, max(n_rates, 1)))
        self.x0 = np.ascontiguousarray(np.atleast_2d(np.asarray(x0, np.int32)).reshape(-1, N))
        self.I = np.ascontiguousarray(np.asarray(I, np.float32).reshape(-1))
        self.E = np.ascontiguousarray(np.asarray(E, np.int32).reshape(-1))
        self.n_sets = self.rates.shape[0]
        if self.x0.shape[0] != self.n_sets:
            raise ValueError("rates has %d rows but x0 has %d" % (self.n_sets, self.x0.shape[0]))
        if sims_per_set is None:
            sims_per_set = max(1, int(n_sims)) if self.n_sets == 1 else None
        if sims_per_set is None:
            raise ValueError("sims_per_set is required with several parameter rows")
        self.hazards = None
        if hz:
            self._hz_bytes = [h.encode() for h in hz]
            self.hazards = (ctypes.c_char_p * len(hz))(*self._hz_bytes)
        self.net = PstNetwork(N, M, self.A.shape[0], _ptr(self.A, ctypes.c_int32), self.V.shape[0],
                              _ptr(self.V, ctypes.c_int32), _ptr(self.H, ctypes.c_int32),
                              _ptr(self.Ht, ctypes.c_int32), _ptr(self.po, ctypes.c_int32),
                              _ptr(self.ps, ctypes.c_int32),
                              ctypes.cast(self.hazards, ctypes.POINTER(ctypes.c_char_p)) if hz else None,
                              n_rates if hz else 0, 0)
        self.cfg = PstConfig(int(method), int(precision), int(n_c), int(ssa_burst), float(eta),
                             float(t_max), self.I.shape[0], self.E.shape[0],
                             _ptr(self.I, ctypes.c_float), _ptr(self.E, ctypes.c_int32),
                             int(n_sims), int(sims_per_set), int(seed), int(sim_offset),
                             1 if specialize else 0, int(layout), int(max_steps))


def simulate(tables, rates, x0, method=METHOD_TAU_LEAPING, precision=PRECISION_FP64, n_c=10,
             eta=0.03, ssa_burst=100, t_max=1.0, I=None, E=(0,), n_sims=1, sims_per_set=None,
             seed=0, sim_offset=0, return_flags=True, specialize=False, max_steps=0):
    """Run simulations on the GPU; returns (O[kappa][ita][n_sims] int32, flags, stats).

    specialize=True compiles (once per network and settings, cached) a kernel
    with the network built in instead of using the table-driven kernel."""
    L = load()
    m = _Marshalled(tables, rates, x0, method, precision, n_c, eta, ssa_burst, t_max, I, E,
                    n_sims, sims_per_set, seed, sim_offset, specialize, max_steps)
    O = np.empty((m.E.shape[0], m.I.shape[0], int(n_sims)), np.int32)
    flags = np.zeros(max(int(n_sims), 1), np.int32) if return_flags else None
    st = PstStats()
    fn = L.pst_gillespie_run if method == METHOD_PURE_SSA else L.pst_tau_leaping_run
    rc = fn(ctypes.byref(m.net), ctypes.byref(m.cfg), _ptr(m.rates, ctypes.c_double),
            _ptr(m.x0, ctypes.c_int32), m.n_sets, _ptr(O, ctypes.c_int32),
            _ptr(flags, ctypes.c_int32) if flags is not None else None, ctypes.byref(st))
    _check(rc, fn.__name__)
    return O, (flags[: int(n_sims)] if flags is not None else None), st.as_dict()


def run_noalloc(tables, rates_dev_ptr, x0_dev_ptr, n_sets, I_dev_ptr, O_dev_ptr, o_stride, flags_dev_ptr=None,
                stream_ptr=None, read_stats=True, method=METHOD_TAU_LEAPING, precision=PRECISION_FP64, n_c=10,
                eta=0.03, ssa_burst=100, t_max=1.0, ita=None, E=(0,), n_sims=1, sims_per_set=None, seed=0,
                sim_offset=0, max_steps=0, layout=LAYOUT_ROWS):
    """pst_run_noalloc: one run on the table-driven kernel with every input
    and output in caller device memory (rates [n_sets][M] float64, x0
    [n_sets][N] int32, sample times [ita] float32, O [kappa][ita][o_stride]
    int32 (or blocks, layout=LAYOUT_BLOCKS), flags) and no device allocation
    by the library.  Returns the counters (waiting for the run) if
    read_stats, else None (enqueued only)."""
    L = load()
    N, M = int(tables["N"]), int(tables["M"])
    if sims_per_set is None:
        if int(n_sets) != 1:
            raise ValueError("sims_per_set is required with several parameter rows")
        sims_per_set = max(1, int(n_sims))
    m = _Marshalled(tables, np.zeros((1, M)), np.zeros((1, N), np.int32), method, precision, n_c, eta, ssa_burst,
                    t_max, np.zeros(int(ita), np.float32), E, n_sims, sims_per_set, seed, sim_offset, False,
                    max_steps, layout)
    st = PstStats() if read_stats else None
    rc = L.pst_run_noalloc(ctypes.byref(m.net), ctypes.byref(m.cfg), rates_dev_ptr, x0_dev_ptr, int(n_sets),
                           I_dev_ptr, O_dev_ptr, int(o_stride), flags_dev_ptr, stream_ptr,
                           ctypes.byref(st) if st is not None else None)
    _check(rc, "pst_run_noalloc")
    return st.as_dict() if st is not None else None


def blocks_to_rows(blocks_dev_ptr, blocks_stride, rows_dev_ptr, rows_stride, kappa, ita, n_sims, stream_ptr=None):
    """pst_blocks_to_rows: device trajectories from LAYOUT_BLOCKS
    [kappa][ceil(ita/8)][blocks_stride][8] to LAYOUT_ROWS
    [kappa][ita][rows_stride] (enqueued on the stream)."""
    _check(load().pst_blocks_to_rows(ctypes.c_void_p(int(blocks_dev_ptr)), int(blocks_stride),
                                     ctypes.c_void_p(int(rows_dev_ptr)), int(rows_stride), int(kappa), int(ita),
                                     int(n_sims), ctypes.c_void_p(int(stream_ptr)) if stream_ptr else None),
           "pst_blocks_to_rows")


def poisson_probe_mixed(lam, words):
    """The mixed-precision kernels' Poisson sampler on given means (float32)
    and Philox words (uint32): returns (counts, branch per draw) -- see
    pst_poisson_probe_mixed in include/pst.h."""
    lam = np.ascontiguousarray(lam, np.float32).reshape(-1)
    words = np.ascontiguousarray(words, np.uint32).reshape(-1)
    if lam.shape != words.shape:
        raise ValueError("lam and words differ in length")
    k = np.empty(lam.shape, np.int32)
    path = np.empty(lam.shape, np.int32)
    rc = load().pst_poisson_probe_mixed(_ptr(lam, ctypes.c_float), _ptr(words, ctypes.c_uint32), lam.size,
                                        _ptr(k, ctypes.c_int32), _ptr(path, ctypes.c_int32))
    _check(rc, "pst_poisson_probe_mixed")
    return k, path


def _ptr_arrays(O_ptrs, strides, flag_ptrs):
    n = len(O_ptrs)
    Op = (ctypes.c_void_p * n)(*[int(p) for p in O_ptrs])
    Sp = (ctypes.c_int64 * n)(*[int(s) for s in strides])
    Fp = (ctypes.c_void_p * n)(*[int(p) if p else None for p in flag_ptrs]) if flag_ptrs else None
    return Op, Sp, Fp


class _RangeMixin:
    """Sample ranges and graph-captured chunks (pst_plan_execute_range,
    pst_plan_graph_*); O_ptrs / strides / flag_ptrs as for a batch execute
    (one entry per job)."""

    def execute_range(self, O_ptrs, strides, f_end, flag_ptrs=None, stream_ptr=None):
        Op, Sp, Fp = _ptr_arrays(O_ptrs, strides, flag_ptrs)
        _check(load().pst_plan_execute_range(self._h, Op, Sp, Fp, int(f_end),
                                             ctypes.c_void_p(int(stream_ptr)) if stream_ptr else None),
               "pst_plan_execute_range")

    def range_position(self):
        out = ctypes.c_int32()
        _check(load().pst_plan_range_position(self._h, ctypes.byref(out)), "pst_plan_range_position")
        return int(out.value)

    def graph_create(self, O_ptrs, strides, f_ends, flag_ptrs=None):
        Op, Sp, Fp = _ptr_arrays(O_ptrs, strides, flag_ptrs)
        ends = (ctypes.c_int32 * len(f_ends))(*[int(f) for f in f_ends])
        _check(load().pst_plan_graph_create(self._h, Op, Sp, Fp, ends, len(f_ends)), "pst_plan_graph_create")

    def graph_launch(self, stream_ptr=None):
        _check(load().pst_plan_graph_launch(self._h, ctypes.c_void_p(int(stream_ptr)) if stream_ptr else None),
               "pst_plan_graph_launch")


def _autotune(h, O_ptrs, strides, flag_ptrs, stream_ptr, fraction):
    n = len(O_ptrs)
    Op = (ctypes.c_void_p * n)(*[int(p) for p in O_ptrs])
    Sp = (ctypes.c_int64 * n)(*[int(s) for s in strides])
    Fp = (ctypes.c_void_p * n)(*[int(p) if p else None for p in flag_ptrs]) if flag_ptrs else None
    out = PstLaunch()
    _check(load().pst_plan_autotune(h, Op, Sp, Fp, ctypes.c_void_p(int(stream_ptr)) if stream_ptr else None,
                                    float(fraction), ctypes.byref(out)), "pst_plan_autotune")
    return out.as_dict()


class Plan(_RangeMixin):
    """Device-resident run: tables uploaded once, trajectories left in device memory."""

    def __init__(self, tables, rates, x0, method=METHOD_TAU_LEAPING, precision=PRECISION_FP64,
                 n_c=10, eta=0.03, ssa_burst=100, t_max=1.0, I=None, E=(0,), n_sims=1,
                 sims_per_set=None, seed=0, sim_offset=0, specialize=False, max_steps=0, layout=LAYOUT_ROWS):
        L = load()
        self._m = _Marshalled(tables, rates, x0, method, precision, n_c, eta, ssa_burst, t_max, I,
                              E, n_sims, sims_per_set, seed, sim_offset, specialize, max_steps, layout)
        self.layout = int(layout)
        self.n_sims = int(n_sims)
        self.kappa = self._m.E.shape[0]
        self.ita = self._m.I.shape[0]
        h = ctypes.c_void_p()
        rc = L.pst_plan_create(ctypes.byref(self._m.net), ctypes.byref(self._m.cfg),
                               _ptr(self._m.rates, ctypes.c_double), _ptr(self._m.x0, ctypes.c_int32),
                               self._m.n_sets, ctypes.byref(h))
        _check(rc, "pst_plan_create")
        self._h = h

    def execute(self, O_dev_ptr, o_stride, flags_dev_ptr=None, stream_ptr=None):
        """Enqueue one run; O_dev_ptr / flags_dev_ptr are device addresses (int)."""
        rc = load().pst_plan_execute(self._h, ctypes.c_void_p(int(O_dev_ptr)), int(o_stride),
                                     ctypes.c_void_p(int(flags_dev_ptr)) if flags_dev_ptr else None,
                                     ctypes.c_void_p(int(stream_ptr)) if stream_ptr else None)
        _check(rc, "pst_plan_execute")

    def set_stream(self, seed, sim_offset=0):
        """Re-key the random stream of the following executions."""
        _check(load().pst_plan_set_stream(self._h, int(seed), int(sim_offset)), "pst_plan_set_stream")

    def set_launch(self, lanes_shift, ssa_slice, waves_trim=0, burst_waves=0):
        """Use this launch shape (e.g. one an autotune returned) for later executions."""
        _check(load().pst_plan_set_launch(self._h, int(lanes_shift), int(ssa_slice), int(waves_trim),
                                          int(burst_waves)),
               "pst_plan_set_launch")

    def autotune(self, O_dev_ptr, o_stride, flags_dev_ptr=None, stream_ptr=None, fraction=1.0):
        """Time the candidate launch shapes on the first `fraction` of the
        sample times (writing into O_dev_ptr) and keep the fastest; returns it.
        The whole horizon by default: the work per sample interval can grow
        over a run (Lotka-Volterra's later samples cost ~3x the first), and a
        shape timed on an early part only can be slower on the whole."""
        return _autotune(self._h, [O_dev_ptr], [o_stride], [flags_dev_ptr] if flags_dev_ptr else None,
                         stream_ptr, fraction)

    def stats(self):
        st = PstStats()
        _check(load().pst_plan_read_stats(self._h, ctypes.byref(st)), "pst_plan_read_stats")
        return st.as_dict()

    def close(self):
        if getattr(self, "_h", None):
            load().pst_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Job:
    """One job of a batch: tables, parameter rows and run settings (the
    arguments of :func:`simulate`)."""

    def __init__(self, tables, rates, x0, method=METHOD_TAU_LEAPING, precision=PRECISION_FP64, n_c=10,
                 eta=0.03, ssa_burst=100, t_max=1.0, I=None, E=(0,), n_sims=1, sims_per_set=None, seed=0,
                 sim_offset=0, specialize=False, max_steps=0, layout=LAYOUT_ROWS):
        self.m = _Marshalled(tables, rates, x0, method, precision, n_c, eta, ssa_burst, t_max, I, E, n_sims,
                             sims_per_set, seed, sim_offset, specialize, max_steps, layout)
        self.n_sims = int(n_sims)
        self.kappa = self.m.E.shape[0]
        self.ita = self.m.I.shape[0]

    def struct(self):
        return PstJob(self.m.net, self.m.cfg, _ptr(self.m.rates, ctypes.c_double),
                      _ptr(self.m.x0, ctypes.c_int32), self.m.n_sets)


def _job_array(jobs):
    arr = (PstJob * len(jobs))()
    for i, j in enumerate(jobs):
        arr[i] = j.struct()
    return arr


def batch_simulate(jobs):
    """Run several jobs in one persistent launch per size class; returns
    ([O_q], [flags_q], stats)."""
    L = load()
    arr = _job_array(jobs)
    Os = [np.empty((j.kappa, j.ita, j.n_sims), np.int32) for j in jobs]
    Fs = [np.zeros(max(j.n_sims, 1), np.int32) for j in jobs]
    O_ptrs = (_P_I32 * len(jobs))(*[_ptr(o, ctypes.c_int32) for o in Os])
    F_ptrs = (_P_I32 * len(jobs))(*[_ptr(f, ctypes.c_int32) for f in Fs])
    st = PstStats()
    _check(L.pst_batch_run(arr, len(jobs), O_ptrs, F_ptrs, ctypes.byref(st)), "pst_batch_run")
    return Os, [f[: j.n_sims] for f, j in zip(Fs, jobs)], st.as_dict()


class BatchPlan(_RangeMixin):
    """Device-resident batch: tables uploaded once; execute() takes device
    addresses of one trajectory array (and optionally one flag array) per job."""

    def __init__(self, jobs):
        self.jobs = list(jobs)
        h = ctypes.c_void_p()
        self._arr = _job_array(self.jobs)
        _check(load().pst_batch_plan_create(self._arr, len(self.jobs), ctypes.byref(h)), "pst_batch_plan_create")
        self._h = h

    def execute(self, O_ptrs, strides, flag_ptrs=None, stream_ptr=None):
        n = len(self.jobs)
        Op = (ctypes.c_void_p * n)(*[int(p) for p in O_ptrs])
        Sp = (ctypes.c_int64 * n)(*[int(s) for s in strides])
        Fp = (ctypes.c_void_p * n)(*[int(p) if p else None for p in flag_ptrs]) if flag_ptrs else None
        _check(load().pst_batch_plan_execute(self._h, Op, Sp, Fp,
                                             ctypes.c_void_p(int(stream_ptr)) if stream_ptr else None),
               "pst_batch_plan_execute")

    def set_launch(self, lanes_shift, ssa_slice, waves_trim=0, burst_waves=0):
        """Use this launch shape (e.g. one an autotune returned) for later executions."""
        _check(load().pst_plan_set_launch(self._h, int(lanes_shift), int(ssa_slice), int(waves_trim),
                                          int(burst_waves)),
               "pst_plan_set_launch")

    def autotune(self, O_ptrs, strides, flag_ptrs=None, stream_ptr=None, fraction=1.0):
        """Time the candidate launch shapes on the first `fraction` of every
        job's sample times (the whole horizon by default) and keep the
        fastest; returns it."""
        return _autotune(self._h, O_ptrs, strides, flag_ptrs, stream_ptr, fraction)

    def set_stream(self, seed, sim_offset=0):
        _check(load().pst_plan_set_stream(self._h, int(seed), int(sim_offset)), "pst_plan_set_stream")

    def stats(self):
        st = PstStats()
        _check(load().pst_plan_read_stats(self._h, ctypes.byref(st)), "pst_plan_read_stats")
        return st.as_dict()

    def close(self):
        if getattr(self, "_h", None):
            load().pst_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def max_network():
    n = ctypes.c_int32()
    m = ctypes.c_int32()
    load().pst_max_network(ctypes.byref(n), ctypes.byref(m))
    return int(n.value), int(m.value)
