"""ctypes front of the CPU oracle (oracle/pst_oracle.c) -- TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module, and only as the checker (or the timed CPU baseline); the
product path in ``py_stoch_amd`` never does.

The oracle restates mod/simulator/cuTauLeaping/P1_P2.py:54-584,
P3.py:29-208, cuTauLeaping.py:48-84 and cuGillespie/gillespie.py:27-193 in
plain C on the same Philox4x32-10 stream as the HIP kernels.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libpst_oracle.so")

METHOD_TAU_LEAPING = 0
METHOD_FORCED_SSA = 1
METHOD_PURE_SSA = 2

FLAG_OVERFLOW = 1
FLAG_HALVING = 2
FLAG_TAU_INF = 4
FLAG_PTRS = 8
FLAG_STEPS = 16
FLAG_STALL = 32


class OrcNetwork(ctypes.Structure):
    _fields_ = [
        ("n_species", ctypes.c_int32),
        ("n_reactions", ctypes.c_int32),
        ("a_size", ctypes.c_int32),
        ("A", ctypes.POINTER(ctypes.c_int32)),
        ("v_size", ctypes.c_int32),
        ("V", ctypes.POINTER(ctypes.c_int32)),
        ("H", ctypes.POINTER(ctypes.c_int32)),
        ("H_type", ctypes.POINTER(ctypes.c_int32)),
        ("prop_order", ctypes.POINTER(ctypes.c_int32)),
        ("prop_step", ctypes.POINTER(ctypes.c_int32)),
    ]


class OrcConfig(ctypes.Structure):
    _fields_ = [
        ("method", ctypes.c_int32),
        ("n_c", ctypes.c_int32),
        ("eta", ctypes.c_double),
        ("ssa_burst", ctypes.c_int32),
        ("t_max", ctypes.c_double),
        ("ita", ctypes.c_int32),
        ("I", ctypes.POINTER(ctypes.c_float)),
        ("kappa", ctypes.c_int32),
        ("E", ctypes.POINTER(ctypes.c_int32)),
        ("n_sims", ctypes.c_int32),
        ("sims_per_set", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("sim_offset", ctypes.c_uint64),
        ("max_steps", ctypes.c_int64),
    ]


class OrcStats(ctypes.Structure):
    _fields_ = [
        ("events", ctypes.c_int64),
        ("leaps", ctypes.c_int64),
        ("rejected_leaps", ctypes.c_int64),
        ("ssa_steps", ctypes.c_int64),
        ("flagged", ctypes.c_int64),
        ("flagged_events", ctypes.c_int64),
    ]

    def as_dict(self):
        return {name: int(getattr(self, name)) for name, _ in self._fields_}


_lib = None


def build(force=False):
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_philox4x32_10.argtypes = [ctypes.POINTER(ctypes.c_uint32)] * 3
        L.orc_philox4x32_10.restype = None
        L.orc_uniform_stream.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int64,
                                         ctypes.POINTER(ctypes.c_double)]
        L.orc_poisson_stream.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double,
                                         ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
        L.orc_log.argtypes = [ctypes.c_double]
        L.orc_log.restype = ctypes.c_double
        L.orc_exp.argtypes = [ctypes.c_double]
        L.orc_exp.restype = ctypes.c_double
        L.orc_lgamma_int.argtypes = [ctypes.c_double]
        L.orc_lgamma_int.restype = ctypes.c_double
        L.orc_simulate.argtypes = [ctypes.POINTER(OrcNetwork), ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(OrcConfig),
                                   ctypes.POINTER(ctypes.c_int32), ctypes.c_int64,
                                   ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(OrcStats),
                                   ctypes.c_int]
        L.orc_simulate.restype = ctypes.c_int
        L.orc_max_threads.restype = ctypes.c_int
        L.orc_set_plant.argtypes = [ctypes.c_int]
        L.orc_set_plant.restype = None
        L.orc_critical_draws.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32),
                                         ctypes.c_int, ctypes.c_double, ctypes.c_uint64, ctypes.c_uint64,
                                         ctypes.c_int64, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_int32)]
        L.orc_critical_draws.restype = None
        _lib = L
    return _lib


def _p(arr, ctype):
    return arr.ctypes.data_as(ctypes.POINTER(ctype))


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return [int(v) for v in o]


def uniforms(seed, sim, n):
    out = np.empty(n, np.float64)
    lib().orc_uniform_stream(seed, sim, n, _p(out, ctypes.c_double))
    return out


def poisson_samples(seed, sim, lam, n):
    out = np.empty(n, np.int64)
    lib().orc_poisson_stream(seed, sim, float(lam), n, _p(out, ctypes.c_int64))
    return out


def log(x):
    return lib().orc_log(float(x))


def exp(x):
    return lib().orc_exp(float(x))


def lgamma_int(x):
    return lib().orc_lgamma_int(float(x))


def max_threads():
    return lib().orc_max_threads()


def critical_draws(a, critical, tau1, seed, n, sim=0):
    """n draws of the critical part of a leap (tau2 and SingleCriticalReaction,
    P1_P2.py:175-220) at propensities a with critical flags `critical`:
    returns (tau[n], picked reaction[n], -1 when tau1 <= tau2)."""
    a = np.ascontiguousarray(a, np.float64)
    crit = np.ascontiguousarray(critical, np.int32)
    tau = np.empty(n, np.float64)
    j = np.empty(n, np.int32)
    lib().orc_critical_draws(_p(a, ctypes.c_double), _p(crit, ctypes.c_int32), a.size, float(tau1), seed, sim, n,
                             _p(tau, ctypes.c_double), _p(j, ctypes.c_int32))
    return tau, j


class planted:
    """Context manager that plants a known mistake in the oracle's critical
    branch for the duration of a block (tests only):
    ``"tau2_a0"`` draws tau2 with a0 instead of a0c, ``"pick_a0"`` normalises
    the critical pick by a0 instead of a0c."""

    _CODES = {"tau2_a0": 1, "pick_a0": 2}

    def __init__(self, what):
        self.code = self._CODES[what]

    def __enter__(self):
        lib().orc_set_plant(self.code)
        return self

    def __exit__(self, *exc):
        lib().orc_set_plant(0)
        return False


def simulate(tables, rates, x0, method=METHOD_TAU_LEAPING, n_c=10, eta=0.03, ssa_burst=100,
             t_max=1.0, I=None, E=(0,), n_sims=1, sims_per_set=None, seed=0, sim_offset=0,
             o_stride=None, threads=0, max_steps=0):
    """Run the oracle.

    ``tables`` is a dict with int32 arrays ``A`` (a_size x 4), ``V`` (v_size x 4),
    ``H``, ``H_type`` (N), ``prop_order``, ``prop_step`` (N x M) and ints ``N``, ``M``.
    ``rates`` is [n_sets][M] float64, ``x0`` is [n_sets][N] int32.
    Returns (O[kappa][ita][o_stride] int32, flags[n_sims] int32, stats dict).
    """
    N, M = int(tables["N"]), int(tables["M"])
    A = np.ascontiguousarray(tables["A"], np.int32).reshape(-1, 4)
    V = np.ascontiguousarray(tables["V"], np.int32).reshape(-1, 4)
    H = np.ascontiguousarray(tables["H"], np.int32)
    Ht = np.ascontiguousarray(tables["H_type"], np.int32)
    po = np.ascontiguousarray(tables["prop_order"], np.int32).reshape(N, M)
    ps = np.ascontiguousarray(tables["prop_step"], np.int32).reshape(N, M)
    rates = np.ascontiguousarray(np.atleast_2d(np.asarray(rates, np.float64)))
    x0 = np.ascontiguousarray(np.atleast_2d(np.asarray(x0, np.int32)))
    I = np.ascontiguousarray(np.asarray(I, np.float32))
    E = np.ascontiguousarray(np.asarray(E, np.int32))
    if sims_per_set is None:
        sims_per_set = max(1, n_sims)
    if o_stride is None:
        o_stride = n_sims
    net = OrcNetwork(N, M, A.shape[0], _p(A, ctypes.c_int32), V.shape[0], _p(V, ctypes.c_int32),
                     _p(H, ctypes.c_int32), _p(Ht, ctypes.c_int32), _p(po, ctypes.c_int32),
                     _p(ps, ctypes.c_int32))
    cfg = OrcConfig(method, n_c, eta, ssa_burst, float(t_max), I.shape[0], _p(I, ctypes.c_float),
                    E.shape[0], _p(E, ctypes.c_int32), n_sims, sims_per_set, seed, sim_offset, int(max_steps))
    O = np.zeros((E.shape[0], I.shape[0], o_stride), np.int32)
    flags = np.zeros(max(n_sims, 1), np.int32)
    st = OrcStats()
    rc = lib().orc_simulate(ctypes.byref(net), _p(rates, ctypes.c_double), _p(x0, ctypes.c_int32),
                            ctypes.byref(cfg), _p(O, ctypes.c_int32), o_stride,
                            _p(flags, ctypes.c_int32), ctypes.byref(st), threads)
    if rc != 0:
        raise ValueError("oracle rejected the configuration (code %d)" % rc)
    return O, flags[:n_sims], st.as_dict()


def _setup_hooks(L):
    if getattr(L, "_hooks", False):
        return
    L.orc_leap_decision.argtypes = [ctypes.POINTER(OrcNetwork), ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_int32), ctypes.c_int32, ctypes.c_double,
                                    ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_int32)]
    L.orc_leap_decision.restype = ctypes.c_int
    L.orc_calc_g.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
    L.orc_calc_g.restype = ctypes.c_double
    L.orc_interpolate.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_float]
    L.orc_interpolate.restype = ctypes.c_int32
    L._hooks = True


def leap_decision(tables, rates, x, n_c=10, eta=0.03):
    """(a, a0, critical flags, tau1, use_ssa) at state x (P1_P2.py:96-173)."""
    L = lib()
    _setup_hooks(L)
    N, M = int(tables["N"]), int(tables["M"])
    A = np.ascontiguousarray(tables["A"], np.int32).reshape(-1, 4)
    V = np.ascontiguousarray(tables["V"], np.int32).reshape(-1, 4)
    H = np.ascontiguousarray(tables["H"], np.int32)
    Ht = np.ascontiguousarray(tables["H_type"], np.int32)
    po = np.ascontiguousarray(tables["prop_order"], np.int32).reshape(N, M)
    ps = np.ascontiguousarray(tables["prop_step"], np.int32).reshape(N, M)
    net = OrcNetwork(N, M, A.shape[0], _p(A, ctypes.c_int32), V.shape[0], _p(V, ctypes.c_int32),
                     _p(H, ctypes.c_int32), _p(Ht, ctypes.c_int32), _p(po, ctypes.c_int32),
                     _p(ps, ctypes.c_int32))
    k = np.ascontiguousarray(rates, np.float64)
    xs = np.ascontiguousarray(x, np.int32)
    a = np.zeros(M, np.float64)
    crit = np.zeros(M, np.int32)
    a0 = ctypes.c_double()
    tau1 = ctypes.c_double()
    use = ctypes.c_int32()
    rc = L.orc_leap_decision(ctypes.byref(net), _p(k, ctypes.c_double), _p(xs, ctypes.c_int32), n_c, eta,
                             _p(a, ctypes.c_double), ctypes.byref(a0), _p(crit, ctypes.c_int32),
                             ctypes.byref(tau1), ctypes.byref(use))
    if rc != 0:
        raise ValueError("bad network")
    return a, a0.value, crit, tau1.value, bool(use.value)


def calc_g(h, h_type, x_i):
    L = lib()
    _setup_hooks(L)
    return L.orc_calc_g(h, h_type, x_i)


def interpolate(x, x_prime, t_old, tau, record_time):
    L = lib()
    _setup_hooks(L)
    return int(L.orc_interpolate(x, x_prime, t_old, tau, record_time))
