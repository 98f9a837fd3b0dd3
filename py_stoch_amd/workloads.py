"""The synthetic workloads the simulator is measured on.

Each workload is a list of jobs (network tables, parameter rows, initial
states and run settings) that ``_native.Job`` / ``_native.Plan`` accept:

  birth_death        X0 = 0, 0 -> X (10), X -> 0 (0.1); pure SSA, 1024
                     simulations, t_end = 100, 101 samples, seed 0
  lotka_volterra     X0 = Y0 = 1000, X -> 2X (10), X + Y -> 2Y (0.01),
                     Y -> 0 (10); tau-leaping, eps 0.03, n_c 10, burst 100,
                     65536 simulations, t_end = 10, 1000 samples, seed 2
  schlogl            X0 = 250 with A = 1e5, B = 2e5 held constant (the
                     examples/schlogl_sbml.xml structure), c = (3e-7, 1e-4,
                     1e-3, 3.5); 262144 simulations, t_end = 10, 100
                     samples, seed 3; tau-leaping or pure SSA
  sweep              4096 parameter sets, 256 simulations each: sets
                     0..2047 the Lotka-Volterra family, 2048..4095 the
                     Schlogl family, every rate constant c_j * 10^u with
                     u ~ U(-1, 1) drawn set by set from
                     numpy.random.Generator(numpy.random.Philox(4));
                     tau-leaping, t_end = 10, 100 samples, seed 4
  random_networks    256 random mass-action networks from
                     numpy.random.default_rng(5) (5-50 species, 10-200
                     reactions), 1024 simulations each, tau-leaping,
                     t_end = 5, 50 samples, every species recorded, seed 5;
                     rate constants span five decades, so many networks are
                     stiff (median ~1e6 leaps per simulation to t_end, some
                     above 1e7): each simulation gets a step budget
                     (max_steps, 20000 by default) and ends flagged when it
                     runs out

Parameter sweeps perturb the kinetic-law parameters (``TlArgs.c``), which
the reference's sim_maker sweeps (sim_maker.py:6-43).
"""
import numpy as np

from . import models
from .mod.parser.tl_parser import tables_of


class Spec:
    """One job: ``tables`` (tables_of), ``rates`` [n_sets][M] per-reaction
    mass-action constants, ``x0`` [n_sets][N], ``kw`` the run settings."""

    def __init__(self, name, tables, rates, x0, **kw):
        self.name = name
        self.tables = tables
        self.rates = np.atleast_2d(np.asarray(rates, np.float64))
        self.x0 = np.atleast_2d(np.asarray(x0, np.int32))
        self.kw = kw

    @property
    def n_sims(self):
        return int(self.kw["n_sims"])

    @property
    def kappa(self):
        return len(self.kw["E"])

    @property
    def ita(self):
        return len(self.kw["I"])


def _single(name, spn, E, t_max, ita, U, seed, method, **extra):
    a = models.make_args(spn, E=E, t_max=t_max, ita=ita, U=U)
    return Spec(name, tables_of(a), a.rate_fn(a.c), a.x_0, method=method, t_max=float(t_max), I=a.I,
                E=a.E, n_sims=U, seed=seed, **extra)


def birth_death():
    return [_single("birth-death", models.birth_death(), [0], 100.0, 101, 1024, 0, method=2)]


def lotka_volterra(U=65536):
    return [_single("lotka-volterra", models.lotka_volterra(), [0, 1], 10.0, 1000, U, 2, method=0,
                    n_c=10, eta=0.03, ssa_burst=100)]


def schlogl(method=0, U=262144):
    return [_single("schlogl", models.schlogl(), [0], 10.0, 100, U, 3, method=method, n_c=10, eta=0.03,
                    ssa_burst=100)]


def sweep(n_sets=4096, sims_per_set=256):
    """Two jobs (one per model family), n_sets/2 parameter rows each."""
    rng = np.random.Generator(np.random.Philox(4))
    fams = []
    for spn, E in ((models.lotka_volterra(), [0, 1]), (models.schlogl(), [0])):
        a = models.make_args(spn, E=E, t_max=10.0, ita=100, U=sims_per_set)
        fams.append(a)
    rows = ([], [])
    for p in range(n_sets):
        f = 0 if p < n_sets // 2 else 1
        a = fams[f]
        c = np.asarray(a.c, np.float64) * 10.0 ** rng.uniform(-1.0, 1.0, size=len(a.c))
        rows[f].append(a.rate_fn(c))
    out = []
    for f, a in enumerate(fams):
        n = len(rows[f])
        out.append(Spec("sweep-" + ("lotka-volterra" if f == 0 else "schlogl"), tables_of(a), np.array(rows[f]),
                        np.tile(a.x_0, (n, 1)), method=0, t_max=10.0, I=a.I, E=a.E, n_sims=n * sims_per_set,
                        sims_per_set=sims_per_set, seed=4, n_c=10, eta=0.03, ssa_burst=100))
    return out


def mass_action_tables(spn):
    """tables_of(TlParser.parse(spn)) for a network whose every hazard is
    c_j / prod_i Pre_ij! * prod_i x_i (x_i - 1) ... (x_i - Pre_ij + 1) (the form
    models.random_network writes), built from Pre / Post directly instead of
    through sympy; returns (tables, per-reaction rates)."""
    from .mod.parser.parser import Parser
    pre = np.asarray(spn.Pre, np.int64)
    post = np.asarray(spn.Post, np.int64)
    N, M = pre.shape
    order = pre.sum(axis=0)                      # reaction orders (tl_parser.py:80-95)
    H = np.zeros(N, np.int32)
    Ht = np.zeros(N, np.int32)
    for i in range(N):
        js = np.nonzero(pre[i])[0]               # reactions whose hazard names species i
        if js.size:
            H[i] = order[js].max()
            Ht[i] = pre[i, js].max()
    fact = np.ones(M)
    for i in range(N):
        for j in range(M):
            for q in range(2, int(pre[i, j]) + 1):
                fact[j] *= q
    A, _ = Parser.flatten_matrix(pre)
    V, _ = Parser.flatten_matrix(post - pre)
    tabs = {"N": N, "M": M, "A": np.asarray(A, np.int32).reshape(-1, 4), "V": np.asarray(V, np.int32).reshape(-1, 4),
            "H": H, "H_type": Ht, "prop_order": pre.astype(np.int32), "prop_step": np.ones((N, M), np.int32),
            "hazard_c": None, "n_params": 0}
    return tabs, np.asarray(spn.c, np.float64) / fact


def random_networks(n_networks=256, sims=1024, max_species=50, max_reactions=200, max_steps=20000):
    rng = np.random.default_rng(5)
    out = []
    for n in range(n_networks):
        spn = models.random_network(rng, max_species=max_species, max_reactions=max_reactions)
        tabs, rates = mass_action_tables(spn)
        N = tabs["N"]
        I = np.linspace(0, 5.0, num=50).astype(np.float32)
        out.append(Spec("random-%d" % n, tabs, rates, np.asarray(spn.M, np.int32), method=0, t_max=5.0, I=I,
                        E=np.arange(N, dtype=np.int32), n_sims=sims, seed=5, n_c=10, eta=0.03, ssa_burst=100,
                        max_steps=max_steps))
    return out


def jobs(specs, precision, specialize=True):
    """_native.Job objects for a list of Spec."""
    from . import _native
    return [_native.Job(s.tables, s.rates, s.x0, precision=precision, specialize=specialize, **s.kw) for s in specs]
