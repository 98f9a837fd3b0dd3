"""Reaction networks of the benchmark workloads, built as SPN objects.

Each builder returns an :class:`SPN` with the same fields an SBML file would
give (Pre, Post, StoichMConst, M, c, P, T, h), so it goes through the same
``TlParser.parse`` as a model read from SBML.

* birth_death       0 -> X (c=10), X -> 0 (c=0.1), X0 = 0
* lotka_volterra    X -> 2X (10), X + Y -> 2Y (0.01), Y -> 0 (10), X0 = Y0 = 1000
* schlogl           the structure of the reference's examples/schlogl_sbml.xml:
                    A + 2X -> A + 3X  c1/2 * A * X(X-1)
                    3X + A -> 2X + A  c2/6 * X(X-1)(X-2)
                    B -> B + X        c3 * B
                    X + B -> B        c4 * X
                    A = 1e5, B = 2e5 constant, X0 = 250
* random_network    random mass-action network (reactant order 0-2, 0-3
                    product species, stoichiometry 1-2, log-uniform rates)
"""
import numpy as np

from .mod.petri_net.SPN import SPN
from .mod.parser import TlArgs, TlParser


def _spn(n_species, reactions, x0, params, constant=()):
    """reactions: list of (pre {i: n}, post {i: n}, hazard string)."""
    spn = SPN()
    M = len(reactions)
    spn.Pre = np.zeros((n_species, M))
    spn.Post = np.zeros((n_species, M))
    spn.StoichMConst = np.zeros((n_species, M))
    spn.h = []
    for j, (pre, post, haz) in enumerate(reactions):
        for i, n in pre.items():
            spn.Pre[i, j] = n
        for i, n in post.items():
            spn.Post[i, j] = n
        spn.h.append(haz)
    for i in constant:
        spn.StoichMConst[i, :] = 1
    spn.M = list(x0)
    spn.c = list(params)
    spn.P = {"species_%d" % i: i for i in range(n_species)}
    spn.T = {"reaction%d" % (j + 1): j for j in range(M)}
    return spn


def birth_death(birth=10.0, death=0.1, x0=0):
    return _spn(1, [({}, {0: 1}, "parameter_0"),
                    ({0: 1}, {}, "parameter_1*species_0")],
                [x0], [birth, death])


def lotka_volterra(c1=10.0, c2=0.01, c3=10.0, x0=1000, y0=1000):
    return _spn(2, [({0: 1}, {0: 2}, "parameter_0*species_0"),
                    ({0: 1, 1: 1}, {1: 2}, "parameter_1*species_0*species_1"),
                    ({1: 1}, {}, "parameter_2*species_1")],
                [x0, y0], [c1, c2, c3])


def schlogl(c1=3e-7, c2=1e-4, c3=1e-3, c4=3.5, x0=250, a=100000, b=200000):
    X, A, B = 0, 1, 2
    return _spn(3, [({A: 1, X: 2}, {X: 3, A: 1}, "parameter_0/2*species_1*species_0*(species_0 - 1)"),
                    ({X: 3, A: 1}, {X: 2, A: 1}, "parameter_1/6*species_0*(species_0 - 1)*(species_0 - 2)"),
                    ({B: 1}, {X: 1, B: 1}, "parameter_2*species_2"),
                    ({X: 1, B: 1}, {B: 1}, "parameter_3*species_0")],
                [x0, a, b], [c1, c2, c3, c4], constant=(A, B))


def michaelis_menten_birth_death(k=20.0, vmax=25.0, km=10.0, x0=0):
    """0 -> X at rate k, X -> 0 at rate vmax X / (km + X): a kinetic law that
    is not a mass-action product (compiled into the kernel)."""
    return _spn(1, [({}, {0: 1}, "parameter_0"),
                    ({0: 1}, {}, "parameter_1*species_0/(parameter_2 + species_0)")], [x0], [k, vmax, km])


def immigration_birth_death(nu=5.0, b=0.5, d=1.0, x0=0):
    """0 -> X with hazard nu + b X (immigration plus births, one polynomial
    that is not a single product), X -> 0 at rate d X."""
    return _spn(1, [({}, {0: 1}, "parameter_0 + parameter_1*species_0"),
                    ({0: 1}, {}, "parameter_2*species_0")], [x0], [nu, b, d])


def random_network(rng, n_species=None, n_reactions=None, max_species=50, max_reactions=200):
    """Random mass-action network: reactant order 0-2 (species drawn with
    replacement), 0-3 product species with stoichiometry 1-2, rates
    log-uniform on [1e-4, 10], propensity c * prod C(x_i, p_i), initial
    counts Poisson with mean uniform on [50, 500]."""
    N = int(rng.integers(5, max_species + 1)) if n_species is None else n_species
    M = int(rng.integers(10, max_reactions + 1)) if n_reactions is None else n_reactions
    reactions, params = [], []
    for j in range(M):
        order = int(rng.integers(0, 3))
        pre = {}
        for _ in range(order):
            s = int(rng.integers(0, N))
            pre[s] = pre.get(s, 0) + 1
        post = {}
        for _ in range(int(rng.integers(0, 4))):
            s = int(rng.integers(0, N))
            post[s] = post.get(s, 0) + int(rng.integers(1, 3))
        c = 10.0 ** rng.uniform(-4, 1)
        denom = 1
        terms = []
        for s, p in sorted(pre.items()):
            f = "*".join(["species_%d" % s] + ["(species_%d - %d)" % (s, q) for q in range(1, p)])
            terms.append(f)
            for q in range(2, p + 1):
                denom *= q
        haz = "parameter_%d" % j
        if denom != 1:
            haz += "/%d" % denom
        if terms:
            haz += "*" + "*".join(terms)
        reactions.append((pre, post, haz))
        params.append(c)
    mu = rng.uniform(50, 500, size=N)
    x0 = rng.poisson(mu).astype(int)
    return _spn(N, reactions, x0, params)


def make_args(spn, E=(0,), t_max=1.0, ita=11, U=128, gillespie=0, precision="fp64", seed=0,
              n_c=10, eta=0.03):
    """Parse an SPN into TlArgs with the given run settings."""
    a = TlArgs()
    a.E = np.array(E, np.int32)
    a.t_max = t_max
    a.ita = ita
    a.U = U
    a.gillespie = gillespie
    a.precision = precision
    a.seed = seed
    a.n_c = n_c
    a.eta = eta
    return TlParser.parse(a, spn)
