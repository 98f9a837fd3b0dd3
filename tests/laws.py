"""Small networks with exact laws, and the checks that compare a sample with
such a law.  Shared by the oracle tests (CPU) and the GPU parity tests.

The critical-reaction networks make tau-leaping exact: with n_c above every
count that a consuming reaction can see, every such reaction is critical
(P1_P2.py:299-328), tau_1 is infinite, the step never switches to SSA, and
each leap is one firing drawn by tau_2 = Exp(a0c) and the cumulative pick
a_j / a0c >= rho (P1_P2.py:175-220, 488-507) -- one exact SSA event per leap,
ending at the firing.  A sample time inside a leap is recorded by linear
interpolation rounded to the nearest integer, halves to even
(P1_P2.py:542-584): a firing shows in the samples from the MIDPOINT of the
inter-event interval that ends with it.  For a pure death with per-molecule
rate k from n molecules the recorded count at time s is therefore
n - #{j : m_j < s}, m_j = T_{j-1} + E_j / 2, where T_j are the event times and
E_j ~ Exp(k (n - j + 1)) the inter-event times: m_j is hypoexponential
(rates k n, ..., k (n - j + 2), then 2 k (n - j + 1)) and its CDF comes from
a small Markov chain (midpoint_death_law).  Once the count is 0 the state is
absorbing and the remaining samples hold it exactly.
"""
import numpy as np
from scipy import linalg, stats

from py_stoch_amd.models import _spn

BIG_NC = 1000000  # n_c above every count: every consuming reaction is critical


def pure_death(k=1.0, x0=50):
    """X -> 0 at rate k x (as a process, X(t) ~ Binomial(x0, exp(-k t)); its
    recorded samples under one critical firing per leap follow
    midpoint_death_law)."""
    return _spn(1, [({0: 1}, {}, "parameter_0*species_0")], [x0], [k])


def competing_decays(k1=1.0, k2=3.0, x0=100):
    """X -> 0 (k1), X -> Y (k2): each firing makes a Y with probability
    k2 / (k1 + k2); once X = 0 the state absorbs with Y ~ Bin(x0, k2/(k1+k2))."""
    return _spn(2, [({0: 1}, {}, "parameter_0*species_0"),
                    ({0: 1}, {1: 1}, "parameter_1*species_0")], [x0, 0], [k1, k2])


def decays_beside_reservoir(k1=1.0, k2=3.0, x0=8, nu=500.0, d=1.0, z0=500):
    """competing_decays on X = species 0, Y = 1 next to an independent
    immigration-death reservoir Z = 2 (0 -> Z at nu, Z -> 0 at d z).  With
    n_c = 10 and x0 < 10 both X reactions are always critical and Z's never:
    a leap draws Poisson counts for Z beside the tau_2 firing of X, with a0
    (~1000) far above a0c (X's propensities)."""
    return _spn(3, [({0: 1}, {}, "parameter_0*species_0"),
                    ({0: 1}, {1: 1}, "parameter_1*species_0"),
                    ({}, {2: 1}, "parameter_2"),
                    ({2: 1}, {}, "parameter_3*species_2")], [x0, 0, z0], [k1, k2, nu, d])


def padded(spn, extra=20, count=7):
    """The same network with `extra` inert species (no reaction names them):
    too many species for the register kernels, so it runs on the
    large-network kernel."""
    n0, M = spn.Pre.shape
    reactions = []
    for j in range(M):
        pre = {i: int(spn.Pre[i, j]) for i in range(n0) if spn.Pre[i, j]}
        post = {i: int(spn.Post[i, j]) for i in range(n0) if spn.Post[i, j]}
        reactions.append((pre, post, spn.h[j]))
    return _spn(n0 + extra, reactions, list(spn.M) + [count] * extra, list(spn.c))


def ks2(x, y):
    """Two-sample KS distance."""
    x, y = np.sort(x), np.sort(y)
    g = np.concatenate([x, y])
    return float(np.abs(np.searchsorted(x, g, side="right") / x.size
                        - np.searchsorted(y, g, side="right") / y.size).max())


def ks_to_law(x, dist):
    """Largest gap between the sample's CDF and the law's CDF at the integers."""
    x = np.sort(np.asarray(x))
    lo, hi = int(min(x.min(), dist.ppf(1e-9))), int(max(x.max(), dist.ppf(1 - 1e-9)))
    k = np.arange(lo - 1, hi + 2)
    return float(np.abs(np.searchsorted(x, k, side="right") / x.size - dist.cdf(k)).max())


def law_gaps(x, dist):
    """(|mean error| in standard errors, KS distance) of sample x against law dist."""
    x = np.asarray(x, np.float64)
    se = np.sqrt(max(float(dist.var()), 1e-300) / x.size)
    return abs(x.mean() - float(dist.mean())) / se, ks_to_law(x, dist)


def assert_law(x, dist, what, z=5.0):
    """Mean within z standard errors and a KS distance within the 99.99% band
    of the law (1.95 / sqrt(n)), plus a small allowance for a discrete law."""
    zz, ks = law_gaps(x, dist)
    n = np.asarray(x).size
    assert zz < z, "%s: mean %.4f vs %.4f (%.1f standard errors)" % (what, np.mean(x), dist.mean(), zz)
    assert ks < 1.95 / np.sqrt(n) + 0.005, "%s: KS %.4f to the exact law" % (what, ks)


def binom(n, p):
    return stats.binom(int(n), float(min(max(p, 0.0), 1.0)))


def midpoint_death_law(n, k, s):
    """Law of the recorded count at time s of a pure death (per-molecule rate
    k, n molecules at t = 0) run as one critical firing per leap, with the
    interpolated samples described in the module docstring."""
    if s <= 0:
        return stats.rv_discrete(values=([n], [1.0]))
    P = np.zeros(n + 2)
    P[0] = 1.0
    for j in range(1, n + 1):
        Q = np.zeros((j, j))
        for i in range(j - 1):
            r = k * (n - i)
            Q[i, i], Q[i, i + 1] = -r, r
        Q[j - 1, j - 1] = -2.0 * k * (n - j + 1)
        P[j] = 1.0 - linalg.expm(Q * s)[0].sum()  # P(m_j < s)
    pk = np.clip(P[:-1] - P[1:], 0.0, 1.0)  # P(n - j midpoints passed) for j = 0..n
    vals = n - np.arange(n + 1)
    return stats.rv_discrete(values=(vals[::-1], (pk / pk.sum())[::-1]))


def midpoint_split_law(n, k, p, s):
    """Law of the recorded count of the second product of two competing
    decays (total rate k per molecule, a firing goes to it with probability
    p): Binomial(D, p) with D = n - (recorded X count) from midpoint_death_law."""
    dX = midpoint_death_law(n, k, s)
    xs = np.arange(n + 1)
    pD = dX.pmf(n - xs)  # P(D = d) for d = xs
    py = np.zeros(n + 1)
    for d in range(n + 1):
        if pD[d] > 0:
            py[:d + 1] += pD[d] * stats.binom.pmf(np.arange(d + 1), d, p)
    return stats.rv_discrete(values=(xs, py / py.sum()))
