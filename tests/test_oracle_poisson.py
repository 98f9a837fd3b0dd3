"""Pinning the oracle's Poisson sampler (replaces curand_poisson, P1_P2.py:239)
against the exact Poisson law: mean, variance and a chi-square test of the pmf
for means in the inversion range, at the branch boundary and in the PTRS range."""
import numpy as np
import pytest
from scipy import stats


@pytest.mark.parametrize("lam", [1e-3, 0.5, 5.0, 9.99, 10.0, 110.0, 1e4])
def test_poisson_law(oracle_mod, lam):
    n = 200000
    x = oracle_mod.poisson_samples(seed=11, sim=3, lam=lam, n=n)
    assert x.min() >= 0
    # mean and variance within 5 standard errors
    se_mean = np.sqrt(lam / n)
    assert abs(x.mean() - lam) < 5 * se_mean + 1e-12
    se_var = np.sqrt((2 * lam ** 2 + lam) / n)
    assert abs(x.var() - lam) < 5 * se_var + 1e-12
    # chi-square on the bins with expected count >= 20
    lo, hi = int(stats.poisson.ppf(1e-6, lam)), int(stats.poisson.ppf(1 - 1e-6, lam))
    ks = np.arange(lo, hi + 1)
    p = stats.poisson.pmf(ks, lam)
    keep = p * n >= 20
    if keep.sum() >= 2:
        obs = np.array([(x == k).sum() for k in ks[keep]], dtype=float)
        exp = p[keep] * n
        rest_o = n - obs.sum()
        rest_e = n - exp.sum()
        chi2 = ((obs - exp) ** 2 / exp).sum() + ((rest_o - rest_e) ** 2 / rest_e if rest_e > 20 else 0.0)
        dof = keep.sum()
        assert stats.chi2.sf(chi2, dof) > 1e-4


def test_poisson_zero_mean(oracle_mod):
    assert (oracle_mod.poisson_samples(0, 0, 0.0, 100) == 0).all()


def test_poisson_planted_mistake_is_caught(oracle_mod):
    """A sampler off by one (a plausible PTRS floor error) must fail the law test."""
    lam = 110.0
    x = oracle_mod.poisson_samples(5, 0, lam, 200000) + 1
    assert abs(x.mean() - lam) > 5 * np.sqrt(lam / x.size)
