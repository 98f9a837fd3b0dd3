"""The Poisson sampler of the mixed-precision kernels on the GPU
(pst_poisson_probe_mixed), draw by draw: every draw of the inversions
(means below 2^16) is the exact inverse Poisson CDF at its uniform, computed
here in double precision with scipy; PTRS (means of 2^16 and above) follows
the Poisson law (every test marked gpu)."""
import numpy as np
import pytest
from scipy import special, stats

pytestmark = pytest.mark.gpu


def _exact_inverse(lam, u):
    """smallest n with P(X <= n) >= u, for float64 arrays lam, u."""
    n = np.floor(np.nan_to_num(special.pdtrik(u, lam), nan=0.0)) - 1
    n = np.maximum(n, 0)
    for _ in range(64):  # step up while the CDF is below u, then down while the one below reaches it
        up = special.pdtr(n, lam) < u
        if not up.any():
            break
        n = n + up
    for _ in range(64):
        down = (n > 0) & (special.pdtr(n - 1, lam) >= u)
        if not down.any():
            break
        n = n - down
    return n.astype(np.int64)


def _u24(words):
    return ((words >> 8).astype(np.float64) + 1.0) * 2.0 ** -24


def test_inversion_is_exact_draw_by_draw(native):
    rng = np.random.default_rng(11)
    lam = np.concatenate([
        np.float32(np.exp(rng.uniform(np.log(1e-3), np.log(10.0), 600_000))),     # search
        np.float32(np.exp(rng.uniform(np.log(10.0), np.log(65535.0), 1_400_000))),  # asymptotic
        np.float32([10.0, 10.0000010, 9.9999990, 65535.99] * 1000),
    ])
    words = rng.integers(0, 2 ** 32, lam.size, dtype=np.uint64).astype(np.uint32)
    # the extreme uniforms: 2^-24, just below 1, and 1.0
    words[:3000] = np.tile(np.uint32([0, 0xFFFFFE00, 0xFFFFFFFF]), 1000)
    print("case: poisson probe, %d draws" % lam.size, flush=True)
    k, path = native.poisson_probe_mixed(lam, words)
    u = _u24(words)
    u_eff = np.minimum(u, 1 - 2.0 ** -25)  # u = 1 is taken as 1 - 2^-25
    exact = _exact_inverse(lam.astype(np.float64), u_eff)
    inv = (path == 2) | (path == 3)
    assert (path[lam < 10] == 1).all() and inv[(lam >= 10) & (lam < 65536)].all()
    bad = np.nonzero(inv & (k != exact))[0]
    assert bad.size == 0, "asymptotic inversion differs at lam=%s u=%s gpu=%s exact=%s" % (
        lam[bad[:3]], u[bad[:3]], k[bad[:3]], exact[bad[:3]])
    # the FP64 decisions ran, and rarely
    n3 = int((path == 3).sum())
    assert 0 < n3 < 2e-3 * inv.sum() + 3000, n3
    # the FP32 search (mean < 10) sums the pmf in FP32: exact except where u
    # lies within the FP32 rounding of that sum (~1e-6) of a CDF value, where
    # it can be off by one (0.1% of draws), and for u within 2^-20 of 1, where
    # FP32 cannot resolve the CDF (probability 1e-6 per draw)
    s = path == 1
    bulk = s & (u_eff < 1 - 2.0 ** -20)
    miss = bulk & (k != exact)
    assert (np.abs(k[miss] - exact[miss]) == 1).all()
    cdf_at = special.pdtr(np.minimum(k[miss], exact[miss]), lam[miss].astype(np.float64))
    assert (np.abs(cdf_at - u_eff[miss]) < 2e-6).all()
    assert miss.sum() < 3e-3 * bulk.sum()
    tail = s & ~bulk
    assert tail.sum() > 1000 and (np.abs(k[tail] - exact[tail]) <= 4).all()


@pytest.mark.parametrize("lam", [12.5, 110.0, 3000.0, 1.5e5])
def test_law_chi_square(native, lam):
    """Chi-square of 400,000 draws against the pmf (PTRS at 1.5e5)."""
    rng = np.random.default_rng(int(lam * 7))
    words = rng.integers(0, 2 ** 32, 400_000, dtype=np.uint64).astype(np.uint32)
    print("case: poisson law lam=%g" % lam, flush=True)
    k, path = native.poisson_probe_mixed(np.full(words.size, lam, np.float32), words)
    if lam < 65536:
        assert ((path == 2) | (path == 3)).all()
    else:
        assert (path == 4).all()
    lo, hi = (int(v) for v in stats.poisson.ppf([1e-4, 1 - 1e-4], lam))
    vals = np.arange(lo, hi + 1)
    obs = np.array([(k < lo).sum()] + [np.count_nonzero(k == v) for v in vals] + [(k > hi).sum()])
    pmf = np.concatenate([[stats.poisson.cdf(lo - 1, lam)], stats.poisson.pmf(vals, lam),
                          [stats.poisson.sf(hi, lam)]])
    expct = pmf * k.size
    keep = expct > 5
    chi2 = (((obs - expct) ** 2) / expct)[keep].sum()
    assert stats.chi2.sf(chi2, keep.sum() - 1) > 1e-4
    assert abs(k.mean() - lam) < 5 * np.sqrt(lam / k.size)
