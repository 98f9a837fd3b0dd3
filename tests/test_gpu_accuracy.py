"""Mixed precision against FP64 on the full throughput workload (Lotka-Volterra
tau-leaping, X0 = Y0 = 1000, c = 10, 0.01, 10, 65,536 simulations, t_end =
10, 1000 samples), both through the C ABI and on independent streams.

The accuracy bar of the mixed path is a two-sample KS distance of at most
0.02 between the two laws at every recorded species and sample time; the
means must also agree within 5 standard errors.  Means are taken over the
simulations that end unflagged, keep both species alive at t_end and stay
below 20 x0 at every sample: while the predator is at or near zero the prey
grows like e^(10 t), and those rare trajectories (about 0.1 %) would dominate
every late-time mean."""
import numpy as np
import pytest

from py_stoch_amd import models
from py_stoch_amd.mod.parser.tl_parser import tables_of

pytestmark = pytest.mark.gpu

U = 65536
CAP = 20 * 1000
EVERY = 10  # compare every 10th of the 1000 samples


def _kept(O, flags):
    return (flags == 0) & (O[:, -1, :] > 0).all(axis=0) & (O.max(axis=(0, 1)) <= CAP)


def _ks(a, b):
    a = np.sort(a)
    b = np.sort(b)
    v = np.concatenate([a, b])
    return float(np.abs(np.searchsorted(a, v, side="right") / a.size
                        - np.searchsorted(b, v, side="right") / b.size).max())


def test_mixed_matches_fp64_on_the_full_workload(native):
    a = models.make_args(models.lotka_volterra(), E=[0, 1], t_max=10.0, ita=1000, U=U)
    tabs, rates = tables_of(a), a.rate_fn(a.c)
    kw = dict(method=native.METHOD_TAU_LEAPING, t_max=10.0, I=a.I, E=a.E, n_sims=U)
    print("case: LV 65536 sims, mixed seed 2 vs FP64 seed 102", flush=True)
    Om, fm, sm = native.simulate(tabs, rates, a.x_0, precision=native.PRECISION_MIXED, seed=2, **kw)
    Of, ff, sf = native.simulate(tabs, rates, a.x_0, precision=native.PRECISION_FP64, seed=102, **kw)
    assert sm["leaps"] > 0 and sm["ssa_steps"] > 0
    x0 = np.asarray(a.x_0)[a.E][:, None]
    assert (Om[:, 0, :] == x0).all() and (Of[:, 0, :] == x0).all()  # sample 0 is x0
    assert (Om >= 0).all()
    # explosive simulations are rare on either path
    assert int((fm != 0).sum()) < 20 and int((ff != 0).sum()) < 20
    km, kf = _kept(Om, fm), _kept(Of, ff)
    assert km.mean() > 0.99 and kf.mean() > 0.99

    ks_worst, z_worst = 0.0, 0.0
    for sp in range(Om.shape[0]):
        for f in range(EVERY - 1, Om.shape[1], EVERY):
            ks_worst = max(ks_worst, _ks(Om[sp, f], Of[sp, f]))
            xm = Om[sp, f, km].astype(np.float64)
            xf = Of[sp, f, kf].astype(np.float64)
            se = np.sqrt(xm.var(ddof=1) / xm.size + xf.var(ddof=1) / xf.size)
            z_worst = max(z_worst, abs(xm.mean() - xf.mean()) / se)
    print("worst KS %.4f, worst mean z %.2f" % (ks_worst, z_worst), flush=True)
    assert ks_worst <= 0.02
    assert z_worst < 5.0
