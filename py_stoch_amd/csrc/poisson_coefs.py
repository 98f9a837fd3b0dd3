"""Generates the coefficient tables of the mixed-precision Poisson inversion
(pst_device.hpp, ``poisson_asym_f32``) and prints them as C++ arrays.

The sampler returns N = min{n : P(X <= n) >= u} for X ~ Poisson(lambda) by
inverting the regularized incomplete gamma function in its shape parameter:
P(X <= n) = Q(n + 1, lambda), so with x* the root of Q(x*, lambda) = u,
N = ceil(x* - 1).  Temme's uniform asymptotic expansion
    Q(a, z) = erfc(eta sqrt(a/2)) / 2 + exp(-a eta^2/2) / sqrt(2 pi a) (C0(eta) + O(1/a)),
    eta^2 / 2 = z/a - 1 - ln(z/a),  sign(eta) = sign(z/a - 1),  C0 = 1/(z/a - 1) - 1/eta
(DLMF 8.12.3-8.12.8), solved for a = x* with s = Phi^{-1}(u) and w = s / sqrt(lambda),
gives
    x* = lambda r0(w) + delta0(w) + d1(w) / lambda + O(lambda^-2),
where r0 solves f(r0) = w, f(r) = sign(r - 1) sqrt(2 (1 - r + r ln r)), and
    delta0(w) = -sqrt(r0) ln(1 - C0 w / sqrt(r0)) / (w D),  D = (r0 - 1)/(r0 eta) + eta/2,
(the first correction, solved without linearising exp(-s delta), which is O(1)
when s is large); d1(w) is the next term, taken here numerically as
lim lambda (x* - lambda r0 - delta0) from exact roots at lambda = 100, 200
(Richardson extrapolation in 1/lambda).  d1(0) = -0.0198 is the known
-1/(50 lambda) of the Poisson median.

Tables (polynomials in monomial form, lowest degree first; float32):
  QW   (r0(w) - 1) / w                 on w in [-0.7, 1.0], degree 10
  D0   delta0(w)                        on w in [-0.7, 1.0], degree 10
  D1   d1(w)                            on w in [-0.7, 1.0], degree 6
  EC   erfinv(x) / x as a polynomial in ln_term - 2.5, ln_term = -ln(1 - x^2) < 5
  ET   erfinv(x) / x as a polynomial in sqrt(ln_term) - 3 for ln_term in [5, 16.7]
(Phi^{-1}(u) = sqrt(2) erfinv(2u - 1), 1 - x^2 = 4 u (1 - u)).

Run:  python3 py_stoch_amd/csrc/poisson_coefs.py   (mpmath, scipy; a few minutes)
"""
import mpmath as mp
import numpy as np
from scipy import special

mp.mp.dps = 40
W_LO, W_HI = -0.7, 1.0


def f_mp(r):
    return mp.sign(r - 1) * mp.sqrt(2 * (1 - r + r * mp.log(r)))


def r0(w):
    w = mp.mpf(w)
    if w == 0:
        return mp.mpf(1)
    return mp.findroot(lambda r: f_mp(r) - w, 1 + w + w * w / 6 if abs(w) < 0.5 else (0.2 if w < 0 else 1 + w))


def delta0(w):
    w = mp.mpf(w)
    if abs(w) < mp.mpf("1e-12"):
        return mp.mpf(1) / 3
    r = r0(w)
    eta = mp.sign(1 - r) * mp.sqrt(2 * (1 / r - 1 + mp.log(r)))
    c0 = r / (1 - r) - 1 / eta
    d = (r - 1) / (r * eta) + eta / 2
    return -mp.sqrt(r) * mp.log(1 - c0 * w / mp.sqrt(r)) / (w * d)


def xstar(lam, s):
    """Root a of Q(a, lam) = Phi(s) (lower form for s > 0, to avoid 1 - u)."""
    lam, s = mp.mpf(lam), mp.mpf(s)
    a0 = lam * r0(s / mp.sqrt(lam)) + delta0(s / mp.sqrt(lam))
    if s <= 0:
        u = mp.ncdf(s)
        return mp.findroot(lambda a: mp.gammainc(a, lam, mp.inf, regularized=True) - u, a0,
                           tol=mp.mpf(10) ** -30)
    v = mp.ncdf(-s)  # 1 - u
    return mp.findroot(lambda a: mp.gammainc(a, 0, lam, regularized=True) - v, a0, tol=mp.mpf(10) ** -30)


def fit(x, y, deg, rel=False):
    P = np.polynomial.polynomial
    c = P.polyfit(x, y, deg, w=1.0 / np.abs(y) if rel else None)
    err = P.polyval(x, c) - y
    return np.float32(c), float(np.abs(err / y if rel else err).max())


def tables():
    ws = np.concatenate([np.cos(np.linspace(0, np.pi, 801)) * (W_HI - W_LO) / 2 + (W_HI + W_LO) / 2,
                         np.linspace(W_LO, W_HI, 1601)])
    ws = ws[np.abs(ws) > 1e-7]
    qw = np.array([float((r0(w) - 1) / w) for w in ws])
    d0 = np.array([float(delta0(w)) for w in ws])
    out = {"QW": fit(ws, qw, 10, rel=True), "D0": fit(ws, d0, 10)}
    w1 = np.linspace(W_LO, W_HI, 35)
    d1 = []
    for w in w1:
        lam_res = [float((xstar(lam, w * np.sqrt(lam)) - lam * r0(w) - delta0(w)) * lam) for lam in (100, 200)]
        d1.append(2 * lam_res[1] - lam_res[0])
    out["D1"] = fit(w1, np.array(d1), 6)

    def h(lw):
        x = np.sqrt(-np.expm1(-lw))
        return special.erfinv(x) / x
    t = np.cos(np.linspace(0, np.pi, 4000))
    lw = t * 2.5 + 2.5 + 1e-12
    out["EC"] = fit(lw - 2.5, h(lw), 9, rel=True)
    lw = t * (16.7 - 5.0) / 2 + (16.7 + 5.0) / 2
    out["ET"] = fit(np.sqrt(lw) - 3.0, h(lw), 9, rel=True)
    return out


if __name__ == "__main__":
    def lit(v):
        t = "%.9g" % v
        return (t if ("." in t or "e" in t) else t + ".0") + "f"
    for name, (c, err) in tables().items():
        print("// %s: max fit error %.2e" % (name, err))
        print("__device__ constexpr float POIS_%s[%d] = {%s};" % (name, len(c), ", ".join(lit(v) for v in c)))
