"""Hazard strings -> mass-action propensity table.

The reference compiles each hazard string into CUDA (parser.py:32-55).  The
gfx950 kernels evaluate a table instead:

    a_j = k_j * prod_i prod_{q < p_ij} (x_i - s_ij * q)

with ``p_ij`` the number of factors of species i in reaction j and ``s_ij``
1 for a falling factorial x(x-1)...(x-p+1) or 0 for a power x^p.  ``k_j`` is a
function of the model parameters only, evaluated in float64 when a run starts.

Every hazard of the reference's example models has this form (simple, cdi,
Lotka-Volterra, Schloegl v1/v2, plasmid stability).  A hazard that does not
(non-polynomial or a polynomial that is not a single falling-factorial /
power product) raises :class:`UnsupportedHazard`.
"""
import itertools

import numpy as np
import sympy


class UnsupportedHazard(ValueError):
    pass


def _falling(x, p):
    r = sympy.Integer(1)
    for q in range(p):
        r = r * (x - q)
    return r


def analyse(hazards, n_species, n_params):
    """Return (prop_order [N x M] int32, prop_step [N x M] int32, coeffs) where
    coeffs[j] is a sympy expression in parameter_0..parameter_{P-1}."""
    S = [sympy.Symbol("species_%d" % i) for i in range(n_species)]
    Pm = [sympy.Symbol("parameter_%d" % k) for k in range(n_params)]
    names = {str(s): s for s in S + Pm}
    M = len(hazards)
    order = np.zeros((n_species, M), np.int32)
    step = np.ones((n_species, M), np.int32)
    coeffs = []
    for j, h in enumerate(hazards):
        expr = sympy.sympify(h.replace("^", "**"), locals=names)
        try:
            poly = sympy.Poly(sympy.expand(expr), *S) if S else None
        except sympy.PolynomialError as exc:
            raise UnsupportedHazard("hazard %d (%s) is not polynomial in the species: %s" % (j, h, exc))
        if poly is None or poly.is_zero:
            raise UnsupportedHazard("hazard %d (%s) is zero or has no species polynomial" % (j, h))
        degs = [poly.degree(s) for s in S]
        mono = tuple(max(d, 0) for d in degs)
        k = poly.coeff_monomial(mono)
        if k == 0:
            raise UnsupportedHazard("hazard %d (%s) has no leading monomial %s" % (j, h, mono))
        k = sympy.simplify(k)
        if any(sym in k.free_symbols for sym in S):
            raise UnsupportedHazard("hazard %d (%s) has a species-dependent coefficient" % (j, h))
        multi = [i for i in range(n_species) if mono[i] >= 2]
        found = None
        for combo in itertools.product((1, 0), repeat=len(multi)):
            st = {i: 1 for i in range(n_species)}
            for i, s in zip(multi, combo):
                st[i] = s
            cand = k
            for i in range(n_species):
                if mono[i] > 0:
                    cand = cand * (_falling(S[i], mono[i]) if st[i] == 1 else S[i] ** mono[i])
            if sympy.expand(expr - cand) == 0:
                found = st
                break
        if found is None:
            raise UnsupportedHazard(
                "hazard %d (%s) is not a mass-action product k * prod x_i(x_i-1)... or k * prod x_i^p" % (j, h))
        for i in range(n_species):
            order[i, j] = mono[i]
            step[i, j] = found[i]
        coeffs.append(k)
    return order, step, coeffs


class RateFunction:
    """Evaluates the per-reaction coefficients k_j from a parameter vector."""

    def __init__(self, coeffs, n_params):
        self.coeffs = list(coeffs)
        self.n_params = n_params
        syms = [sympy.Symbol("parameter_%d" % k) for k in range(n_params)]
        self._fns = [sympy.lambdify(syms, c, modules="math") for c in self.coeffs]

    def __call__(self, c):
        c = np.atleast_1d(np.asarray(c, np.float64)).reshape(-1)
        if c.shape[0] < self.n_params:
            raise ValueError("%d parameter values given, %d needed" % (c.shape[0], self.n_params))
        vals = [float(v) for v in c[: self.n_params]]
        return np.array([float(f(*vals)) for f in self._fns], np.float64)


# ---------------------------------------------------------------------------
# hazards that are not a mass-action product: compiled as C expressions
# ---------------------------------------------------------------------------
from sympy.printing.c import C99CodePrinter  # noqa: E402


class _HazardPrinter(C99CodePrinter):
    """C for one hazard, in the kernel's arithmetic type R: species as X(i)
    (the count converted to R), parameters as C(k), every number as R(...),
    small integer powers as products (the reference splices the hazard
    string into its kernel instead, parser.py:32-55)."""

    def _print_Symbol(self, expr):
        name = expr.name
        if name.startswith("species_"):
            return "X(%d)" % int(name[8:])
        if name.startswith("parameter_"):
            return "C(%d)" % int(name[10:])
        raise UnsupportedHazard("unknown symbol %s in a hazard" % name)

    def _print_Integer(self, expr):
        return "R(%d.0)" % int(expr)

    def _print_Rational(self, expr):
        return "(R(%d.0) / R(%d.0))" % (int(expr.p), int(expr.q))

    def _print_Float(self, expr):
        return "R(%s)" % repr(float(expr))

    def _print_Pow(self, expr):
        b, e = expr.as_base_exp()
        if e.is_Integer and 1 <= int(e) <= 6:
            return "(" + "*".join([self.parenthesize(b, 100)] * int(e)) + ")"
        if e.is_Integer and -6 <= int(e) <= -1:
            return "(R(1.0) / (" + "*".join([self.parenthesize(b, 100)] * (-int(e))) + "))"
        if e == sympy.Rational(1, 2):
            return "sqrt(%s)" % self._print(b)
        return "pow(%s, %s)" % (self._print(b), self._print(e))


def hazard_code(hazards, n_species, n_params):
    """One C expression per hazard (see _HazardPrinter)."""
    S = [sympy.Symbol("species_%d" % i) for i in range(n_species)]
    Pm = [sympy.Symbol("parameter_%d" % k) for k in range(n_params)]
    names = {str(s): s for s in S + Pm}
    pr = _HazardPrinter()
    out = []
    for j, h in enumerate(hazards):
        expr = sympy.sympify(h.replace("^", "**"), locals=names)
        bad = [s for s in expr.free_symbols if s not in S and s not in Pm]
        if bad:
            raise UnsupportedHazard("hazard %d (%s) uses unknown symbols %s" % (j, h, bad))
        out.append(pr.doprint(expr))
    return out


class ParameterRows:
    """Rates function of a network with compiled hazards: the kernel reads
    the parameter vector itself."""

    def __init__(self, n_params):
        self.n_params = n_params

    def __call__(self, c):
        c = np.atleast_1d(np.asarray(c, np.float64)).reshape(-1)
        if c.shape[0] < self.n_params:
            raise ValueError("%d parameter values given, %d needed" % (c.shape[0], self.n_params))
        return c[: self.n_params].copy() if self.n_params else np.zeros(1, np.float64)


def propensity_tables(hazards, n_species, n_params):
    """(prop_order, prop_step, rate_fn, hazard_c): the mass-action table when
    every hazard has that form (hazard_c None), else the C expressions."""
    try:
        order, step, coeffs = analyse(hazards, n_species, n_params)
        return order, step, RateFunction(coeffs, n_params), None
    except UnsupportedHazard:
        M = len(hazards)
        return (np.zeros((n_species, M), np.int32), np.ones((n_species, M), np.int32), ParameterRows(n_params),
                hazard_code(hazards, n_species, n_params))
