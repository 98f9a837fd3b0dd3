"""Kinetic laws that are not a mass-action product: the parser keeps them as
C expressions (as the reference splices hazards into its kernel,
mod/parser/parser.py:32-55) and the C ABI validates them before they reach
the compiler.  No GPU needed."""
import numpy as np
import pytest

from py_stoch_amd import models
from py_stoch_amd.mod.parser import mass_action
from py_stoch_amd.mod.parser.g_parser import GArgs, GParser
from py_stoch_amd.mod.parser.tl_parser import tables_of


def test_printer_forms():
    code = mass_action.hazard_code(
        ["parameter_0*species_0^2", "parameter_1*species_0/(parameter_2 + species_0)",
         "parameter_0 + parameter_3*species_1", "exp(-parameter_0*species_0)*0.5", "species_1^(1/2)/3"], 2, 4)
    assert code[0] == "C(0)*(X(0)*X(0))"
    assert code[1] == "C(1)*X(0)/(C(2) + X(0))"
    assert code[2] == "C(0) + C(3)*X(1)"
    assert code[3] == "R(0.5)*exp(-C(0)*X(0))"
    assert "sqrt(X(1))" in code[4] and "R(3.0)" in code[4]


def test_mass_action_networks_keep_the_table():
    a = models.make_args(models.lotka_volterra(), E=[0, 1])
    assert a.hazard_c is None and tables_of(a)["hazard_c"] is None


def test_michaelis_menten_is_compiled():
    a = GArgs()
    a.E, a.t_max, a.ita, a.U = [0], 10.0, 11, 64
    a = GParser.parse(a, models.michaelis_menten_birth_death())
    assert a.hazard_c == ["C(0)", "C(1)*X(0)/(C(2) + X(0))"]
    assert list(a.rate_fn(a.c)) == [20.0, 25.0, 10.0]  # rows carry the parameters


def test_polynomial_non_product_hazard_keeps_reaction_orders():
    a = models.make_args(models.immigration_birth_death(), E=[0])
    assert a.hazard_c == ["C(0) + C(1)*X(0)", "C(2)*X(0)"]
    assert list(a.H) == [1] and list(a.H_type) == [1]


def _plan_error(native, hazards, specialize=True):
    a = GArgs()
    a.E, a.t_max, a.ita, a.U = [0], 10.0, 11, 64
    a = GParser.parse(a, models.michaelis_menten_birth_death())
    tabs = tables_of(a)
    tabs["hazard_c"] = hazards
    with pytest.raises(native.NativeError) as e:
        native.Plan(tabs, a.rate_fn(a.c), a.x_0, method=2, t_max=10.0, I=a.I, E=a.E, n_sims=64,
                    specialize=specialize)
    return str(e.value)


def test_c_abi_validates_hazards(native):
    assert "specialize = 1" in _plan_error(native, ["C(0)", "C(1)*X(0)"], specialize=False)
    assert "not allowed" in _plan_error(native, ["C(0); while(1)", "C(1)"])
    assert "unknown name 'printf'" in _plan_error(native, ["printf(0)", "C(1)"])
    assert "unbalanced" in _plan_error(native, ["(C(0)", "C(1)"])
    assert "not allowed" in _plan_error(native, ["C(0)", "#include"])
    assert "empty" in _plan_error(native, ["", "C(1)"])
    # indexes are integer literals inside the network's species / parameter counts
    assert "X(5) is out of range" in _plan_error(native, ["C(0)*X(5)", "C(1)"])
    assert "C(7) is out of range" in _plan_error(native, ["C(7)", "C(1)"])
    assert "integer literal" in _plan_error(native, ["C(0)*X(0+1)", "C(1)"])
    assert "integer literal" in _plan_error(native, ["C(0)*X(-1)", "C(1)"])
    # a comment could swallow the next hazard's assignment
    assert "comments" in _plan_error(native, ["C(0) /* ", "*/ C(1)"])
    assert "comments" in _plan_error(native, ["C(0) // x", "C(1)"])
