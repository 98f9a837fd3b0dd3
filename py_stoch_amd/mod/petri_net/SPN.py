"""Stochastic Petri net built from an SBML model.

Same fields and semantics as the reference's SPN (mod/petri_net/SPN.py:5-150):

* ``Pre``, ``Post``, ``StoichMConst``: float64 [N x M] reactant, product and
  constant-species matrices (SPN.py:71-97);
* ``M``: initial amounts (SPN.py:99-104); ``c``: model parameter values
  (SPN.py:144-150);
* ``P``: {'species_i': i}; ``T``: {reaction id: j} (SPN.py:47-62);
* ``h``: hazard strings over ``species_i`` / ``parameter_k`` names
  (SPN.py:106-142).  The reference formats them with
  libsbml.formulaToL3String; here they are sympy's rendering of the same
  expression (same names, same mathematics, possibly different term order).

Compartment identifiers in a kinetic law are replaced by the compartment size,
and kinetic-law local parameters by their values.
"""
import numpy as np
import sympy

from . import sbml as _sbml


class SPN:
    def __init__(self):
        self.Pre = []
        self.Post = []
        self.M = []
        self.h = []
        self.c = []
        self.P = {}
        self.T = {}
        self.StoichMConst = []

    def sbml_2_stochastic_petri_net(self, sbml_model):
        self.Pre, self.Post, self.StoichMConst = self.get_stoichiometries(sbml_model)
        self.M = self.get_initial_state_vector(sbml_model)
        self.c = self.get_rate_constants_vector(sbml_model)
        constants = self.get_rate_constants(sbml_model)
        self.P = self.get_species_vector(sbml_model)
        self.T = self.get_reactions_vector(sbml_model)
        self.h = self.get_hazards_vector(sbml_model, self.P, constants)

    @staticmethod
    def get_species_vector(sbml_model):
        return {"species_" + str(i): i for i, _ in enumerate(sbml_model.getListOfSpecies())}

    @staticmethod
    def get_reactions_vector(sbml_model):
        return {r.getId(): j for j, r in enumerate(sbml_model.getListOfReactions())}

    @staticmethod
    def get_rate_constants(sbml_model):
        return {p.getId(): k for k, p in enumerate(sbml_model.getListOfParameters())}

    @staticmethod
    def get_stoichiometries(sbml_model):
        species = sbml_model.getListOfSpecies()
        reactions = sbml_model.getListOfReactions()
        pre = np.zeros([len(species), len(reactions)])
        post = np.zeros_like(pre)
        const = np.zeros_like(pre)
        for j, reaction in enumerate(reactions):
            reactants = {}
            for r in reaction.getListOfReactants():
                reactants[r.getSpecies()] = reactants.get(r.getSpecies(), 0) + r.getStoichiometry()
            products = {}
            for p in reaction.getListOfProducts():
                products[p.getSpecies()] = products.get(p.getSpecies(), 0) + p.getStoichiometry()
            for i, s in enumerate(species):
                sid = s.getId()
                if sid in reactants:
                    pre[i, j] = int(reactants[sid])
                if sid in products:
                    post[i, j] = int(products[sid])
                if s.getConstant():
                    const[i, j] = 1
        return pre, post, const

    @staticmethod
    def get_initial_state_vector(sbml_model):
        return [s.getInitialAmount() for s in sbml_model.getListOfSpecies()]

    @staticmethod
    def get_hazards_vector(sbml_model, species_dict, param_dict):
        species_ids = [s.getId() for s in sbml_model.getListOfSpecies()]
        param_ids = [p.getId() for p in sbml_model.getListOfParameters()]
        comp = sbml_model.getCompartmentSizes() if hasattr(sbml_model, "getCompartmentSizes") else {}
        h = []
        for reaction in sbml_model.getListOfReactions():
            kl = reaction.getKineticLaw()
            expr = kl.getMath().expr
            subs = {}
            # species first, then parameters, as SPN.py:111-114 renames them
            for i, sid in enumerate(species_ids):
                subs[sympy.Symbol(sid)] = sympy.Symbol("species_" + str(i))
            for pid in param_ids:
                if sympy.Symbol(pid) not in subs:
                    subs[sympy.Symbol(pid)] = sympy.Symbol("parameter_" + str(param_dict[pid]))
            locals_ = kl.getLocalParameters() if hasattr(kl, "getLocalParameters") else {}
            for name, val in locals_.items():
                subs[sympy.Symbol(name)] = sympy.Float(val)
            for name, size in comp.items():
                if sympy.Symbol(name) not in subs:
                    subs[sympy.Symbol(name)] = sympy.Float(size) if size != int(size) else sympy.Integer(int(size))
            renamed = expr.xreplace(subs)
            unknown = [str(s) for s in renamed.free_symbols
                       if not (str(s).startswith("species_") or str(s).startswith("parameter_"))]
            if unknown:
                raise ValueError("kinetic law of reaction %s uses undefined names %s"
                                 % (reaction.getId(), sorted(unknown)))
            h.append(str(renamed))
        return h

    @staticmethod
    def get_rate_constants_vector(sbml_model):
        return [p.getValue() for p in sbml_model.getListOfParameters()]


def from_sbml_file(path):
    """Convenience: read an SBML file into an SPN."""
    spn = SPN()
    spn.sbml_2_stochastic_petri_net(_sbml.read_model(path))
    return spn
