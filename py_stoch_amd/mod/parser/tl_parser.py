"""Tau-leaping arguments and parser (reference: mod/parser/tl_parser.py:8-120)."""
import re

import numpy as np
import sympy

from .parser import Parser
from . import mass_action


class TlArgs:
    def __init__(self):
        # Definitions follow the cuTauLeaping paper (Nobile et al. 2014), as in
        # tl_parser.py:8-47.
        self.c = []       # parameter values
        self.x_0 = []     # initial species amounts

        self.A = []       # flattened Pre
        self.V = []       # flattened Post - Pre
        self.V_t = []     # flattened transpose of V
        self.V_bar = []   # flattened |V * StoichMConst|

        self.M = 0        # reactions
        self.N = 0        # species
        self.A_size = 0
        self.V_size = 0
        self.V_t_size = 0
        self.V_bar_size = 0

        self.hazards = []  # C source of UpdatePropensities (compatibility only)

        self.H = []        # highest order of reaction per species
        self.H_type = []

        self.E = np.array([0]).astype(np.int32)  # recorded species
        self.t_max = 1     # end time
        self.ita = 11      # number of sample times
        self.U = 128       # number of simulations

        self.kappa = len(self.E)
        self.I = np.array([])  # sample times (float32)

        self.n_c = 10      # critical-reaction threshold
        self.eta = 0.03    # tau-selection error control

        self.gillespie = 0  # 1 forces SSA

        # additions of this implementation (defaults keep the reference's behaviour)
        self.precision = "fp64"   # "fp64" (plain) or "mixed" (FP32 propensities / tau bound)
        self.seed = 0             # Philox key; the reference seeds by thread id only
        self.sim_offset = 0       # global id of simulation 0
        self.ssa_burst = 100      # SSA steps per fallback burst (P3.py:71)
        self.prop_order = None    # mass-action table, see mass_action.py
        self.prop_step = None
        self.rate_fn = None       # parameters c -> per-reaction coefficients k_j


class TlParser(Parser):
    @staticmethod
    def parse(args_out, stochastic_petri_net):
        spn = stochastic_petri_net
        args_out.c = np.array(spn.c).astype(np.float64)
        args_out.x_0 = np.array(spn.M).astype(np.int32)

        ma = np.asarray(spn.Pre)
        mb = np.asarray(spn.Post)
        mv = mb - ma
        mv_t = mv.transpose()
        mv_bar = abs(mv * np.asarray(spn.StoichMConst))

        args_out.A, args_out.A_size = TlParser.flatten_matrix(ma)
        args_out.V, args_out.V_size = TlParser.flatten_matrix(mv)
        args_out.V_t, args_out.V_t_size = TlParser.flatten_matrix(mv_t)
        args_out.V_bar, args_out.V_bar_size = TlParser.flatten_matrix(mv_bar)

        args_out.M = len(spn.T)
        args_out.N = len(spn.P)

        args_out.H, args_out.H_type = TlParser.get_hors(spn)
        args_out.hazards = TlParser.define_hazards(spn)

        n_params = len(np.atleast_1d(np.asarray(spn.c)))
        # mass-action table, or C expressions compiled into the kernel
        order, step, rate_fn, hazard_c = mass_action.propensity_tables(spn.h, args_out.N, n_params)
        args_out.prop_order = order
        args_out.prop_step = step
        args_out.rate_fn = rate_fn
        args_out.hazard_c = hazard_c
        args_out.n_params = n_params

        args_out.kappa = len(args_out.E)
        args_out.I = np.linspace(0, args_out.t_max, num=args_out.ita).astype(np.float32)
        return args_out

    @staticmethod
    def get_reaction_orders(hazards, species):
        """Total degree of each hazard in the species symbols (tl_parser.py:80-95)."""
        syms = sympy.symbols(list(species.keys()))
        if not isinstance(syms, (list, tuple)):
            syms = [syms]
        orders = []
        for hazard in hazards:
            py_hazard = str(sympy.expand(hazard.replace("^", "**")))
            orders.append(sympy.Poly(py_hazard, *syms).total_degree())
        return orders

    @staticmethod
    def get_hors(spn):
        """Highest order of reaction H and its type H_type per species
        (tl_parser.py:97-120): for every reaction whose hazard mentions the
        species, H is the largest reaction order and H_type the largest Pre
        stoichiometry of the species over all those reactions."""
        hors = np.zeros(len(spn.P))
        hors_type = np.zeros(len(spn.P))
        orders = TlParser.get_reaction_orders(spn.h, spn.P)
        for reaction, r_idx in spn.T.items():
            for species, s_idx in spn.P.items():
                hazard = spn.h[r_idx]
                if re.search("(?<![a-zA-Z0-9_])" + species + "(?![a-zA-Z0-9_])", hazard) is not None:
                    if hors[s_idx] <= orders[r_idx]:
                        hors[s_idx] = orders[r_idx]
                    stoich = spn.Pre[s_idx][r_idx]
                    if stoich > hors_type[s_idx]:
                        hors_type[s_idx] = int(stoich)
        return np.array(hors).astype(np.int32), np.array(hors_type).astype(np.int32)


def tables_of(args):
    """The network tables of a parsed TlArgs/GArgs in the C ABI's layout."""
    N, M = int(args.N), int(args.M)
    A = np.asarray(args.A, np.int32).reshape(-1, 4) if getattr(args, "A", None) is not None and len(args.A) else np.zeros((0, 4), np.int32)
    V = np.asarray(args.V, np.int32).reshape(-1, 4)
    H = np.asarray(getattr(args, "H", np.zeros(N)), np.int32).reshape(N) if len(getattr(args, "H", [])) else np.zeros(N, np.int32)
    Ht = np.asarray(getattr(args, "H_type", np.zeros(N)), np.int32).reshape(N) if len(getattr(args, "H_type", [])) else np.zeros(N, np.int32)
    return {"N": N, "M": M, "A": A, "V": V, "H": H, "H_type": Ht,
            "prop_order": np.asarray(args.prop_order, np.int32).reshape(N, M),
            "prop_step": np.asarray(args.prop_step, np.int32).reshape(N, M),
            "hazard_c": getattr(args, "hazard_c", None),
            "n_params": int(getattr(args, "n_params", 0))}
