"""Pure-SSA arguments and parser (reference: mod/parser/g_parser.py:5-52)."""
import numpy as np

from .parser import Parser
from . import mass_action


class GArgs:
    def __init__(self):
        self.c = []
        self.x_0 = []
        self.V = []
        self.M = 0
        self.N = 0
        self.V_size = 0
        self.hazards = []
        self.E = np.array([0]).astype(np.int32)
        self.t_max = 1
        self.ita = 11
        self.U = 128
        self.kappa = len(self.E)
        self.I = np.array([])
        # additions of this implementation
        self.precision = "fp64"
        self.seed = 0
        self.sim_offset = 0
        self.prop_order = None
        self.prop_step = None
        self.rate_fn = None


class GParser(Parser):
    @staticmethod
    def parse(args_out, stochastic_petri_net):
        spn = stochastic_petri_net
        args_out.c = np.array(spn.c).astype(np.float64)
        args_out.x_0 = np.array(spn.M).astype(np.int32)
        mv = np.asarray(spn.Post) - np.asarray(spn.Pre)
        args_out.V, args_out.V_size = GParser.flatten_matrix(mv)
        args_out.M = len(spn.T)
        args_out.N = len(spn.P)
        args_out.hazards = GParser.define_hazards(spn)
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
