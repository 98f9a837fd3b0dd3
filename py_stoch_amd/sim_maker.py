"""Parameter sweeps -> list of TlArgs (reference: sim_maker.py:6-70)."""
import numpy

from .mod.parser import TlParser, TlArgs
from .mod.petri_net import SPN


class TLArgsList:
    @staticmethod
    def make_list(sbml_model, sweep_species_init, sweep_params, output_species, duration,
                  num_output_time_points, num_simulations, gillespie=0):
        """One TlArgs per combination of initial amounts and parameter values
        (the cartesian product of the per-position lists, sim_maker.py:8-43)."""
        species_init_lists = rest_list(sweep_species_init)
        param_lists = rest_list(sweep_params)

        sim_objects = []
        for species_init_list in species_init_lists:
            for param_list in param_lists:
                so = SimObj()
                so.species_init = species_init_list
                so.params = param_list
                sim_objects.append(so)

        base = SPN.SPN()
        base.sbml_2_stochastic_petri_net(sbml_model)
        tl_args_list = []
        for so in sim_objects:
            spn = SPN.SPN()
            spn.Pre, spn.Post, spn.StoichMConst = base.Pre, base.Post, base.StoichMConst
            spn.P, spn.T, spn.h = dict(base.P), dict(base.T), list(base.h)
            spn.M = numpy.array(so.species_init).astype(numpy.int32)
            spn.c = numpy.array(so.params).astype(numpy.float64)

            tl_args = TlArgs()
            tl_args.E = numpy.array(output_species).astype(numpy.int32)
            tl_args.t_max = duration
            tl_args.ita = num_output_time_points
            tl_args.U = num_simulations
            tl_args.gillespie = gillespie
            tl_args_list.append(TlParser.parse(tl_args, spn))
        return tl_args_list


def rest_list(list_in, index=0):
    """Cartesian product of a list of lists, as sim_maker.py:46-64 builds it
    (a single inner list is returned as is)."""
    if index == len(list_in) - 1:
        return list_in[index]
    sub_lists = rest_list(list_in, index + 1)
    out = []
    for i in range(len(list_in[index])):
        for j in range(len(sub_lists)):
            out.append([list_in[index][i]] + sub_lists[j] if isinstance(sub_lists[j], list)
                       else [list_in[index][i]] + [sub_lists[j]])
    return out


class SimObj:
    def __init__(self):
        self.species_init = []
        self.params = []
