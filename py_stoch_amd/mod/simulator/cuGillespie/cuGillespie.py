"""Single-launch pure SSA on gfx950.

Drop-in for the reference's CuGillespie (mod/simulator/cuGillespie/
cuGillespie.py:15-46, kernel gillespie.py:27-149): every simulation starts
from ``args.x_0`` and takes direct-method SSA steps while ``t < t_max``;
sample points not reached before t_max stay 0, as in the reference.
"""
from ...simulator import StochasticSimulator
from .... import _native


class CuGillespie(StochasticSimulator):
    def run(self, my_args):
        self.calculate_sizes(my_args)
        return self._run(my_args, _native.METHOD_PURE_SSA)
