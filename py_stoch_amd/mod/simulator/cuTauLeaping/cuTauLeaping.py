"""Tau-leaping simulator with SSA fallback on gfx950.

Drop-in for the reference's CuTauLeaping (mod/simulator/cuTauLeaping/
cuTauLeaping.py:17-84): ``CuTauLeaping().run(tl_args)`` returns a fresh
numpy int32 array ``[kappa][ita][U]`` and rounds ``tl_args.U`` up as the
reference does.  ``tl_args.gillespie = 1`` forces SSA (P1_P2.py:67-74).

Behind it, libpst_hip.so runs every simulation to the end in one kernel
launch (kernel_P1_P2 + kernel_P3 + the host loop fused per lane).  Counters
of the last run (events, leaps, rejected leaps, SSA steps, flagged
simulations, kernel time) are in ``self.last_stats``; per-simulation flag
bits in ``self.last_flags``.
"""
from ...simulator import StochasticSimulator
from .... import _native


def _method(args):
    return _native.METHOD_FORCED_SSA if int(getattr(args, "gillespie", 0)) else _native.METHOD_TAU_LEAPING


class CuTauLeaping(StochasticSimulator):
    def run(self, my_args):
        self.calculate_sizes(my_args)
        return self._run(my_args, _method(my_args))

    def run_batch(self, args_list):
        """Run a list of TlArgs (e.g. ``sim_maker.TLArgsList.make_list``) in
        one batched launch; returns the list of ``[kappa][ita][U]`` arrays
        (each ``U`` rounded up as by ``run``)."""
        return self._run_batch(list(args_list), _method)
