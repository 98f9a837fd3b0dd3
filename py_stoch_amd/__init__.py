"""py_stoch_amd -- gfx950 (MI355X) implementation of py_stoch's simulator hot path.

Exact Gillespie SSA and cuTauLeaping-style tau-leaping, one simulation per GPU
lane, behind the reference's Python API:

    from py_stoch_amd import sim_maker
    from py_stoch_amd.mod.simulator import cuTauLeaping
    args_list = sim_maker.TLArgsList.make_list(model, species_init, params, ...)
    O = cuTauLeaping.CuTauLeaping().run(args_list[0])   # int32 [kappa][ita][U]

The module tree mirrors the reference's (``mod.parser``, ``mod.petri_net``,
``mod.simulator.cuTauLeaping``, ``mod.simulator.cuGillespie``, ``mod.utils``,
``sim_maker``) under the ``py_stoch_amd`` package prefix.
"""
__version__ = "0.1.0"
