"""The reference's own import lines run unchanged with py_stoch_amd/dropin on
sys.path (examples/lotkaVolterra_sim.py:1-4, 13-21, 57-61, sim_maker.py:1-2),
and resolve to this package's module objects -- not copies."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN = os.path.join(ROOT, "py_stoch_amd", "dropin")

SCRIPT = r'''
import libsbml
from mod.simulator import cuTauLeaping
import sim_maker
from mod.utils.saver import save_results
from mod.parser import TlArgs, TlParser, GArgs, GParser, Parser
from mod.petri_net import SPN
from mod.simulator.cuGillespie.cuGillespie import CuGillespie
import mod.simulator.simulator

import py_stoch_amd
import py_stoch_amd.mod.simulator.cuTauLeaping as real_tl
import py_stoch_amd.mod.petri_net.sbml as real_sbml
import py_stoch_amd.sim_maker as real_sm
import py_stoch_amd.mod.utils.saver as real_saver
import py_stoch_amd.mod.simulator.cuGillespie.cuGillespie as real_g

assert cuTauLeaping is real_tl and cuTauLeaping.CuTauLeaping is real_tl.CuTauLeaping
assert libsbml is real_sbml and sim_maker is real_sm and save_results is real_saver.save_results
assert CuGillespie is real_g.CuGillespie
assert mod.simulator.simulator.StochasticSimulator is py_stoch_amd.mod.simulator.simulator.StochasticSimulator
assert real_tl.__spec__.name == "py_stoch_amd.mod.simulator.cuTauLeaping"

reader = libsbml.SBMLReader()
document = reader.readSBML(SBML)
assert document.getNumErrors() == 0
sbml_model = document.getModel()
tl_args_list = sim_maker.TLArgsList.make_list(sbml_model, [[1000], [1000]], [[10], [0.01, 0.02], [10]],
                                              [0, 1], 10, 1001, 1280)
assert len(tl_args_list) == 2 and isinstance(tl_args_list[0], TlArgs)
assert list(tl_args_list[1].c) == [10.0, 0.02, 10.0]
print("DROPIN_OK", py_stoch_amd.__file__)
'''


def test_reference_imports_resolve_to_this_package(tmp_path):
    env = dict(os.environ)
    env["PYTHONPATH"] = DROPIN  # the repository root is not on the path
    sbml = os.path.join(ROOT, "tests", "golden", "lotka_volterra.xml")
    r = subprocess.run([sys.executable, "-c", "SBML = %r\n" % sbml + SCRIPT], cwd=str(tmp_path), env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "DROPIN_OK " + os.path.join(ROOT, "py_stoch_amd", "__init__.py") in r.stdout
