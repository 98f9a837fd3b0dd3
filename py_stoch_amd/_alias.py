"""The reference's top-level import names, served by this package.

The reference's scripts import ``libsbml``, ``mod.*`` and ``sim_maker`` as
top-level names (examples/lotkaVolterra_sim.py:1-4, sim_maker.py:1-2,
mod/simulator/__init__.py:1).  With ``py_stoch_amd/dropin`` on ``sys.path``
those names resolve to this package: the first such import runs one of the
small files in that directory, which calls :func:`install`.  install() puts
an import hook first on ``sys.meta_path`` that answers every ``mod.X.Y``
with the module object ``py_stoch_amd.mod.X.Y`` itself (the same object, so
classes and module state are shared, not copied), ``sim_maker`` with
``py_stoch_amd.sim_maker`` and ``libsbml`` with the package's SBML reader
``py_stoch_amd.mod.petri_net.sbml`` (which SPN reads models from).
"""
import importlib
import importlib.abc
import importlib.util
import sys

ALIASES = {
    "mod": "py_stoch_amd.mod",
    "sim_maker": "py_stoch_amd.sim_maker",
    "libsbml": "py_stoch_amd.mod.petri_net.sbml",
}


def real_name(fullname):
    """The module of this package behind a reference import name, or None."""
    head, dot, rest = fullname.partition(".")
    if head not in ALIASES or (dot and head != "mod"):
        return None
    return ALIASES[head] + dot + rest


class _AliasFinder(importlib.abc.MetaPathFinder, importlib.abc.Loader):
    def find_spec(self, fullname, path=None, target=None):
        real = real_name(fullname)
        if real is None:
            return None
        try:
            module = importlib.import_module(real)
        except ModuleNotFoundError:
            return None
        spec = importlib.util.spec_from_loader(fullname, self, is_package=hasattr(module, "__path__"))
        spec.loader_state = (module, module.__spec__)
        return spec

    def create_module(self, spec):
        return spec.loader_state[0]

    def exec_module(self, module):
        # the module already ran under its own name; keep that name's spec
        alias_spec = module.__spec__
        module.__spec__ = alias_spec.loader_state[1]


_FINDER = _AliasFinder()


def install(name=None):
    """Install the import hook (once).  Called with the name of a top-level
    alias module that is being imported, also make that name refer to the
    package's module."""
    if _FINDER not in sys.meta_path:
        sys.meta_path.insert(0, _FINDER)
    if name in ALIASES:
        sys.modules[name] = importlib.import_module(ALIASES[name])
