from .cuTauLeaping import CuTauLeaping
