"""Top-level `mod` of the reference's scripts, served by py_stoch_amd (see py_stoch_amd/_alias.py)."""
import os
import sys

try:
    import py_stoch_amd._alias as _alias
except ImportError:  # only this directory is on sys.path: add the repository root
    sys.path.append(os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
    import py_stoch_amd._alias as _alias

_alias.install(__name__)
