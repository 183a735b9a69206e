"""The reference's Python package name, ``spirit``, over the HIP engine.

Reference scripts run unchanged: ``from spirit import state, simulation``,
``import spirit.parameters.llg``, ``from spirit.scalar import scalar`` and the like
resolve to the modules of :mod:`spirit_amd`, which bind the same C API
(core/python/spirit/*.py) to ``spirit_amd/libSpirit.so``.  Every submodule below is
the very module object of ``spirit_amd``, so both names share one library handle
and one state.  ``hip`` holds the engine-only extensions.  Not provided (outside
this engine): ``htst``, ``parameters.ema`` and ``parameters.mmf``.
"""
import importlib
import sys

import spirit_amd

_MODULES = ("spiritlib", "scalar", "version", "constants", "log", "state", "system", "geometry", "hamiltonian",
            "configuration", "chain", "transition", "simulation", "quantities", "io", "hip", "parameters",
            "parameters.llg", "parameters.gneb", "parameters.mc")

for _name in _MODULES:
    _mod = importlib.import_module("spirit_amd." + _name)
    sys.modules[__name__ + "." + _name] = _mod
    if "." not in _name:
        globals()[_name] = _mod

__version__ = spirit_amd.__version__
