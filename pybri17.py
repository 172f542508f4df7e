"""Drop-in ``pybri17`` module.

The reference builds a compiled module named ``pybri17``
(``PYBIND11_MODULE(pybri17, m)``, python/pybri17.cpp:98-107) and its callers
write ``import pybri17`` (python/demo.py:3).  This module is that name for the
implementation in ``bri17_amd.pybri17``: the same classes
(CartesianGrid2f64 / CartesianGrid3f64 / Hooke2f64 / Hooke3f64, plus the
``...f32`` variants), methods, layout checks (ValueError) and the module
attributes ``__author__`` and ``__version__``.  The per-frequency functions
run in libbri17_amd.so on the host and do not need a GPU.
"""
from bri17_amd.pybri17 import *  # noqa: F401,F403
from bri17_amd.pybri17 import __all__, __author__, __implementation__, __version__  # noqa: F401
