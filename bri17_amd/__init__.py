"""bri17_amd: MI355X (gfx950) matrix-free periodic stiffness operator.

The compute path is libbri17_amd.so (hand-written HIP kernels behind the C
ABI of include/bri17_amd.h).  Importing the package does not touch the GPU.
"""
from . import _lib
from .operator import StiffnessMatrix
from .batch import Batch
from .server import Server
from . import pybri17

__all__ = ["StiffnessMatrix", "Batch", "Server", "pybri17"]
__version__ = "0.1.0"
