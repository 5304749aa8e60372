"""doug_amd: DOUG's two-level Schwarz PCG solve path on AMD Instinct MI355X.

The compute path is libdoug_hip.so (HIP kernels for gfx950) behind the C-ABI
declared in include/doug.h; this package is its Python binding plus the
DOUG.ctl / file-format readers and synthetic problem generators.
"""
from . import problems  # noqa: F401
from ._lib import LIB_PATH, load  # noqa: F401
from .solver import DougError, Solver, default_params, make_params, partition_graph  # noqa: F401

__all__ = ["Solver", "DougError", "make_params", "default_params", "partition_graph", "problems", "load",
           "LIB_PATH"]
