"""specter_amd -- MI355X-native implementation of SPECTER's per-RK-substep hot
path (FC-Gram continuation, 3D/1D FFTs, pseudospectral operators,
pressure/Poisson projection and wall BCs) behind the libspx C ABI
(include/spx.h).

The compute path is libspx.so (hand-written HIP for gfx950); this package is
the host-side mirror of the reference interface.  Nothing here falls back to
host computation: a missing library raises SpxError.
"""
from .lib import SpxError, load_library, header_symbols  # noqa: F401
from .tables import load_dirichlet_tables  # noqa: F401
from .fftp import FCPlan  # noqa: F401
from .solver import Solver, SolverParams  # noqa: F401
from . import ic  # noqa: F401

__all__ = ["SpxError", "load_library", "header_symbols", "load_dirichlet_tables", "FCPlan",
           "Solver", "SolverParams", "ic"]
