"""aspect_amd — MI355X-native matrix-free Stokes solver (ASPECT's block GMG path).

The HIP library ``libaspect_mfs.so`` holds every kernel and the solver; this
package is the Python mirror of the StokesMatrixFreeHandler interface on top
of its C-ABI (``include/aspect_mfs.h``).
"""
from .mfs import (AVERAGE, BC_OPEN, BC_TANGENTIAL, BC_ZERO, Box, Ensemble, NoConvergence, SolverError,
                  SolverOutputs, StokesMatrixFreeHandler, TriaMultigrid, TriaOperators, load_library, make_box)

__all__ = ["AVERAGE", "BC_OPEN", "BC_TANGENTIAL", "BC_ZERO", "Box", "Ensemble", "NoConvergence", "SolverError",
           "SolverOutputs", "StokesMatrixFreeHandler", "TriaMultigrid", "TriaOperators", "load_library", "make_box"]
