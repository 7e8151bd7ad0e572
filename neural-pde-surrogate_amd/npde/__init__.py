"""npde: MI355X-native 2D Euler finite-volume step (HIP kernels behind a C-ABI, reached by ctypes).

The compute path is libnpde.so only; importing npde.solver without it raises.
"""
from .solver import (GAMMA, LIMITERS, FLUXES, PERIODIC, TRANSMISSIVE, Batch, Grid, Solver, apply_bcs,
                     builtin_problem, make_grid, to_conserved, to_primitive)
from ._lib import NonPhysicalStateError, NpdeError

__all__ = ["GAMMA", "LIMITERS", "FLUXES", "PERIODIC", "TRANSMISSIVE", "Batch", "Grid", "Solver", "apply_bcs",
           "builtin_problem", "make_grid", "to_conserved", "to_primitive", "NonPhysicalStateError", "NpdeError"]
