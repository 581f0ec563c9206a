"""rfnetwork_amd: MI355X (gfx950) FDTD engine for rfnetwork's field solver.

The solver API (``FDTD_Solver``) follows rfnetwork.solver.FDTD_Solver; the time stepping runs in
hand-written HIP kernels (librfdtd.so, C ABI in include/rfdtd.h) reached through
``hip_func.solver_run_hip``, the drop-in for the reference's ``cuda_func.solver_run_cu``.
"""
from . import geometry, utils
from .geometry import Box, Cube, Line, Rectangle
from .solver import FDTD_Solver
from .units import conv, const

__all__ = ["FDTD_Solver", "Box", "Cube", "Line", "Rectangle", "const", "conv", "geometry", "utils"]
