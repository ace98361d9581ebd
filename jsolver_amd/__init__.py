"""jsolver_amd: jsolver's geometric multigrid solver for AMD MI355X (gfx950).

Public modules mirror the reference package: grid_1d (grid_1D),
boundary_handler_2d / boundary_handler_3d (do_BCs), solver_2d (solve_2d and
wrappers), solver_3d (solve_3d and wrappers).  Compute runs in the HIP library
jsolver_amd/_lib/libjsolver_amd.so, loaded on first use; there is no CPU path.
"""
from .grid_1d import grid_1D  # noqa: F401

__all__ = ["grid_1D"]
