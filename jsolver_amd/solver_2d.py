"""2D multigrid solver: the reference's public API (jsolver/solver_2d.py:10-236)
on MI355X.

Every function keeps the reference's name, positional/keyword signature and
defaults, and returns ``(phi, count)`` with ``phi`` a float64 torch tensor on
the GPU (ghost-padded ``(my+3, mx+3)``) and ``count`` the number of cycles.
Inputs may be numpy arrays, torch tensors (CPU or GPU) or Python floats; they
are never modified.

Keyword-only additions to ``solve_2d`` (no counterpart in the reference):
  precision="fp64" | "mixed"  -- FP32 cycles inside FP64 iterative refinement
  stop="distance" | "rel_residual", rtol=1e-10 -- stop on ||rhs - L phi|| / ||rhs||
  inner_cycles=0      -- FP32 cycles per refinement step ("mixed"; 0 = predicted
                         from the measured contraction)
  check_every=1       -- "fp64" + rel_residual: cycles between residual checks
                         (0 = spaced by the predicted cycles-to-tolerance)
  return_history=False -- also return the per-cycle stop quantity (numpy)
  return_info=False   -- also return a dict (count, outer steps, last distance ...)
  boundary="zero" | "phi" -- "phi": Dirichlet values are the boundary layers of
                         the incoming phi (set them with do_BCs(phi, vals_...));
                         the result keeps them.  The reference's solver always
                         uses zeros (README.md:90).  rel_residual is then
                         relative to the right-hand side with the boundary
                         values moved into it.
"""
import numpy as np

from . import _solver
from ._solver import solve_batched_2d  # noqa: F401  (batched extension, no reference counterpart)

INT32_MAX = int(np.iinfo(np.int32).max)


def solve_2d(grid_x, grid_y, phi, rhs, lam, D_xx, D_yy, *, D_xy=None, fixed=False, epsilon=1e-12,
             gamma=1, fcycle=False, skip_pre=False, nu1=1, nu2=1, max_iters_outer=INT32_MAX,
             max_iters_inner=INT32_MAX, **extra):
    """Solve div(D grad phi) - lam phi = rhs with zero Dirichlet boundaries
    (solver_2d.py:196-236).  grid_x: columns, grid_y: rows.  D_xx/D_yy either
    both scalars or both arrays; D_xy needs array D.  Raises ValueError for a
    non-linear grid, a resolution that is not a power of 2 or not efficient for
    multigrid, a dtype other than float64, or incompatible shapes."""
    return _solver.solve(2, (grid_x, grid_y), phi, rhs, lam, (D_xx, D_yy), D_xy, fixed=fixed,
                         epsilon=epsilon, gamma=gamma, fcycle=fcycle, skip_pre=skip_pre, nu1=nu1,
                         nu2=nu2, max_iters_outer=max_iters_outer,
                         max_iters_inner=max_iters_inner, **extra)


class Solver_2D:
    """Configuration holder with the reference's constructor
    (solver_2d.py:247-277); ``solve`` runs the GPU solve."""

    def __init__(self, grid_x, grid_y, fixed=False, epsilon=1e-12, gamma=1, fcycle=False,
                 skip_pre=False, nu1=1, nu2=1, max_iters_outer=INT32_MAX, max_iters_inner=INT32_MAX):
        self.levels, (self.mx_level, self.my_level) = _solver.level_layout((grid_x.mx, grid_y.mx))
        _solver.check_grids((grid_x, grid_y))
        self.grid_x, self.grid_y = grid_x, grid_y
        self.arr_shape = [(int(self.my_level[l]) + 3, int(self.mx_level[l]) + 3) for l in range(self.levels)]
        self.idel_x = grid_x.idel * 0.5 ** np.arange(self.levels)
        self.idel_y = grid_y.idel * 0.5 ** np.arange(self.levels)
        self.kw = dict(fixed=fixed, epsilon=epsilon, gamma=gamma, fcycle=fcycle, skip_pre=skip_pre,
                       nu1=nu1, nu2=nu2, max_iters_outer=max_iters_outer,
                       max_iters_inner=max_iters_inner)

    def solve(self, phi, rhs, lam, D_xx, D_yy, D_xy=None):
        return solve_2d(self.grid_x, self.grid_y, phi, rhs, lam, D_xx, D_yy, D_xy=D_xy, **self.kw)


# ---- positional wrappers (solver_2d.py:10-194) ----------------------------
def solve_2d_simple(grid_x, grid_y, phi, rhs, lam, D_xx, D_yy, D_xy=None):
    return solve_2d(grid_x=grid_x, grid_y=grid_y, phi=phi, rhs=rhs, lam=lam, D_xx=D_xx, D_yy=D_yy, D_xy=D_xy)


def solve_2d_fixed_simple(grid_x, grid_y, iters, phi, rhs, lam, D_xx, D_yy, D_xy=None):
    return solve_2d_fixed(grid_x, grid_y, iters, 1, phi, rhs, lam, D_xx, D_yy, D_xy)


def solve_2d_fixed(grid_x, grid_y, iters_outer, iters_inner, phi, rhs, lam, D_xx, D_yy, D_xy=None):
    return solve_2d(grid_x, grid_y, phi, rhs, lam, D_xx, D_yy, D_xy=D_xy, fixed=True,
                    max_iters_outer=iters_outer, max_iters_inner=iters_inner)


def solve_2d_skip_pre_simple(grid_x, grid_y, phi, rhs, lam, D_xx, D_yy, D_xy=None):
    return solve_2d(grid_x, grid_y, phi, rhs, lam, D_xx, D_yy, D_xy=D_xy, skip_pre=True)


def solve_2d_fixed_skip_pre_simple(grid_x, grid_y, iters, phi, rhs, lam, D_xx, D_yy, D_xy=None):
    return solve_2d_fixed_skip_pre(grid_x, grid_y, iters, 1, phi, rhs, lam, D_xx, D_yy, D_xy)


def solve_2d_fixed_skip_pre(grid_x, grid_y, iters_outer, iters_inner, phi, rhs, lam, D_xx, D_yy, D_xy=None):
    return solve_2d(grid_x, grid_y, phi, rhs, lam, D_xx, D_yy, D_xy=D_xy, fixed=True,
                    max_iters_outer=iters_outer, max_iters_inner=iters_inner, skip_pre=True)


def solve_2d_fcycle_simple(grid_x, grid_y, phi, rhs, lam, D_xx, D_yy, D_xy=None):
    return solve_2d(grid_x, grid_y, phi, rhs, lam, D_xx, D_yy, D_xy=D_xy, fcycle=True)


def solve_2d_fixed_fcycle_simple(grid_x, grid_y, iters, phi, rhs, lam, D_xx, D_yy, D_xy=None):
    return solve_2d_fixed_fcycle(grid_x, grid_y, iters, 1, phi, rhs, lam, D_xx, D_yy, D_xy)


def solve_2d_fixed_fcycle(grid_x, grid_y, iters_outer, iters_inner, phi, rhs, lam, D_xx, D_yy, D_xy=None):
    return solve_2d(grid_x, grid_y, phi, rhs, lam, D_xx, D_yy, D_xy=D_xy, fixed=True,
                    max_iters_outer=iters_outer, max_iters_inner=iters_inner, fcycle=True)


def solve_2d_fcycle_skip_pre_simple(grid_x, grid_y, phi, rhs, lam, D_xx, D_yy, D_xy=None):
    return solve_2d(grid_x, grid_y, phi, rhs, lam, D_xx, D_yy, D_xy=D_xy, fcycle=True, skip_pre=True)


def solve_2d_fixed_fcycle_skip_pre_simple(grid_x, grid_y, iters, phi, rhs, lam, D_xx, D_yy, D_xy=None):
    return solve_2d_fixed_fcycle_skip_pre(grid_x, grid_y, iters, 1, phi, rhs, lam, D_xx, D_yy, D_xy)


def solve_2d_fixed_fcycle_skip_pre(grid_x, grid_y, iters_outer, iters_inner, phi, rhs, lam, D_xx, D_yy, D_xy=None):
    return solve_2d(grid_x, grid_y, phi, rhs, lam, D_xx, D_yy, D_xy=D_xy, fixed=True,
                    max_iters_outer=iters_outer, max_iters_inner=iters_inner, skip_pre=True, fcycle=True)


def solve_2d_wcycle_simple(grid_x, grid_y, phi, rhs, lam, D_xx, D_yy, D_xy=None):
    return solve_2d(grid_x, grid_y, phi, rhs, lam, D_xx, D_yy, D_xy=D_xy, gamma=2)


def solve_2d_fixed_wcycle_simple(grid_x, grid_y, iters, phi, rhs, lam, D_xx, D_yy, D_xy=None):
    return solve_2d_fixed_wcycle(grid_x, grid_y, iters, 1, phi, rhs, lam, D_xx, D_yy, D_xy)


def solve_2d_fixed_wcycle(grid_x, grid_y, iters_outer, iters_inner, phi, rhs, lam, D_xx, D_yy, D_xy=None):
    return solve_2d(grid_x, grid_y, phi, rhs, lam, D_xx, D_yy, D_xy=D_xy, fixed=True,
                    max_iters_outer=iters_outer, max_iters_inner=iters_inner, gamma=2)


def solve_2d_wcycle_skip_pre_simple(grid_x, grid_y, phi, rhs, lam, D_xx, D_yy, D_xy=None):
    return solve_2d(grid_x, grid_y, phi, rhs, lam, D_xx, D_yy, D_xy=D_xy, gamma=2, skip_pre=True)


def solve_2d_fixed_wcycle_skip_pre_simple(grid_x, grid_y, iters, phi, rhs, lam, D_xx, D_yy, D_xy=None):
    return solve_2d_fixed_wcycle_skip_pre(grid_x, grid_y, iters, 1, phi, rhs, lam, D_xx, D_yy, D_xy)


def solve_2d_fixed_wcycle_skip_pre(grid_x, grid_y, iters_outer, iters_inner, phi, rhs, lam, D_xx, D_yy, D_xy=None):
    return solve_2d(grid_x, grid_y, phi, rhs, lam, D_xx, D_yy, D_xy=D_xy, fixed=True,
                    max_iters_outer=iters_outer, max_iters_inner=iters_inner, skip_pre=True, gamma=2)
