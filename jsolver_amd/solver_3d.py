"""3D multigrid solver: the reference's public API (jsolver/solver_3d.py:10-237)
on MI355X.

Same conventions as solver_2d: the reference's names, signatures and defaults;
returns ``(phi, count)`` with ``phi`` a float64 torch tensor on the GPU
(ghost-padded ``(mz+3, my+3, mx+3)``).  The keyword-only additions of
``solve_2d`` (precision, stop, rtol, inner_cycles, check_every, return_history,
return_info, boundary) apply here too.  There is no D_xy in 3D.
"""
import numpy as np

from . import _solver

INT32_MAX = int(np.iinfo(np.int32).max)


def solve_3d(grid_x, grid_y, grid_z, phi, rhs, lam, D_xx, D_yy, D_zz, *, fixed=False, epsilon=1e-12,
             gamma=1, fcycle=False, skip_pre=False, nu1=1, nu2=1, max_iters_outer=INT32_MAX,
             max_iters_inner=INT32_MAX, **extra):
    """Solve div(D grad phi) - lam phi = rhs in 3D (solver_3d.py:196-237).
    grid_x: third index, grid_y: second, grid_z: first.  D_xx/D_yy/D_zz all
    scalars or all arrays."""
    return _solver.solve(3, (grid_x, grid_y, grid_z), phi, rhs, lam, (D_xx, D_yy, D_zz), None,
                         fixed=fixed, epsilon=epsilon, gamma=gamma, fcycle=fcycle,
                         skip_pre=skip_pre, nu1=nu1, nu2=nu2, max_iters_outer=max_iters_outer,
                         max_iters_inner=max_iters_inner, **extra)


class Solver_3D:
    """Configuration holder with the reference's constructor
    (solver_3d.py:248-279); ``solve`` runs the GPU solve."""

    def __init__(self, grid_x, grid_y, grid_z, fixed=False, epsilon=1e-12, gamma=1, fcycle=False,
                 skip_pre=False, nu1=1, nu2=1, max_iters_outer=INT32_MAX, max_iters_inner=INT32_MAX):
        self.levels, (self.mx_level, self.my_level, self.mz_level) = _solver.level_layout(
            (grid_x.mx, grid_y.mx, grid_z.mx))
        _solver.check_grids((grid_x, grid_y, grid_z))
        self.grids = (grid_x, grid_y, grid_z)
        self.arr_shape = [(int(self.mz_level[l]) + 3, int(self.my_level[l]) + 3, int(self.mx_level[l]) + 3)
                          for l in range(self.levels)]
        self.idel_x = grid_x.idel * 0.5 ** np.arange(self.levels)
        self.idel_y = grid_y.idel * 0.5 ** np.arange(self.levels)
        self.idel_z = grid_z.idel * 0.5 ** np.arange(self.levels)
        self.kw = dict(fixed=fixed, epsilon=epsilon, gamma=gamma, fcycle=fcycle, skip_pre=skip_pre,
                       nu1=nu1, nu2=nu2, max_iters_outer=max_iters_outer,
                       max_iters_inner=max_iters_inner)

    def solve(self, phi, rhs, lam, D_xx, D_yy, D_zz):
        return solve_3d(*self.grids, phi, rhs, lam, D_xx, D_yy, D_zz, **self.kw)


# ---- positional wrappers (solver_3d.py:10-194) ----------------------------
def solve_3d_simple(grid_x, grid_y, grid_z, phi, rhs, lam, D_xx, D_yy, D_zz):
    return solve_3d(grid_x=grid_x, grid_y=grid_y, grid_z=grid_z, phi=phi, rhs=rhs, lam=lam, D_xx=D_xx, D_yy=D_yy, D_zz=D_zz)


def solve_3d_fixed_simple(grid_x, grid_y, grid_z, iters, phi, rhs, lam, D_xx, D_yy, D_zz):
    return solve_3d_fixed(grid_x, grid_y, grid_z, iters, 1, phi, rhs, lam, D_xx, D_yy, D_zz)


def solve_3d_fixed(grid_x, grid_y, grid_z, iters_outer, iters_inner, phi, rhs, lam, D_xx, D_yy, D_zz):
    return solve_3d(grid_x, grid_y, grid_z, phi, rhs, lam, D_xx, D_yy, D_zz, fixed=True,
                    max_iters_outer=iters_outer, max_iters_inner=iters_inner)


def solve_3d_skip_pre_simple(grid_x, grid_y, grid_z, phi, rhs, lam, D_xx, D_yy, D_zz):
    return solve_3d(grid_x, grid_y, grid_z, phi, rhs, lam, D_xx, D_yy, D_zz, skip_pre=True)


def solve_3d_fixed_skip_pre_simple(grid_x, grid_y, grid_z, iters, phi, rhs, lam, D_xx, D_yy, D_zz):
    return solve_3d_fixed_skip_pre(grid_x, grid_y, grid_z, iters, 1, phi, rhs, lam, D_xx, D_yy, D_zz)


def solve_3d_fixed_skip_pre(grid_x, grid_y, grid_z, iters_outer, iters_inner, phi, rhs, lam, D_xx, D_yy, D_zz):
    return solve_3d(grid_x, grid_y, grid_z, phi, rhs, lam, D_xx, D_yy, D_zz, fixed=True,
                    max_iters_outer=iters_outer, max_iters_inner=iters_inner, skip_pre=True)


def solve_3d_fcycle_simple(grid_x, grid_y, grid_z, phi, rhs, lam, D_xx, D_yy, D_zz):
    return solve_3d(grid_x, grid_y, grid_z, phi, rhs, lam, D_xx, D_yy, D_zz, fcycle=True)


def solve_3d_fixed_fcycle_simple(grid_x, grid_y, grid_z, iters, phi, rhs, lam, D_xx, D_yy, D_zz):
    return solve_3d_fixed_fcycle(grid_x, grid_y, grid_z, iters, 1, phi, rhs, lam, D_xx, D_yy, D_zz)


def solve_3d_fixed_fcycle(grid_x, grid_y, grid_z, iters_outer, iters_inner, phi, rhs, lam, D_xx, D_yy, D_zz):
    return solve_3d(grid_x, grid_y, grid_z, phi, rhs, lam, D_xx, D_yy, D_zz, fixed=True,
                    max_iters_outer=iters_outer, max_iters_inner=iters_inner, fcycle=True)


def solve_3d_fcycle_skip_pre_simple(grid_x, grid_y, grid_z, phi, rhs, lam, D_xx, D_yy, D_zz):
    return solve_3d(grid_x, grid_y, grid_z, phi, rhs, lam, D_xx, D_yy, D_zz, fcycle=True, skip_pre=True)


def solve_3d_fixed_fcycle_skip_pre_simple(grid_x, grid_y, grid_z, iters, phi, rhs, lam, D_xx, D_yy, D_zz):
    return solve_3d_fixed_fcycle_skip_pre(grid_x, grid_y, grid_z, iters, 1, phi, rhs, lam, D_xx, D_yy, D_zz)


def solve_3d_fixed_fcycle_skip_pre(grid_x, grid_y, grid_z, iters_outer, iters_inner, phi, rhs, lam, D_xx, D_yy, D_zz):
    return solve_3d(grid_x, grid_y, grid_z, phi, rhs, lam, D_xx, D_yy, D_zz, fixed=True,
                    max_iters_outer=iters_outer, max_iters_inner=iters_inner, skip_pre=True, fcycle=True)


def solve_3d_wcycle_simple(grid_x, grid_y, grid_z, phi, rhs, lam, D_xx, D_yy, D_zz):
    return solve_3d(grid_x, grid_y, grid_z, phi, rhs, lam, D_xx, D_yy, D_zz, gamma=2)


def solve_3d_fixed_wcycle_simple(grid_x, grid_y, grid_z, iters, phi, rhs, lam, D_xx, D_yy, D_zz):
    return solve_3d_fixed_wcycle(grid_x, grid_y, grid_z, iters, 1, phi, rhs, lam, D_xx, D_yy, D_zz)


def solve_3d_fixed_wcycle(grid_x, grid_y, grid_z, iters_outer, iters_inner, phi, rhs, lam, D_xx, D_yy, D_zz):
    return solve_3d(grid_x, grid_y, grid_z, phi, rhs, lam, D_xx, D_yy, D_zz, fixed=True,
                    max_iters_outer=iters_outer, max_iters_inner=iters_inner, gamma=2)


def solve_3d_wcycle_skip_pre_simple(grid_x, grid_y, grid_z, phi, rhs, lam, D_xx, D_yy, D_zz):
    return solve_3d(grid_x, grid_y, grid_z, phi, rhs, lam, D_xx, D_yy, D_zz, gamma=2, skip_pre=True)


def solve_3d_fixed_wcycle_skip_pre_simple(grid_x, grid_y, grid_z, iters, phi, rhs, lam, D_xx, D_yy, D_zz):
    return solve_3d_fixed_wcycle_skip_pre(grid_x, grid_y, grid_z, iters, 1, phi, rhs, lam, D_xx, D_yy, D_zz)


def solve_3d_fixed_wcycle_skip_pre(grid_x, grid_y, grid_z, iters_outer, iters_inner, phi, rhs, lam, D_xx, D_yy, D_zz):
    return solve_3d(grid_x, grid_y, grid_z, phi, rhs, lam, D_xx, D_yy, D_zz, fixed=True,
                    max_iters_outer=iters_outer, max_iters_inner=iters_inner, skip_pre=True, gamma=2)
