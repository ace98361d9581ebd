"""Shared host logic of solve_2d / solve_3d: argument checks with the
reference's error types and messages, plan caching, and the call into the
C-ABI solve loop.

Reference: Solver_2D / Solver_3D (jsolver/solver_2d.py:238-479,
jsolver/solver_3d.py:239-486).  The level recursion, smoothing, restriction
and prolongation run on the GPU inside libjsolver_amd.so; this module only
validates, converts and dispatches.
"""
import ctypes
from collections import OrderedDict

import numpy as np
import torch

from . import _lib
from ._device import (device, dtype_name, is_scalar, ptr, scalar_value, shape_of, stream_ptr,
                      to_device)

INT32_MAX = int(np.iinfo(np.int32).max)
_NAMES = "xyz"


# ---------------------------------------------------------------------------
# level layout checks (solver_2d.py:279-324, solver_3d.py:281-336)
# ---------------------------------------------------------------------------
def level_layout(ms):
    """ms per axis in (x, y[, z]) order -> (levels, m per level per axis).

    Raises ValueError exactly where the reference does: non-power-of-2
    resolution first (axis order x, y, z), then a coarsest level above 5."""
    for m, n in zip(ms, _NAMES):
        if np.bitwise_and(m, m - 1) > 0:
            raise ValueError(f"Resolution of {n}-grid is not a power of 2. Value: {m}")
    per_axis = [(np.ceil(np.log2(m)) + 1).astype(int) for m in ms]
    levels = int(min(per_axis))
    m_level = [2 ** np.arange(L - 1, L - levels - 1, -1, dtype=int) for L in per_axis]
    for ml, n in zip(m_level, _NAMES):
        if ml[-1] > 5:
            raise ValueError(f"Resolution of {n}-grid not efficient for multigrid configuration. "
                             f"Resolution of coarsest level: {ml[-1]}")
    return levels, m_level


def check_grids(grids):
    if not all(g.is_linear for g in grids):
        raise ValueError("Only linear grids are supported!")


# ---------------------------------------------------------------------------
# plans
# ---------------------------------------------------------------------------
class Plan:
    """Owns a jsv_plan (per-level device buffers of one grid and precision;
    2D plans may hold a batch of independent problems)."""

    def __init__(self, dim, ms, idels, precision, batch=1):
        lib = _lib.load()
        self.dim = dim
        self.ms = tuple(int(m) for m in ms)
        self.precision = precision
        self.batch = int(batch)
        self.dev = device()
        h = ctypes.c_void_p()
        mz = self.ms[2] if dim == 3 else 1
        idz = float(idels[2]) if dim == 3 else 1.0
        _lib.check(lib.jsv_plan_create_batched(dim, self.ms[0], self.ms[1], mz, self.batch,
                                               float(idels[0]), float(idels[1]), idz,
                                               _lib.JSV_MIXED if precision == "mixed" else _lib.JSV_FP64,
                                               ctypes.byref(h)))
        self.h = h
        lv = ctypes.c_int()
        _lib.check(lib.jsv_plan_levels(self.h, ctypes.byref(lv)))
        self.levels = lv.value
        self.op_key = None

    def level_shape(self, level):
        s = (ctypes.c_int * 3)()
        _lib.check(_lib.load().jsv_plan_level_shape(self.h, level, s))
        return (s[1], s[2]) if self.dim == 2 else (s[0], s[1], s[2])

    def device_bytes(self):
        b = ctypes.c_int64()
        _lib.check(_lib.load().jsv_plan_device_bytes(self.h, ctypes.byref(b)))
        return b.value

    def set_operator(self, lam, D, D_xy=None):
        """lam, D: device tensors (float64, fine shape) or Python floats."""
        lib = _lib.load()
        lam_t = lam if torch.is_tensor(lam) else None
        lam_s = 0.0 if lam_t is not None else float(lam)
        spatial = torch.is_tensor(D[0])
        dsc = (ctypes.c_double * 3)(*([0.0] * 3 if spatial else [float(d) for d in D] + [0.0] * (3 - len(D))))
        dz = D[2] if (self.dim == 3 and spatial) else None
        _lib.check(lib.jsv_plan_set_operator(
            self.h, ptr(lam_t), lam_s, ptr(D[0] if spatial else None), ptr(D[1] if spatial else None),
            ptr(dz), dsc, ptr(D_xy), stream_ptr()))
        self._keep = (lam_t, D, D_xy)   # inputs stay alive until the setup kernels ran

    def __del__(self):
        try:
            if getattr(self, "h", None) is not None and _lib._lib is not None:
                torch.cuda.synchronize()
                _lib._lib.jsv_plan_destroy(self.h)
                self.h = None
        except Exception:
            pass


_PLANS = OrderedDict()
_PLAN_CACHE_SIZE = 4


def get_plan(dim, ms, idels, precision, batch=1, role="primal"):
    """Cached plan; role "tangent" is a second plan of the same shape that
    holds the tangent operator of a forward-mode solve."""
    key = (dim, tuple(int(m) for m in ms), tuple(float(i) for i in idels), precision, int(batch), role)
    p = _PLANS.get(key)
    if p is None:
        p = Plan(dim, ms, idels, precision, batch)
        _PLANS[key] = p
        while len(_PLANS) > _PLAN_CACHE_SIZE:
            _PLANS.popitem(last=False)
    else:
        _PLANS.move_to_end(key)
    return p


def clear_plans():
    _PLANS.clear()


# ---------------------------------------------------------------------------
# the solve (Solver_2D.solve / Solver_3D.solve)
# ---------------------------------------------------------------------------
def make_config(fixed, epsilon, gamma, fcycle, skip_pre, nu1, nu2, max_iters_outer, stop, rtol,
                inner_cycles, history=None, check_every=1):
    c = _lib.JsvConfig()
    c.fixed = 1 if fixed else 0
    c.epsilon = float(epsilon)
    c.gamma = int(gamma)
    c.fcycle = 1 if fcycle else 0
    c.skip_pre = 1 if skip_pre else 0
    c.nu1 = int(nu1)
    c.nu2 = int(nu2)
    c.max_iters = int(max_iters_outer)
    c.stop = {"distance": _lib.JSV_STOP_DISTANCE, "rel_residual": _lib.JSV_STOP_REL_RESIDUAL}[stop]
    c.rtol = float(rtol)
    c.inner_cycles = int(inner_cycles)
    c.check_every = int(check_every)
    if history is not None:
        c.history = history.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        c.history_len = history.size
    return c


def validate_inputs(dim, phi, rhs, lam, D, D_xy):
    """dtype and shape checks of solver_2d.py:444-453 / solver_3d.py:452-460."""
    scalar_lambda = is_scalar(lam)
    scalar_diffusion = all(is_scalar(d) for d in D)
    spatial = not scalar_diffusion
    off_diag = spatial and D_xy is not None
    dts = [dtype_name(phi), dtype_name(rhs), dtype_name(lam)] + [dtype_name(d) for d in D]
    if D_xy is not None:
        dts.append(dtype_name(D_xy))
    if not all(t == "float64" for t in dts):
        raise ValueError("Incorrect configuration: dtype must be 'float64' for all arguments.")
    shp = [shape_of(d) for d in D]
    ps, rs, ls = shape_of(phi), shape_of(rhs), shape_of(lam)
    bad = (not all(s == shp[0] for s in shp)
           or (scalar_lambda and ps != rs)
           or (not scalar_lambda and not ps == rs == ls)
           or (spatial and not all(ps == rs == s for s in shp))
           or (off_diag and shape_of(D_xy) != ps))
    if bad:
        names = ["D_xx", "D_yy", "D_zz"][:dim]
        dpart = ", ".join(f"{n}={s}" for n, s in zip(names, shp))
        extra = "" if dim == 3 else f", D_xy={'None' if D_xy is None else shape_of(D_xy)}"
        raise ValueError(f"Shape mismatch: phi={ps}, rhs={rs}, lam={ls}, {dpart}{extra}.")
    return scalar_lambda, spatial, off_diag


def solve(dim, grids, phi, rhs, lam, D, D_xy=None, *, fixed=False, epsilon=1e-12, gamma=1,
          fcycle=False, skip_pre=False, nu1=1, nu2=1, max_iters_outer=INT32_MAX,
          max_iters_inner=INT32_MAX, precision="fp64", stop="distance", rtol=1e-10,
          inner_cycles=0, check_every=1, return_history=False, return_info=False, boundary="zero"):
    ms = [g.mx for g in grids]
    levels, m_level = level_layout(ms)                  # Solver.__init__ -> compute_shapes
    check_grids(grids)                                  # compute_idel
    scalar_lambda, spatial, off_diag = validate_inputs(dim, phi, rhs, lam, D, D_xy)
    if precision not in ("fp64", "mixed"):
        raise ValueError(f"precision must be 'fp64' or 'mixed', got {precision!r}")
    if stop not in ("distance", "rel_residual"):
        raise ValueError(f"stop must be 'distance' or 'rel_residual', got {stop!r}")
    if boundary not in ("zero", "phi"):
        raise ValueError(f"boundary must be 'zero' or 'phi', got {boundary!r}")
    # max_iters_inner only feeds the reference's unused coarsest_level_solve
    # (solver_2d.py:547-580); it is accepted and has no effect, as there.
    del max_iters_inner
    full_shape = tuple(int(m) + 3 for m in reversed(ms))
    if tuple(shape_of(phi)) != full_shape:
        raise ValueError(f"Shape mismatch: phi={shape_of(phi)} does not match the grids "
                         f"(expected {full_shape}).")
    dev = device()
    args = [phi, rhs, lam, *D] + ([D_xy] if D_xy is not None else [])
    from . import tangent
    differentiated = any(tangent.has_tangent(v) for v in args) or (
        torch.is_grad_enabled() and any(torch.is_tensor(v) and v.requires_grad for v in args))
    if boundary == "phi" and differentiated:
        raise NotImplementedError("boundary='phi' is not differentiable; use zero boundaries")
    if any(tangent.has_tangent(v) for v in args):
        if return_history:
            raise ValueError("return_history is not available for differentiated solves")
        out, count, info = tangent.solve_forward_ad(
            dim, grids, phi, rhs, lam, D, D_xy,
            dict(fixed=fixed, epsilon=epsilon, gamma=gamma, fcycle=fcycle, skip_pre=skip_pre, nu1=nu1,
                 nu2=nu2, max_iters_outer=max_iters_outer, precision=precision, stop=stop, rtol=rtol,
                 inner_cycles=inner_cycles, check_every=check_every),
            (dev, tuple(int(m) for m in ms), tuple(float(g.idel) for g in grids)))
        return (out, count, info) if return_info else (out, count)
    plan = get_plan(dim, ms, [g.idel for g in grids], precision)
    if torch.is_grad_enabled() and any(torch.is_tensor(v) and v.requires_grad for v in args):
        return _solve_differentiable(dim, plan, dev, full_shape, phi, rhs, lam, D, D_xy, fixed,
                                     precision, dict(fixed=True, epsilon=epsilon, gamma=gamma,
                                                     fcycle=fcycle, skip_pre=skip_pre, nu1=nu1,
                                                     nu2=nu2, max_iters_outer=max_iters_outer,
                                                     stop="distance", rtol=rtol,
                                                     inner_cycles=inner_cycles),
                                     return_history, return_info, levels)
    phi_d = to_device(phi, dev)
    rhs_d = to_device(rhs, dev)
    lam_v = scalar_value(lam) if scalar_lambda else to_device(lam, dev)
    if spatial:
        D_v = [to_device(d, dev) for d in D]
        Dxy_v = to_device(D_xy, dev) if off_diag else None
    else:
        D_v = [scalar_value(d) for d in D]
        Dxy_v = None
    plan.set_operator(lam_v, D_v, Dxy_v)
    if boundary == "phi":
        rhs_d = dirichlet_lift(plan, phi_d, rhs_d)
    hist = None
    if return_history:
        cap = max_iters_outer if (fixed or max_iters_outer < 100000) else 100000
        hist = np.full(int(min(cap, 10_000_000)), np.nan)
    cfg = make_config(fixed, epsilon, gamma, fcycle, skip_pre, nu1, nu2, max_iters_outer, stop,
                      rtol, inner_cycles, hist, check_every)
    out = torch.empty_like(phi_d)
    res = _lib.JsvResult()
    _lib.check(_lib.load().jsv_solve(plan.h, ptr(phi_d), ptr(rhs_d), ptr(out), ctypes.byref(cfg),
                                     ctypes.byref(res), stream_ptr()))
    count = int(res.count)
    if boundary == "phi":
        out = with_boundary(out, phi_d, dim)
    ret = (out, count)
    if return_history:
        n = int(res.outer_steps) if precision == "mixed" else count
        ret = ret + (hist[:n].copy(),)
    if return_info:
        ret = ret + ({"count": count, "outer_steps": int(res.outer_steps),
                      "distance": float(res.distance), "rel_residual": float(res.rel_residual),
                      "nonfinite": bool(res.nonfinite), "levels": levels},)
    return ret


def dirichlet_lift(plan, phi_b, rhs):
    """rhs with the Dirichlet values phi_b holds on the boundary nodes moved to
    the neighbouring unknowns (jsv_dirichlet_lift; the plan's operator)."""
    out = torch.empty_like(rhs)
    _lib.check(_lib.load().jsv_dirichlet_lift(plan.h, ptr(phi_b), ptr(rhs), ptr(out), stream_ptr()))
    return out


def with_boundary(out, phi_b, dim, batched=False):
    """The solution on the unknowns, phi_b's values (boundary nodes and
    ghosts) everywhere else."""
    inner = (slice(None),) if batched else ()
    inner = inner + (slice(2, -2),) * dim
    res = phi_b.clone()
    res[inner] = out[inner]
    return res


def _solve_differentiable(dim, plan, dev, full_shape, phi, rhs, lam, D, D_xy, fixed, precision,
                          cfg_kw, return_history, return_info, levels):
    """Fixed-mode solve recorded for torch.autograd (backward = the hand-written
    adjoint).  Gradients flow to tensor arguments: phi, rhs, lam and D (arrays
    or 0-d scalars) and D_xy."""
    from . import adjoint
    if not fixed:
        raise ValueError("Reverse-mode differentiation needs fixed=True (a fixed number of cycles); "
                         "the dynamic stop rule is not differentiable.")
    if precision != "fp64":
        raise ValueError("reverse-mode differentiation runs in precision='fp64'")
    if return_history:
        raise ValueError("return_history is not available for differentiated solves")

    def dev_t(v):
        """float64 device tensor; 0-d for scalars; the autograd graph is kept."""
        if torch.is_tensor(v):
            return v.to(device=dev, dtype=torch.float64).contiguous()
        return torch.from_numpy(np.ascontiguousarray(np.asarray(v, dtype=np.float64))).to(dev)

    phi_d, rhs_d = dev_t(phi), dev_t(rhs)
    lam_v = dev_t(lam).reshape(()) if is_scalar(lam) else dev_t(lam)
    D_v = [dev_t(d).reshape(()) if is_scalar(d) else dev_t(d) for d in D]
    Dxy_v = dev_t(D_xy) if D_xy is not None else None
    out = adjoint.solve_fixed(plan, phi_d, rhs_d, lam_v, D_v, cfg_kw, Dxy_v)
    count = int(cfg_kw["max_iters_outer"])
    ret = (out, count)
    if return_info:
        ret = ret + ({"count": count, "outer_steps": count, "distance": float("nan"),
                      "rel_residual": -1.0, "nonfinite": False, "levels": levels},)
    return ret


def solve_batched_2d(grid_x, grid_y, phi, rhs, lam, D_xx, D_yy, *, D_xy=None, fixed=False,
                     epsilon=1e-12, gamma=1, fcycle=False, skip_pre=False, nu1=1, nu2=1,
                     max_iters_outer=INT32_MAX, precision="fp64", stop="distance", rtol=1e-10,
                     inner_cycles=0, check_every=1, boundary="zero"):
    """B independent 2D problems on one grid, solved together (one launch per
    operation for the whole batch).  phi, rhs and array coefficients have shape
    (B, my+3, mx+3); lam / D may be Python floats shared by all problems.
    Returns (phi (B, my+3, mx+3) float64 GPU tensor, counts int64[B],
    last stop quantity float64[B]); every problem's result equals its own
    solve_2d result."""
    ms = [grid_x.mx, grid_y.mx]
    level_layout(ms)
    check_grids((grid_x, grid_y))
    dev = device()
    phi_d, rhs_d = to_device(phi, dev), to_device(rhs, dev)
    if phi_d.ndim != 3 or tuple(phi_d.shape[1:]) != (grid_y.mx + 3, grid_x.mx + 3) \
            or phi_d.shape != rhs_d.shape:
        raise ValueError(f"Shape mismatch: phi={tuple(phi_d.shape)}, rhs={tuple(rhs_d.shape)}; expected "
                         f"(B, {grid_y.mx + 3}, {grid_x.mx + 3}).")
    B = phi_d.shape[0]
    plan = get_plan(2, ms, [grid_x.idel, grid_y.idel], precision, B)

    def arr_or_scalar(v):
        if is_scalar(v):
            return scalar_value(v)
        t = to_device(v, dev)
        if t.shape != phi_d.shape:
            raise ValueError(f"Shape mismatch: {tuple(t.shape)} vs phi {tuple(phi_d.shape)}.")
        return t
    lam_v = arr_or_scalar(lam)
    D_v = [arr_or_scalar(D_xx), arr_or_scalar(D_yy)]
    if torch.is_tensor(D_v[0]) != torch.is_tensor(D_v[1]):
        raise ValueError("D_xx and D_yy must both be scalars or both be arrays")
    Dxy_v = arr_or_scalar(D_xy) if D_xy is not None else None
    plan.set_operator(lam_v, D_v, Dxy_v)
    if boundary not in ("zero", "phi"):
        raise ValueError(f"boundary must be 'zero' or 'phi', got {boundary!r}")
    if boundary == "phi":
        rhs_d = dirichlet_lift(plan, phi_d, rhs_d)
    cfg = make_config(fixed, epsilon, gamma, fcycle, skip_pre, nu1, nu2, max_iters_outer, stop, rtol,
                      inner_cycles, check_every=check_every)
    out = torch.empty_like(phi_d)
    res = _lib.JsvResult()
    counts = np.zeros(B, dtype=np.int64)
    stopv = np.zeros(B, dtype=np.float64)
    _lib.check(_lib.load().jsv_solve_batched(
        plan.h, ptr(phi_d), ptr(rhs_d), ptr(out), ctypes.byref(cfg), ctypes.byref(res),
        counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
        stopv.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), stream_ptr()))
    if boundary == "phi":
        out = with_boundary(out, phi_d, 2, batched=True)
    return out, counts, stopv
