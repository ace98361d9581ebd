"""Forward-mode differentiation of solves (torch forward AD and torch.func.jvp).

The reference differentiates its solves forward with jax.jvp, in fixed and in
dynamic mode (examples/benchmark.py:1146-1182; tests/test_solver_2d.py:173-176
runs check_grads(modes=["fwd"]) on the dynamic solve).  Here solve_2d /
solve_3d route to this module whenever an argument carries a forward-mode
tangent (a dual tensor of torch.autograd.forward_ad, or a tensor inside
torch.func.jvp); the tangent of phi is computed by the hand-written tangent
cycles of libjsolver_amd.so (jsv_tangent):

    phi_dot = torch.func.jvp(lambda r: solve_2d(gx, gy, phi0, r, lam, Dxx, Dyy)[0],
                             (rhs,), (v,))[1]

A dynamic solve's tangent is the tangent of its cycle count: the stop rule
only decides how many cycles run (as jax.jvp does through the while loop).
Tangents may be attached to phi, rhs, lam and the D components (arrays or
0-d tensors), and to D_xy.  FP64 only.
"""
import ctypes

import numpy as np
import torch
import torch.autograd.forward_ad as fwAD

from . import _lib
from ._device import is_scalar, ptr, scalar_value, stream_ptr


def has_tangent(v):
    """True for a tensor that carries a forward-mode tangent."""
    if not torch.is_tensor(v):
        return False
    if torch._C._functorch.is_functorch_wrapped_tensor(v):
        return True
    return fwAD.unpack_dual(v).tangent is not None


class _Box:
    """Carries the cycle count from forward() to jvp() (a plain object: the
    transforms of torch.func pass it through unchanged, unlike a dict)."""

    def __init__(self, tangent_fn):
        self.tangent_fn = tangent_fn
        self.count = None
        self.info = None


class _TangentSolve(torch.autograd.Function):
    """forward: the solve itself (fast path); jvp: the tangent cycles."""

    @staticmethod
    def forward(box, solve_fn, phi0, rhs, lam, *D_all):
        out, count, info = solve_fn(phi0, rhs, lam, *D_all)
        box.count, box.info = count, info
        return out

    @staticmethod
    def setup_context(ctx, inputs, output):
        box, solve_fn, phi0, rhs, lam, *D_all = inputs
        ctx.set_materialize_grads(False)
        ctx.box = box
        ctx.lam_t = torch.is_tensor(lam)
        ctx.lam_s = None if torch.is_tensor(lam) else lam
        D_t = [torch.is_tensor(d) for d in D_all]
        ctx.D_t = D_t
        ctx.D_s = [None if t else d for d, t in zip(D_all, D_t)]
        tensors = [phi0, rhs] + ([lam] if ctx.lam_t else []) + [d for d, t in zip(D_all, D_t) if t]
        ctx.save_for_forward(*tensors)

    @staticmethod
    def jvp(ctx, _box, _fn, dphi0, drhs, dlam, *dD_all):
        saved = list(ctx.saved_tensors)
        phi0, rhs = saved[0], saved[1]
        rest = saved[2:]
        lam = rest.pop(0) if ctx.lam_t else ctx.lam_s
        D_all = [rest.pop(0) if t else s for t, s in zip(ctx.D_t, ctx.D_s)]
        # the kernels run inside a second Function: under torch.func.jvp the
        # tensors a jvp rule receives are still wrapped (no data pointer),
        # while a Function's forward receives plain ones
        return _TangentCycles.apply(ctx.box, len(D_all), phi0, rhs, lam, *D_all, dphi0, drhs, dlam,
                                    *dD_all)

    @staticmethod
    def backward(ctx, *grads):
        raise RuntimeError("this solve was recorded for forward-mode differentiation; use "
                           "requires_grad inputs (fixed=True) for the reverse pass")


class _TangentCycles(torch.autograd.Function):
    """The tangent cycles as an op of their own (not differentiable again)."""

    @staticmethod
    def forward(box, n, *arrs):
        phi0, rhs, lam = arrs[:3]
        D_all = list(arrs[3:3 + n])
        dphi0, drhs, dlam = arrs[3 + n:6 + n]
        return box.tangent_fn(phi0, rhs, lam, D_all, dphi0, drhs, dlam, list(arrs[6 + n:]), box.count)

    @staticmethod
    def setup_context(ctx, inputs, output):
        pass

    @staticmethod
    def backward(ctx, *grads):
        raise RuntimeError("the tangent of a solve is not differentiable again")


def solve_forward_ad(dim, grids, phi, rhs, lam, D, D_xy, solve_kw, plan_args):
    """solve_2d / solve_3d with forward-mode tangents on some arguments.
    Returns (phi, count, info) where phi carries the tangent."""
    from . import _solver
    precision = solve_kw.get("precision", "fp64")
    if precision != "fp64":
        raise ValueError("forward-mode differentiation runs in precision='fp64'")
    dev, ms, idels = plan_args

    def dev_t(v):
        # keeps the tangent: .to() and .contiguous() are differentiable
        if torch.is_tensor(v):
            return v.to(device=dev, dtype=torch.float64).contiguous()
        return torch.from_numpy(np.ascontiguousarray(np.asarray(v, dtype=np.float64))).to(dev)

    phi_d, rhs_d = dev_t(phi), dev_t(rhs)
    lam_d = dev_t(lam) if torch.is_tensor(lam) or not is_scalar(lam) else float(lam)
    D_d = [dev_t(d) if torch.is_tensor(d) or not is_scalar(d) else float(d) for d in D]
    Dxy_d = dev_t(D_xy) if D_xy is not None else None
    D_all = D_d + ([Dxy_d] if Dxy_d is not None else [])

    def primal(phi0, rhs_, lam_, *D_all_):
        Dp = list(D_all_[:dim])
        Dxy = D_all_[dim] if len(D_all_) > dim else None
        kw = dict(solve_kw)
        kw["return_info"] = True
        out, count, info = _solver.solve(dim, grids, phi0, rhs_, lam_, Dp, Dxy, **kw)
        return out, count, info

    def tangent(phi0, rhs_, lam_, D_all_, dphi0, drhs, dlam, dD_all, count):
        return _tangent_cycles(dim, ms, idels, dev, phi0, rhs_, lam_, D_all_, dphi0, drhs, dlam,
                               dD_all, count, solve_kw)

    box = _Box(tangent)
    out = _TangentSolve.apply(box, primal, phi_d, rhs_d, lam_d, *D_all)
    return out, box.count, box.info


def _tangent_cycles(dim, ms, idels, dev, phi0, rhs, lam, D_all, dphi0, drhs, dlam, dD_all, count,
                    solve_kw):
    """phi_dot after `count` cycles (jsv_tangent)."""
    from . import _solver
    lib = _lib.load()
    D = list(D_all[:dim])
    Dxy = D_all[dim] if len(D_all) > dim else None
    dD = list(dD_all[:dim])
    dDxy = dD_all[dim] if len(dD_all) > dim else None
    spatial = not is_scalar(D[0])

    def val(v):
        return scalar_value(v) if is_scalar(v) else v.contiguous()

    P = _solver.get_plan(dim, ms, idels, "fp64")
    P.set_operator(val(lam), [val(d) for d in D], Dxy.contiguous() if Dxy is not None else None)
    T = _solver.get_plan(dim, ms, idels, "fp64", role="tangent")
    # tangent operator: the coefficient setup applied to the tangents
    if is_scalar(lam):
        lam_dot = scalar_value(dlam) if dlam is not None else 0.0
    else:
        lam_dot = dlam.to(torch.float64).contiguous() if dlam is not None else 0.0
    if spatial:
        D_dot = [dd.to(torch.float64).contiguous() if dd is not None else torch.zeros_like(d)
                 for d, dd in zip(D, dD)]
    else:
        D_dot = [scalar_value(dd) if dd is not None else 0.0 for dd in dD]
    Dxy_dot = None
    if Dxy is not None:
        Dxy_dot = dDxy.to(torch.float64).contiguous() if dDxy is not None else torch.zeros_like(Dxy)
    T.set_operator(lam_dot, D_dot, Dxy_dot)
    cfg = _solver.make_config(True, solve_kw.get("epsilon", 1e-12), solve_kw.get("gamma", 1),
                              solve_kw.get("fcycle", False), solve_kw.get("skip_pre", False),
                              solve_kw.get("nu1", 1), solve_kw.get("nu2", 1), int(count), "distance",
                              1e-10, 0)
    out_dot = torch.empty_like(phi0)
    c = lambda t: None if t is None else t.to(torch.float64).contiguous()
    dp, dr = c(dphi0), c(drhs)
    _lib.check(lib.jsv_tangent(P.h, T.h, ptr(phi0), ptr(rhs), ptr(dp), ptr(dr), ctypes.byref(cfg),
                               ptr(None), ptr(out_dot), stream_ptr()))
    return out_dot
