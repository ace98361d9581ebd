"""Reverse-mode differentiation of fixed-mode solves (torch.autograd).

The reference differentiates ``solve_*_fixed*`` with jax.vjp
(examples/benchmark.py:1184-1226; tests/test_solver_2d.py:200-207).  Here
the backward pass is the hand-written adjoint in libjsolver_amd.so
(jsv_forward_checkpoints / jsv_adjoint): the forward stores the iterate at
every cycle boundary ((K+1) fine arrays) and the backward recomputes each
cycle's interior states into a tape right before reversing that cycle.

solve_2d / solve_3d route here automatically when ``fixed=True`` and any
tensor argument requires grad; dynamic mode is not reverse-differentiable
(as with the reference's while_loop, README.md:20).  Gradients reach every
operator input the reference's jax.vjp reaches: phi0, rhs, lambda and D as
arrays or scalars (tests/test_solver_2d.py:200-207 differentiates scalar
lambda and D), and the 2D D_xy array (examples/benchmark.py:1223-1229).
"""
import ctypes

import torch

from . import _lib
from ._device import ptr, stream_ptr


class _FixedSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, plan, cfg_kw, phi0, rhs, lam, Dxy, *D):
        # lam and each D: a 0-d tensor (the scalar path, solver_2d.py:417-421)
        # or an array of the fine shape; Dxy: an array or None
        from ._solver import make_config
        K = int(cfg_kw["max_iters_outer"])
        cfg = make_config(**cfg_kw)
        lam_v = float(lam) if lam.ndim == 0 else lam
        D_v = [float(d) for d in D] if D[0].ndim == 0 else list(D)
        plan.set_operator(lam_v, D_v, Dxy)
        shape = tuple(phi0.shape)
        ckpt = torch.empty((K + 1,) + shape, dtype=torch.float64, device=phi0.device)
        _lib.check(_lib.load().jsv_forward_checkpoints(plan.h, ptr(phi0), ptr(rhs), ctypes.byref(cfg),
                                                       ptr(ckpt), stream_ptr()))
        ctx.plan, ctx.cfg_kw, ctx.has_dxy = plan, cfg_kw, Dxy is not None
        ctx.save_for_backward(phi0, rhs, ckpt, lam, Dxy, *D)
        return ckpt[K].clone()

    @staticmethod
    def backward(ctx, g_out):
        from ._solver import make_config
        phi0, rhs, ckpt, lam, Dxy, *D = ctx.saved_tensors
        plan = ctx.plan
        lam_scalar, D_scalar = lam.ndim == 0, D[0].ndim == 0
        # the plan may have served another operator since the forward pass
        plan.set_operator(float(lam) if lam_scalar else lam,
                          [float(d) for d in D] if D_scalar else list(D), Dxy)
        cfg = make_config(**ctx.cfg_kw)
        g = g_out.detach().to(torch.float64).contiguous()
        # inputs: plan, cfg_kw, phi0, rhs, lam, Dxy, *D
        need = ctx.needs_input_grad
        mk = lambda flag: torch.empty_like(phi0) if flag else None
        g_phi0, g_rhs = mk(need[2]), mk(need[3])
        g_lam = mk(need[4] and not lam_scalar)
        g_dxy = mk(ctx.has_dxy and need[5])
        g_D = [mk(need[6 + a] and not D_scalar) for a in range(plan.dim)]
        want_sc = (lam_scalar and need[4]) or (D_scalar and any(need[6:]))
        g_sc = torch.zeros(4, dtype=torch.float64, device=phi0.device) if want_sc else None
        gD3 = g_D + [None] * (3 - len(g_D))
        _lib.check(_lib.load().jsv_adjoint_full(
            plan.h, ptr(phi0), ptr(rhs), ptr(ckpt), ctypes.byref(cfg), ptr(g), ptr(g_phi0), ptr(g_rhs),
            ptr(g_lam), ptr(gD3[0]), ptr(gD3[1]), ptr(gD3[2]), ptr(g_dxy), ptr(g_sc), stream_ptr()))
        if lam_scalar and need[4]:
            g_lam = g_sc[0].clone()
        if D_scalar:
            g_D = [g_sc[1 + a].clone() if need[6 + a] else None for a in range(plan.dim)]
        return (None, None, g_phi0, g_rhs, g_lam, g_dxy, *g_D)


def solve_fixed(plan, phi0, rhs, lam, D, cfg_kw, Dxy=None):
    """Differentiable fixed solve.  phi0/rhs: device tensors; lam and D:
    device tensors of the fine shape or 0-d device tensors (scalars, which
    take the reference's scalar path); Dxy: device tensor or None."""
    return _FixedSolve.apply(plan, cfg_kw, phi0, rhs, lam, Dxy, *D)


def adjoint_device_bytes(plan):
    b = ctypes.c_int64()
    _lib.check(_lib.load().jsv_adjoint_device_bytes(plan.h, ctypes.byref(b)))
    return b.value
