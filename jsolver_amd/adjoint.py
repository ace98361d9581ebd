"""Reverse-mode differentiation of fixed-mode solves (torch.autograd).

The reference differentiates ``solve_*_fixed*`` with jax.vjp
(examples/benchmark.py:1184-1226; tests/test_solver_2d.py:200-207).  Here
the backward pass is the hand-written adjoint in libjsolver_amd.so
(jsv_forward_checkpoints / jsv_adjoint): the forward stores the iterate at
every cycle boundary ((K+1) fine arrays) and the backward recomputes each
cycle's interior states into a tape right before reversing that cycle.

solve_2d / solve_3d route here automatically when ``fixed=True`` and any
tensor argument requires grad; dynamic mode is not reverse-differentiable
(as with the reference's while_loop, README.md:20).
"""
import ctypes

import torch

from . import _lib
from ._device import ptr, stream_ptr


class _FixedSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, plan, cfg_kw, lam_scalar, d_scalars, phi0, rhs, lam, *D):
        # lam / D: device tensors, or None where the scalar is used
        from ._solver import make_config
        K = int(cfg_kw["max_iters_outer"])
        cfg = make_config(**cfg_kw)
        lam_v = lam if lam is not None else lam_scalar
        D_v = list(D) if D[0] is not None else list(d_scalars)
        plan.set_operator(lam_v, D_v, None)
        shape = tuple(phi0.shape)
        ckpt = torch.empty((K + 1,) + shape, dtype=torch.float64, device=phi0.device)
        _lib.check(_lib.load().jsv_forward_checkpoints(plan.h, ptr(phi0), ptr(rhs), ctypes.byref(cfg),
                                                       ptr(ckpt), stream_ptr()))
        ctx.plan, ctx.cfg_kw, ctx.lam_scalar, ctx.d_scalars = plan, cfg_kw, lam_scalar, d_scalars
        ctx.has_lam, ctx.has_D = lam is not None, D[0] is not None
        ctx.save_for_backward(phi0, rhs, ckpt, *([lam] if lam is not None else []),
                              *([d for d in D] if D[0] is not None else []))
        return ckpt[K].clone()

    @staticmethod
    def backward(ctx, g_out):
        from ._solver import make_config
        saved = list(ctx.saved_tensors)
        phi0, rhs, ckpt = saved[:3]
        rest = saved[3:]
        lam = rest.pop(0) if ctx.has_lam else None
        D = rest if ctx.has_D else [None] * ctx.plan.dim
        plan = ctx.plan
        # the plan may have served another operator since the forward pass
        plan.set_operator(lam if lam is not None else ctx.lam_scalar,
                          list(D) if ctx.has_D else list(ctx.d_scalars), None)
        cfg = make_config(**ctx.cfg_kw)
        g = g_out.detach().to(torch.float64).contiguous()
        need = ctx.needs_input_grad
        # inputs: plan, cfg_kw, lam_scalar, d_scalars, phi0, rhs, lam, *D
        mk = lambda flag: torch.empty_like(phi0) if flag else None
        g_phi0 = mk(need[4])
        g_rhs = mk(need[5])
        g_lam = mk(ctx.has_lam and need[6])
        g_D = [mk(ctx.has_D and need[7 + a]) for a in range(plan.dim)]
        gD3 = g_D + [None] * (3 - len(g_D))
        _lib.check(_lib.load().jsv_adjoint(plan.h, ptr(phi0), ptr(rhs), ptr(ckpt), ctypes.byref(cfg),
                                           ptr(g), ptr(g_phi0), ptr(g_rhs), ptr(g_lam), ptr(gD3[0]),
                                           ptr(gD3[1]), ptr(gD3[2]), stream_ptr()))
        return (None, None, None, None, g_phi0, g_rhs, g_lam, *g_D)


def solve_fixed(plan, phi0, rhs, lam, D, cfg_kw):
    """Differentiable fixed solve.  phi0/rhs: device tensors; lam/D: device
    tensors or Python floats (floats get no gradient)."""
    lam_t = lam if torch.is_tensor(lam) else None
    lam_s = 0.0 if lam_t is not None else float(lam)
    spatial = torch.is_tensor(D[0])
    D_t = list(D) if spatial else [None] * len(D)
    d_s = [0.0] * len(D) if spatial else [float(d) for d in D]
    return _FixedSolve.apply(plan, cfg_kw, lam_s, d_s, phi0, rhs, lam_t, *D_t)


def adjoint_device_bytes(plan):
    b = ctypes.c_int64()
    _lib.check(_lib.load().jsv_adjoint_device_bytes(plan.h, ctypes.byref(b)))
    return b.value
