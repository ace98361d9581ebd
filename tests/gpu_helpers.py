"""Helpers for GPU tests: drive single operations of a plan through the C ABI
and build matching oracle operators."""
import ctypes

import numpy as np
import torch

from jsolver_amd import _lib, _solver
from jsolver_amd._device import ptr, stream_ptr
from oracle import gmg


def rand_problem(dim, ms, seed, spatial=True, lam_array=True, dxy=False):
    """Random operator inputs on the padded fine shape; ms in (x, y[, z]) order."""
    rng = np.random.default_rng(seed)
    shape = tuple(m + 3 for m in reversed(ms))
    if spatial:
        D = [np.exp(0.5 * rng.standard_normal(shape)) for _ in range(dim)]
    else:
        D = [float(v) for v in rng.uniform(0.5, 2.0, dim)]
    lam = rng.uniform(0.0, 2.0, shape) if lam_array else float(rng.uniform(0, 2))
    Dxy = 0.1 * rng.standard_normal(shape) if (dxy and dim == 2 and spatial) else None
    rhs = rng.standard_normal(shape)
    phi = rng.standard_normal(shape)
    return phi, rhs, lam, D, Dxy


def grids_of(ms):
    return [gmg.OracleGrid(m) for m in ms]


def oracle_operator(dim, ms, lam, D, Dxy=None):
    gr = grids_of(ms)
    return gmg.build_operator(dim, tuple(ms), tuple(g.idel for g in gr), lam, list(D), Dxy)


class DevPlan:
    """A plan with its operator set, plus upload/download of level buffers."""

    def __init__(self, dim, ms, lam, D, Dxy=None, precision="fp64"):
        gr = grids_of(ms)
        self.dim = dim
        self.plan = _solver.Plan(dim, ms, [g.idel for g in gr], precision)
        dev = self.plan.dev
        lam_v = torch.from_numpy(np.ascontiguousarray(lam)).to(dev) if np.ndim(lam) else float(lam)
        if np.ndim(D[0]):
            D_v = [torch.from_numpy(np.ascontiguousarray(d)).to(dev) for d in D]
        else:
            D_v = [float(d) for d in D]
        Dxy_v = torch.from_numpy(np.ascontiguousarray(Dxy)).to(dev) if Dxy is not None else None
        self.plan.set_operator(lam_v, D_v, Dxy_v)
        self.lib = _lib.load()

    def shape(self, level):
        return self.plan.level_shape(level)

    def upload(self, level, which, arr, prec=_lib.JSV_FP64):
        dt = torch.float64 if prec == _lib.JSV_FP64 else torch.float32
        t = torch.as_tensor(np.ascontiguousarray(arr)).to(self.plan.dev, dt).contiguous()
        assert tuple(t.shape) == tuple(self.shape(level))
        _lib.check(self.lib.jsv_level_upload(self.plan.h, prec, level, which, ptr(t), stream_ptr()))
        torch.cuda.synchronize()

    def download(self, level, which, prec=_lib.JSV_FP64):
        dt = torch.float64 if prec == _lib.JSV_FP64 else torch.float32
        t = torch.empty(self.shape(level), dtype=dt, device=self.plan.dev)
        _lib.check(self.lib.jsv_level_download(self.plan.h, prec, level, which, ptr(t), stream_ptr()))
        torch.cuda.synchronize()
        return t.double().cpu().numpy()

    def smooth(self, level, sweeps=1, prec=_lib.JSV_FP64):
        _lib.check(self.lib.jsv_smooth(self.plan.h, prec, level, sweeps, stream_ptr()))

    def restrict(self, level, prec=_lib.JSV_FP64):
        _lib.check(self.lib.jsv_restrict(self.plan.h, prec, level, stream_ptr()))

    def prolong(self, level, nu2, prec=_lib.JSV_FP64):
        _lib.check(self.lib.jsv_prolong_add(self.plan.h, prec, level, nu2, stream_ptr()))

    def cycle(self, cfg, skip_pre=False, prec=_lib.JSV_FP64):
        _lib.check(self.lib.jsv_cycle(self.plan.h, prec, ctypes.byref(cfg), int(skip_pre), stream_ptr()))

    def residual(self, phi, rhs):
        dev = self.plan.dev
        p = torch.as_tensor(phi).to(dev, torch.float64).contiguous()
        r = torch.as_tensor(rhs).to(dev, torch.float64).contiguous()
        out = torch.empty_like(p)
        n2 = ctypes.c_double()
        _lib.check(self.lib.jsv_residual(self.plan.h, ptr(p), ptr(r), ptr(out), ctypes.byref(n2),
                                         stream_ptr()))
        return out.cpu().numpy(), n2.value


def cfg(**kw):
    base = dict(fixed=True, epsilon=1e-12, gamma=1, fcycle=False, skip_pre=False, nu1=1, nu2=1,
                max_iters_outer=1, stop="distance", rtol=1e-10, inner_cycles=4)
    base.update(kw)
    return _solver.make_config(**base)
