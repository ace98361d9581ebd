"""Measurements of the benchmark configurations beside the headline one
(bench.py times the 3D 256^3 V-cycle solve).  One JSON line per
configuration, on one GPU:

    python -m jsolver_amd.configs [--only 1,2,3,4]

1. 2D 128^2, D = 1, lambda = 0, phi_ana = x(1-x)^3 y(1-y): 10 V-cycles from
   zero; RMSE against phi_ana over [1:-1, 1:-1] (the reference's measure) and
   the L2 and relative residual after every cycle; the 13-cycle RMSE next to
   the reference test's value (tests/test_solver_2d.py:190-193).
2. 2D 1024^2 anisotropic (r = 100), W-cycles to relative residual 1e-10,
   FP64 against FP32 cycles with FP64 refinement.
3. 3D 256^3 anisotropic, F-cycles to relative residual 1e-10, FP64 against
   mixed (the V-cycle case is bench.py's).
4. 2D 512^2 anisotropic (r = 10), 8 fixed V-cycles, L = 0.5 ||phi - phi_t||^2:
   time of the forward solve and of the hand-written adjoint, their ratio,
   and the peak device memory of each; the forward-mode directional
   derivative of L (tangent cycles), its time, and its agreement with the
   adjoint gradient.
"""
import argparse
import json
import sys
import time

import numpy as np
import torch

from . import _solver, adjoint, fields
from .grid_1d import grid_1D
from .solver_2d import solve_2d, solve_2d_fixed_simple
from .solver_3d import solve_3d


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def _timed(fn, reps=1, warm=True):
    if warm:
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        out = fn()
    torch.cuda.synchronize()
    return out, (time.perf_counter() - t0) / reps


def config1():
    m = 128
    g = grid_1D(m)
    x = g.centers[None, :]
    y = g.centers[:, None]
    ana = x * (1 - x) ** 3 * y * (1 - y)
    rhs = -2 * x * (1 - x) ** 3 - 6 * (1 - x) * ((1 - x) - x) * y * (1 - y)
    # dynamic rel_residual mode that never meets rtol = 0: the same iterates as
    # 10 fixed cycles, with the stop quantity recorded after every cycle
    phi, n, hist = solve_2d(g, g, np.zeros_like(rhs), rhs, 0.0, 1.0, 1.0, stop="rel_residual", rtol=0.0,
                            max_iters_outer=10, return_history=True)
    fixed, _ = solve_2d_fixed_simple(g, g, 10, np.zeros_like(rhs), rhs, 0.0, 1.0, 1.0)
    # the reference's error measure: RMSE over [1:-1, 1:-1] (boundary nodes included)
    rmse = lambda a: float(np.sqrt(np.mean((a.cpu().numpy()[1:-1, 1:-1] - ana[1:-1, 1:-1]) ** 2)))
    phi13, _ = solve_2d_fixed_simple(g, g, 13, np.zeros_like(rhs), rhs, 0.0, 1.0, 1.0)
    rhs_norm = float(np.linalg.norm(rhs[2:-2, 2:-2]))
    return {"config": "2D 128^2 plain, 10 V-cycles, FP64", "cycles": n,
            "same_as_fixed_mode": bool(torch.equal(phi, fixed)),
            "rmse_vs_analytic": rmse(phi),
            "max_abs_error_vs_analytic": float(np.abs(phi.cpu().numpy() - ana)[1:-1, 1:-1].max()),
            "l2_residual_per_cycle": [float(v) * rhs_norm for v in hist],
            "rel_residual_per_cycle": [float(v) for v in hist],
            "rmse_after_13_fixed_cycles": rmse(phi13),
            "reference_rmse_after_13_fixed_cycles": 1.010315424011278e-06}


def config2(max_cycles=20000):
    m = 1024
    (gx, gy), rhs, lam, dxx, dyy = fields.anisotropic_2d(m)
    dev = torch.device("cuda", 0)
    t = [torch.from_numpy(v).to(dev) for v in (rhs, lam, dxx, dyy)]
    z = torch.zeros_like(t[0])
    out = {"config": "2D 1024^2 anisotropic r=100, W-cycles to rel. residual 1e-10"}
    res = {}
    for prec in ("fp64", "mixed"):
        _log(f"config 2: {prec}")
        run = lambda k: solve_2d(gx, gy, z, *t, gamma=2, precision=prec, stop="rel_residual", rtol=1e-10,
                                 max_iters_outer=k, return_info=True)
        run(2)                                            # plan set-up outside the timed solve
        (phi, n, info), dt = _timed(lambda: run(max_cycles), reps=1, warm=False)
        res[prec] = phi
        out[prec] = {"time_s": round(dt, 4), "cycles": n, "rel_residual": info["rel_residual"],
                     "finite": not info["nonfinite"], "converged": info["rel_residual"] <= 1e-10}
    out["speedup_mixed"] = round(out["fp64"]["time_s"] / out["mixed"]["time_s"], 3)
    out["rel_error_mixed_vs_fp64"] = float(torch.linalg.norm(res["mixed"] - res["fp64"]) /
                                           torch.linalg.norm(res["fp64"]))
    return out


def config3():
    m = 256
    grids, rhs, lam, dxx, dyy, dzz = fields.anisotropic_3d(m)
    dev = torch.device("cuda", 0)
    t = [torch.from_numpy(v).to(dev) for v in (rhs, lam, dxx, dyy, dzz)]
    z = torch.zeros_like(t[0])
    out = {"config": "3D 256^3 anisotropic, F-cycles to rel. residual 1e-10"}
    res = {}
    for prec, ce in (("fp64", 1), ("fp64_adaptive_checks", 0), ("mixed", 1)):
        p = "mixed" if prec == "mixed" else "fp64"
        _log(f"config 3: {prec}")
        (phi, n, info), dt = _timed(lambda: solve_3d(*grids, z, *t, fcycle=True, precision=p, stop="rel_residual",
                                                     rtol=1e-10, check_every=ce, return_info=True))
        res[prec] = phi
        out[prec] = {"time_s": round(dt, 4), "cycles": n, "rel_residual": info["rel_residual"]}
    out["speedup_mixed"] = round(out["fp64"]["time_s"] / out["mixed"]["time_s"], 3)
    out["speedup_mixed_vs_adaptive"] = round(out["fp64_adaptive_checks"]["time_s"] / out["mixed"]["time_s"], 3)
    out["rel_error_mixed_vs_fp64"] = float(torch.linalg.norm(res["mixed"] - res["fp64"]) /
                                           torch.linalg.norm(res["fp64"]))
    return out


def config4():
    m = 512
    (gx, gy), rhs, lam, dxx, dyy = fields.anisotropic_2d(m, config_id=4, ratio=10.0)
    phi_t = fields.grf(rhs.shape, (gx.del_, gy.del_), 0.1, 1.0, fields.rng_for(4, 50))
    dev = torch.device("cuda", 0)
    base = [torch.from_numpy(v).to(dev) for v in (rhs, lam, dxx, dyy)]
    pt = torch.from_numpy(phi_t).to(dev)
    phi0 = torch.zeros_like(base[0])
    inner = (slice(2, -2), slice(2, -2))

    def forward(args):
        phi, _ = solve_2d_fixed_simple(gx, gy, 8, phi0, *args)
        return 0.5 * torch.sum((phi[inner] - pt[inner]) ** 2)

    with torch.no_grad():
        _, t_plain = _timed(lambda: forward(base), reps=3)
    ts = [b.clone().requires_grad_(True) for b in base]
    forward(ts).backward()                                # warm-up (plans, tapes)
    for x in ts:
        x.grad = None
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats()
    mem0 = torch.cuda.memory_allocated()     # tensors alive before (inputs, earlier configurations)
    t0 = time.perf_counter()
    L = forward(ts)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    mem_fwd = torch.cuda.max_memory_allocated()
    t0 = time.perf_counter()
    L.backward()
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t0
    mem_all = torch.cuda.max_memory_allocated()
    # forward mode: directional derivative of L along seeded directions of
    # (rhs, lambda, D_xx, D_yy) by the tangent cycles, against <grad L, v>
    gen = torch.Generator(device="cpu").manual_seed(4)
    dirs = [torch.randn(b.shape, generator=gen, dtype=torch.float64).to(dev) for b in base]
    f = lambda *a: forward(list(a))
    (_, dL), t_jvp = _timed(lambda: torch.func.jvp(f, tuple(base), tuple(dirs)), reps=3)
    dL_rev = float(sum(torch.sum(x.grad * v) for x, v in zip(ts, dirs)))
    plan = _solver.get_plan(2, (m, m), (gx.idel, gy.idel), "fp64")
    return {"config": "2D 512^2 anisotropic r=10, 8 fixed V-cycles, adjoint of L = 0.5||phi - phi_t||^2",
            "loss": float(L.detach()), "forward_s": round(t_fwd, 5), "forward_plain_s": round(t_plain, 5),
            "backward_s": round(t_bwd, 5), "backward_over_forward": round(t_bwd / t_fwd, 3),
            "backward_over_plain_forward": round(t_bwd / t_plain, 3),
            "peak_torch_bytes_forward": int(mem_fwd - mem0), "peak_torch_bytes_with_backward": int(mem_all - mem0),
            "plan_device_bytes": plan.device_bytes(), "adjoint_device_bytes": adjoint.adjoint_device_bytes(plan),
            "jvp_s": round(t_jvp, 5), "jvp_over_plain_forward": round(t_jvp / t_plain, 3),
            "jvp_directional_derivative": float(dL),
            "jvp_vs_adjoint_rel_diff": abs(float(dL) - dL_rev) / abs(dL_rev),
            "note": "torch peaks are above what was allocated before the forward pass (the inputs were): the "
                    "per-cycle checkpoints, the output and the gradients; the plan's level buffers and the "
                    "adjoint's work arrays are the library's own allocations, listed apart"}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default="1,2,3,4")
    ap.add_argument("--max-cycles", type=int, default=20000, help="cycle cap of the 2D W-cycle solves")
    a = ap.parse_args(argv)
    todo = {"1": config1, "2": lambda: config2(a.max_cycles), "3": config3, "4": config4}
    for k in a.only.split(","):
        print(json.dumps(todo[k.strip()]()), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
