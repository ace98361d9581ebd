"""A cached plan serves many solves: state left by one call must not leak
into the next.

* A checkpointed (differentiable) forward solve after an ordinary solve of
  another operator on the same plan: the coarse program staged the first
  operator's coefficients; the forward value must be the second operator's
  (bit for bit vs the oracle) and its gradient must pass the transpose
  identity.
* A configuration that stages more coarse levels regrows the staged
  coefficient image; cycles captured before then must not replay against the
  old image after the operator changes.
* Mixed precision in fixed mode runs exactly max_iters FP32 cycles in a few
  refinement steps (not one cycle per step) and follows the FP64 cycles.
"""

import numpy as np
import pytest
import torch

from jsolver_amd import _solver
from jsolver_amd.grid_1d import grid_1D
from jsolver_amd.solver_2d import solve_2d
from jsolver_amd.solver_3d import solve_3d
from oracle import gmg

pytestmark = pytest.mark.gpu


def _inputs(dim, ms, seed, scale=0.4):
    rng = np.random.default_rng(seed)
    shape = tuple(m + 3 for m in reversed(ms))
    D = [np.exp(scale * rng.standard_normal(shape)) for _ in range(dim)]
    lam = rng.uniform(0, 2, shape)
    rhs = rng.standard_normal(shape)
    phi0 = rng.standard_normal(shape)
    return phi0, rhs, lam, D


def test_checkpointed_forward_after_solve_of_other_operator():
    ms, K = (32, 32), 3
    g = [grid_1D(m) for m in ms]
    og = [gmg.OracleGrid(m) for m in ms]
    dev = torch.device("cuda")
    # ordinary solve of operator A: builds and stages the coarse program
    pA, rA, lA, DA = _inputs(2, ms, 1)
    outA, _ = solve_2d(*g, pA, rA, lA, *DA, fixed=True, max_iters_outer=K)
    np.testing.assert_array_equal(outA.cpu().numpy(),
                                  gmg.solve_2d(*og, pA, rA, lA, *DA, fixed=True, max_iters_outer=K)[0])
    # differentiable solve of operator B, same grid and configuration
    pB, rB, lB, DB = _inputs(2, ms, 2)
    xs = [pB, rB, lB] + DB
    ts = [torch.tensor(x, device=dev, requires_grad=True) for x in xs]
    phi, count = solve_2d(*g, *ts, fixed=True, max_iters_outer=K)
    assert count == K
    want = gmg.solve_2d(*og, pB, rB, lB, *DB, fixed=True, max_iters_outer=K)[0]
    np.testing.assert_array_equal(phi.detach().cpu().numpy(), want)
    rng = np.random.default_rng(3)
    vs = [rng.standard_normal(pB.shape) for _ in xs]
    w = rng.standard_normal(pB.shape)
    h = 1e-30
    xc = [x + 1j * h * v for x, v in zip(xs, vs)]
    jv = np.imag(gmg.solve_2d(*og, *xc, fixed=True, max_iters_outer=K)[0]) / h
    lhs = float(np.sum(w * jv))
    phi.backward(torch.tensor(w, device=dev))
    rhs_side = sum(float(torch.sum(t.grad * torch.tensor(v, device=dev))) for t, v in zip(ts, vs))
    assert rhs_side == pytest.approx(lhs, rel=1e-11, abs=1e-11 * abs(lhs))


def test_image_regrow_then_new_operator_matches_fresh_plan():
    """32^3 mixed: V(100,100) leaves the 16^3 level out of LDS (its op list
    takes the room), V(1,1) stages it (the image grows), then V(100,100) on a
    new operator must equal the same solve on a fresh plan."""
    ms = (32, 32, 32)
    g = [grid_1D(m) for m in ms]
    kwA = dict(precision="mixed", fixed=True, max_iters_outer=2, nu1=100, nu2=100)
    kwB = dict(precision="mixed", fixed=True, max_iters_outer=2)
    p1, r1, l1, D1 = _inputs(3, ms, 11)
    p2, r2, l2, D2 = _inputs(3, ms, 12)
    _solver.clear_plans()
    solve_3d(*g, p1, r1, l1, *D1, **kwA)
    plan = _solver.get_plan(3, ms, [x.idel for x in g], "mixed")
    b_a = plan.device_bytes()
    solve_3d(*g, p1, r1, l1, *D1, **kwB)
    b_b = plan.device_bytes()
    # the 16^3 level joined the image: its coefficient and phi arrays over
    # 17^3 points (the level's rhs sits after the image), 98,260 bytes in FP32
    assert b_b - b_a >= 5 * 17 ** 3 * 4, (b_a, b_b)
    shared, _ = solve_3d(*g, p2, r2, l2, *D2, **kwA)
    _solver.clear_plans()
    fresh, _ = solve_3d(*g, p2, r2, l2, *D2, **kwA)
    np.testing.assert_array_equal(shared.cpu().numpy(), fresh.cpu().numpy())


@pytest.mark.parametrize("K", [20, 40])
def test_mixed_fixed_runs_exact_cycle_count(K):
    ms = (64, 64)
    g = [grid_1D(m) for m in ms]
    _, rhs, lam, D = _inputs(2, ms, 21, scale=1.0)
    zero = np.zeros(rhs.shape)
    mix, n, info = solve_2d(*g, zero, rhs, lam, *D, precision="mixed", fixed=True,
                            max_iters_outer=K, return_info=True)
    assert n == K and info["count"] == K
    # 8 cycles first, then steps sized by the measured contraction (never one
    # cycle per step)
    assert 2 <= info["outer_steps"] <= 1 + (K - 8) // 2, info
    ref, _ = solve_2d(*g, zero, rhs, lam, *D, fixed=True, max_iters_outer=K)
    rel = float(torch.linalg.norm(mix - ref) / torch.linalg.norm(ref))
    assert rel < 1e-5, rel
