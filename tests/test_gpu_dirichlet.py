"""Non-zero Dirichlet values (boundary="phi"): the boundary layers of the
incoming phi, as do_BCs(phi, vals_...) sets them (boundary_handler_2d.py:3-29),
are the boundary values of the solve.  The reference's solver always applies
zeros (README.md:90: its calls "would need to be changed").

  * the GPU lift (jsv_dirichlet_lift) equals the oracle's, bit for bit;
  * fixed solves equal the oracle's lift + zero-boundary solve + restored
    boundary, bit for bit (2D, D_xy, 3D, batched);
  * a quadratic solution (which the 5/7-point operator reproduces exactly)
    is recovered from its boundary values alone; the oracle's lift itself is
    pinned against a direct solve of the Dirichlet system
    (tests/test_oracle_pins.py).
"""
import numpy as np
import pytest
import torch

from gpu_helpers import DevPlan, rand_problem
from jsolver_amd import boundary_handler_2d, boundary_handler_3d
from jsolver_amd._solver import solve_batched_2d
from jsolver_amd.grid_1d import grid_1D
from jsolver_amd.solver_2d import solve_2d
from jsolver_amd.solver_3d import solve_3d
from oracle import gmg

pytestmark = pytest.mark.gpu

CASES = [
    ("2d", 2, (16, 16), True, False),
    ("2d_32x16_scalarD", 2, (32, 16), False, False),
    ("2d_dxy", 2, (16, 32), True, True),
    ("3d", 3, (8, 16, 8), True, False),
    ("3d_scalarD", 3, (16, 16, 16), False, False),
]


def _case(case, seed):
    _, dim, ms, spatial, dxy = case
    phi, rhs, lam, D, Dxy = rand_problem(dim, ms, seed, spatial, True, dxy)
    rng = np.random.default_rng(seed + 100)
    pb = rng.standard_normal(phi.shape)      # boundary values on every layer
    return dim, ms, phi, rhs, lam, D, Dxy, pb


def _oracle_solve(dim, ms, pb, rhs, lam, D, Dxy, K, **kw):
    og = [gmg.OracleGrid(m) for m in ms]
    op = gmg.build_operator(dim, ms, tuple(g.idel for g in og), lam, list(D), Dxy)
    rl = gmg.dirichlet_lift(op, pb, rhs)
    out, _ = gmg.solve(op, gmg.Config(fixed=True, max_iters_outer=K, **kw), pb, rl)
    inner = (slice(2, -2),) * dim
    res = pb.copy()
    res[inner] = out[inner]
    return res, rl


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_lift_bitexact(case):
    dim, ms, phi, rhs, lam, D, Dxy, pb = _case(case, 3)
    dp = DevPlan(dim, ms, lam, D, Dxy)
    from jsolver_amd import _solver
    dev = dp.plan.dev
    got = _solver.dirichlet_lift(dp.plan, torch.tensor(pb, device=dev), torch.tensor(rhs, device=dev))
    _, want = _oracle_solve(dim, ms, pb, rhs, lam, D, Dxy, 0)
    np.testing.assert_array_equal(got.cpu().numpy(), want)
    assert np.count_nonzero(want != rhs) > 0


@pytest.mark.parametrize("kw", [dict(), dict(gamma=2), dict(fcycle=True, skip_pre=True)], ids=["V", "W", "F_skip"])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fixed_solve_bitexact(case, kw):
    dim, ms, phi, rhs, lam, D, Dxy, pb = _case(case, 5)
    K = 3
    g = [grid_1D(m) for m in ms]
    if dim == 2:
        out, n = solve_2d(*g, pb, rhs, lam, *D, D_xy=Dxy, fixed=True, max_iters_outer=K, boundary="phi", **kw)
    else:
        out, n = solve_3d(*g, pb, rhs, lam, *D, fixed=True, max_iters_outer=K, boundary="phi", **kw)
    want, _ = _oracle_solve(dim, ms, pb, rhs, lam, D, Dxy, K, **kw)
    assert n == K
    np.testing.assert_array_equal(out.cpu().numpy(), want)


def test_batched_equals_single():
    ms = (32, 32)
    g = [grid_1D(m) for m in ms]
    B = 3
    probs = [_case(("b", 2, ms, True, False), 20 + b) for b in range(B)]
    stack = lambda i: np.stack([p[i] for p in probs])
    pbs, rhss, lams = stack(7), stack(3), stack(4)
    Ds = [np.stack([p[5][a] for p in probs]) for a in range(2)]
    out, counts, _ = solve_batched_2d(*g, pbs, rhss, lams, *Ds, boundary="phi", epsilon=1e-10)
    for b in range(B):
        one, n = solve_2d(*g, pbs[b], rhss[b], lams[b], Ds[0][b], Ds[1][b], boundary="phi", epsilon=1e-10)
        assert n == counts[b]
        np.testing.assert_array_equal(out[b].cpu().numpy(), one.cpu().numpy())


@pytest.mark.parametrize("dim", [2, 3])
def test_quadratic_solution_from_boundary_values(dim):
    """phi = x^2 + 2 y^2 (+ 3 z^2) solves div grad phi = 6 (12) exactly on the
    grid (the 5/7-point operator is exact for quadratics); with D = 1, lambda =
    0 and only the boundary values given, the converged solve reproduces it.
    The boundary values are set with do_BCs(phi, vals_...), the reference's
    own way of applying them."""
    m = 32 if dim == 2 else 16
    g = grid_1D(m)
    x = np.asarray(g.centers)
    if dim == 2:
        Y, X = np.meshgrid(x, x, indexing="ij")
        ana = X ** 2 + 2 * Y ** 2
        rhs = np.full(ana.shape, 6.0)
        p0 = boundary_handler_2d.do_BCs(np.zeros(ana.shape), vals_xb=ana[:, 1], vals_xe=ana[:, -2],
                                        vals_yb=ana[1, :], vals_ye=ana[-2, :])
        p0 = np.asarray(p0.cpu() if torch.is_tensor(p0) else p0)
        # do_BCs sets the ghost layer and the boundary node to the same value;
        # only the boundary node (index 1, m+1) enters the stencil
        out, n = solve_2d(g, g, p0, rhs, 0.0, 1.0, 1.0, boundary="phi", stop="rel_residual", rtol=1e-13)
    else:
        Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
        ana = X ** 2 + 2 * Y ** 2 + 3 * Z ** 2
        rhs = np.full(ana.shape, 12.0)
        p0 = ana.copy()
        p0[2:-2, 2:-2, 2:-2] = 0.0          # boundary layers only
        out, n = solve_3d(g, g, g, p0, rhs, 0.0, 1.0, 1.0, 1.0, boundary="phi", stop="rel_residual",
                          rtol=1e-13)
    got = out.cpu().numpy()
    inner = (slice(2, -2),) * dim
    assert np.max(np.abs(got[inner] - ana[inner])) < 1e-10 * np.max(np.abs(ana))
    # the boundary layers are the given values
    np.testing.assert_array_equal(got[(slice(1, 2),) + (slice(None),) * (dim - 1)],
                                  p0[(slice(1, 2),) + (slice(None),) * (dim - 1)])
    _ = boundary_handler_3d


def test_zero_boundary_values_equal_default():
    dim, ms, phi, rhs, lam, D, Dxy, pb = _case(CASES[0], 9)
    g = [grid_1D(m) for m in ms]
    p0 = gmg.zero_rim(phi)
    a, na = solve_2d(*g, p0, rhs, lam, *D, boundary="phi")
    b, nb = solve_2d(*g, p0, rhs, lam, *D)
    assert na == nb
    np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())


def test_errors():
    dim, ms, phi, rhs, lam, D, Dxy, pb = _case(CASES[0], 1)
    g = [grid_1D(m) for m in ms]
    with pytest.raises(ValueError, match="boundary"):
        solve_2d(*g, pb, rhs, lam, *D, boundary="one")
    t = torch.tensor(rhs, device="cuda", requires_grad=True)
    with pytest.raises(NotImplementedError, match="boundary='phi'"):
        solve_2d(*g, pb, t, lam, *D, fixed=True, max_iters_outer=2, boundary="phi")
