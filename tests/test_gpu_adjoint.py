"""Reverse pass (hand-written adjoint) of fixed-mode solves.

The reference's check is jax.test_util.check_grads(modes=["rev"])
(tests/test_solver_2d.py:200-207), an FD consistency check.  Here:
  * transpose identity <w, J v> = <J^T w, v> with J v computed EXACTLY (to
    ~1e-15) by complex-step differentiation of the CPU oracle, J^T w by the
    GPU adjoint, on every input (phi0, rhs, lambda, D components) at once;
  * central finite differences (step 1e-6) of L = 0.5 ||phi - phi_t||^2
    through the GPU forward at 515^2 on 32 seeded random cells;
  * torch.autograd integration (phi.backward through solve_2d).
"""
import numpy as np
import pytest
import torch

from jsolver_amd.grid_1d import grid_1D
from jsolver_amd.solver_2d import solve_2d, solve_2d_fixed_simple
from jsolver_amd.solver_3d import solve_3d
from oracle import gmg

pytestmark = pytest.mark.gpu

CASES = [
    ("2d_V", 2, (16, 16), dict()),
    ("2d_V_16x32", 2, (16, 32), dict()),
    ("2d_W", 2, (16, 16), dict(gamma=2)),
    ("2d_F", 2, (16, 16), dict(fcycle=True)),
    ("2d_skip", 2, (16, 16), dict(skip_pre=True)),
    ("2d_nu2_0", 2, (16, 16), dict(nu1=2, nu2=0)),
    ("2d_nu_2_3", 2, (16, 16), dict(nu1=2, nu2=3)),
    ("3d_V", 3, (8, 8, 8), dict()),
    ("3d_V_8x16x8", 3, (8, 16, 8), dict()),
    ("3d_F", 3, (8, 8, 8), dict(fcycle=True)),
    ("3d_W_nu2_0", 3, (8, 8, 8), dict(gamma=2, nu1=1, nu2=0)),
]


def _inputs(dim, ms, seed):
    rng = np.random.default_rng(seed)
    shape = tuple(m + 3 for m in reversed(ms))
    D = [np.exp(0.4 * rng.standard_normal(shape)) for _ in range(dim)]
    lam = rng.uniform(0, 2, shape)
    rhs = rng.standard_normal(shape)
    phi0 = rng.standard_normal(shape)
    return phi0, rhs, lam, D


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_transpose_identity_vs_complex_step(case):
    _, dim, ms, kw = case
    K = 3
    phi0, rhs, lam, D = _inputs(dim, ms, 17)
    rng = np.random.default_rng(99)
    shape = rhs.shape
    vs = [rng.standard_normal(shape) for _ in range(3 + dim)]   # phi0, rhs, lam, D...
    w = rng.standard_normal(shape)
    # J v by complex step on the oracle
    h = 1e-30
    og = [gmg.OracleGrid(m) for m in ms]
    xs = [phi0, rhs, lam] + D
    xc = [x + 1j * h * v for x, v in zip(xs, vs)]
    f = gmg.solve_2d if dim == 2 else gmg.solve_3d
    out = f(*og, xc[0], xc[1], xc[2], *xc[3:], fixed=True, max_iters_outer=K, **kw)[0]
    jv = np.imag(out) / h
    lhs = float(np.sum(w * jv))
    # J^T w by the GPU adjoint through torch.autograd
    dev = torch.device("cuda")
    ts = [torch.tensor(x, device=dev, requires_grad=True) for x in xs]
    g = [grid_1D(m) for m in ms]
    if dim == 2:
        phi, count = solve_2d(*g, ts[0], ts[1], ts[2], ts[3], ts[4], fixed=True, max_iters_outer=K, **kw)
    else:
        phi, count = solve_3d(*g, ts[0], ts[1], ts[2], ts[3], ts[4], ts[5], fixed=True,
                              max_iters_outer=K, **kw)
    assert count == K
    # the forward value equals the oracle's (bit for bit)
    want = f(*og, *xs[:3], *xs[3:], fixed=True, max_iters_outer=K, **kw)[0]
    np.testing.assert_array_equal(phi.detach().cpu().numpy(), want)
    phi.backward(torch.tensor(w, device=dev))
    rhs_side = sum(float(torch.sum(t.grad * torch.tensor(v, device=dev))) for t, v in zip(ts, vs))
    assert rhs_side == pytest.approx(lhs, rel=1e-11, abs=1e-11 * abs(lhs))
    # and per input (each cotangent separately)
    for i, (t, v) in enumerate(zip(ts, vs)):
        xc1 = [x + (1j * h * vv if j == i else 0) for j, (x, vv) in enumerate(zip(xs, vs))]
        jv1 = np.imag(f(*og, xc1[0], xc1[1], xc1[2], *xc1[3:], fixed=True, max_iters_outer=K, **kw)[0]) / h
        a = float(np.sum(w * jv1))
        b = float(torch.sum(t.grad * torch.tensor(v, device=dev)))
        assert b == pytest.approx(a, rel=1e-10, abs=1e-12 * max(1.0, abs(lhs))), f"input {i}"


def test_central_fd_515():
    """C4-like configuration: 512^2 interior, log-normal D (r = 10), 8 fixed
    V-cycles, L = 0.5 ||phi - phi_t||^2; gradient vs central FD (step 1e-6)
    on 32 random cells for rhs, lambda, D_xx, D_yy."""
    from jsolver_amd import fields
    m = 512
    (gx, gy), rhs, lam, dxx, dyy = fields.anisotropic_2d(m, config_id=4, ratio=10.0)
    rng = fields.rng_for(4, 50)
    phi_t = fields.grf(rhs.shape, (gx.del_, gy.del_), 0.1, 1.0, rng)
    dev = torch.device("cuda")
    base = [torch.tensor(a, device=dev) for a in (rhs, lam, dxx, dyy)]
    ts = [b.clone().requires_grad_(True) for b in base]
    phi0 = torch.zeros_like(base[0])
    pt = torch.tensor(phi_t, device=dev)
    inner = (slice(2, -2), slice(2, -2))

    def loss(r, l, a, b):
        phi, _ = solve_2d_fixed_simple(gx, gy, 8, phi0, r, l, a, b)
        return 0.5 * torch.sum((phi[inner] - pt[inner]) ** 2)

    L = loss(*ts)
    L.backward()
    cells = np.stack([rng.integers(2, m + 1, 32), rng.integers(2, m + 1, 32)], axis=1)
    eps = 1e-6
    with torch.no_grad():
        for k in range(4):
            for (iy, ix) in cells:
                args_p = [b.clone() for b in base]
                args_m = [b.clone() for b in base]
                args_p[k][iy, ix] += eps
                args_m[k][iy, ix] -= eps
                fd = (float(loss(*args_p)) - float(loss(*args_m))) / (2 * eps)
                gr = float(ts[k].grad[iy, ix])
                assert gr == pytest.approx(fd, rel=2e-5, abs=1e-7 * max(1.0, abs(float(L)))), (k, iy, ix)


def test_dynamic_mode_not_differentiable():
    phi0, rhs, lam, D = _inputs(2, (16, 16), 3)
    g = [grid_1D(16), grid_1D(16)]
    t = torch.tensor(rhs, device="cuda", requires_grad=True)
    with pytest.raises(ValueError):
        solve_2d(*g, phi0, t, lam, *D)


@pytest.mark.parametrize("case", [CASES[0], CASES[2], CASES[3], CASES[7], CASES[9]], ids=lambda c: c[0])
def test_graph_replayed_adjoint_equals_eager(case):
    """The reversed cycles replay one captured graph per configuration: the
    gradients equal the launch-by-launch pass bit for bit, on first capture
    and on later replays with other inputs."""
    from jsolver_amd import _lib, _solver
    _, dim, ms, kw = case
    g = [grid_1D(m) for m in ms]
    dev = torch.device("cuda")
    plan = _solver.get_plan(dim, ms, [gg.idel for gg in g], "fp64")
    f = solve_2d if dim == 2 else solve_3d

    def grads(seed, graphs):
        _lib.check(_lib.load().jsv_plan_set_tuning(plan.h, _lib.JSV_TUNE_GRAPHS, graphs))
        phi0, rhs, lam, D = _inputs(dim, ms, seed)
        ts = [torch.tensor(x, device=dev, requires_grad=True) for x in [phi0, rhs, lam] + D]
        phi, _ = f(*g, *ts, fixed=True, max_iters_outer=3, **kw)
        phi.backward(torch.tensor(np.random.default_rng(seed + 1).standard_normal(rhs.shape), device=dev))
        return [t.grad.cpu().numpy() for t in ts]

    try:
        for seed in (5, 6, 7):
            eager = grads(seed, 0)
            replay = grads(seed, 1)
            for a, b in zip(eager, replay):
                np.testing.assert_array_equal(a, b)
    finally:
        _lib.check(_lib.load().jsv_plan_set_tuning(plan.h, _lib.JSV_TUNE_GRAPHS, 1))


# ---------------------------------------------------------------------------
# every operator input of the reference's jax.vjp: scalar lambda / D (the
# reference's own reverse check, tests/test_solver_2d.py:200-207) and the 2D
# D_xy array (examples/benchmark.py:1223-1229)
# ---------------------------------------------------------------------------
def _vjp_check(dim, ms, xs, vs, K, kw, rel=1e-11):
    """<w, J v> (complex step of the oracle) against sum <J^T w, v> (GPU
    adjoint through torch.autograd), all inputs at once and one by one.
    xs: [phi0, rhs, lam, D..., (Dxy)]; lam and D may be scalars."""
    og = [gmg.OracleGrid(m) for m in ms]
    f = gmg.solve_2d if dim == 2 else gmg.solve_3d
    nd = dim
    has_dxy = len(xs) == 3 + nd + 1

    def oracle(args):
        extra = dict(D_xy=args[3 + nd]) if has_dxy else {}
        return f(*og, args[0], args[1], args[2], *args[3:3 + nd], fixed=True, max_iters_outer=K,
                 **extra, **kw)[0]

    h = 1e-30
    shape = np.shape(xs[0])
    w = np.random.default_rng(123).standard_normal(shape)
    dev = torch.device("cuda")
    ts = [torch.tensor(x, device=dev, dtype=torch.float64, requires_grad=True) for x in xs]
    g = [grid_1D(m) for m in ms]
    extra = dict(D_xy=ts[3 + nd]) if has_dxy else {}
    solve = solve_2d if dim == 2 else solve_3d
    phi, count = solve(*g, ts[0], ts[1], ts[2], *ts[3:3 + nd], fixed=True, max_iters_outer=K, **extra, **kw)
    assert count == K
    np.testing.assert_array_equal(phi.detach().cpu().numpy(), oracle(xs))
    phi.backward(torch.tensor(w, device=dev))
    for i, t in enumerate(ts):
        assert t.grad is not None, i
        assert tuple(t.grad.shape) == tuple(np.shape(xs[i])), i
    total_l, total_r = 0.0, 0.0
    for i, (t, v) in enumerate(zip(ts, vs)):
        xc = [x + (1j * h * vv if j == i else 0) for j, (x, vv) in enumerate(zip(xs, vs))]
        a = float(np.sum(w * np.imag(oracle(xc)) / h))
        b = float(torch.sum(t.grad * torch.tensor(v, device=dev, dtype=torch.float64)))
        scale = max(abs(a), 1e-300)
        assert b == pytest.approx(a, rel=rel, abs=1e-12 * scale), (i, a, b)
        total_l += a
        total_r += b
    assert total_r == pytest.approx(total_l, rel=rel)


def test_reverse_problem1_scalar_lambda_and_D_128():
    """The reference's reverse-mode check: problem 1 at 128^2 (D_xx = D_yy = 1,
    lambda = 0, all scalars), 13 fixed V-cycles, gradients with respect to
    phi, rhs, lambda, D_xx and D_yy."""
    from problems import problem_2d
    m = 128
    gx = grid_1D(m)
    _, rhs, lam, dxx, dyy, _ = problem_2d(1, gx.centers, gx.centers)
    phi0 = np.zeros(rhs.shape)
    rng = np.random.default_rng(5)
    xs = [phi0, rhs, lam, dxx, dyy]
    vs = [rng.standard_normal(rhs.shape), rng.standard_normal(rhs.shape), 0.7, -0.4, 1.3]
    _vjp_check(2, (m, m), xs, vs, 13, {}, rel=1e-10)


@pytest.mark.parametrize("dim,ms,kw", [
    (2, (16, 32), dict()),
    (2, (16, 16), dict(gamma=2)),
    (3, (8, 8, 16), dict()),
    (3, (8, 8, 8), dict(fcycle=True, nu1=2, nu2=1)),
])
@pytest.mark.parametrize("which", ["scalar_both", "scalar_D", "scalar_lam"])
def test_reverse_scalar_parts(dim, ms, kw, which):
    """Scalar D takes the reference's scalar path (no face averaging; in 3D it
    differs from a constant D array on coarse levels), so its cotangent is its
    own, not the sum over a broadcast array's cells."""
    rng = np.random.default_rng(31)
    shape = tuple(m + 3 for m in reversed(ms))
    phi0, rhs = rng.standard_normal(shape), rng.standard_normal(shape)
    lam = 0.8 if which in ("scalar_both", "scalar_lam") else rng.uniform(0, 2, shape)
    if which in ("scalar_both", "scalar_D"):
        D = [1.3, 0.6, 2.0][:dim]
        vD = [0.5, -1.1, 0.9][:dim]
    else:
        D = [np.exp(0.4 * rng.standard_normal(shape)) for _ in range(dim)]
        vD = [rng.standard_normal(shape) for _ in range(dim)]
    vlam = -0.3 if np.ndim(lam) == 0 else rng.standard_normal(shape)
    xs = [phi0, rhs, lam] + D
    vs = [rng.standard_normal(shape), rng.standard_normal(shape), vlam] + vD
    _vjp_check(dim, ms, xs, vs, 3, kw)


@pytest.mark.parametrize("ms,kw", [
    ((16, 16), dict()),
    ((16, 32), dict()),
    ((16, 16), dict(gamma=2)),
    ((16, 16), dict(fcycle=True)),
    ((16, 16), dict(skip_pre=True)),
    ((16, 16), dict(nu1=2, nu2=0)),
    ((32, 16), dict(nu1=1, nu2=2)),
])
def test_reverse_dxy(ms, kw):
    """2D with D_xy (the 9-point stencil, four sequential sub-colours): the
    gradient with respect to every input, D_xy included."""
    rng = np.random.default_rng(47)
    shape = tuple(m + 3 for m in reversed(ms))
    phi0, rhs = rng.standard_normal(shape), rng.standard_normal(shape)
    lam = rng.uniform(0, 2, shape)
    D = [np.exp(0.3 * rng.standard_normal(shape)) + 1.0 for _ in range(2)]
    Dxy = 0.2 * rng.standard_normal(shape)
    xs = [phi0, rhs, lam] + D + [Dxy]
    vs = [rng.standard_normal(shape) for _ in xs]
    _vjp_check(2, ms, xs, vs, 3, kw)


def test_reverse_dxy_problem7_fd():
    """Problem 7 of the reference's tests (spatial D with D_xy) at 64^2, 6 fixed
    V-cycles: the D_xy gradient of L = 0.5 ||phi||^2 against central finite
    differences on 8 cells."""
    from problems import problem_2d
    m = 64
    gx = grid_1D(m)
    _, rhs, lam, dxx, dyy, dxy = problem_2d(7, gx.centers, gx.centers)
    dev = torch.device("cuda")
    base = torch.tensor(dxy, device=dev)
    args = [torch.tensor(np.broadcast_to(a, rhs.shape).copy(), device=dev) for a in (rhs, lam, dxx, dyy)]
    phi0 = torch.zeros_like(args[0])
    inner = (slice(2, -2), slice(2, -2))

    def loss(q):
        phi, _ = solve_2d(gx, gx, phi0, *args, D_xy=q, fixed=True, max_iters_outer=6)
        return 0.5 * torch.sum(phi[inner] ** 2)

    t = base.clone().requires_grad_(True)
    L = loss(t)
    L.backward()
    rng = np.random.default_rng(9)
    eps = 1e-6
    with torch.no_grad():
        for (iy, ix) in zip(rng.integers(2, m + 1, 8), rng.integers(2, m + 1, 8)):
            qp, qm = base.clone(), base.clone()
            qp[iy, ix] += eps
            qm[iy, ix] -= eps
            fd = (float(loss(qp)) - float(loss(qm))) / (2 * eps)
            # FD rounding noise ~ 1e-16 |L| / eps
            assert float(t.grad[iy, ix]) == pytest.approx(fd, rel=2e-5, abs=1e-7 * max(1.0, abs(float(L))))
