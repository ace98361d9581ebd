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
