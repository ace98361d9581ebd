"""Tangent (forward-mode) pass of solves: torch.func.jvp / forward AD through
solve_2d and solve_3d, computed by the hand-written tangent cycles.

The reference's check is jax.test_util.check_grads(modes=["fwd"]) on its
solves (tests/test_solver_2d.py:173-176), an FD consistency check.  Here:
  * J v against complex-step differentiation of the CPU oracle (exact to
    ~1e-15), every input (phi0, rhs, lambda, D components, D_xy) at once, for
    V/W/F cycles, skip_pre and nu variants, 2D and 3D;
  * with tangents only on (phi0, rhs) the cycle is linear, so J v equals the
    solve run on the tangents, bit for bit (GPU solve and oracle);
  * dynamic mode: the tangent of the cycle count the primal stop rule chose;
  * scalar D / lambda tangents, central finite differences at 255^2, errors.
"""
import numpy as np
import pytest
import torch
import torch.autograd.forward_ad as fwAD

from jsolver_amd.grid_1d import grid_1D
from jsolver_amd.solver_2d import solve_2d, solve_2d_fixed_simple
from jsolver_amd.solver_3d import solve_3d
from oracle import gmg

pytestmark = pytest.mark.gpu

CASES = [
    ("2d_V", 2, (16, 16), dict(), False),
    ("2d_V_16x32", 2, (16, 32), dict(), False),
    ("2d_W", 2, (16, 16), dict(gamma=2), False),
    ("2d_F", 2, (16, 16), dict(fcycle=True), False),
    ("2d_skip", 2, (16, 16), dict(skip_pre=True), False),
    ("2d_nu2_0", 2, (16, 16), dict(nu1=2, nu2=0), False),
    ("2d_nu_2_3", 2, (16, 16), dict(nu1=2, nu2=3), False),
    ("2d_Dxy_V", 2, (16, 16), dict(), True),
    ("2d_Dxy_F_skip", 2, (32, 16), dict(fcycle=True, skip_pre=True), True),
    ("3d_V", 3, (8, 8, 8), dict(), False),
    ("3d_V_8x16x8", 3, (8, 16, 8), dict(), False),
    ("3d_F", 3, (8, 8, 8), dict(fcycle=True), False),
    ("3d_W_nu2_0", 3, (8, 8, 8), dict(gamma=2, nu1=1, nu2=0), False),
    ("3d_skip_16", 3, (16, 16, 16), dict(skip_pre=True), False),
]


def _inputs(dim, ms, seed, dxy=False):
    rng = np.random.default_rng(seed)
    shape = tuple(m + 3 for m in reversed(ms))
    D = [np.exp(0.4 * rng.standard_normal(shape)) for _ in range(dim)]
    lam = rng.uniform(0, 2, shape)
    rhs = rng.standard_normal(shape)
    phi0 = rng.standard_normal(shape)
    Dxy = 0.1 * rng.standard_normal(shape) if dxy else None
    return phi0, rhs, lam, D, Dxy


def _oracle(dim, ms, xs, Dxy, K, **kw):
    og = [gmg.OracleGrid(m) for m in ms]
    if dim == 2:
        return gmg.solve_2d(*og, xs[0], xs[1], xs[2], xs[3], xs[4], D_xy=Dxy, fixed=True,
                            max_iters_outer=K, **kw)[0]
    return gmg.solve_3d(*og, *xs, fixed=True, max_iters_outer=K, **kw)[0]


def _gpu_solve(dim, ms):
    g = [grid_1D(m) for m in ms]
    return (lambda *a, **kw: solve_2d(*g, *a, **kw)) if dim == 2 else (lambda *a, **kw: solve_3d(*g, *a, **kw))


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_jvp_vs_complex_step(case):
    _, dim, ms, kw, dxy = case
    K = 3
    phi0, rhs, lam, D, Dxy = _inputs(dim, ms, 17, dxy)
    rng = np.random.default_rng(99)
    xs = [phi0, rhs, lam] + D + ([Dxy] if dxy else [])
    vs = [rng.standard_normal(phi0.shape) for _ in xs]
    # J v by complex step on the oracle, all directions at once
    h = 1e-30
    xc = [x + 1j * h * v for x, v in zip(xs, vs)]
    jv = np.imag(_oracle(dim, ms, xc[:3 + dim], xc[3 + dim] if dxy else None, K, **kw)) / h
    want = _oracle(dim, ms, xs[:3 + dim], Dxy, K, **kw)
    dev = torch.device("cuda")
    solve = _gpu_solve(dim, ms)
    ts = [torch.tensor(x, device=dev) for x in xs]
    tv = [torch.tensor(v, device=dev) for v in vs]

    def f(*a):
        extra = dict(D_xy=a[3 + dim]) if dxy else {}
        return solve(*a[:3 + dim], fixed=True, max_iters_outer=K, **extra, **kw)[0]

    out, out_dot = torch.func.jvp(f, tuple(ts), tuple(tv))
    np.testing.assert_array_equal(out.cpu().numpy(), want)
    got = out_dot.cpu().numpy()
    scale = np.max(np.abs(jv))
    np.testing.assert_allclose(got, jv, rtol=1e-9, atol=1e-11 * scale)
    # per input: each tangent direction alone (forward AD dual tensors)
    for i in range(len(xs)):
        xc1 = [x + (1j * h * v if j == i else 0) for j, (x, v) in enumerate(zip(xs, vs))]
        jv1 = np.imag(_oracle(dim, ms, xc1[:3 + dim], xc1[3 + dim] if dxy else None, K, **kw)) / h
        with fwAD.dual_level():
            args = [fwAD.make_dual(t, v) if j == i else t for j, (t, v) in enumerate(zip(ts, tv))]
            o = f(*args)
            t1 = fwAD.unpack_dual(o).tangent
        assert t1 is not None
        np.testing.assert_allclose(t1.cpu().numpy(), jv1, rtol=1e-9,
                                   atol=1e-11 * max(scale, np.max(np.abs(jv1)))), f"input {i}"


@pytest.mark.parametrize("dim,ms,kw", [(2, (32, 32), dict(gamma=2)), (3, (16, 16, 16), dict(fcycle=True))])
def test_linear_tangent_equals_solve_on_tangents(dim, ms, kw):
    """Tangents on (phi0, rhs) only: the cycle is affine in (phi, rhs), so J v
    is the solve on (phi0_dot, rhs_dot), bit for bit."""
    K = 4
    phi0, rhs, lam, D, _ = _inputs(dim, ms, 5)
    rng = np.random.default_rng(6)
    dphi, drhs = rng.standard_normal(phi0.shape), rng.standard_normal(phi0.shape)
    dev = torch.device("cuda")
    solve = _gpu_solve(dim, ms)
    lam_t = torch.tensor(lam, device=dev)
    D_t = [torch.tensor(d, device=dev) for d in D]
    f = lambda p, r: solve(p, r, lam_t, *D_t, fixed=True, max_iters_outer=K, **kw)[0]
    _, jv = torch.func.jvp(f, (torch.tensor(phi0, device=dev), torch.tensor(rhs, device=dev)),
                           (torch.tensor(dphi, device=dev), torch.tensor(drhs, device=dev)))
    lin = solve(torch.tensor(dphi, device=dev), torch.tensor(drhs, device=dev), lam_t, *D_t, fixed=True,
                max_iters_outer=K, **kw)[0]
    np.testing.assert_array_equal(jv.cpu().numpy(), lin.cpu().numpy())
    np.testing.assert_array_equal(jv.cpu().numpy(), _oracle(dim, ms, [dphi, drhs, lam] + D, None, K, **kw))


@pytest.mark.parametrize("dim,ms,stop", [(2, (32, 32), "distance"), (3, (16, 16, 16), "distance"),
                                         (3, (16, 16, 16), "rel_residual")])
def test_dynamic_mode_tangent(dim, ms, stop):
    """Dynamic solve: same count as without tangents; the tangent is that of
    the fixed solve with that many cycles (jax.jvp through the while loop)."""
    phi0, rhs, lam, D, _ = _inputs(dim, ms, 8)
    phi0 = np.zeros_like(phi0)
    rng = np.random.default_rng(9)
    vs = [rng.standard_normal(rhs.shape) for _ in range(2 + dim)]   # rhs, lam, D...
    dev = torch.device("cuda")
    solve = _gpu_solve(dim, ms)
    base = [torch.tensor(x, device=dev) for x in [rhs, lam] + D]
    p0 = torch.tensor(phi0, device=dev)
    ref, count = solve(p0, *base, stop=stop)
    with fwAD.dual_level():
        args = [fwAD.make_dual(b, torch.tensor(v, device=dev)) for b, v in zip(base, vs)]
        out, c2 = solve(p0, *args, stop=stop)
        prim, tan = fwAD.unpack_dual(out)
    assert c2 == count and count > 3
    np.testing.assert_array_equal(prim.cpu().numpy(), ref.cpu().numpy())
    h = 1e-30
    xc = [phi0] + [x + 1j * h * v for x, v in zip([rhs, lam] + D, vs)]
    jv = np.imag(_oracle(dim, ms, xc, None, count)) / h
    np.testing.assert_allclose(tan.cpu().numpy(), jv, rtol=1e-8, atol=1e-11 * np.max(np.abs(jv)))


def test_scalar_coefficient_tangents():
    """0-d tensor D components and lambda carry tangents (scalar branch of the
    coefficient setup, solver_2d.py:417-421)."""
    ms = (32, 32)
    K = 3
    rng = np.random.default_rng(3)
    shape = (35, 35)
    rhs, phi0 = rng.standard_normal(shape), rng.standard_normal(shape)
    dx, dy, lam = 1.3, 0.7, 0.4
    vx, vy, vl = 0.3, -1.1, 2.0
    dev = torch.device("cuda")
    g = [grid_1D(m) for m in ms]
    f = lambda a, b, c: solve_2d(*g, torch.tensor(phi0, device=dev), torch.tensor(rhs, device=dev), c, a, b,
                                 fixed=True, max_iters_outer=K)[0]
    sc = lambda v: torch.tensor(v, dtype=torch.float64, device=dev)
    _, jv = torch.func.jvp(f, (sc(dx), sc(dy), sc(lam)), (sc(vx), sc(vy), sc(vl)))
    h = 1e-30
    og = [gmg.OracleGrid(m) for m in ms]
    want = np.imag(gmg.solve_2d(*og, phi0, rhs, lam + 1j * h * vl, dx + 1j * h * vx, dy + 1j * h * vy,
                                fixed=True, max_iters_outer=K)[0]) / h
    np.testing.assert_allclose(jv.cpu().numpy(), want, rtol=1e-9, atol=1e-11 * np.max(np.abs(want)))


def test_central_fd_255():
    """Anisotropic 256^2 problem, 8 fixed V-cycles: directional derivative in a
    random direction of (rhs, lambda, D_xx, D_yy) against central finite
    differences of the GPU solve (step 1e-6)."""
    from jsolver_amd import fields
    m = 256
    (gx, gy), rhs, lam, dxx, dyy = fields.anisotropic_2d(m, config_id=4, ratio=10.0)
    rng = np.random.default_rng(41)
    dev = torch.device("cuda")
    base = [torch.tensor(a, device=dev) for a in (rhs, lam, dxx, dyy)]
    dirs = [torch.tensor(rng.standard_normal(rhs.shape), device=dev) for _ in range(4)]
    dirs[2] *= base[2]   # relative perturbation of the positive D fields
    dirs[3] *= base[3]
    phi0 = torch.zeros_like(base[0])
    f = lambda r, l, a, b: solve_2d_fixed_simple(gx, gy, 8, phi0, r, l, a, b)[0]
    _, jv = torch.func.jvp(f, tuple(base), tuple(dirs))
    eps = 1e-6
    with torch.no_grad():
        fp = f(*[b + eps * d for b, d in zip(base, dirs)])
        fm = f(*[b - eps * d for b, d in zip(base, dirs)])
    fd = (fp - fm) / (2 * eps)
    err = float(torch.linalg.norm(jv - fd) / torch.linalg.norm(fd))
    assert err < 1e-6, err


def test_errors():
    ms = (16, 16)
    phi0, rhs, lam, D, _ = _inputs(2, ms, 1)
    dev = torch.device("cuda")
    g = [grid_1D(m) for m in ms]
    ts = [torch.tensor(x, device=dev) for x in [phi0, rhs, lam] + D]
    with fwAD.dual_level():
        r = fwAD.make_dual(ts[1], torch.ones_like(ts[1]))
        with pytest.raises(ValueError, match="precision='fp64'"):
            solve_2d(*g, ts[0], r, ts[2], ts[3], ts[4], precision="mixed")
        with pytest.raises(ValueError, match="return_history"):
            solve_2d(*g, ts[0], r, ts[2], ts[3], ts[4], fixed=True, max_iters_outer=2,
                     return_history=True)
