"""The dynamic stop rule evaluated on the device (JSV_TUNE_DEVICE_STOP).

A dynamic solve of one problem that tests every cycle (the reference's
distance rule, solver_2d.py:464-477, or the relative residual checked every
cycle) runs its test inside each replayed cycle and reads it back once per
block of cycles.  Against the host reading every test (the knob off, itself
checked against the oracle by the parity tests): the same count and the same
phi bit for bit, the same stop values and history -- bit for bit where the
residual is summed by the separate residual pass, to rounding where the first
sweep of a cycle sums it (the FP64 pair sweep at level 0: a different
summation order of the same per-point residuals).
"""
import numpy as np
import pytest

from gpu_helpers import rand_problem
from jsolver_amd import _lib, _solver
from jsolver_amd.grid_1d import grid_1D
from jsolver_amd.solver_2d import solve_2d
from jsolver_amd.solver_3d import solve_3d
from oracle import gmg

pytestmark = pytest.mark.gpu

KWS = [
    pytest.param(dict(), id="v"),
    pytest.param(dict(gamma=2, fcycle=True), id="fw"),
    pytest.param(dict(skip_pre=True), id="vskip"),
    pytest.param(dict(nu1=2, nu2=1), id="v21"),
    pytest.param(dict(max_iters_outer=3), id="limit3"),
]
# every solve here converges in well under LIMIT cycles: a wrong stop value
# shows as a count mismatch, not as a solve running to the int32 limit
LIMIT = 400
STOPS = [pytest.param(dict(), id="distance"),
         pytest.param(dict(stop="rel_residual", rtol=1e-10), id="relres")]


def _plan(dim, ms):
    return _solver.get_plan(dim, list(ms), [grid_1D(m).idel for m in ms], "fp64")


def _tune(dim, ms, key, value):
    lib = _lib.load()
    _lib.check(lib.jsv_plan_set_tuning(_plan(dim, ms).h, key, value))


def _both(dim, ms, kw, fused_min_m=None, rows=0):
    """(device stop, host stop) results of the same solve; fused_min_m / rows:
    level 0 smoothed by the pair sweep with that many rows per tile."""
    phi, rhs, lam, D, _ = rand_problem(dim, ms, seed=31)
    g = [grid_1D(m) for m in ms]
    solve = solve_2d if dim == 2 else solve_3d
    kw = {"max_iters_outer": LIMIT, **kw}
    res = []
    try:
        if fused_min_m is not None:
            _tune(dim, ms, _lib.JSV_TUNE_FUSED_MIN_M, fused_min_m)
            _tune(dim, ms, _lib.JSV_TUNE_SWEEP_ROWS, rows)
        for on in (1, 0):
            _tune(dim, ms, _lib.JSV_TUNE_DEVICE_STOP, on)
            out, n, hist, info = solve(*g, phi, rhs, lam, *D, return_history=True, return_info=True, **kw)
            res.append((out.cpu().numpy(), n, hist, info))
    finally:
        _tune(dim, ms, _lib.JSV_TUNE_DEVICE_STOP, 1)
        if fused_min_m is not None:
            _tune(dim, ms, _lib.JSV_TUNE_FUSED_MIN_M, 256)
            _tune(dim, ms, _lib.JSV_TUNE_SWEEP_ROWS, 0)
    return res


def _check(res, exact_values):
    (od, nd, hd, idd), (oh, nh, hh, ih) = res
    assert nd == nh
    np.testing.assert_array_equal(od, oh)
    assert idd["nonfinite"] == ih["nonfinite"]
    assert len(hd) == len(hh) == nd
    if exact_values:
        np.testing.assert_array_equal(hd, hh)
        assert idd["rel_residual"] == ih["rel_residual"] and idd["distance"] == ih["distance"]
    else:
        np.testing.assert_allclose(hd, hh, rtol=1e-11, atol=0)
        assert idd["rel_residual"] == pytest.approx(ih["rel_residual"], rel=1e-11)


@pytest.mark.parametrize("stop", STOPS)
@pytest.mark.parametrize("kw", KWS)
def test_device_stop_3d_pair_sweep(kw, stop):
    """64^3 with the fused pair sweep at level 0 (5-row tiles): the relative
    residual is summed inside each cycle's first sweep (out of place, its
    input intact); v21 leaves level 0's phi in its second buffer each cycle
    (eager cycles); vskip has no first sweep (separate residual pass)."""
    res = _both(3, (64, 64, 64), {**kw, **stop}, fused_min_m=64, rows=5)
    relres = stop.get("stop") == "rel_residual"
    _check(res, exact_values=not relres or kw.get("skip_pre", False))
    if "max_iters_outer" in kw:
        assert res[0][1] == 3


@pytest.mark.parametrize("rows", [2, 4, 6])
def test_device_stop_3d_pair_sweep_rows(rows):
    """The residual-summing sweep with 2- and 4-row tiles (6 rows: no such
    variant; the residual pass runs after the out-of-place sweep instead)."""
    res = _both(3, (64, 32, 64), dict(stop="rel_residual", rtol=1e-10, gamma=2), fused_min_m=32, rows=rows)
    _check(res, exact_values=rows == 6)


@pytest.mark.parametrize("stop", STOPS)
@pytest.mark.parametrize("kw", KWS)
@pytest.mark.parametrize("dim,ms", [(3, (32, 16, 32)), (2, (128, 64))], ids=["3d", "2d"])
def test_device_stop_residual_pass(dim, ms, kw, stop):
    """Levels smoothed in place: the test's residual is a separate pass before
    the cycle, summed as the host's check sums it, so every value is bit-equal."""
    _check(_both(dim, ms, {**kw, **stop}), exact_values=True)


def test_device_stop_matches_oracle_count_and_relres():
    """Distance: the oracle's count and phi through the device test; relative
    residual: the value reported is the oracle's residual norm of the phi
    returned, and one cycle fewer does not reach the tolerance."""
    ms = (64, 64, 64)
    phi, rhs, lam, D, _ = rand_problem(3, ms, seed=5)
    g = [grid_1D(m) for m in ms]
    og = [gmg.OracleGrid(m) for m in ms]
    zero = np.zeros(rhs.shape)
    try:
        _tune(3, ms, _lib.JSV_TUNE_FUSED_MIN_M, 64)
        out, n = solve_3d(*g, zero, rhs, lam, *D, epsilon=1e-9, max_iters_outer=LIMIT)
        want, wn = gmg.solve_3d(*og, zero, rhs, lam, *D, epsilon=1e-9)
        assert n == wn
        np.testing.assert_array_equal(out.cpu().numpy(), want)
        out, n, info = solve_3d(*g, zero, rhs, lam, *D, stop="rel_residual", rtol=1e-9, return_info=True,
                                max_iters_outer=LIMIT)
        assert n < LIMIT
        op = gmg.build_operator(3, ms, tuple(x.idel for x in og), lam, list(D))
        rel = gmg.residual_norm(op, out.cpu().numpy(), rhs) / gmg.rhs_norm(rhs)
        assert rel <= 1e-9
        assert info["rel_residual"] == pytest.approx(rel, rel=1e-9)
        prev, _ = solve_3d(*g, zero, rhs, lam, *D, fixed=True, max_iters_outer=n - 1)
        assert gmg.residual_norm(op, prev.cpu().numpy(), rhs) / gmg.rhs_norm(rhs) > 1e-9
    finally:
        _tune(3, ms, _lib.JSV_TUNE_FUSED_MIN_M, 256)


def test_device_stop_nan_and_repeat():
    """A NaN in rhs stops after one cycle through the device test (NaN is not
    above the tolerance); the plan then solves a finite problem again with
    the flags reset."""
    ms = (64, 64, 64)
    phi, rhs, lam, D, _ = rand_problem(3, ms, seed=8)
    g = [grid_1D(m) for m in ms]
    bad = rhs.copy()
    bad[10, 20, 30] = np.nan
    try:
        _tune(3, ms, _lib.JSV_TUNE_FUSED_MIN_M, 64)
        for stop in (dict(max_iters_outer=LIMIT), dict(stop="rel_residual", rtol=1e-10, max_iters_outer=LIMIT)):
            out, n, info = solve_3d(*g, np.zeros(rhs.shape), bad, lam, *D, return_info=True, **stop)
            assert n == 1 and info["nonfinite"]
            ok1, n1 = solve_3d(*g, np.zeros(rhs.shape), rhs, lam, *D, **stop)
            _tune(3, ms, _lib.JSV_TUNE_DEVICE_STOP, 0)
            ok0, n0 = solve_3d(*g, np.zeros(rhs.shape), rhs, lam, *D, **stop)
            _tune(3, ms, _lib.JSV_TUNE_DEVICE_STOP, 1)
            assert LIMIT > n1 == n0 > 1
            np.testing.assert_array_equal(ok1.cpu().numpy(), ok0.cpu().numpy())
    finally:
        _tune(3, ms, _lib.JSV_TUNE_DEVICE_STOP, 1)
        _tune(3, ms, _lib.JSV_TUNE_FUSED_MIN_M, 256)


@pytest.mark.parametrize("kw", [dict(skip_pre=True), dict(nu1=2, nu2=1), dict(nu1=1, nu2=2),
                                dict(skip_pre=True, gamma=2, fcycle=True)],
                         ids=["vskip", "v21", "v12", "fwskip"])
def test_odd_sweeps_on_fused_level_match_oracle(kw):
    """Cycles that leave an out-of-place (fused) level's phi in its second
    buffer are replayed eagerly; the capture attempt must leave both buffer
    lists as they were.  Fixed and dynamic solves give the oracle's bits."""
    ms = (64, 64, 64)
    phi, rhs, lam, D, _ = rand_problem(3, ms, seed=17)
    g = [grid_1D(m) for m in ms]
    og = [gmg.OracleGrid(m) for m in ms]
    try:
        _tune(3, ms, _lib.JSV_TUNE_FUSED_MIN_M, 64)
        for rep in range(2):
            out, n = solve_3d(*g, phi, rhs, lam, *D, fixed=True, max_iters_outer=3, **kw)
            want, _ = gmg.solve_3d(*og, phi, rhs, lam, *D, fixed=True, max_iters_outer=3, **kw)
            np.testing.assert_array_equal(out.cpu().numpy(), want, err_msg=f"fixed, call {rep}")
        out, n = solve_3d(*g, phi, rhs, lam, *D, max_iters_outer=60, epsilon=1e-8, **kw)
        want, wn = gmg.solve_3d(*og, phi, rhs, lam, *D, max_iters_outer=60, epsilon=1e-8, **kw)
        assert n == wn
        np.testing.assert_array_equal(out.cpu().numpy(), want)
    finally:
        _tune(3, ms, _lib.JSV_TUNE_FUSED_MIN_M, 256)
