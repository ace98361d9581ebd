"""Host-side behaviour of the public API, without a GPU.

The reference's own unit and error tests restated for the drop-in API:
grid_1D values (tests/test_grid_1d.py:9-52), the ValueError messages of bad
grids (tests/test_solver_3d.py:85-133, tests/test_solver_2d.py:77-99), plus
the level layout against the oracle's restatement for every power-of-two
shape, and the loud failure of the compute path when no GPU is visible.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest
import torch
from hypothesis import given, settings
from hypothesis import strategies as st

from jsolver_amd import _lib, _solver
from jsolver_amd.grid_1d import grid_1D
from jsolver_amd.solver_2d import solve_2d, solve_2d_simple
from jsolver_amd.solver_3d import solve_3d_simple
from oracle import gmg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ---- grid_1D (tests/test_grid_1d.py:5-52) -----------------------------------
def test_grid_values():
    g = grid_1D(mx=128, xb=0., xe=1., rim=1, grid_type=0, centered=True)
    assert (g.mx, g.nx, g.centered, g.rim, g.xb, g.xe, g.grid_type) == (128, 129, True, 1, 0.0, 1.0, 0)
    assert g.length == 1.0
    assert g.del_ == pytest.approx(0.0078125)
    assert g.idel == pytest.approx(128.0)
    assert g.centers[0] == pytest.approx(-0.0078125)
    assert g.centers[-1] == pytest.approx(1.0078125)
    assert g.edges[0] == pytest.approx(-0.01171875)
    assert g.edges[-1] == pytest.approx(1.01171875)
    assert g.widths[0] == pytest.approx(0.0078125)
    assert g.widths[-1] == pytest.approx(0.0078125)
    assert g.centers.size == g.widths.size == g.edges.size - 1


def test_grid_sinusoidal_values():
    g = grid_1D(mx=128, xb=0., xe=1., rim=1, grid_type=1)
    assert g.centers[0] == pytest.approx(-0.00428662)
    assert g.centers[-1] == pytest.approx(1.0121)
    assert g.edges[0] == pytest.approx(-0.00857324)
    assert g.edges[-1] == pytest.approx(1.01639, rel=1e-5)
    assert g.widths[0] == pytest.approx(0.00857324)
    assert g.widths[-1] == pytest.approx(0.00857324)


def test_grid_unknown_type_and_value_semantics():
    with pytest.raises(ValueError):
        grid_1D(mx=128, grid_type=42)
    a, b = grid_1D(64), grid_1D(64)
    assert a == b and hash(a) == hash(b) and a is not b
    assert grid_1D(64) != grid_1D(64, xe=2.0)
    assert len({grid_1D(32), grid_1D(32), grid_1D(64)}) == 2


# ---- errors before any compute (tests/test_solver_3d.py:85-133) -----------
Z = np.zeros(131)


@pytest.mark.parametrize("mx,my,mz,msg", [
    (100, 128, 128, "Resolution of x-grid is not a power of 2"),
    (128, 100, 32, "Resolution of y-grid is not a power of 2"),
    (128, 32, 55, "Resolution of z-grid is not a power of 2"),
    (32, 4, 4, "Resolution of x-grid not efficient for multigrid configuration"),
    (4, 32, 4, "Resolution of y-grid not efficient for multigrid configuration"),
    (4, 4, 32, "Resolution of z-grid not efficient for multigrid configuration"),
])
def test_3d_grid_errors(mx, my, mz, msg):
    with pytest.raises(ValueError) as e:
        solve_3d_simple(grid_1D(mx=mx), grid_1D(mx=my), grid_1D(mx=mz), Z, Z, Z, 1.0, 1.0, 1.0)
    assert msg in str(e.value)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_3d_nonlinear_grid_error(axis):
    gs = [grid_1D(mx=32, grid_type=1 if a == axis else 0) for a in range(3)]
    with pytest.raises(ValueError) as e:
        solve_3d_simple(*gs, Z, Z, Z, 1.0, 1.0, 1.0)
    assert "Only linear grids are supported" in str(e.value)


@pytest.mark.parametrize("mx,my", [(128, 100), (100, 128), (128, 4), (4, 128)])
def test_2d_grid_errors(mx, my):
    with pytest.raises(ValueError):
        solve_2d_simple(grid_1D(mx=mx), grid_1D(mx=my), Z, Z, Z, 1.0, 1.0)


@pytest.mark.parametrize("axis", [0, 1])
def test_2d_nonlinear_grid_error(axis):
    gs = [grid_1D(mx=128, grid_type=1 if a == axis else 0) for a in range(2)]
    with pytest.raises(ValueError):
        solve_2d_simple(*gs, Z, Z, Z, 1.0, 1.0)


def test_shape_and_argument_errors():
    g = grid_1D(16)
    ok = np.zeros((19, 19))
    with pytest.raises(ValueError, match="Shape mismatch"):
        solve_2d(g, g, np.zeros((18, 19)), ok, 0.0, 1.0, 1.0)
    with pytest.raises(ValueError, match="precision"):
        solve_2d(g, g, ok, ok, 0.0, 1.0, 1.0, precision="fp16")
    with pytest.raises(ValueError, match="stop"):
        solve_2d(g, g, ok, ok, 0.0, 1.0, 1.0, stop="never")


# ---- level layout against the oracle, every power-of-two shape ------------
@settings(max_examples=200, deadline=None)
@given(st.lists(st.integers(min_value=1, max_value=12), min_size=2, max_size=3))
def test_level_layout_matches_oracle(log2s):
    ms = [1 << k for k in log2s]
    names = ["x", "y", "z"][: len(ms)]
    try:
        want = gmg.level_layout(ms, names)
    except ValueError as e:
        with pytest.raises(ValueError) as got:
            _solver.level_layout(ms)
        assert str(got.value) == str(e)
        return
    levels, per_axis = _solver.level_layout(ms)
    assert levels == want[0]
    # oracle: per level, array-axis order (slowest first); here: per axis (x, y, z)
    for lvl in range(levels):
        assert tuple(int(per_axis[a][lvl]) for a in reversed(range(len(ms)))) == want[1][lvl]


# ---- the compute path has no CPU fallback ---------------------------------
@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU failure")
def test_compute_fails_loudly_without_gpu():
    g = grid_1D(16)
    a = np.zeros((19, 19))
    with pytest.raises(_lib.JsvError, match="GPU"):
        solve_2d(g, g, a, a, 0.0, 1.0, 1.0)


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU failure")
def test_forward_mode_fails_loudly_without_gpu():
    """A dual tensor routes the solve to the tangent path, which also needs the GPU."""
    import torch.autograd.forward_ad as fwAD
    g = grid_1D(16)
    a = torch.zeros((19, 19), dtype=torch.float64)
    with fwAD.dual_level():
        r = fwAD.make_dual(a, torch.ones_like(a))
        with pytest.raises(_lib.JsvError, match="GPU"):
            solve_2d(g, g, a, r, 0.0, 1.0, 1.0)


def test_tangent_detection():
    """Forward-mode tangents are recognised on dual tensors and inside
    torch.func.jvp, and only there."""
    import torch.autograd.forward_ad as fwAD
    from jsolver_amd.tangent import has_tangent
    assert not has_tangent(np.zeros(3)) and not has_tangent(1.0) and not has_tangent(torch.zeros(3))
    with fwAD.dual_level():
        assert has_tangent(fwAD.make_dual(torch.zeros(3), torch.ones(3)))
        assert not has_tangent(torch.zeros(3))
    seen = []
    torch.func.jvp(lambda x: (seen.append(has_tangent(x)), 2 * x)[1], (torch.zeros(2),), (torch.ones(2),))
    assert seen == [True]


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]
