"""Bit-exact parity at the benchmark configurations' full sizes.

3D 256^3 (the headline problem): GPU FP64 V- and F-cycles against the
oracle's C restatement (bit-identical to the NumPy oracle, tests/
test_oracle_c.py), which runs a 256^3 cycle in a fraction of a second.
2D 1024^2 (the W-cycle configuration) and 512^2 (the adjoint configuration):
against the NumPy oracle, one or two cycles.
"""
import os

import numpy as np
import pytest
import torch

from jsolver_amd import fields
from jsolver_amd.solver_2d import solve_2d
from jsolver_amd.solver_3d import solve_3d
from oracle import gmg, gmg_c

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


@pytest.mark.parametrize("kw,ncyc", [(dict(), 3), (dict(fcycle=True), 2)], ids=["V3", "F2"])
def test_3d_256_cycles_bitexact(kw, ncyc):
    grids, rhs, lam, dxx, dyy, dzz = fields.anisotropic_3d(256)
    op = gmg.build_operator(3, (256,) * 3, tuple(g.idel for g in grids), lam, [dxx, dyy, dzz])
    cfg = gmg.Config(fixed=True, **kw)
    want = gmg_c.cycles3d(op, cfg, np.zeros(rhs.shape), rhs, ncyc, threads=THREADS)
    dev = torch.device("cuda", 0)
    t = [torch.from_numpy(v).to(dev) for v in (rhs, lam, dxx, dyy, dzz)]
    got, n = solve_3d(*grids, torch.zeros_like(t[0]), *t, fixed=True, max_iters_outer=ncyc, **kw)
    assert n == ncyc
    np.testing.assert_array_equal(got.cpu().numpy(), want)


@pytest.mark.parametrize("m,kw,ncyc,make", [
    (1024, dict(gamma=2), 1, lambda m: fields.anisotropic_2d(m)),
    (512, dict(), 2, lambda m: fields.anisotropic_2d(m, config_id=4, ratio=10.0)),
], ids=["W1024", "V512"])
def test_2d_full_size_cycles_bitexact(m, kw, ncyc, make):
    (gx, gy), rhs, lam, dxx, dyy = make(m)
    op = gmg.build_operator(2, (m, m), (gx.idel, gy.idel), lam, [dxx, dyy])
    cfg = gmg.Config(fixed=True, **kw)
    want = gmg.zero_rim(np.zeros(rhs.shape))
    for _ in range(ncyc):
        want = gmg.cycle(op, cfg, want, rhs, False, 0, False)
    dev = torch.device("cuda", 0)
    t = [torch.from_numpy(v).to(dev) for v in (rhs, lam, dxx, dyy)]
    got, n = solve_2d(gx, gy, torch.zeros_like(t[0]), *t, fixed=True, max_iters_outer=ncyc, **kw)
    assert n == ncyc
    np.testing.assert_array_equal(got.cpu().numpy(), want)
