"""The C restatement of the oracle's 3D cycle (oracle/gmg3.c, used as the
timed CPU baseline) against the NumPy oracle: bit for bit, for every cycle
type and thread count."""
import numpy as np
import pytest

from oracle import gmg, gmg_c

CASES = [((8, 8, 8), {}), ((16, 8, 32), {}), ((16, 16, 16), dict(fcycle=True)), ((16, 16, 16), dict(gamma=2)),
         ((16, 16, 16), dict(nu1=2, nu2=0)), ((16, 16, 16), dict(nu1=2, nu2=3)), ((16, 16, 16), dict(skip_pre=True))]


@pytest.mark.parametrize("ms,kw", CASES, ids=[f"{m}-{k}" for m, k in CASES])
@pytest.mark.parametrize("spatial", [True, False])
def test_c_cycles_equal_numpy_oracle(ms, kw, spatial):
    rng = np.random.default_rng(11)
    shape = tuple(m + 3 for m in reversed(ms))
    D = [np.exp(0.5 * rng.standard_normal(shape)) for _ in range(3)] if spatial else [1.3, 0.7, 2.1]
    lam = rng.uniform(0, 2, shape) if spatial else 0.5
    rhs = rng.standard_normal(shape)
    phi = rng.standard_normal(shape)
    op = gmg.build_operator(3, ms, [float(m) for m in ms], lam, D)
    cfg = gmg.Config(fixed=True, **kw)
    want = gmg.zero_rim(phi)
    for _ in range(3):
        want = gmg.cycle(op, cfg, want, rhs, cfg.fcycle, 0, cfg.skip_pre)
    for threads in (1, 3):
        got = gmg_c.cycles3d(op, cfg, phi, rhs, 3, threads=threads)
        np.testing.assert_array_equal(got, want)


def test_c_oracle_converges_like_numpy_solve():
    """13 V-cycles of a constant-coefficient problem (scalar D and lambda)
    from zero: the C cycles equal the NumPy oracle's."""
    ms = (16, 16, 16)
    g = gmg.OracleGrid(16)
    x = g.centers[None, None, :]
    y = g.centers[None, :, None]
    z = g.centers[:, None, None]
    rhs = np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z) * np.ones((19, 19, 19))
    op = gmg.build_operator(3, ms, [16.0] * 3, 0.0, [1.0, 1.0, 1.0])
    cfg = gmg.Config(fixed=True)
    want = np.zeros(rhs.shape)
    for _ in range(13):
        want = gmg.cycle(op, cfg, want, rhs)
    np.testing.assert_array_equal(gmg_c.cycles3d(op, cfg, np.zeros(rhs.shape), rhs, 13, threads=2), want)
