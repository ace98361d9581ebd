"""GPU parity: every HIP operation of the multigrid path against the CPU oracle.

FP64 kernels evaluate the reference's formulas term by term with no FMA
contraction, so per-operation results are compared BIT-EXACTLY with the
oracle (tolerance 0).  Whole solves are compared the same way, plus the
reference's own goldens (exact iteration counts, RMSE within 1e-9 rel).
"""
import json
import os

import numpy as np
import pytest
import torch

from gpu_helpers import DevPlan, cfg, oracle_operator, rand_problem
from oracle import gmg
from problems import problem_2d, problem_3d

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
G = json.load(open(os.path.join(HERE, "golden", "reference_goldens.json")))

CASES = [
    # id, dim, ms (x,y[,z]), spatial, lam_array, dxy
    ("2d_spatial_16", 2, (16, 16), True, True, False),
    ("2d_spatial_16x32", 2, (16, 32), True, True, False),
    ("2d_spatial_32x8", 2, (32, 8), True, False, False),
    ("2d_scalarD_lamarr", 2, (16, 16), False, True, False),
    ("2d_scalar", 2, (16, 16), False, False, False),
    ("2d_dxy_16", 2, (16, 16), True, True, True),
    ("2d_dxy_8x16", 2, (8, 16), True, True, True),
    ("3d_spatial_8", 3, (8, 8, 8), True, True, False),
    ("3d_spatial_8x16x8", 3, (8, 16, 8), True, True, False),
    ("3d_spatial_16x8x32", 3, (16, 8, 32), True, False, False),
    ("3d_scalar_lamarr", 3, (8, 8, 8), False, True, False),
]


def _setup(case, seed=1):
    _, dim, ms, spatial, lam_array, dxy = case
    phi, rhs, lam, D, Dxy = rand_problem(dim, ms, seed, spatial, lam_array, dxy)
    op = oracle_operator(dim, ms, lam, D, Dxy)
    dp = DevPlan(dim, ms, lam, D, Dxy)
    return dp, op, phi, rhs


def _level_arrays(op, level, seed):
    rng = np.random.default_rng(seed)
    shape = op.lv[level].shape
    return gmg.zero_rim(rng.standard_normal(shape)), rng.standard_normal(shape)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_smooth_bitexact(case):
    dp, op, _, _ = _setup(case)
    for level in range(op.levels - 1):
        phi, rhs = _level_arrays(op, level, 10 + level)
        dp.upload(level, 0, phi)
        dp.upload(level, 1, rhs)
        dp.smooth(level, 2)
        got = dp.download(level, 0)
        want = gmg.smooth(op, level, gmg.smooth(op, level, phi, rhs), rhs)
        np.testing.assert_array_equal(got, want, err_msg=f"level {level}")


@pytest.mark.parametrize("slab_bytes", [1, 3 * 8 * 6 * 35 * 19])
def test_smooth_3d_slab_schedule_bitexact(slab_bytes):
    """The z-slab launch schedule of the 3D smoother (red on slab k, black on
    slab k-1) gives the same bits as one red and one black launch."""
    from jsolver_amd import _lib
    case = ("3d_spatial_16x32x32", 3, (16, 32, 32), True, True, False)
    dp, op, _, _ = _setup(case)
    _lib.check(_lib.load().jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_SLAB_BYTES, slab_bytes))
    for level in range(2):
        phi, rhs = _level_arrays(op, level, 60 + level)
        dp.upload(level, 0, phi)
        dp.upload(level, 1, rhs)
        dp.smooth(level, 2)
        got = dp.download(level, 0)
        want = gmg.smooth(op, level, gmg.smooth(op, level, phi, rhs), rhs)
        np.testing.assert_array_equal(got, want, err_msg=f"level {level}")


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_restrict_bitexact(case):
    dp, op, _, _ = _setup(case)
    for level in range(op.levels - 2):
        phi, rhs = _level_arrays(op, level, 20 + level)
        dp.upload(level, 0, phi)
        dp.upload(level, 1, rhs)
        dp.upload(level + 1, 1, np.zeros(op.lv[level + 1].shape))
        dp.restrict(level)
        got = dp.download(level + 1, 1)
        want = gmg.restrict(op, level, phi, rhs)
        np.testing.assert_array_equal(got, want, err_msg=f"level {level}")


FUSED_VARIANTS = [(1, 0, 0), (2, 0, 1), (2, 2, 1), (2, 4, 1), (2, 5, 1), (2, 6, 1), (2, 14, 1),
                  (2, 0, 2), (2, 2, 2), (2, 5, 2), (2, 14, 2)]


@pytest.mark.parametrize("variant,rows,pairs", FUSED_VARIANTS,
                         ids=[f"v{v}_rows{r}_pairs{p}" for v, r, p in FUSED_VARIANTS])
@pytest.mark.parametrize("zchunk", [1, 2, 3, 5, 0])
@pytest.mark.parametrize("ms", [(16, 32, 32), (32, 16, 8), (8, 8, 8)])
def test_fused_sweep_bitexact(ms, zchunk, variant, rows, pairs):
    """The fused out-of-place z-marching sweep (forced on every level, several
    y-tiles and z-chunk sizes, both kernels) gives the oracle's bits for
    sweeps and cycles."""
    from jsolver_amd import _lib
    case = ("3d_fused", 3, ms, True, True, False)
    dp, op, phi0, rhs0 = _setup(case, seed=13)
    lib = _lib.load()
    _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_FUSED_MIN_M, 2))
    _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_ZCHUNK, zchunk))
    _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_SWEEP_VARIANT, variant))
    _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_SWEEP_ROWS, rows))
    _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_SWEEP_PAIRS, pairs))
    for level in range(op.levels - 1):
        phi, rhs = _level_arrays(op, level, 80 + level)
        dp.upload(level, 0, phi)
        dp.upload(level, 1, rhs)
        dp.smooth(level, 3)
        want = phi
        for _ in range(3):
            want = gmg.smooth(op, level, want, rhs)
        np.testing.assert_array_equal(dp.download(level, 0), want, err_msg=f"level {level}")
    for kw in (dict(), dict(gamma=2, fcycle=True), dict(nu1=2, nu2=0)):
        dp.upload(0, 0, phi0)
        dp.upload(0, 1, rhs0)
        dp.cycle(cfg(**kw))
        c = gmg.Config(fixed=True, max_iters_outer=1, **kw)
        np.testing.assert_array_equal(dp.download(0, 0), gmg.cycle(op, c, gmg.zero_rim(phi0), rhs0, c.fcycle))


@pytest.mark.parametrize("zigzag", [0, 1, 2])
@pytest.mark.parametrize("zchunk", [1, 3, 0])
@pytest.mark.parametrize("pairs", [1, 2])
@pytest.mark.parametrize("ms", [(16, 32, 32), (32, 16, 64), (8, 8, 8)])
def test_pair_sweep_zigzag_bitexact(ms, pairs, zchunk, zigzag):
    """Pair sweeps that march downward (every other plain sweep of a level
    with zig-zag on; zig-zag 2: every other z-chunk of each sweep) give the
    oracle's bits; so do 1, 2 and 3 consecutive sweeps, which start in either
    direction, and the cycles around them."""
    from jsolver_amd import _lib
    case = ("3d_zigzag", 3, ms, True, True, False)
    dp, op, phi0, rhs0 = _setup(case, seed=17)
    lib = _lib.load()
    for key, v in ((_lib.JSV_TUNE_FUSED_MIN_M, 2), (_lib.JSV_TUNE_ZCHUNK, zchunk),
                   (_lib.JSV_TUNE_SWEEP_PAIRS, pairs), (_lib.JSV_TUNE_ZIGZAG, zigzag)):
        _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, key, v))
    phi, rhs = _level_arrays(op, 0, 70)
    want = phi
    for sweeps in (1, 2, 3):
        want = gmg.smooth(op, 0, want, rhs)
        dp.upload(0, 0, phi)
        dp.upload(0, 1, rhs)
        dp.smooth(0, sweeps)
        np.testing.assert_array_equal(dp.download(0, 0), want, err_msg=f"{sweeps} sweeps")
    for kw in (dict(), dict(nu1=2, nu2=2), dict(gamma=2, fcycle=True)):
        dp.upload(0, 0, phi0)
        dp.upload(0, 1, rhs0)
        dp.cycle(cfg(**kw))
        c = gmg.Config(fixed=True, max_iters_outer=1, **kw)
        np.testing.assert_array_equal(dp.download(0, 0), gmg.cycle(op, c, gmg.zero_rim(phi0), rhs0, c.fcycle))


FR_SHAPES = [(0, 0), (4, 3), (5, 5), (6, 2), (14, 0), (14, 1)]


@pytest.mark.parametrize("zigzag", [1, 0, 2])
@pytest.mark.parametrize("fr", [1, 0])
@pytest.mark.parametrize("rows,zchunk", FR_SHAPES, ids=lambda v: str(v))
@pytest.mark.parametrize("pairs", [1, 2])
@pytest.mark.parametrize("ms", [(16, 32, 32), (32, 16, 64), (64, 32, 16), (128, 32, 32)])
def test_fused_first_half_residuals_bitexact(ms, pairs, rows, zchunk, fr, zigzag):
    """The last pre-smoothing pair sweep storing the restriction's first-half
    residuals at its tiles' inner rows and planes, k_resid_edges3 the tile
    edge rows and chunk edge planes (JSV_TUNE_FUSE_RESID 1; 0: one residual
    pass): whole cycles give the oracle's bits.  Tile rows 4 .. 14, z-chunks
    of 1 .. all planes; 128 wide: rows of whole waves; marching downward
    (zig-zag after an upward sweep), upward, and in chunks of alternating
    direction (zig-zag 2)."""
    from jsolver_amd import _lib
    case = ("3d_fr", 3, ms, True, True, False)
    dp, op, phi0, rhs0 = _setup(case, seed=37)
    lib = _lib.load()
    for key, v in ((_lib.JSV_TUNE_FUSED_MIN_M, 2), (_lib.JSV_TUNE_TWO_PASS_MIN, 1),
                   (_lib.JSV_TUNE_SWEEP_PAIRS, pairs), (_lib.JSV_TUNE_SWEEP_ROWS, rows),
                   (_lib.JSV_TUNE_ZCHUNK, zchunk), (_lib.JSV_TUNE_FUSE_RESID, fr), (_lib.JSV_TUNE_ZIGZAG, zigzag)):
        _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, key, v))
    for kw in (dict(), dict(nu1=2, nu2=1), dict(gamma=2, fcycle=True), dict(nu2=0)):
        c = gmg.Config(fixed=True, max_iters_outer=1, **kw)
        want = gmg.cycle(op, c, gmg.zero_rim(phi0), rhs0, c.fcycle)
        for rep in range(2):
            dp.upload(0, 0, phi0)
            dp.upload(0, 1, rhs0)
            dp.cycle(cfg(**kw))
            np.testing.assert_array_equal(dp.download(0, 0), want, err_msg=f"{kw} call {rep}")


@pytest.mark.parametrize("pairs", [1, 2])
@pytest.mark.parametrize("fuse", [2, 0])
@pytest.mark.parametrize("rows", [0, 2, 5])
@pytest.mark.parametrize("ms", [(16, 32, 32), (32, 16, 16), (8, 8, 8), (32, 32, 64)])
def test_fused_prolongation_cycles_bitexact(ms, rows, fuse, pairs):
    """Post-smoothing pair sweeps that apply the prolongation themselves (a
    coarse-plane ring in LDS) give the oracle's bits for whole cycles: V, W,
    F, nu2 = 2 and 3, skipped pre-smoothing."""
    from jsolver_amd import _lib
    case = ("3d_fused_pro", 3, ms, True, True, False)
    dp, op, phi0, rhs0 = _setup(case, seed=31)
    lib = _lib.load()
    _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_FUSED_MIN_M, 2))
    _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_SWEEP_VARIANT, 2))
    _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_SWEEP_ROWS, rows))
    _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_FUSE_PROLONG, fuse))
    _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_SWEEP_PAIRS, pairs))
    for kw in (dict(), dict(gamma=2), dict(fcycle=True), dict(nu1=1, nu2=2), dict(gamma=2, nu1=2, nu2=3),
               dict(skip_pre=True)):
        dp.upload(0, 0, phi0)
        dp.upload(0, 1, rhs0)
        dp.cycle(cfg(**kw), skip_pre=kw.get("skip_pre", False))
        c = gmg.Config(fixed=True, max_iters_outer=1, **kw)
        want = gmg.cycle(op, c, gmg.zero_rim(phi0), rhs0, c.fcycle, 0, c.skip_pre)
        np.testing.assert_array_equal(dp.download(0, 0), want, err_msg=str(kw))


@pytest.mark.parametrize("ms", [(16, 32, 32), (32, 16, 16)])
def test_fused_prolongation_fp32_matches_separate(ms):
    """FP32 cycles (the mixed solver's inner cycles): the fused prolongation
    (default for FP32) and the two-pairs-per-thread kernel give the same bits
    as the separate prolongation pass and one pair per thread."""
    from jsolver_amd import _lib
    phi, rhs, lam, D, _ = rand_problem(3, ms, seed=37)
    outs = []
    for fuse, pairs in ((1, 0), (0, 1), (1, 2), (0, 2)):
        dp = DevPlan(3, ms, lam, D, precision="mixed")
        lib = _lib.load()
        _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_FUSED_MIN_M, 2))
        _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_FUSE_PROLONG, fuse))
        _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_SWEEP_PAIRS, pairs))
        dp.upload(0, 0, gmg.zero_rim(phi).astype(np.float32), prec=_lib.JSV_MIXED)
        dp.upload(0, 1, rhs.astype(np.float32), prec=_lib.JSV_MIXED)
        for kw in (dict(), dict(gamma=2, nu2=2)):
            dp.cycle(cfg(**kw), prec=_lib.JSV_MIXED)
        outs.append(dp.download(0, 0, prec=_lib.JSV_MIXED))
    for o in outs[1:]:
        np.testing.assert_array_equal(outs[0], o)
    assert np.abs(outs[0]).max() > 0


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_restrict_two_pass_bitexact(case):
    """The two-pass restriction of large levels (residuals of the first RBGS
    half stored, then weighted) gives the oracle's bits; forced on every level."""
    from jsolver_amd import _lib
    dp, op, _, _ = _setup(case)
    _lib.check(_lib.load().jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_TWO_PASS_MIN, 1))
    for level in range(op.levels - 2):
        phi, rhs = _level_arrays(op, level, 20 + level)
        dp.upload(level, 0, phi)
        dp.upload(level, 1, rhs)
        dp.upload(level + 1, 1, np.zeros(op.lv[level + 1].shape))
        dp.restrict(level)
        got = dp.download(level + 1, 1)
        want = gmg.restrict(op, level, phi, rhs)
        np.testing.assert_array_equal(got, want, err_msg=f"level {level}")
    # and whole cycles with it
    phi, rhs = _level_arrays(op, 0, 70)
    dp.upload(0, 0, phi)
    dp.upload(0, 1, rhs)
    dp.cycle(cfg(gamma=2))
    c = gmg.Config(fixed=True, max_iters_outer=1, gamma=2)
    np.testing.assert_array_equal(dp.download(0, 0), gmg.cycle(op, c, phi, rhs))


@pytest.mark.parametrize("nu2", [1, 0])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_prolong_bitexact(case, nu2):
    dp, op, _, _ = _setup(case)
    for level in range(op.levels - 2):
        phi, _ = _level_arrays(op, level, 30 + level)
        pc, _ = _level_arrays(op, level + 1, 40 + level)
        dp.upload(level, 0, phi)
        dp.upload(level + 1, 0, pc)
        dp.prolong(level, nu2)
        got = dp.download(level, 0)
        want = phi + gmg.prolong(op, level, pc, nu2)
        np.testing.assert_array_equal(got, want, err_msg=f"level {level}")


CYCLES = [
    ("V", dict()),
    ("W", dict(gamma=2)),
    ("F", dict(fcycle=True)),
    ("V_skip", dict(skip_pre=True)),
    ("V_nu2_0", dict(nu1=2, nu2=0)),
    ("W_nu1_2_nu2_3", dict(gamma=2, nu1=2, nu2=3)),
]
# the coarse program's fused ops change with nu1 and nu2 (no first half in the
# restriction when nu1 = 0 above the coarsest level)
FUSED_PROG_CYCLES = [CYCLES[0], CYCLES[1], CYCLES[2], CYCLES[4], CYCLES[5],
                     ("W_nu1_0_nu2_2", dict(gamma=2, nu1=0, nu2=2))]


@pytest.mark.parametrize("cyc", CYCLES, ids=[c[0] for c in CYCLES])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_cycle_bitexact(case, cyc):
    dp, op, phi, rhs = _setup(case, seed=5)
    kw = cyc[1]
    c = gmg.Config(fixed=True, max_iters_outer=1, **kw)
    phi_raw = phi.copy()                      # non-zero rim: the cycle must zero it
    dp.upload(0, 0, phi_raw)
    dp.upload(0, 1, rhs)
    dp.cycle(cfg(**kw), skip_pre=kw.get("skip_pre", False))
    got = dp.download(0, 0)
    want = gmg.cycle(op, c, gmg.zero_rim(phi_raw), rhs, c.fcycle, 0, c.skip_pre)
    np.testing.assert_array_equal(got, want)


COARSE_CASES = [c for c in CASES if c[0] in ("2d_spatial_16x32", "2d_dxy_16", "3d_spatial_16x8x32",
                                               "3d_scalar_lamarr")] + [
    ("2d_spatial_128x64", 2, (128, 64), True, True, False),
    ("2d_dxy_64", 2, (64, 64), True, True, True),
    ("3d_spatial_32", 3, (32, 32, 32), True, True, False),
]


@pytest.mark.parametrize("lds", [1, 0])
@pytest.mark.parametrize("coarse_m", [0, 4, 1 << 20])
@pytest.mark.parametrize("cyc", [CYCLES[0], CYCLES[1], CYCLES[2], CYCLES[4]], ids=lambda c: c[0])
@pytest.mark.parametrize("case", COARSE_CASES, ids=[c[0] for c in COARSE_CASES])
def test_cycle_coarse_program_bitexact(case, cyc, coarse_m, lds):
    """The single-workgroup coarse program (off, from the 4-point levels, and
    from level 1 down; its coarsest levels staged in LDS or not) gives the
    oracle's bits for whole cycles."""
    from jsolver_amd import _lib
    dp, op, phi, rhs = _setup(case, seed=17)
    _lib.check(_lib.load().jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_COARSE_M, coarse_m))
    _lib.check(_lib.load().jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_COARSE_LDS, lds))
    kw = cyc[1]
    c = gmg.Config(fixed=True, max_iters_outer=1, **kw)
    for rep in range(2):                      # second call replays the cached program
        dp.upload(0, 0, phi)
        dp.upload(0, 1, rhs)
        dp.cycle(cfg(**kw))
        want = gmg.cycle(op, c, gmg.zero_rim(phi), rhs, c.fcycle)
        np.testing.assert_array_equal(dp.download(0, 0), want, err_msg=f"call {rep}")


HALVES_CASES = [c for c in CASES if c[0] in ("2d_spatial_16x32", "2d_scalar", "3d_spatial_16x8x32",
                                               "3d_scalar_lamarr")] + [
    ("2d_spatial_128x64", 2, (128, 64), True, True, False),
    ("3d_spatial_32", 3, (32, 32, 32), True, True, False),
]


@pytest.mark.parametrize("fuse", [1, 0])
@pytest.mark.parametrize("coarse_m", [0, None])
@pytest.mark.parametrize("cyc", [CYCLES[0], CYCLES[1], CYCLES[2], CYCLES[4], CYCLES[5]], ids=lambda c: c[0])
@pytest.mark.parametrize("case", HALVES_CASES, ids=[c[0] for c in HALVES_CASES])
def test_cycle_fused_halves_bitexact(case, cyc, coarse_m, fuse):
    """Levels smoothed one launch per half: the restriction running the
    coarse level's first half (its phi is zero) and the prolongation applied
    inside the first post-smoothing half (JSV_TUNE_FUSE_HALVES on and off,
    every level launched or the default coarse program) give the oracle's
    bits for whole cycles."""
    from jsolver_amd import _lib
    lib = _lib.load()
    dp, op, phi, rhs = _setup(case, seed=23)
    _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_FUSE_HALVES, fuse))
    if coarse_m is not None:
        _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_COARSE_M, coarse_m))
    kw = cyc[1]
    c = gmg.Config(fixed=True, max_iters_outer=1, **kw)
    want = gmg.cycle(op, c, gmg.zero_rim(phi), rhs, c.fcycle)
    for rep in range(2):
        dp.upload(0, 0, phi)
        dp.upload(0, 1, rhs)
        dp.cycle(cfg(**kw))
        np.testing.assert_array_equal(dp.download(0, 0), want, err_msg=f"call {rep}")


@pytest.mark.parametrize("dim,ms", [(3, (64, 32, 32)), (2, (256, 128))])
def test_cycle_fused_halves_fp32_matches_separate(dim, ms):
    """FP32 cycles (the mixed solver's inner cycles): the fused halves give the
    bits of the separate launches."""
    from jsolver_amd import _lib
    lib = _lib.load()
    phi, rhs, lam, D, _ = rand_problem(dim, ms, seed=41)
    outs = []
    for fuse in (1, 0):
        dp = DevPlan(dim, ms, lam, D, precision="mixed")
        _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_FUSE_HALVES, fuse))
        _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_COARSE_M, 0))
        dp.upload(0, 0, gmg.zero_rim(phi).astype(np.float32), prec=_lib.JSV_MIXED)
        dp.upload(0, 1, rhs.astype(np.float32), prec=_lib.JSV_MIXED)
        for kw in (dict(), dict(gamma=2, nu2=2), dict(fcycle=True)):
            dp.cycle(cfg(**kw), prec=_lib.JSV_MIXED)
        outs.append(dp.download(0, 0, prec=_lib.JSV_MIXED))
    np.testing.assert_array_equal(outs[0], outs[1])
    assert np.abs(outs[0]).max() > 0


@pytest.mark.parametrize("fuse", [1, 0])
@pytest.mark.parametrize("solo,threads", [(0, 1024), (64, 256), (64, 512), (16, 1024)])
@pytest.mark.parametrize("cyc", [CYCLES[0], CYCLES[1], CYCLES[2], CYCLES[4], FUSED_PROG_CYCLES[5]],
                         ids=lambda c: c[0])
@pytest.mark.parametrize("case", COARSE_CASES, ids=[c[0] for c in COARSE_CASES])
def test_cycle_coarse_program_launch_shapes(case, cyc, solo, threads, fuse):
    """The coarse program from level 1 down with its small ops run by one wave
    (or never) and other workgroup sizes, its fused ops on or off: still the
    oracle's bits."""
    from jsolver_amd import _lib
    lib = _lib.load()
    dp, op, phi, rhs = _setup(case, seed=19)
    _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_FUSE_HALVES, fuse))
    _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_COARSE_M, 1 << 20))
    _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_COARSE_SOLO, solo))
    _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_COARSE_THREADS, threads))
    kw = cyc[1]
    c = gmg.Config(fixed=True, max_iters_outer=1, **kw)
    want = gmg.cycle(op, c, gmg.zero_rim(phi), rhs, c.fcycle)
    for rep in range(2):
        dp.upload(0, 0, phi)
        dp.upload(0, 1, rhs)
        dp.cycle(cfg(**kw))
        np.testing.assert_array_equal(dp.download(0, 0), want, err_msg=f"call {rep}")


REG_CASES = [c for c in CASES if c[0] in ("2d_spatial_16x32", "2d_spatial_32x8", "2d_scalar")] + [
    ("2d_spatial_128x64", 2, (128, 64), True, True, False),
    ("2d_spatial_64", 2, (64, 64), True, True, False),
    ("2d_spatial_128x32", 2, (128, 32), True, False, False)]


@pytest.mark.parametrize("fuse", [1, 0])
@pytest.mark.parametrize("reg", [2, 1, 0])
@pytest.mark.parametrize("cyc", FUSED_PROG_CYCLES, ids=lambda c: c[0])
@pytest.mark.parametrize("case", REG_CASES, ids=[c[0] for c in REG_CASES])
def test_cycle_coarse_program_register_runs(case, cyc, reg, fuse):
    """2D one-wave runs with their levels in registers (neighbours by lane
    shuffles; square and rectangular levels; reg 2: also runs from a level of
    65 to 256 points, a 2 x 2 block per lane, down through up to three levels
    below it) and in LDS, with the restriction
    running the coarse level's first half and the prolongation inside the first
    post-smoothing half (JSV_TUNE_FUSE_HALVES) or as separate ops: the
    oracle's bits."""
    from jsolver_amd import _lib
    lib = _lib.load()
    dp, op, phi, rhs = _setup(case, seed=29)
    _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_COARSE_M, 1 << 20))
    _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_COARSE_REG, reg))
    _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_FUSE_HALVES, fuse))
    kw = cyc[1]
    c = gmg.Config(fixed=True, max_iters_outer=1, **kw)
    want = gmg.cycle(op, c, gmg.zero_rim(phi), rhs, c.fcycle)
    for rep in range(2):
        dp.upload(0, 0, phi)
        dp.upload(0, 1, rhs)
        dp.cycle(cfg(**kw))
        np.testing.assert_array_equal(dp.download(0, 0), want, err_msg=f"call {rep}")


def test_coarse_tuning_limits():
    from jsolver_amd import _lib
    lib = _lib.load()
    dp, _, _, _ = _setup(CASES[0], seed=1)
    assert lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_COARSE_SOLO, 65) != 0
    assert lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_COARSE_THREADS, 128) != 0
    assert lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_COARSE_REG, 3) != 0


def test_zigzag_tuning_limit():
    """JSV_TUNE_ZIGZAG takes 0, 1 or 2; a larger value is an error, and the
    plan keeps working with its previous setting."""
    from jsolver_amd import _lib
    lib = _lib.load()
    case = ("3d_zigzag", 3, (16, 32, 32), True, True, False)
    dp, op, phi0, rhs0 = _setup(case, seed=3)
    assert lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_ZIGZAG, 3) != 0
    assert b"JSV_TUNE_ZIGZAG" in lib.jsv_last_error()
    _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_FUSED_MIN_M, 2))
    dp.upload(0, 0, phi0)
    dp.upload(0, 1, rhs0)
    dp.cycle(cfg())
    c = gmg.Config(fixed=True, max_iters_outer=1)
    np.testing.assert_array_equal(dp.download(0, 0), gmg.cycle(op, c, gmg.zero_rim(phi0), rhs0, c.fcycle))


@pytest.mark.parametrize("dim,ms", [(2, (64, 64)), (3, (16, 16, 16))])
def test_cycle_coarse_program_new_operator(dim, ms):
    """Same plan, a second operator of the same kind: the coarse program's
    LDS coefficient image must follow the new coefficients."""
    dp = None
    for seed in (31, 32):
        phi, rhs, lam, D, _ = rand_problem(dim, ms, seed=seed)
        if dp is None:
            dp = DevPlan(dim, ms, lam, D)
        else:
            dev = dp.plan.dev
            t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
            dp.plan.set_operator(t(lam), [t(d) for d in D])
        op = oracle_operator(dim, ms, lam, D)
        for kw in (dict(), dict(gamma=2)):
            c = gmg.Config(fixed=True, max_iters_outer=1, **kw)
            dp.upload(0, 0, phi)
            dp.upload(0, 1, rhs)
            dp.cycle(cfg(**kw))
            np.testing.assert_array_equal(dp.download(0, 0), gmg.cycle(op, c, gmg.zero_rim(phi), rhs, c.fcycle),
                                          err_msg=f"seed {seed} {kw}")


def test_cycle_coarse_program_operator_switch():
    """One plan, operator switched D_xy -> 5-point -> D_xy: each cycle still
    gives the oracle's bits (the D_xy arrays stay allocated in between)."""
    ms = (64, 64)
    phi, rhs, lam, D, Dxy = rand_problem(2, ms, seed=23, dxy=True)
    dp = DevPlan(2, ms, lam, D, Dxy)
    dev = dp.plan.dev
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    for kw in (dict(), dict(gamma=2), dict(fcycle=True)):
        for use_dxy in (True, False, True):
            dp.plan.set_operator(t(lam), [t(d) for d in D], t(Dxy) if use_dxy else None)
            op = oracle_operator(2, ms, lam, D, Dxy if use_dxy else None)
            c = gmg.Config(fixed=True, max_iters_outer=1, **kw)
            dp.upload(0, 0, phi)
            dp.upload(0, 1, rhs)
            dp.cycle(cfg(**kw))
            np.testing.assert_array_equal(dp.download(0, 0), gmg.cycle(op, c, gmg.zero_rim(phi), rhs, c.fcycle),
                                          err_msg=f"{kw} dxy={use_dxy}")


@pytest.mark.parametrize("coarse_m", [0, 16])
@pytest.mark.parametrize("case", [c for c in CASES if c[0] in ("2d_spatial_16x32", "2d_dxy_16", "3d_spatial_16x8x32")]
                         + [("3d_spatial_32", 3, (32, 32, 32), True, True, False),
                            ("2d_spatial_128x64", 2, (128, 64), True, True, False)], ids=lambda c: c[0])
def test_cycle_clears_stale_coarse_phi(case, coarse_m):
    """Coarse phi arrays full of junk (rims included) before a cycle: the
    restriction's zeroing of the next level's phi must clear every element,
    so the cycle still gives the oracle's bits."""
    from jsolver_amd import _lib
    dp, op, phi, rhs = _setup(case, seed=41)
    _lib.check(_lib.load().jsv_plan_set_tuning(dp.plan.h, _lib.JSV_TUNE_COARSE_M, coarse_m))
    rng = np.random.default_rng(5)
    for kw in (dict(), dict(gamma=2, fcycle=True)):
        for level in range(1, op.levels):
            dp.upload(level, 0, rng.standard_normal(op.lv[level].shape))
        dp.upload(0, 0, phi)
        dp.upload(0, 1, rhs)
        dp.cycle(cfg(**kw))
        c = gmg.Config(fixed=True, max_iters_outer=1, **kw)
        want = gmg.cycle(op, c, gmg.zero_rim(phi), rhs, c.fcycle)
        np.testing.assert_array_equal(dp.download(0, 0), want, err_msg=str(kw))


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_residual_matches_oracle(case):
    dp, op, phi, rhs = _setup(case, seed=7)
    got, n2 = dp.residual(phi, rhs)
    want = gmg.residual(op, phi, rhs)
    np.testing.assert_array_equal(got, want)
    assert n2 == pytest.approx(float(np.sum(want * want)), rel=1e-13)


# ---------------------------------------------------------------------------
# whole solves through the public API
# ---------------------------------------------------------------------------
from jsolver_amd.grid_1d import grid_1D  # noqa: E402
from jsolver_amd.solver_2d import (solve_2d, solve_2d_fcycle_simple, solve_2d_fixed_simple,  # noqa: E402
                                   solve_2d_simple, solve_2d_skip_pre_simple,
                                   solve_2d_wcycle_simple)
from jsolver_amd.solver_3d import solve_3d, solve_3d_simple  # noqa: E402

REL, ABS = G["tolerance"]["rel"], G["tolerance"]["abs"]
SMALL_2D = [g for g in G["tests_2d"] if g["mx"] <= 64 and g["my"] <= 128]


@pytest.mark.parametrize("g", SMALL_2D, ids=lambda g: f"p{g['problem']}_{g['mx']}x{g['my']}")
def test_solve_2d_goldens(g):
    gx, gy = grid_1D(g["mx"]), grid_1D(g["my"])
    ana, rhs, lam, dx, dy, dxy = problem_2d(g["problem"], gx.centers, gy.centers)
    phi0 = np.zeros(rhs.shape)
    out, count = solve_2d_simple(gx, gy, phi0, rhs, lam, dx, dy, dxy)
    res = out.cpu().numpy()
    assert count == g["iterations"]
    assert gmg.rmse(ana, res) == pytest.approx(g["rmse"], abs=ABS, rel=REL)
    ogx, ogy = gmg.OracleGrid(g["mx"]), gmg.OracleGrid(g["my"])
    want, wc = gmg.solve_2d(ogx, ogy, phi0, rhs, lam, dx, dy, dxy)
    assert wc == count
    np.testing.assert_array_equal(res, want)


@pytest.mark.parametrize("g", G["tests_2d_p1_variants"], ids=lambda g: g["kind"])
def test_solve_2d_p1_variants(g):
    gx, gy = grid_1D(128), grid_1D(128)
    ana, rhs, lam, dx, dy, _ = problem_2d(1, gx.centers, gy.centers)
    phi0 = np.zeros(rhs.shape)
    fn = {"fcycle": lambda: solve_2d_fcycle_simple(gx, gy, phi0, rhs, lam, dx, dy),
          "wcycle": lambda: solve_2d(gx, gy, phi0, rhs, lam, dx, dy, gamma=2),
          "fixed13": lambda: solve_2d_fixed_simple(gx, gy, 13, phi0, rhs, lam, dx, dy),
          "fixed13_fcycle": lambda: solve_2d(gx, gy, phi0, rhs, lam, dx, dy, fixed=True,
                                             max_iters_outer=13, fcycle=True)}[g["kind"]]
    out, count = fn()
    if g["exact"]:
        assert count == g["iterations"]
    else:
        assert count <= g["iterations"]
    assert gmg.rmse(ana, out.cpu().numpy()) == pytest.approx(g["rmse"], rel=g["rel"])


@pytest.mark.parametrize("g", G["tests_3d"], ids=lambda g: f"p{g['problem']}_{g['mx']}x{g['my']}x{g['mz']}")
def test_solve_3d_goldens(g):
    gx, gy, gz = grid_1D(g["mx"]), grid_1D(g["my"]), grid_1D(g["mz"])
    ana, rhs, lam, dx, dy, dz = problem_3d(g["problem"], gx.centers, gy.centers, gz.centers)
    phi0 = np.zeros(rhs.shape)
    out, count = solve_3d_simple(gx, gy, gz, phi0, rhs, lam, dx, dy, dz)
    assert count == g["iterations"]
    assert gmg.rmse(ana, out.cpu().numpy()) == pytest.approx(g["rmse"], abs=ABS, rel=REL)


TABLE_KW = {"reference_v": dict(), "reference_vskip": dict(skip_pre=True),
            "reference_f": dict(fcycle=True), "reference_fskip": dict(fcycle=True, skip_pre=True),
            "reference_w": dict(gamma=2)}


def _table_rows(res_set):
    out = []
    for name in TABLE_KW:
        for p, res, it, val in G["tables_2d"][name]["rows"]:
            if res in res_set and np.isfinite(val) and not (name == "reference_v" and p == 4 and res == 64):
                out.append(pytest.param(name, p, res, it, val, id=f"{name}-p{p}-{res}"))
    return out


# The tables were produced on the reference author's platform (a JAX/XLA
# evaluation with its own operation fusion); at 256^2 one row differs from the
# IEEE evaluation here by 1.06e-9 relative RMSE with the same iteration count,
# so table rows use 5e-9 (the reference's tests themselves use 1e-9, which the
# test-suite goldens above meet).
TABLE_REL = 5e-9


@pytest.mark.parametrize("name,p,res,it,val", _table_rows({8, 16, 32, 256}))
def test_solve_2d_table_goldens(name, p, res, it, val):
    gx = grid_1D(res)
    ana, rhs, lam, dx, dy, dxy = problem_2d(p, gx.centers, gx.centers)
    out, count = solve_2d(gx, gx, np.zeros(rhs.shape), rhs, lam, dx, dy, D_xy=dxy, **TABLE_KW[name])
    assert count == it
    assert gmg.rmse(ana, out.cpu().numpy()) == pytest.approx(val, abs=ABS, rel=TABLE_REL)


def test_solve_3d_spatial_matches_oracle_per_cycle():
    """3D spatially varying D (no reference golden): every cycle's iterate
    equals the oracle's bit for bit, and the dynamic count matches."""
    ms = (16, 16, 32)
    phi, rhs, lam, D, _ = rand_problem(3, ms, seed=11)
    gr = [grid_1D(m) for m in ms]
    ogr = [gmg.OracleGrid(m) for m in ms]
    phi0 = np.zeros(rhs.shape)
    hist_o = []
    want, wc = gmg.solve_3d(*ogr, phi0, rhs, lam, *D, history=hist_o, fixed=True, max_iters_outer=6)
    for k in range(1, 7):
        out, c = solve_3d(*gr, phi0, rhs, lam, *D, fixed=True, max_iters_outer=k)
        np.testing.assert_array_equal(out.cpu().numpy(), hist_o[k - 1], err_msg=f"cycle {k}")
    want, wc = gmg.solve_3d(*ogr, phi0, rhs, lam, *D)
    out, count = solve_3d(*gr, phi0, rhs, lam, *D)
    assert count == wc
    np.testing.assert_array_equal(out.cpu().numpy(), want)


def test_solve_nonzero_rim_and_warm_start():
    """A phi with non-zero boundary/ghost values: the first distance includes
    them (solver_2d.py:471-473) and the result matches the oracle."""
    ms = (32, 32)
    phi, rhs, lam, D, _ = rand_problem(2, ms, seed=3)
    g = [grid_1D(m) for m in ms]
    og = [gmg.OracleGrid(m) for m in ms]
    for kw in (dict(), dict(skip_pre=True), dict(gamma=2, fcycle=True)):
        out, count = solve_2d(*g, phi, rhs, lam, *D, **kw)
        want, wc = gmg.solve_2d(*og, phi, rhs, lam, *D, **kw)
        assert count == wc
        np.testing.assert_array_equal(out.cpu().numpy(), want)


def test_inputs_not_mutated_and_torch_inputs():
    ms = (16, 16)
    phi, rhs, lam, D, _ = rand_problem(2, ms, seed=4)
    g = [grid_1D(m) for m in ms]
    t = [torch.from_numpy(a.copy()).cuda() for a in (phi, rhs, lam, D[0], D[1])]
    keep = [x.clone() for x in t]
    out, count = solve_2d(*g, *t)
    for a, b in zip(t, keep):
        assert torch.equal(a, b)
    out2, count2 = solve_2d(*g, phi, rhs, lam, *D)
    assert count == count2 and torch.equal(out, out2)


def test_nan_input_stops_after_one_cycle():
    """NaN in rhs: the reference's `dist > eps` is False for NaN, so the loop
    ends after one cycle with a NaN phi (solver_2d.py:465-477)."""
    ms = (16, 16)
    phi, rhs, lam, D, _ = rand_problem(2, ms, seed=8)
    rhs[5, 7] = np.nan
    g = [grid_1D(m) for m in ms]
    out, count, info = solve_2d(*g, np.zeros(rhs.shape), rhs, lam, *D, return_info=True)
    assert count == 1
    assert info["nonfinite"]
    assert torch.isnan(out).any()


# ---------------------------------------------------------------------------
# do_BCs
# ---------------------------------------------------------------------------
def test_do_bcs_2d():
    from jsolver_amd.boundary_handler_2d import do_BCs
    rng = np.random.default_rng(0)
    a = rng.standard_normal((19, 23))
    for kw in (dict(), dict(rim=2), dict(direction=0), dict(direction=1),
               dict(vals_xb=rng.standard_normal(19), vals_xe=rng.standard_normal(19),
                    vals_yb=rng.standard_normal(23), vals_ye=rng.standard_normal(23)),
               dict(direction=1, rim=0, vals_yb=np.full(23, 2.0))):
        np.testing.assert_array_equal(do_BCs(a, **kw).cpu().numpy(), gmg.bcs_2d(a, **kw))
    # the reference's own test: ones(131, 131) -> four zero rims
    z = do_BCs(np.ones((131, 131))).cpu().numpy()
    assert (z[:2] == 0).all() and (z[-2:] == 0).all() and (z[:, :2] == 0).all() and (z[:, -2:] == 0).all()
    assert (z[2:-2, 2:-2] == 1).all()


def test_do_bcs_3d():
    from jsolver_amd.boundary_handler_3d import do_BCs
    rng = np.random.default_rng(1)
    a = rng.standard_normal((9, 9, 12))
    for kw in (dict(), dict(rim=2), dict(direction=0), dict(direction=1), dict(direction=2),
               dict(vals_xb=rng.standard_normal((9, 9)), vals_xe=rng.standard_normal((9, 9)),
                    vals_yb=rng.standard_normal((9, 12)), vals_ye=rng.standard_normal((9, 12)))):
        np.testing.assert_array_equal(do_BCs(a, **kw).cpu().numpy(), gmg.bcs_3d(a, **kw))


# ---------------------------------------------------------------------------
# batched 2D solves: every problem equals its own solve (bit for bit)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("kw", [dict(), dict(stop="rel_residual", rtol=1e-9), dict(gamma=2, fcycle=True),
                                dict(stop="rel_residual", rtol=1e-9, check_every=0),
                                dict(stop="rel_residual", rtol=1e-9, check_every=3, max_iters_outer=40),
                                dict(precision="mixed", stop="rel_residual", rtol=1e-10),
                                dict(precision="mixed", stop="rel_residual", rtol=1e-10, inner_cycles=5,
                                     max_iters_outer=33),
                                dict(precision="mixed", gamma=2),
                                dict(fixed=True, max_iters_outer=5)],
                         ids=["distance", "relres", "W_F", "relres_adaptive", "relres_every3_cap",
                              "mixed", "mixed_inner5_cap", "mixed_distance", "fixed"])
def test_batched_equals_single(kw):
    from jsolver_amd.solver_2d import solve_batched_2d
    ms = (32, 32)
    g = [grid_1D(m) for m in ms]
    B = 5
    probs = [rand_problem(2, ms, seed=100 + b) for b in range(B)]
    # spread the coefficient scales so problems stop after different counts
    for b, pr in enumerate(probs):
        pr[3][1] *= (1.0 + 3 * b)
    phi0 = np.stack([np.zeros(pr[1].shape) for pr in probs])
    rhs = np.stack([pr[1] for pr in probs])
    lam = np.stack([pr[2] for pr in probs])
    dx = np.stack([pr[3][0] for pr in probs])
    dy = np.stack([pr[3][1] for pr in probs])
    out, counts, _ = solve_batched_2d(*g, phi0, rhs, lam, dx, dy, **kw)
    singles = [solve_2d(*g, phi0[b], rhs[b], lam[b], dx[b], dy[b], **kw) for b in range(B)]
    for b in range(B):
        assert counts[b] == singles[b][1], b
        np.testing.assert_array_equal(out[b].cpu().numpy(), singles[b][0].cpu().numpy())
    if "max_iters_outer" not in kw:
        assert len(set(counts.tolist())) > 1      # the problems really stopped at different cycles
