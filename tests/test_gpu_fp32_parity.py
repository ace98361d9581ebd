"""FP32 kernels (the mixed solver's inner cycles) against the oracle run in
float32 on the FP32-cast operator (gmg.cast_operator): every operation
rounds once per term in the reference's order in both, with no FMA
contraction on the GPU and IEEE division, so the comparison is bit for bit
(tolerance 0), per operation and per whole cycle, for every kernel variant
the FP32 path uses: one launch per half, the fused pair sweep (one and two
column pairs per thread, with and without the prolongation inside it), the
one- and two-pass restriction, the fused halves and the coarse program."""
import numpy as np
import pytest

from gpu_helpers import DevPlan, cfg, oracle_operator, rand_problem
from jsolver_amd import _lib
from oracle import gmg

pytestmark = pytest.mark.gpu

F32 = _lib.JSV_MIXED

CASES = [
    # id, dim, ms (x,y[,z]), spatial, lam_array, dxy
    ("2d_spatial_32", 2, (32, 32), True, True, False),
    ("2d_spatial_16x64", 2, (16, 64), True, False, False),
    ("2d_scalar", 2, (32, 32), False, False, False),
    ("2d_dxy_32", 2, (32, 32), True, True, True),
    ("3d_spatial_16", 3, (16, 16, 16), True, True, False),
    ("3d_spatial_16x32x16", 3, (16, 32, 16), True, False, False),
    ("3d_scalar_lamarr", 3, (16, 16, 16), False, True, False),
]


def _setup(case, seed=3):
    _, dim, ms, spatial, lam_array, dxy = case
    phi, rhs, lam, D, Dxy = rand_problem(dim, ms, seed, spatial, lam_array, dxy)
    op32 = gmg.cast_operator(oracle_operator(dim, ms, lam, D, Dxy), np.float32)
    dp = DevPlan(dim, ms, lam, D, Dxy, precision="mixed")
    return dp, op32, phi.astype(np.float32), rhs.astype(np.float32)


def _level_arrays(op, level, seed):
    rng = np.random.default_rng(seed)
    shape = op.lv[level].shape
    phi = gmg.zero_rim(rng.standard_normal(shape).astype(np.float32))
    return phi, rng.standard_normal(shape).astype(np.float32)


def _tune(dp, **kw):
    lib = _lib.load()
    for k, v in kw.items():
        _lib.check(lib.jsv_plan_set_tuning(dp.plan.h, getattr(_lib, "JSV_TUNE_" + k), v))


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fp32_smooth_bitexact(case):
    dp, op, _, _ = _setup(case)
    for level in range(op.levels - 1):
        phi, rhs = _level_arrays(op, level, 10 + level)
        dp.upload(level, 0, phi, prec=F32)
        dp.upload(level, 1, rhs, prec=F32)
        dp.smooth(level, 2, prec=F32)
        want = gmg.smooth(op, level, gmg.smooth(op, level, phi, rhs), rhs)
        np.testing.assert_array_equal(dp.download(level, 0, prec=F32), want, err_msg=f"level {level}")


@pytest.mark.parametrize("threads", [0, 1024])
@pytest.mark.parametrize("zigzag", [0, 1, 2])
@pytest.mark.parametrize("pairs", [1, 2])
@pytest.mark.parametrize("ms", [(32, 16, 16), (64, 32, 16), (256, 64, 64)])
def test_fp32_fused_pair_sweep_bitexact(ms, pairs, zigzag, threads):
    """The fused out-of-place pair sweep (the 256^3 fine level's kernel),
    forced onto small levels; with zig-zag on, one of each level's two
    sweeps marches downward.  Two pairs per thread: 10-row tiles of 768
    threads (default) or 14-row tiles of 1024 (256 wide: rows of whole
    waves)."""
    case = ("3d", 3, ms, True, True, False)
    dp, op, _, _ = _setup(case, seed=5)
    _tune(dp, FUSED_MIN_M=2, SWEEP_PAIRS=pairs, ZIGZAG=zigzag, SWEEP_THREADS=threads)
    for level in range(2):
        phi, rhs = _level_arrays(op, level, 20 + level)
        dp.upload(level, 0, phi, prec=F32)
        dp.upload(level, 1, rhs, prec=F32)
        dp.smooth(level, 2, prec=F32)
        want = gmg.smooth(op, level, gmg.smooth(op, level, phi, rhs), rhs)
        np.testing.assert_array_equal(dp.download(level, 0, prec=F32), want, err_msg=f"level {level}")


@pytest.mark.parametrize("two_pass", [0, 1])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fp32_restrict_bitexact(case, two_pass):
    dp, op, _, _ = _setup(case)
    _tune(dp, TWO_PASS_MIN=1 if two_pass else 1 << 40)
    for level in range(op.levels - 2):
        phi, rhs = _level_arrays(op, level, 30 + level)
        phi = gmg.smooth(op, level, phi, rhs)      # restriction assumes a smoothed iterate
        dp.upload(level, 0, phi, prec=F32)
        dp.upload(level, 1, rhs, prec=F32)
        dp.restrict(level, prec=F32)
        want = gmg.restrict(op, level, phi, rhs)
        np.testing.assert_array_equal(dp.download(level + 1, 1, prec=F32), want, err_msg=f"level {level}")


@pytest.mark.parametrize("nu2", [0, 1])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fp32_prolong_bitexact(case, nu2):
    dp, op, _, _ = _setup(case)
    for level in range(op.levels - 2):
        phi, _ = _level_arrays(op, level, 40 + level)
        pc, _ = _level_arrays(op, level + 1, 50 + level)
        dp.upload(level, 0, phi, prec=F32)
        dp.upload(level + 1, 0, pc, prec=F32)
        dp.prolong(level, nu2, prec=F32)
        want = phi + gmg.prolong(op, level, pc, nu2)
        got = dp.download(level, 0, prec=F32)
        inner = (slice(2, -2),) * op.dim
        np.testing.assert_array_equal(got[inner], want[inner], err_msg=f"level {level}")


CYCLES = [dict(), dict(gamma=2), dict(fcycle=True), dict(nu1=2, nu2=1), dict(gamma=2, nu1=1, nu2=2)]


@pytest.mark.parametrize("kw", CYCLES, ids=lambda k: ",".join(f"{a}={b}" for a, b in k.items()) or "V")
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fp32_cycle_bitexact(case, kw):
    """One FP32 cycle with the production defaults (fused halves, coarse
    program, graph replay), twice in a row."""
    dp, op, phi, rhs = _setup(case, seed=7)
    c = gmg.Config(fixed=True, max_iters_outer=1, **kw)
    want = gmg.cycle(op, c, gmg.zero_rim(phi), rhs, c.fcycle)
    for rep in range(2):
        dp.upload(0, 0, gmg.zero_rim(phi), prec=F32)
        dp.upload(0, 1, rhs, prec=F32)
        dp.cycle(cfg(**kw), prec=F32)
        np.testing.assert_array_equal(dp.download(0, 0, prec=F32), want, err_msg=f"call {rep}")


@pytest.mark.parametrize("reg", [2, 1])
@pytest.mark.parametrize("kw", CYCLES, ids=lambda k: ",".join(f"{a}={b}" for a, b in k.items()) or "V")
@pytest.mark.parametrize("case", [c for c in CASES if c[1] == 2 and not c[5]], ids=lambda c: c[0])
def test_fp32_cycle_coarse_program_register_runs(case, kw, reg):
    """One FP32 cycle whose coarse program starts at level 1, its one-wave
    runs in registers: one point per lane (reg 1) or also from a level of up
    to 256 points, a 2 x 2 block per lane (reg 2)."""
    dp, op, phi, rhs = _setup(case, seed=11)
    _tune(dp, COARSE_M=1 << 20, COARSE_REG=reg)
    c = gmg.Config(fixed=True, max_iters_outer=1, **kw)
    want = gmg.cycle(op, c, gmg.zero_rim(phi), rhs, c.fcycle)
    for rep in range(2):
        dp.upload(0, 0, gmg.zero_rim(phi), prec=F32)
        dp.upload(0, 1, rhs, prec=F32)
        dp.cycle(cfg(**kw), prec=F32)
        np.testing.assert_array_equal(dp.download(0, 0, prec=F32), want, err_msg=f"call {rep}")


@pytest.mark.parametrize("fr", [1, 0])
@pytest.mark.parametrize("rows,zchunk", [(0, 0), (5, 3), (14, 0), (6, 1)], ids=lambda v: str(v))
@pytest.mark.parametrize("ms,pairs", [((256, 64, 64), 2), ((64, 32, 16), 2), ((32, 16, 64), 1)],
                         ids=lambda v: str(v))
@pytest.mark.parametrize("zigzag", [1, 0, 2])
def test_fp32_fused_first_half_residuals_bitexact(ms, pairs, rows, zchunk, fr, zigzag):
    """FP32 cycles whose last pre-smoothing pair sweep also stores the
    restriction's first-half residuals (k_resid_edges3 adds the tile edges),
    against the oracle run in float32; 256 wide with two pairs per thread is
    the 256^3 solve's sweep shape (rows of whole waves)."""
    case = ("3d_fr32", 3, ms, True, True, False)
    dp, op, phi, rhs = _setup(case, seed=41)
    _tune(dp, FUSED_MIN_M=2, TWO_PASS_MIN=1, SWEEP_PAIRS=pairs, SWEEP_ROWS=rows, ZCHUNK=zchunk, FUSE_RESID=fr,
          ZIGZAG=zigzag)
    for kw in (dict(), dict(gamma=2, nu2=2)):
        c = gmg.Config(fixed=True, max_iters_outer=1, **kw)
        want = gmg.cycle(op, c, gmg.zero_rim(phi), rhs, c.fcycle)
        for rep in range(2):
            dp.upload(0, 0, gmg.zero_rim(phi), prec=F32)
            dp.upload(0, 1, rhs, prec=F32)
            dp.cycle(cfg(**kw), prec=F32)
            np.testing.assert_array_equal(dp.download(0, 0, prec=F32), want, err_msg=f"{kw} call {rep}")


@pytest.mark.parametrize("threads", [0, 768, 1024])
@pytest.mark.parametrize("fuse_prolong", [0, 2])
@pytest.mark.parametrize("kw", [dict(), dict(gamma=2, nu2=2), dict(fcycle=True)], ids=["V", "W", "F"])
def test_fp32_cycle_fused_sweep_bitexact(kw, fuse_prolong, threads):
    """A 3D FP32 cycle with the fine level on the fused pair sweep (two pairs
    per thread, prolongation inside the post-smoothing sweep or separate;
    JSV_TUNE_SWEEP_THREADS 768: every two-pair sweep in 10-row tiles of 768
    threads, 1024: in 14-row tiles of 1024 where that fits)."""
    case = ("3d", 3, (64, 32, 32), True, True, False)
    dp, op, phi, rhs = _setup(case, seed=9)
    _tune(dp, FUSED_MIN_M=32, SWEEP_PAIRS=2, FUSE_PROLONG=fuse_prolong, SWEEP_THREADS=threads)
    c = gmg.Config(fixed=True, max_iters_outer=1, **kw)
    want = gmg.cycle(op, c, gmg.zero_rim(phi), rhs, c.fcycle)
    dp.upload(0, 0, gmg.zero_rim(phi), prec=F32)
    dp.upload(0, 1, rhs, prec=F32)
    dp.cycle(cfg(**kw), prec=F32)
    np.testing.assert_array_equal(dp.download(0, 0, prec=F32), want)
