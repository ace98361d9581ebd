"""Pins for the oracle where the reference holds no golden vector.

* 3D spatially varying D (solver_3d.py:338-372): closed form of the averaging
  quirk for constant D arrays; brute-force sparse direct solve of the level-0
  operator on 8^3 equals the converged multigrid solution.
* 2D: constant D arrays reproduce the scalar path; brute force on 8^2 and
  16x8 with and without D_xy.
* Discrete sine eigenvectors are reproduced exactly (closed form).
* do_BCs: the reference's own 2D tests (tests/test_boundary_handler_2d.py),
  and a point-by-point statement of the 3D faces (direction, rim, values,
  the z faces reading vals_y*).
* Prolongation: affine fields are reproduced exactly wherever the
  reference interpolates (3D bilinear faces only when nu2 <= 0).
* The operator is symmetric without D_xy; ghost values of D, lambda and rhs
  have no effect.
* The FP32 restatement (cast_operator, float32 phi and rhs) computes in
  float32, within FP32 rounding of the FP64 oracle.
Each check is chosen so a plausible slip (a dropped face, a shifted index, a
missing average, a sign) fails it.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import gmg


def _grids(ms):
    return [gmg.OracleGrid(m) for m in ms]


def _idels(ms):
    return tuple(g.idel for g in _grids(ms))


# ---------------------------------------------------------------------------
# independent assembly of the level-0 operator from the documented face rules
# ---------------------------------------------------------------------------
def assemble_level0(dim, ms, lam, D, Dxy=None):
    """Sparse matrix of L (sum A_nb phi_nb - A_c phi) on the unknowns 2..m.
    2D faces: 0.5 (D[p] + D[p+e]) idel^2 (averaged); 3D faces at level 0:
    D[p] idel^2 (not averaged).  D_xy: 9-point cross terms with coefficient
    idel_x idel_y / 4 built from D_xy at the face-adjacent nodes."""
    shape = tuple(m + 3 for m in reversed(ms))
    idel = [g.idel for g in _grids(ms)]      # x, y, z
    inner = [np.arange(2, n - 2) for n in shape]
    idx = -np.ones(shape, dtype=np.int64)
    mesh = np.meshgrid(*inner, indexing="ij")
    idx[tuple(mesh)] = np.arange(mesh[0].size).reshape(mesh[0].shape)
    N = mesh[0].size
    rows, cols, vals = [], [], []
    Darr = [np.broadcast_to(np.asarray(d, dtype=float), shape) for d in D]   # D_xx, D_yy[, D_zz]
    lam_a = np.broadcast_to(np.asarray(lam, dtype=float), shape)
    for p in zip(*[m.ravel() for m in mesh]):
        i = idx[p]
        diag = 0.0
        for a in range(dim):            # array axis a <-> physical axis dim-1-a
            phys = dim - 1 - a
            s = idel[phys] ** 2
            Dp = Darr[phys]
            for sgn in (+1, -1):
                q = list(p)
                q[a] += sgn
                q = tuple(q)
                face_lo = p if sgn > 0 else q
                face_hi = q if sgn > 0 else p
                if dim == 2:
                    f = 0.5 * (Dp[face_lo] + Dp[face_hi]) * s
                else:
                    f = Dp[face_lo] * s
                diag += f
                if idx[q] >= 0:
                    rows.append(i); cols.append(idx[q]); vals.append(f)
        diag += lam_a[p]
        rows.append(i); cols.append(i); vals.append(-diag)
        if Dxy is not None:
            c = idel[0] * idel[1] * 0.25
            y, x = p
            Q = Dxy
            terms = {(-1, -1): (Q[y, x - 1] + Q[y - 1, x]) * c,
                     (+1, -1): -(Q[y, x - 1] + Q[y + 1, x]) * c,
                     (-1, +1): -(Q[y, x + 1] + Q[y - 1, x]) * c,
                     (+1, +1): (Q[y, x + 1] + Q[y + 1, x]) * c}
            for (dy, dx), v in terms.items():
                q = (y + dy, x + dx)
                if idx[q] >= 0:
                    rows.append(i); cols.append(idx[q]); vals.append(v)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(N, N))
    return A, idx, mesh


def _direct_solve(dim, ms, lam, D, rhs, Dxy=None):
    A, idx, mesh = assemble_level0(dim, ms, lam, D, Dxy)
    b = rhs[tuple(mesh)].ravel()
    x = spla.spsolve(A.tocsc(), b)
    out = np.zeros(rhs.shape)
    out[tuple(mesh)] = x.reshape(mesh[0].shape)
    return out, A


def _mg_converged(dim, ms, lam, D, rhs, Dxy=None, cycles=200):
    op = gmg.build_operator(dim, ms, _idels(ms), lam, list(D), Dxy)
    phi, _ = gmg.solve(op, gmg.Config(fixed=True, max_iters_outer=cycles), np.zeros(rhs.shape), rhs)
    return phi, op


@pytest.mark.parametrize("ms,dxy", [((8, 8), False), ((16, 8), False), ((8, 16), True), ((8, 8), True)])
def test_2d_brute_force(ms, dxy):
    rng = np.random.default_rng(sum(ms) + dxy)
    shape = tuple(m + 3 for m in reversed(ms))
    D = [np.exp(0.3 * rng.standard_normal(shape)) for _ in range(2)]
    lam = rng.uniform(0, 2, shape)
    Dxy = 0.1 * rng.standard_normal(shape) if dxy else None
    rhs = rng.standard_normal(shape)
    want, _ = _direct_solve(2, ms, lam, D, rhs, Dxy)
    got, op = _mg_converged(2, ms, lam, D, rhs, Dxy)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-10 * np.abs(want).max())
    # the oracle's operator application agrees with the assembled matrix
    A, idx, mesh = assemble_level0(2, ms, lam, D, Dxy)
    v = gmg.zero_rim(rng.standard_normal(shape))
    Lv = gmg.apply_operator(op, v)
    np.testing.assert_allclose(Lv[tuple(mesh)].ravel(), A @ v[tuple(mesh)].ravel(), rtol=1e-13, atol=1e-9)


@pytest.mark.parametrize("ms", [(8, 8, 8), (8, 16, 8)])
def test_3d_brute_force_spatial(ms):
    rng = np.random.default_rng(7 + sum(ms))
    shape = tuple(m + 3 for m in reversed(ms))
    D = [np.exp(0.3 * rng.standard_normal(shape)) for _ in range(3)]
    lam = rng.uniform(0, 2, shape)
    rhs = rng.standard_normal(shape)
    want, A = _direct_solve(3, ms, lam, D, rhs)
    got, op = _mg_converged(3, ms, lam, D, rhs)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-10 * np.abs(want).max())
    # a 2D-style (averaged) level-0 face would not reproduce this solution
    Dav = [0.5 * (d + np.roll(d, -1, axis=a)) for a, d in zip((2, 1, 0), D)]
    wrong, _ = _direct_solve(3, ms, lam, Dav, rhs)
    assert np.abs(wrong - want).max() > 1e-4 * np.abs(want).max()


@pytest.mark.parametrize("level", [0, 1, 2, 3])
def test_3d_constant_D_face_quirk(level):
    """Constant D arrays: level 0/1 faces equal the scalar faces; at level
    l >= 2 the face next to the upper boundary is (1/2 + 2^-l) D, because
    level l is injected from the already averaged level l-1 whose ghost is 0
    (solver_3d.py:363-370).  Lower faces are unaffected."""
    ms = (32, 32, 32)
    c = 1.7
    shape = (35, 35, 35)
    D = [np.full(shape, c)] * 3
    op = gmg.build_operator(3, ms, _idels(ms), 0.0, D)
    lv = op.lv[level]
    m = ms[0] >> level
    sq = (_idels(ms)[0] * 0.5 ** level) ** 2
    for a in range(3):
        plus = lv.a_plus[a]          # (m+1)^3, entry k <-> phi index k+1
        upper = [slice(1, m)] * 3
        upper[a] = m - 1             # phi index m: face (m, m+1)
        lower = [slice(1, m)] * 3
        lower[a] = 1                 # phi index 2: face (2, 3)
        inner = [slice(1, m)] * 3
        inner[a] = slice(1, m - 1)
        expect_upper = (0.5 + 2.0 ** -level) * c * sq if level >= 2 else c * sq
        np.testing.assert_allclose(plus[tuple(upper)], expect_upper, rtol=1e-15)
        np.testing.assert_allclose(plus[tuple(lower)], c * sq, rtol=1e-15)
        np.testing.assert_allclose(plus[tuple(inner)], c * sq, rtol=1e-15)


def test_2d_constant_D_arrays_equal_scalar_path():
    ms = (32, 32)
    shape = (35, 35)
    rng = np.random.default_rng(3)
    rhs = rng.standard_normal(shape)
    lam = rng.uniform(0, 1, shape)
    gx, gy = _grids(ms)
    a, ca = gmg.solve_2d(gx, gy, np.zeros(shape), rhs, lam, np.full(shape, 2.0), np.full(shape, 0.5),
                         fixed=True, max_iters_outer=8)
    b, cb = gmg.solve_2d(gx, gy, np.zeros(shape), rhs, lam, 2.0, 0.5, fixed=True, max_iters_outer=8)
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-12 * np.abs(b).max())


@pytest.mark.parametrize("dim,ms,k", [(2, (32, 32), (1, 2)), (2, (16, 32), (3, 1)), (3, (16, 16, 16), (1, 2, 1))])
def test_discrete_sine_eigenvector(dim, ms, k):
    """D scalar, lambda = 0, rhs = -mu s_k with s_k the discrete sine mode:
    the converged solution is s_k, mu = sum_a 4 D_a idel_a^2 sin^2(k_a pi h_a / 2)."""
    gr = _grids(ms)
    Dv = [1.0, 0.7, 1.3][:dim]
    shape = tuple(m + 3 for m in reversed(ms))
    s = np.ones(shape)
    mu = 0.0
    for phys in range(dim):
        a = dim - 1 - phys
        h = gr[phys].del_
        i = np.arange(shape[a])
        mode = np.sin(k[phys] * np.pi * (i - 1) * h)
        sh = [1] * dim
        sh[a] = shape[a]
        s = s * mode.reshape(sh)
        mu += 4 * Dv[phys] * gr[phys].idel ** 2 * np.sin(k[phys] * np.pi * h / 2) ** 2
    s = gmg.zero_rim(s)
    rhs = -mu * s
    op = gmg.build_operator(dim, ms, _idels(ms), 0.0, Dv)
    phi, _ = gmg.solve(op, gmg.Config(fixed=True, max_iters_outer=60), np.zeros(shape), rhs)
    np.testing.assert_allclose(phi, s, rtol=0, atol=1e-11)


@pytest.mark.parametrize("dim,ms", [(2, (8, 8)), (3, (4, 4, 8))])
def test_operator_symmetric(dim, ms):
    rng = np.random.default_rng(11)
    shape = tuple(m + 3 for m in reversed(ms))
    D = [np.exp(0.5 * rng.standard_normal(shape)) for _ in range(dim)]
    lam = rng.uniform(0, 1, shape)
    op = gmg.build_operator(dim, ms, _idels(ms), lam, D)
    inner = tuple(slice(2, n - 2) for n in shape)
    n = int(np.prod([n - 4 for n in shape]))
    M = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(shape)
        e[inner].flat[j] = 1.0
        M[:, j] = gmg.apply_operator(op, e)[inner].ravel()
    np.testing.assert_allclose(M, M.T, rtol=1e-14, atol=0)
    assert np.all(np.linalg.eigvalsh(M) < 0)    # L is negative definite


@pytest.mark.parametrize("dim,ms", [(2, (16, 16)), (3, (8, 8, 8))])
def test_ghost_values_have_no_effect(dim, ms):
    rng = np.random.default_rng(5)
    shape = tuple(m + 3 for m in reversed(ms))
    D = [np.exp(0.5 * rng.standard_normal(shape)) for _ in range(dim)]
    lam = rng.uniform(0, 1, shape)
    rhs = rng.standard_normal(shape)
    gr = _grids(ms)
    base = (gmg.solve_2d if dim == 2 else gmg.solve_3d)(*gr, np.zeros(shape), rhs, lam, *D,
                                                        fixed=True, max_iters_outer=4)[0]
    ghost = np.zeros(shape, dtype=bool)
    for a in range(dim):
        sl = [slice(None)] * dim
        sl[a] = 0
        ghost[tuple(sl)] = True
        sl[a] = -1
        ghost[tuple(sl)] = True

    def perturb(x):
        y = x.copy()
        y[ghost] = rng.uniform(5, 9, ghost.sum())
        return y
    got = (gmg.solve_2d if dim == 2 else gmg.solve_3d)(*gr, np.zeros(shape), perturb(rhs), perturb(lam),
                                                       *[perturb(d) for d in D],
                                                       fixed=True, max_iters_outer=4)[0]
    inner = tuple(slice(1, n - 1) for n in shape)
    np.testing.assert_array_equal(got[inner], base[inner])


# ---------------------------------------------------------------------------
# non-zero Dirichlet values (gmg.dirichlet_lift): the converged multigrid
# solution of the lifted right-hand side equals a direct solve of the level-0
# system whose boundary nodes hold the values, with the boundary couplings
# taken from the same independent face rules as assemble_level0
# ---------------------------------------------------------------------------
def _boundary_rhs(dim, ms, D, g, rhs, Dxy=None):
    shape = tuple(m + 3 for m in reversed(ms))
    idel = [gg.idel for gg in _grids(ms)]
    inner = [np.arange(2, n - 2) for n in shape]
    mesh = np.meshgrid(*inner, indexing="ij")
    Darr = [np.broadcast_to(np.asarray(d, dtype=float), shape) for d in D]
    b = rhs[tuple(mesh)].astype(float).copy()
    it = np.nditer(mesh[0], flags=["multi_index"])
    for _ in it:
        p = tuple(mm[it.multi_index] for mm in mesh)
        k = it.multi_index
        for a in range(dim):
            phys = dim - 1 - a
            s_ = idel[phys] ** 2
            Dp = Darr[phys]
            for sgn in (+1, -1):
                q = list(p)
                q[a] += sgn
                q = tuple(q)
                if not (q[a] == 1 or q[a] == shape[a] - 2):
                    continue
                lo, hi = (p, q) if sgn > 0 else (q, p)
                f = 0.5 * (Dp[lo] + Dp[hi]) * s_ if dim == 2 else Dp[lo] * s_
                b[k] -= f * g[q]
        if Dxy is not None:
            c = idel[0] * idel[1] * 0.25
            y, x = p
            Q = Dxy
            terms = {(-1, -1): (Q[y, x - 1] + Q[y - 1, x]) * c,
                     (+1, -1): -(Q[y, x - 1] + Q[y + 1, x]) * c,
                     (-1, +1): -(Q[y, x + 1] + Q[y - 1, x]) * c,
                     (+1, +1): (Q[y, x + 1] + Q[y + 1, x]) * c}
            for (dy, dx), v in terms.items():
                q = (y + dy, x + dx)
                if q[0] in (1, shape[0] - 2) or q[1] in (1, shape[1] - 2):
                    b[k] -= v * g[q]
    return b, mesh


@pytest.mark.parametrize("dim,ms,dxy,spatial", [(2, (8, 8), False, True), (2, (16, 8), True, True),
                                                (2, (8, 16), False, False), (3, (8, 8, 8), False, True),
                                                (3, (8, 16, 8), False, False)])
def test_dirichlet_lift_matches_direct_solve(dim, ms, dxy, spatial):
    rng = np.random.default_rng(11)
    shape = tuple(m + 3 for m in reversed(ms))
    D = [np.exp(0.3 * rng.standard_normal(shape)) for _ in range(dim)] if spatial else [1.3, 0.7, 2.0][:dim]
    lam = rng.uniform(0, 1, shape)
    rhs = rng.standard_normal(shape)
    Dxy = 0.1 * rng.standard_normal(shape) if dxy else None
    g = rng.standard_normal(shape)
    A, _, _ = assemble_level0(dim, ms, lam, D, Dxy)
    b, mesh = _boundary_rhs(dim, ms, D, g, rhs, Dxy)
    x = spla.spsolve(A.tocsc(), b.ravel()).reshape(mesh[0].shape)
    op = gmg.build_operator(dim, ms, _idels(ms), lam, list(D), Dxy)
    phi, _ = gmg.solve(op, gmg.Config(fixed=True, max_iters_outer=300), np.zeros(shape),
                       gmg.dirichlet_lift(op, g, rhs))
    got = phi[tuple(mesh)]
    assert np.max(np.abs(got - x)) <= 1e-10 * np.max(np.abs(x))
    # a lift that dropped a term (here the last one applied) fails the pin
    if dim == 2 and not dxy:
        bad = rhs.copy()
        rl = gmg.dirichlet_lift(op, g, rhs)
        bad[2:-2, -3] = rl[2:-2, -3] + 0.5 * (rl[2:-2, -3] - rhs[2:-2, -3])
        pb, _ = gmg.solve(op, gmg.Config(fixed=True, max_iters_outer=300), np.zeros(shape), bad)
        assert np.max(np.abs(pb[tuple(mesh)] - x)) > 1e-3 * np.max(np.abs(x))


# ---------------------------------------------------------------------------
# do_BCs: the reference's own tests (tests/test_boundary_handler_2d.py:13-64)
# restated against the oracle, and hand-built 3D expectations
# ---------------------------------------------------------------------------
def _pad(core, n_edge, dim):
    """jnp.pad(ones(dim - 2 n_edge), (n_edge, n_edge)) as in the reference's tests."""
    return np.pad(np.full(dim - 2 * n_edge, core), (n_edge, n_edge))


def test_bcs_2d_reference_default():
    dim = 131
    new = gmg.bcs_2d(np.ones((dim, dim)))
    for row in (new[0], new[1], new[-1], new[-2], new[:, 0], new[:, 1], new[:, -1], new[:, -2]):
        assert np.all(row == np.zeros(dim))
    # and nothing else changed
    assert np.all(new[2:-2, 2:-2] == 1.0)


def test_bcs_2d_reference_vals():
    dim = 131
    twos = np.ones(dim) * 2
    new = gmg.bcs_2d(np.ones((dim, dim)), vals_xb=twos, vals_xe=twos, vals_yb=twos, vals_ye=twos)
    for row in (new[0], new[1], new[-1], new[-2], new[:, 0], new[:, 1], new[:, -1], new[:, -2]):
        assert np.all(row == twos)
    assert np.all(new[2:-2, 2:-2] == 1.0)


def test_bcs_2d_reference_xdir():
    dim = 131
    new = gmg.bcs_2d(np.ones((dim, dim)), direction=0)
    for row in (new[0], new[1], new[-1], new[-2]):
        assert np.all(row == _pad(1.0, 2, dim))
    for col in (new[:, 0], new[:, 1], new[:, -1], new[:, -2]):
        assert np.all(col == np.zeros(dim))


def test_bcs_2d_reference_ydir():
    dim = 131
    new = gmg.bcs_2d(np.ones((dim, dim)), direction=1)
    for row in (new[0], new[1], new[-1], new[-2]):
        assert np.all(row == np.zeros(dim))
    for col in (new[:, 0], new[:, 1], new[:, -1], new[:, -2]):
        assert np.all(col == _pad(1.0, 2, dim))


def _bcs_3d_by_points(arr, rim, direction, vxb, vxe, vyb, vye):
    """Point-by-point statement of boundary_handler_3d.py:21-35: the x faces
    are set first, then the y faces, then the z faces, each over the whole
    face (later faces win on shared edges).  x faces take vals_x*[z, y],
    y faces vals_y*[z, x], and the z faces also vals_y*, read as [y, x]
    (boundary_handler_3d.py:31-33)."""
    nz, ny, nx = arr.shape
    w = rim + 1
    out = arr.copy()
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                v = arr[z, y, x]
                if direction in (-1, 0):
                    if x < w:
                        v = 0.0 if vxb is None else vxb[z, y]
                    elif x >= nx - w:
                        v = 0.0 if vxe is None else vxe[z, y]
                if direction in (-1, 1):
                    if y < w:
                        v = 0.0 if vyb is None else vyb[z, x]
                    elif y >= ny - w:
                        v = 0.0 if vye is None else vye[z, x]
                if direction in (-1, 2):
                    if z < w:
                        v = 0.0 if vyb is None else vyb[y, x]
                    elif z >= nz - w:
                        v = 0.0 if vye is None else vye[y, x]
                out[z, y, x] = v
    return out


@pytest.mark.parametrize("rim", [1, 2])
@pytest.mark.parametrize("direction", [-1, 0, 1, 2])
@pytest.mark.parametrize("with_vals", [False, True])
def test_bcs_3d_hand_built(rim, direction, with_vals):
    rng = np.random.default_rng(40 + rim)
    n = 9                       # nz == ny: the z faces read vals_y* as (ny, nx)
    arr = rng.standard_normal((n, n, 11))
    vals = [None] * 4
    if with_vals:
        # distinct values per face, so the order of the faces and the axes
        # each face reads show up
        vals = [rng.standard_normal((n, n)) + 10 * k for k in range(2)] + \
               [rng.standard_normal((n, 11)) + 10 * k for k in range(2, 4)]
    got = gmg.bcs_3d(arr, rim=rim, direction=direction, vals_xb=vals[0], vals_xe=vals[1],
                     vals_yb=vals[2], vals_ye=vals[3],
                     vals_zb=None if not with_vals else np.full((n, 11), -99.0),
                     vals_ze=None if not with_vals else np.full((n, 11), -99.0))
    want = _bcs_3d_by_points(arr, rim, direction, *vals)
    np.testing.assert_array_equal(got, want)
    if with_vals and direction in (-1, 2):
        # the z faces hold vals_yb / vals_ye, never vals_zb / vals_ze
        assert not np.any(got == -99.0)
        np.testing.assert_array_equal(got[0], vals[2])
        np.testing.assert_array_equal(got[-1], vals[3])


def test_bcs_3d_pointwise_statement_has_power():
    """The point-by-point statement separates the plausible misreadings: z
    faces from vals_z*, and x faces set last."""
    rng = np.random.default_rng(7)
    n = 7
    arr = rng.standard_normal((n, n, n))
    v = [rng.standard_normal((n, n)) + 10 * k for k in range(4)]
    want = _bcs_3d_by_points(arr, 1, -1, *v)
    wrong_z = gmg.bcs_3d(arr, vals_xb=v[0], vals_xe=v[1], vals_yb=v[2], vals_ye=v[3])
    wrong_z[:2] = v[0][None]          # z faces from another array
    assert not np.array_equal(wrong_z, want)
    x_last = gmg.bcs_3d(arr, vals_xb=v[0], vals_xe=v[1], vals_yb=v[2], vals_ye=v[3])
    x_last[:, :, :2] = v[0][:, :, None]
    x_last[:, :, -2:] = v[1][:, :, None]
    assert not np.array_equal(x_last, want)


# ---------------------------------------------------------------------------
# prolongation: linear interpolation reproduces affine fields exactly
# (solver_2d.py:649-668, solver_3d.py:654-677)
# ---------------------------------------------------------------------------
def _affine(shape, coef):
    """coef[0] + sum_a coef[a+1] * (index along array axis a), dyadic so every
    interpolation is exact in FP64."""
    idx = np.meshgrid(*[np.arange(n, dtype=float) for n in shape], indexing="ij")
    return coef[0] + sum(c * i for c, i in zip(coef[1:], idx))


def _coarse_from_fine_affine(op, level, coef):
    """Coarse level values whose fine point 2I-1 sits at coarse index I: the
    affine function of the fine index, sampled at the coarse points."""
    fine_shape = op.lv[level].shape
    coarse_shape = op.lv[level + 1].shape
    pc = np.zeros(coarse_shape)
    inner = tuple(slice(1, n - 1) for n in coarse_shape)
    # coarse index I (1..mc+1) <-> fine index 2I - 1
    idx = np.meshgrid(*[2.0 * np.arange(1, n - 1) - 1.0 for n in coarse_shape], indexing="ij")
    pc[inner] = coef[0] + sum(c * i for c, i in zip(coef[1:], idx))
    return pc, _affine(fine_shape, coef)


@pytest.mark.parametrize("dim,ms", [(2, (16, 8)), (3, (8, 8, 16))])
@pytest.mark.parametrize("nu2", [0, -1, 1])
def test_prolongation_reproduces_affine_fields(dim, ms, nu2):
    lam = 0.0
    D = [1.0] * dim
    op = gmg.build_operator(dim, ms, _idels(ms), lam, D)
    coef = [0.5, 0.25, -0.125, 0.375][:dim + 1]
    for level in range(op.levels - 2):
        pc, want = _coarse_from_fine_affine(op, level, coef)
        c = gmg.prolong(op, level, pc, nu2)
        inner = (slice(1, -1),) * dim
        got, ref = c[inner], want[inner]
        # points with exactly two even coordinates (3D) / both even (2D) are
        # interpolated only when nu2 <= 0; otherwise they stay 0 (the red
        # half of the post-smoothing sweep overwrites them)
        ev = np.meshgrid(*[(np.arange(n) % 2 == 0) for n in c.shape], indexing="ij")
        neven = sum(e.astype(int) for e in ev)[inner]
        skipped = (neven == 2) if nu2 > 0 else np.zeros_like(neven, dtype=bool)
        np.testing.assert_array_equal(got[~skipped], ref[~skipped])
        assert np.all(got[skipped] == 0.0)
        # ghost layers untouched
        assert np.all(c[(0,) + (slice(None),) * (dim - 1)] == 0.0)


def test_prolongation_affine_pin_has_power():
    """A variant with one weight of the 3D bilinear faces changed fails the
    affine check."""
    ms = (8, 8, 8)
    op = gmg.build_operator(3, ms, _idels(ms), 0.0, [1.0] * 3)
    coef = [0.5, 0.25, -0.125, 0.375]
    pc, want = _coarse_from_fine_affine(op, 0, coef)
    c = gmg.prolong(op, 0, pc, 0)
    O, E, Pp, Pm = slice(1, -1, 2), slice(2, -1, 2), slice(3, None, 2), slice(1, -2, 2)
    bad = c.copy()
    bad[E, O, E] = 0.25 * (c[Pp, O, Pp] + c[Pm, O, Pp] + c[Pp, O, Pm] + c[Pp, O, Pm])
    inner = (slice(1, -1),) * 3
    assert np.array_equal(c[inner], want[inner])
    assert not np.array_equal(bad[inner], want[inner])


def test_oracle_runs_in_float32():
    """The FP32 restatement really computes in float32: float32 results, within
    FP32 rounding of the FP64 oracle and not equal to its rounding."""
    ms = (16, 16, 16)
    rng = np.random.default_rng(3)
    shape = tuple(m + 3 for m in reversed(ms))
    D = [np.exp(0.5 * rng.standard_normal(shape)) for _ in range(3)]
    lam = rng.uniform(0.0, 2.0, shape)
    rhs = rng.standard_normal(shape)
    phi = rng.standard_normal(shape)
    op64 = gmg.build_operator(3, ms, tuple(g.idel for g in _grids(ms)), lam, D)
    op32 = gmg.cast_operator(op64, np.float32)
    p32, r32 = phi.astype(np.float32), rhs.astype(np.float32)
    c = gmg.Config(fixed=True, max_iters_outer=1)
    a = gmg.cycle(op32, c, gmg.zero_rim(p32), r32)
    b = gmg.cycle(op64, c, gmg.zero_rim(p32.astype(np.float64)), r32.astype(np.float64))
    assert a.dtype == np.float32
    assert np.max(np.abs(a - b)) <= 1e-4 * np.max(np.abs(b))
    assert not np.array_equal(a, b.astype(np.float32))
