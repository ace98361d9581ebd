"""CPU oracle: an FP64 NumPy restatement of jsolver's geometric multigrid solver.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this module;
only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may use it, and only as the checker (or the timed CPU baseline), never as
the thing shipped.

The reference is pure Python/JAX and cannot run in this environment (no jax),
so this file restates its arithmetic from the source text.  Each function cites
the reference lines it follows.  Array layout follows the reference exactly:
ghost-padded ``(m+3)`` per axis, axis order (y, x) in 2D and (z, y, x) in 3D,
and per-level coefficient arrays of shape ``(m+1)`` per axis whose entry ``k``
belongs to phi index ``k+1`` (jsolver/solver_2d.py:403-409,
jsolver/solver_3d.py:413-422).  The order of every floating-point sum is kept
as written in the reference, so the oracle matches it to rounding of fused
multiply-adds (the reference's own tests accept 1e-9 relative).

Pinning (what checks each function, see tests/test_oracle_*.py):
  * level_layout / grids: the reference's error-message tests and grid tests.
  * build_operator 2D (scalar and spatial D, D_xy), smooth, restrict, prolong, cycle,
    solve (V/F/W, fixed/dynamic, skip_pre): the exact iteration counts and
    RMSE goldens of tests/test_solver_2d.py and examples/benchmark.py tables.
  * build_operator 3D scalar path and the 3D cycle: goldens of tests/test_solver_3d.py.
  * build_operator 3D spatial path (no golden exists): closed forms -- constant D
    arrays reproduce the scalar operator at level 0 and 1 and give the
    (1/2 + 2^-l) D upper faces at level l >= 2; the brute-force sparse direct
    solve of the level-0 operator equals the converged multigrid solution;
    operator symmetry; ghost values of D/lambda/rhs have no effect.
  * residual_norm / apply_operator (new, no reference counterpart): compared
    with a scipy sparse assembly of the same operator.
  * bcs_2d: the reference's own do_BCs tests (default, vals, x and y direction)
    restated; bcs_3d (no reference test): hand-built expected arrays for every
    direction and rim, including the z faces taking vals_yb / vals_ye.
  * prolong, including the nu2 <= 0 (even, even[, even]) interpolations: every
    fine unknown reproduces an affine field sampled on the coarse grid exactly,
    and a variant with one face weight changed fails that check.
  * dirichlet_lift (boundary="phi", no reference counterpart): equals a direct
    sparse solve with the boundary values moved to the right-hand side.
  * cast_operator (the FP32 hierarchy): the float32 results stay within FP32
    rounding of the FP64 oracle.
Every function has such a check; no part of the oracle is unpinned.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

INT32_MAX = int(np.iinfo(np.int32).max)

# Sub-lattice (colour) orders of the RBGS sweep, given per axis in (x, y[, z])
# order as in jsolver/solver_2d.py:684 and jsolver/solver_3d.py:693.
SUBCOLORS_2D = ((0, 0), (1, 1), (0, 1), (1, 0))
SUBCOLORS_3D = ((0, 0, 0), (0, 1, 1), (1, 1, 0), (1, 0, 1),
                (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1))


# ----------------------------------------------------------------------------
# grid_1D (jsolver/grid_1d.py:56-114)
# ----------------------------------------------------------------------------
class OracleGrid:
    """Node grid: index i <-> x = xb + (i - 1 - 0.5 + 0.5) * del (centered)."""

    def __init__(self, mx, xb=0.0, xe=1.0, rim=1, grid_type=0, centered=True):
        # grid_1d.py:82-89
        self.mx, self.xb, self.xe, self.rim = mx, xb, xe, rim
        self.grid_type, self.centered = grid_type, centered
        self.nx = mx + 1
        self.length = xe - xb
        self.del_ = self.length / (mx if centered else self.nx)
        self.idel = 1 / self.del_
        # grid_1d.py:91-114
        pos = np.arange(-rim, mx + rim + 2)
        if grid_type == 0:
            e = (pos - 0.5 if centered else pos) * self.del_ + xb
        elif grid_type == 1:
            e = (pos + 2 * np.sin(2 * np.pi * pos / self.nx)) * self.del_ + xb
        else:
            raise ValueError(f"No such grid grid_type: {grid_type}")
        self.edges = e
        self.centers = 0.5 * (e[:-1] + e[1:])
        self.widths = e[1:] - e[:-1]
        self.is_linear = grid_type == 0


# ----------------------------------------------------------------------------
# do_BCs (jsolver/boundary_handler_2d.py:3-29, boundary_handler_3d.py:3-35)
# ----------------------------------------------------------------------------
def _ft(*arrs):
    """float64; float32 when every input is float32 (the FP32 restatement of
    the mixed solver's inner cycles, see cast_operator); complex128 when any
    input is complex (complex-step JVPs)."""
    dts = [np.asarray(a).dtype for a in arrs]
    if dts and all(d == np.float32 for d in dts):
        return np.dtype(np.float32)
    return np.result_type(np.float64, *dts)


def zero_rim(a: np.ndarray, rim: int = 1) -> np.ndarray:
    """Return a copy with the first/last rim+1 layers of every axis zeroed."""
    out = np.array(a, dtype=_ft(a), copy=True)
    w = rim + 1
    for ax in range(out.ndim):
        idx = [slice(None)] * out.ndim
        idx[ax] = slice(0, w)
        out[tuple(idx)] = 0.0
        idx[ax] = slice(out.shape[ax] - w, out.shape[ax])
        out[tuple(idx)] = 0.0
    return out


def bcs_2d(arr, rim=1, direction=-1, vals_xb=None, vals_xe=None, vals_yb=None, vals_ye=None):
    """boundary_handler_2d.py:19-29: columns (x, direction 0), then rows (y, 1)."""
    a = np.array(arr, dtype=np.float64, copy=True)
    w = rim + 1
    if direction in (-1, 0):
        a[:, :w] = 0.0 if vals_xb is None else np.asarray(vals_xb)[:, None]
        a[:, a.shape[1] - w:] = 0.0 if vals_xe is None else np.asarray(vals_xe)[:, None]
    if direction in (-1, 1):
        a[:w] = 0.0 if vals_yb is None else np.asarray(vals_yb)[None, :]
        a[a.shape[0] - w:] = 0.0 if vals_ye is None else np.asarray(vals_ye)[None, :]
    return a


def bcs_3d(arr, rim=1, direction=-1, vals_xb=None, vals_xe=None, vals_yb=None,
           vals_ye=None, vals_zb=None, vals_ze=None):
    """boundary_handler_3d.py:21-35.  Faithful to the reference, the z faces
    take ``vals_yb``/``vals_ye`` (boundary_handler_3d.py:31-33); ``vals_z*``
    are accepted and unused."""
    a = np.array(arr, dtype=np.float64, copy=True)
    w = rim + 1
    if direction in (-1, 0):
        a[:, :, :w] = 0.0 if vals_xb is None else np.asarray(vals_xb)[:, :, None]
        a[:, :, a.shape[2] - w:] = 0.0 if vals_xe is None else np.asarray(vals_xe)[:, :, None]
    if direction in (-1, 1):
        a[:, :w, :] = 0.0 if vals_yb is None else np.asarray(vals_yb)[:, None, :]
        a[:, a.shape[1] - w:, :] = 0.0 if vals_ye is None else np.asarray(vals_ye)[:, None, :]
    if direction in (-1, 2):
        a[:w] = 0.0 if vals_yb is None else np.asarray(vals_yb)[None, :, :]
        a[a.shape[0] - w:] = 0.0 if vals_ye is None else np.asarray(vals_ye)[None, :, :]
    return a


# ----------------------------------------------------------------------------
# level layout (solver_2d.py:279-324, solver_3d.py:281-336)
# ----------------------------------------------------------------------------
def level_layout(ms: Sequence[int], names: Sequence[str]):
    """ms, names in (x, y[, z]) order.  Returns (levels, per-axis m per level
    in array-axis order (slowest first))."""
    for m, n in zip(ms, names):
        if (m & (m - 1)) > 0:
            raise ValueError(f"Resolution of {n}-grid is not a power of 2. Value: {m}")
    lv = [int(np.ceil(np.log2(m))) + 1 for m in ms]
    levels = min(lv)
    per_axis = []
    for m_levels in lv:
        per_axis.append([2 ** (m_levels - 1 - l) for l in range(levels)])
    for m_l, n in zip(per_axis, names):
        if m_l[-1] > 5:
            raise ValueError(f"Resolution of {n}-grid not efficient for multigrid configuration. "
                             f"Resolution of coarsest level: {m_l[-1]}")
    # array-axis order: reverse (x,y,z) -> (z,y,x)
    m_level = [tuple(per_axis[a][l] for a in reversed(range(len(ms)))) for l in range(levels)]
    return levels, m_level


# ----------------------------------------------------------------------------
# slicing helpers.  "win(o)" is the reference's [1:-1] window shifted by o.
# ----------------------------------------------------------------------------
def _win(o: int) -> slice:
    return slice(1 + o, (-1 + o) if (-1 + o) != 0 else None)


def _sub(l: int, o: int) -> slice:
    """phi indices l+1, l+3, ... <= m+1 shifted by o (the [l+1:-1:2] family)."""
    stop = -1 + o
    return slice(l + 1 + o, stop if stop != 0 else None, 2)


def _unit(d: int, axis: int, o: int):
    t = [0] * d
    t[axis] = o
    return t


# ----------------------------------------------------------------------------
# per-level operator
# ----------------------------------------------------------------------------
@dataclass
class Level:
    shape: tuple                 # full padded shape (m+3 per axis)
    # spatial operator: minus/plus face coefficient per array axis, centre
    a_minus: Optional[List[np.ndarray]] = None
    a_plus: Optional[List[np.ndarray]] = None
    a_c: object = None           # scalar or (m+1)^d array
    # scalar operator: one coefficient per array axis
    a_axis: Optional[List[float]] = None
    # 2D off-diagonal coefficients (ll, rl, lr, rr), solver_2d.py:411-416
    a_diag: Optional[dict] = None

    @property
    def spatial(self):
        return self.a_minus is not None


@dataclass
class Operator:
    dim: int
    levels: int
    lv: List[Level] = field(default_factory=list)
    off_diag: bool = False
    dtype: object = np.float64     # complex128 when a coefficient input is complex


def cast_operator(op: "Operator", dtype=np.float32) -> "Operator":
    """The operator with every coefficient rounded to `dtype`: what the mixed
    solver's FP32 hierarchy holds (coefficients set up in FP64, then cast).
    With float32 phi and rhs, smooth / restrict / prolong / cycle then run in
    float32, one rounding per operation in the same order as in FP64."""
    def c(v):
        if v is None:
            return None
        if _is_scalar(v):
            return dtype(v)
        return np.asarray(v).astype(dtype)
    lv = []
    for L in op.lv:
        lv.append(Level(shape=L.shape,
                        a_minus=None if L.a_minus is None else [c(a) for a in L.a_minus],
                        a_plus=None if L.a_plus is None else [c(a) for a in L.a_plus],
                        a_c=c(L.a_c),
                        a_axis=None if L.a_axis is None else [c(a) for a in L.a_axis],
                        a_diag=None if L.a_diag is None else {k: c(v) for k, v in L.a_diag.items()}))
    return Operator(dim=op.dim, levels=op.levels, lv=lv, off_diag=op.off_diag, dtype=np.dtype(dtype))


def _inject(prev: np.ndarray, shape) -> np.ndarray:
    """zeros(shape) with [1:-1]^d <- prev[1::2]^d (solver_2d.py:353, solver_3d.py:364)."""
    out = np.zeros(shape, dtype=_ft(prev))
    d = len(shape)
    out[(slice(1, -1),) * d] = prev[(slice(1, None, 2),) * d]
    return out


def _average_along(a: np.ndarray, axis: int) -> np.ndarray:
    """Face average along `axis` over the reference's index window: axis range
    [0, n-1), other axes [1, n-1) (solver_2d.py:360-361, solver_3d.py:368-370)."""
    out = a.copy()
    d = a.ndim
    dst = [slice(1, -1)] * d
    hi = [slice(1, -1)] * d
    lo = [slice(1, -1)] * d
    dst[axis] = slice(None, -1)
    hi[axis] = slice(1, None)
    lo[axis] = slice(None, -1)
    out[tuple(dst)] = 0.5 * (a[tuple(hi)] + a[tuple(lo)])
    return out


def _is_scalar(v) -> bool:
    return np.ndim(v) == 0


def build_operator(dim, ms, idel, lam, D, D_xy=None) -> Operator:
    """Per-level coefficients.

    ms, idel: per axis in (x, y[, z]) order.  D: list (D_xx, D_yy[, D_zz]).
    2D: solver_2d.py:326-421.  3D: solver_3d.py:338-429.
    """
    names = "xyz"[:dim]
    levels, m_level = level_layout(ms, names)
    shapes = [tuple(m + 3 for m in ml) for ml in m_level]
    # sq_idel per axis (x,y,z) per level: idel * 0.5**arange(levels), squared
    sq = [(np.float64(i) * 0.5 ** np.arange(levels)) ** 2 for i in idel]
    # array axis a corresponds to physical axis (dim-1-a)
    sq_ax = [sq[dim - 1 - a] for a in range(dim)]
    op = Operator(dim=dim, levels=levels,
                  dtype=_ft(lam, *D, *([] if D_xy is None else [D_xy])))
    scalar_lam = _is_scalar(lam)
    spatial = not all(_is_scalar(x) for x in D)
    op.off_diag = spatial and D_xy is not None

    lam_l = None
    if not scalar_lam:
        # compute_lambda: 2d:365-383 / 3d:374-392
        lam_l = [np.asarray(lam, dtype=_ft(lam))]
        for l in range(1, levels):
            lam_l.append(_inject(lam_l[-1], shapes[l]))

    if spatial:
        # physical axis order D = (D_xx, D_yy[, D_zz]); array axis a <-> D[dim-1-a]
        Dl = [[np.asarray(D[dim - 1 - a], dtype=_ft(D[dim - 1 - a]))] for a in range(dim)]
        if dim == 2:
            # inject all levels from raw values first, then average every level
            for a in range(dim):
                for l in range(1, levels):
                    Dl[a].append(_inject(Dl[a][l - 1], shapes[l]))
                Dl[a] = [_average_along(Dl[a][l], a) for l in range(levels)]
        else:
            # level 0 never averaged; level l injected from level l-1 after
            # that level was averaged (solver_3d.py:363-370)
            for a in range(dim):
                for l in range(1, levels):
                    Dl[a].append(_average_along(_inject(Dl[a][l - 1], shapes[l]), a))
        for l in range(levels):
            minus, plus = [], []
            for a in range(dim):
                s = sq_ax[a][l]
                w_m = [slice(1, -1)] * dim
                w_p = [slice(1, -1)] * dim
                w_m[a] = slice(0, -2)
                minus.append(Dl[a][l][tuple(w_m)] * s)
                plus.append(Dl[a][l][tuple(w_p)] * s)
            # centre: fastest axis first, minus then plus, then lambda
            acc = None
            for a in reversed(range(dim)):
                acc = minus[a] if acc is None else acc + minus[a]
                acc = acc + plus[a]
            lam_term = lam if scalar_lam else lam_l[l][(slice(1, -1),) * dim]
            acc = acc + lam_term
            lev = Level(shape=shapes[l], a_minus=minus, a_plus=plus, a_c=acc)
            op.lv.append(lev)
        if op.off_diag:
            Dxy = [np.asarray(D_xy, dtype=_ft(D_xy))]
            for l in range(1, levels):
                Dxy.append(_inject(Dxy[-1], shapes[l]))
            idx = np.float64(idel[0]) * 0.5 ** np.arange(levels)
            idy = np.float64(idel[1]) * 0.5 ** np.arange(levels)
            cxy = idx * idy * 0.25
            for l in range(levels):
                q = Dxy[l]
                c = cxy[l]
                # q[r, c-1], q[r-1, c], q[r+1, c], q[r, c+1] on the [1:-1] window
                w_xm = q[1:-1, 0:-2]
                w_ym = q[0:-2, 1:-1]
                w_yp = q[2:, 1:-1]
                w_xp = q[1:-1, 2:]
                op.lv[l].a_diag = {
                    "ll": (w_xm + w_ym) * c,
                    "lr": -(w_xm + w_yp) * c,
                    "rl": -(w_xp + w_ym) * c,
                    "rr": (w_xp + w_yp) * c,
                }
    else:
        # np.float64 scalars (complex128 under a complex-step perturbation)
        Dv = [np.asarray(D[dim - 1 - a], dtype=_ft(D[dim - 1 - a]))[()] for a in range(dim)]
        for l in range(levels):
            a_axis = [Dv[a] * sq_ax[a][l] for a in range(dim)]
            # add = 2 * (A_x + A_y [+ A_z]), fastest axis first
            s = None
            for a in reversed(range(dim)):
                s = a_axis[a] if s is None else s + a_axis[a]
            add = 2 * s
            a_c = add + lam if scalar_lam else add + lam_l[l][(slice(1, -1),) * dim]
            op.lv.append(Level(shape=shapes[l], a_axis=a_axis, a_c=a_c))
    return op


# ----------------------------------------------------------------------------
# smoother (solver_2d.py:670-692, solver_3d.py:679-706)
# ----------------------------------------------------------------------------
def _coef_sub(arr, ls):
    """coefficient arrays are indexed [l::2] per axis."""
    if _is_scalar(arr):
        return arr
    return arr[tuple(slice(l, None, 2) for l in ls)]


def smooth(op: Operator, level: int, phi: np.ndarray, rhs: np.ndarray) -> np.ndarray:
    """One red-black Gauss-Seidel sweep over all sub-colours, in place on a copy."""
    lv = op.lv[level]
    d = op.dim
    phi = zero_rim(np.asarray(phi, dtype=np.result_type(_ft(phi, rhs), op.dtype)))
    orders = SUBCOLORS_2D if d == 2 else SUBCOLORS_3D
    for col in orders:
        # col is (lx, ly[, lz]); array axis a uses col[d-1-a]
        ls = [col[d - 1 - a] for a in range(d)]
        tgt = tuple(_sub(ls[a], 0) for a in range(d))

        def nb(axis, o):
            return phi[tuple(_sub(ls[a], o if a == axis else 0) for a in range(d))]

        # sum over neighbours, slowest axis first, plus side before minus side,
        # accumulated left to right
        terms = []
        for a in range(d):
            if lv.spatial:
                terms.append(_coef_sub(lv.a_plus[a], ls) * nb(a, +1))
                terms.append(_coef_sub(lv.a_minus[a], ls) * nb(a, -1))
            else:
                terms.append(lv.a_axis[a] * nb(a, +1))
                terms.append(lv.a_axis[a] * nb(a, -1))
        acc = terms[0]
        for t in terms[1:]:
            acc = acc + t
        acc = acc - rhs[tgt]
        if op.off_diag:
            dg = lv.a_diag
            ly, lx = ls

            def dn(oy, ox):
                return phi[_sub(ly, oy), _sub(lx, ox)]
            off = (_coef_sub(dg["ll"], ls) * dn(-1, -1) + _coef_sub(dg["rl"], ls) * dn(-1, +1)
                   + _coef_sub(dg["lr"], ls) * dn(+1, -1) + _coef_sub(dg["rr"], ls) * dn(+1, +1))
            acc = acc + off
        with np.errstate(divide="ignore", invalid="ignore"):
            # boundary nodes may divide by a zero centre coefficient; they are
            # overwritten by the rim zeroing right after, as in the reference
            phi[tgt] = acc / _coef_sub(lv.a_c, ls)
        phi = zero_rim(phi)
    return phi


# ----------------------------------------------------------------------------
# residual + restriction (solver_2d.py:616-647, solver_3d.py:622-652)
# ----------------------------------------------------------------------------
def residual_window(op: Operator, level: int, phi: np.ndarray, rhs: np.ndarray) -> np.ndarray:
    """rhs - L phi on the [1:-1]^d window (boundary nodes included, as the
    reference computes them before do_BCs)."""
    lv = op.lv[level]
    d = op.dim

    def sh(axis, o):
        return phi[tuple(_win(o if a == axis else 0) for a in range(d))]

    c = phi[(slice(1, -1),) * d]
    r = lv.a_c * c
    for a in reversed(range(d)):  # fastest axis first, minus then plus
        if lv.spatial:
            r = r - lv.a_minus[a] * sh(a, -1)
            r = r - lv.a_plus[a] * sh(a, +1)
        else:
            r = r - lv.a_axis[a] * sh(a, -1)
            r = r - lv.a_axis[a] * sh(a, +1)
    r = r + rhs[(slice(1, -1),) * d]
    if op.off_diag:
        dg = lv.a_diag
        off = (dg["ll"] * phi[:-2, :-2] + dg["rl"] * phi[:-2, 2:]
               + dg["lr"] * phi[2:, :-2] + dg["rr"] * phi[2:, 2:])
        r = r - off
    return r


def restrict(op: Operator, level: int, phi: np.ndarray, rhs: np.ndarray) -> np.ndarray:
    d = op.dim
    rw = residual_window(op, level, phi, rhs)
    full = np.zeros(rhs.shape, dtype=rw.dtype)
    full[(slice(1, -1),) * d] = rw
    full = zero_rim(full)
    cshape = op.lv[level + 1].shape
    rc = np.zeros(cshape, dtype=full.dtype)
    P, M, C = slice(2, None, 2), slice(None, -2, 2), slice(1, -1, 2)   # +1, -1, centre
    if d == 2:
        rc[1:-1, 1:-1] = 0.0625 * (full[P, P] + full[P, M] + full[M, P] + full[M, M]) + 0.25 * full[C, C]
    else:
        edge = (full[C, P, P] + full[C, P, M] + full[C, M, P] + full[C, M, M]
                + full[P, C, P] + full[P, C, M] + full[M, C, P] + full[M, C, M]
                + full[P, P, C] + full[P, M, C] + full[M, P, C] + full[M, M, C])
        rc[1:-1, 1:-1, 1:-1] = 0.03125 * edge + 0.125 * full[C, C, C]
    return zero_rim(rc)


# ----------------------------------------------------------------------------
# prolongation (solver_2d.py:649-668, solver_3d.py:654-677)
# ----------------------------------------------------------------------------
def prolong(op: Operator, level: int, phic: np.ndarray, nu2: int) -> np.ndarray:
    d = op.dim
    shape = op.lv[level].shape
    c = np.zeros(shape, dtype=_ft(phic))
    O = slice(1, -1, 2)         # odd indices 1..m+1 (coarse points)
    E = slice(2, -1, 2)         # even indices 2..m
    Pp = slice(3, None, 2)      # even + 1
    Pm = slice(1, -2, 2)        # even - 1
    c[(O,) * d] = phic[(slice(1, -1),) * d]
    if d == 2:
        c[O, E] = 0.5 * (c[O, Pp] + c[O, Pm])
        c[E, O] = 0.5 * (c[Pp, O] + c[Pm, O])
        if nu2 <= 0 or op.off_diag:
            c[E, E] = 0.25 * (c[Pp, Pp] + c[Pm, Pp] + c[Pp, Pm] + c[Pm, Pm])
    else:
        c[O, O, E] = 0.5 * (c[O, O, Pp] + c[O, O, Pm])
        c[O, E, O] = 0.5 * (c[O, Pp, O] + c[O, Pm, O])
        c[E, O, O] = 0.5 * (c[Pp, O, O] + c[Pm, O, O])
        if nu2 <= 0:
            c[O, E, E] = 0.25 * (c[O, Pp, Pp] + c[O, Pm, Pp] + c[O, Pp, Pm] + c[O, Pm, Pm])
            c[E, O, E] = 0.25 * (c[Pp, O, Pp] + c[Pm, O, Pp] + c[Pp, O, Pm] + c[Pm, O, Pm])
            c[E, E, O] = 0.25 * (c[Pp, Pp, O] + c[Pm, Pp, O] + c[Pp, Pm, O] + c[Pm, Pm, O])
        c[E, E, E] = 0.125 * (c[Pp, Pp, Pp] + c[Pm, Pp, Pp] + c[Pp, Pm, Pp] + c[Pm, Pm, Pp]
                              + c[Pp, Pp, Pm] + c[Pm, Pp, Pm] + c[Pp, Pm, Pm] + c[Pm, Pm, Pm])
    return c


# ----------------------------------------------------------------------------
# cycle and solve (solver_2d.py:423-479, 582-614)
# ----------------------------------------------------------------------------
@dataclass
class Config:
    fixed: bool = False
    epsilon: float = 1e-12
    gamma: int = 1
    fcycle: bool = False
    skip_pre: bool = False
    nu1: int = 1
    nu2: int = 1
    max_iters_outer: int = INT32_MAX


def cycle(op: Operator, cfg: Config, phi, rhs, fcyc=False, level=0, skip_pre=False, trace=None):
    if level == op.levels - 2:
        for _ in range(cfg.nu1 + cfg.nu2):
            phi = smooth(op, level, phi, rhs)
        return phi
    if not skip_pre:
        for _ in range(cfg.nu1):
            phi = smooth(op, level, phi, rhs)
    rc = restrict(op, level, phi, rhs)
    pc = np.zeros(op.lv[level + 1].shape, dtype=rc.dtype)
    if fcyc:
        for _ in range(cfg.gamma):
            pc = cycle(op, cfg, pc, rc, True, level + 1)
    for _ in range(cfg.gamma):
        pc = cycle(op, cfg, pc, rc, False, level + 1)
    phi = phi + prolong(op, level, pc, cfg.nu2)
    for _ in range(cfg.nu2):
        phi = smooth(op, level, phi, rhs)
    return phi


def solve(op: Operator, cfg: Config, phi, rhs, history=None):
    """Returns (phi, count).  history: optional list receiving phi after each cycle."""
    phi = np.asarray(phi, dtype=_ft(phi))
    rhs = np.asarray(rhs, dtype=_ft(rhs))
    if cfg.skip_pre:
        for _ in range(cfg.nu1):
            phi = smooth(op, 0, phi, rhs)
    if cfg.fixed:
        for _ in range(cfg.max_iters_outer):
            phi = cycle(op, cfg, zero_rim(phi), rhs, cfg.fcycle, 0, cfg.skip_pre)
            if history is not None:
                history.append(phi.copy())
        return phi, cfg.max_iters_outer
    count, dist = 0, math.inf
    while dist > cfg.epsilon and count < cfg.max_iters_outer:
        old = phi
        phi = cycle(op, cfg, zero_rim(phi), rhs, cfg.fcycle, 0, cfg.skip_pre)
        dist = float(np.max(np.abs(old - phi)))
        count += 1
        if history is not None:
            history.append(phi.copy())
    return phi, count


# ----------------------------------------------------------------------------
# new quantities with no live counterpart in the reference
# ----------------------------------------------------------------------------
def apply_operator(op: Operator, phi: np.ndarray, level: int = 0) -> np.ndarray:
    """L phi = sum A_nb phi_nb - A_c phi (+ offdiag) on unknowns [2:-2]^d,
    returned on the full padded shape (zeros elsewhere).  Basis: the
    reference's (dead) compute_residual, solver_2d.py:509-515, with the 3D
    z-neighbour coefficients taken from the z faces (the reference's
    solver_3d.py:517 uses the y faces there, a bug not reproduced)."""
    d = op.dim
    z = np.zeros(phi.shape)
    r = residual_window(op, level, phi, z)   # = A_c phi - sum A_nb phi_nb - off
    out = np.zeros(phi.shape, dtype=r.dtype)
    inner = (slice(1, -1),) * d
    out[inner] = -r
    return zero_rim(out)


def residual(op: Operator, phi: np.ndarray, rhs: np.ndarray, level: int = 0) -> np.ndarray:
    """rhs - L phi on the unknowns (zero on boundary nodes and ghosts)."""
    d = op.dim
    rw = residual_window(op, level, zero_rim(phi), rhs)
    out = np.zeros(phi.shape, dtype=rw.dtype)
    out[(slice(1, -1),) * d] = rw
    return zero_rim(out)


def dirichlet_lift(op: Operator, phi_b: np.ndarray, rhs: np.ndarray) -> np.ndarray:
    """Non-zero Dirichlet values (an extension: the reference's solver always
    calls do_BCs with zeros, README.md:90; do_BCs(arr, vals_...) is
    boundary_handler_2d.py:3-29).  The values phi_b holds on the level-0
    boundary nodes (index 1 and m+1 per axis) move to the right-hand side of
    the unknowns next to them: rhs - sum over boundary-node neighbours q of
    A(p, q) phi_b[q], with A the smoother's neighbour coefficients (face
    coefficients and D_xy corners).  A zero-boundary solve of the result
    equals a solve whose boundary nodes hold phi_b, sweep by sweep up to
    rounding.  Subtraction order per point: axes fastest first (x, y[, z]),
    minus side then plus side, then the corners ll, rl, lr, rr."""
    d = op.dim
    lv = op.lv[0]
    out = np.array(rhs, dtype=np.result_type(_ft(rhs, phi_b), op.dtype), copy=True)
    shape = rhs.shape
    m = [n - 3 for n in shape]                 # per array axis
    inner = tuple(slice(2, n - 2) for n in shape)
    grids = np.meshgrid(*[np.arange(2, n - 2) for n in shape], indexing="ij")

    def bnd(a, idx):
        return (idx == 1) | (idx == m[a] + 1)

    def coef(arr, offs):
        # coefficient array on the [1:-1] window, at unknown + offs (window index = array index - 1)
        if np.ndim(arr) == 0:
            return arr
        sl = tuple(slice(1 + o, n - 3 + o) for o, n in zip(offs, shape))
        return arr[sl]

    def nbval(offs):
        sl = tuple(slice(2 + o, n - 2 + o) for o, n in zip(offs, shape))
        return phi_b[sl]

    def sub(mask, c, v):
        o = out[inner]
        o[mask] = o[mask] - (c * v)[mask] if np.ndim(c) else o[mask] - (c * v[mask])
        out[inner] = o

    for a in reversed(range(d)):
        for sgn in (-1, +1):
            offs = [0] * d
            offs[a] = sgn
            mask = bnd(a, grids[a] + sgn)
            if lv.spatial:
                c = coef(lv.a_minus[a] if sgn < 0 else lv.a_plus[a], [0] * d)
            else:
                c = lv.a_axis[a]
            sub(mask, c, nbval(offs))
    if op.off_diag:
        dg = lv.a_diag
        for key, (oy, ox) in (("ll", (-1, -1)), ("rl", (-1, 1)), ("lr", (1, -1)), ("rr", (1, 1))):
            mask = bnd(0, grids[0] + oy) | bnd(1, grids[1] + ox)
            sub(mask, coef(dg[key], [0, 0]), nbval([oy, ox]))
    return out


def residual_norm(op: Operator, phi, rhs) -> float:
    return float(np.linalg.norm(residual(op, phi, rhs)))


def rhs_norm(rhs) -> float:
    d = np.ndim(rhs)
    return float(np.linalg.norm(np.asarray(rhs)[(slice(2, -2),) * d]))


# ----------------------------------------------------------------------------
# convenience front ends mirroring solve_2d / solve_3d
# ----------------------------------------------------------------------------
def solve_2d(grid_x, grid_y, phi, rhs, lam, D_xx, D_yy, D_xy=None, history=None, **kw):
    if not grid_x.is_linear or not grid_y.is_linear:
        raise ValueError("Only linear grids are supported!")
    op = build_operator(2, (grid_x.mx, grid_y.mx), (grid_x.idel, grid_y.idel), lam, [D_xx, D_yy], D_xy)
    return solve(op, Config(**kw), phi, rhs, history)


def solve_3d(grid_x, grid_y, grid_z, phi, rhs, lam, D_xx, D_yy, D_zz, history=None, **kw):
    if not grid_x.is_linear or not grid_y.is_linear or not grid_z.is_linear:
        raise ValueError("Only linear grids are supported!")
    op = build_operator(3, (grid_x.mx, grid_y.mx, grid_z.mx), (grid_x.idel, grid_y.idel, grid_z.idel),
                        lam, [D_xx, D_yy, D_zz])
    return solve(op, Config(**kw), phi, rhs, history)


def rmse(phi_ana, phi_res) -> float:
    """tests/test_solver_2d.py:20-23 (window [1:-1]^d)."""
    d = np.ndim(phi_ana)
    e = (np.asarray(phi_ana) - np.asarray(phi_res))[(slice(1, -1),) * d]
    return float(np.linalg.norm(e) / np.sqrt(e.size))
