"""Manufactured problems of the reference's test-suite, restated as data
generators (numpy, FP64).  Each returns (phi_ana, rhs, lam, D_xx, D_yy[, D_xy])
in 2D or (phi_ana, rhs, lam, D_xx, D_yy, D_zz) in 3D on node centres.

2D problems: tests/test_solver_2d.py:131-144 (p1), 243-252 (p2), 320-331 (p3),
399-411 (p4), 479-490 (p5), 558-573 (p7), 641-656 (p9), 724-739 (p10),
807-836 (p11).  3D: tests/test_solver_3d.py:168-179 (p1), 248-262 (p2),
330-344 (p3); the polynomial 3D problem of examples/example3D.py:29-41 is "p0".
"""
import numpy as np

PI = np.pi


def problem_2d(pid, xc, yc):
    shape = (yc.size, xc.size)
    x = xc[np.newaxis]
    y = yc[:, np.newaxis]
    if pid == 1:
        ox, oy = 1. - x, 1. - y
        px3 = x * ox ** 3
        py = y * oy
        return px3 * py, -2 * px3 - 6 * ox * (ox - x) * py, 0.0, 1.0, 1.0, None
    if pid == 2:
        sy = np.sin(PI * yc)[:, np.newaxis]
        sx = np.sin(PI * xc)[np.newaxis]
        ana = sx * sy
        return ana, -(PI ** 2) * (1.0 + 1.0) * ana, 0.0, 1.0, 1.0, None
    if pid == 3:
        sy = np.sin(PI * yc)[:, np.newaxis]
        sx, cx = np.sin(PI * x), np.cos(PI * x)
        dxx = np.full(shape, x)
        dyy = np.full(shape, x)
        ana = sx * sy
        return ana, PI * cx * sy - 2 * PI ** 2 * x * ana, 0.0, dxx, dyy, None
    sy, cy = np.sin(PI * y), np.cos(PI * y)
    sx, cx = np.sin(PI * x), np.cos(PI * x)
    ana = sx * sy
    lam = 0.2 * x * y ** 2
    if pid == 4:
        dxx = np.full(shape, x)
        dyy = np.zeros(shape) + y
        s = x + y
        return ana, PI * np.sin(PI * s) - (s * PI ** 2 + lam) * ana, lam, dxx, dyy, None
    if pid == 5:
        dxx, dyy = 1.0, 1e-2
        return ana, -((PI ** 2) * (dxx + dyy) + lam) * ana, lam, dxx, dyy, None
    if pid == 7:
        c = 0.1
        dxx = np.full(shape, x)
        dyy = np.zeros(shape) + y
        dxy = np.full(shape, c * x)
        rhs = ((1. + c) * PI * sx * cy + PI * cx * sy + c * 2. * x * PI ** 2 * cx * cy
               - ((x + y) * PI ** 2 + lam) * ana)
        return ana, rhs, lam, dxx, dyy, dxy
    if pid == 9:
        c = 0.5
        dxx = np.full(shape, x)
        dyy = np.zeros(shape) + y
        dxy = c * x * y
        rhs = ((1 + c * x) * PI * cx * sy + (1 + c * y) * PI * sx * cy
               + c * 2 * x * y * PI ** 2 * cx * cy - ((x + y) * PI ** 2 + lam) * ana)
        return ana, rhs, lam, dxx, dyy, dxy
    if pid == 10:
        c = 0.1
        dxx = np.zeros(shape) + y
        dyy = np.zeros(shape) + y
        dxy = c * x * y
        rhs = (PI * (1 + c * y) * sx * cy + c * x * PI * cx * sy
               + c * 2 * x * y * PI ** 2 * cx * cy - (2 * y * PI ** 2 + lam) * ana)
        return ana, rhs, lam, dxx, dyy, dxy
    if pid == 11:
        dpar, dperp = 1., 0.1
        ang = np.arctan2(y, x)
        sa, ca = np.sin(ang), np.cos(ang)
        sca, s2, c2 = sa * ca, sa ** 2, ca ** 2
        dxx = dpar * s2 + dperp * c2
        dyy = dpar * c2 + dperp * s2
        dxy = (dperp - dpar) * sca
        r2 = x ** 2 + y ** 2
        with np.errstate(divide="ignore", invalid="ignore"):
            dpx = -y / r2
            dpy = x / r2
        dDxx_dx = 2 * (dpar - dperp) * sca * dpx
        dDyy_dy = 2 * (dperp - dpar) * sca * dpy
        dDxy_dx = (dperp - dpar) * (c2 - s2) * dpx
        dDxy_dy = (dperp - dpar) * (c2 - s2) * dpy
        rhs = ((dDyy_dy + dDxy_dx) * PI * sx * cy + (dDxx_dx + dDxy_dy) * PI * cx * sy
               + 2 * dxy * PI ** 2 * cx * cy - ((dxx + dyy) * PI ** 2 + lam) * ana)
        return ana, rhs, lam, dxx, dyy, dxy
    raise ValueError(pid)


def problem_3d(pid, xc, yc, zc):
    x = xc[np.newaxis, np.newaxis, :]
    y = yc[np.newaxis, :, np.newaxis]
    z = zc[:, np.newaxis, np.newaxis]
    if pid == 0:
        ox, oy, oz = 1.0 - x, 1.0 - y, 1.0 - z
        px3, py, pz = x * ox ** 3, y * oy, z * oz
        ana = px3 * py * pz
        rhs = -6 * ox * (ox - x) * py * pz - 2 * px3 * pz - 2 * px3 * py
        return ana, rhs, 0.0, 1.0, 1.0, 1.0
    sx, sy, sz = np.sin(PI * x), np.sin(PI * y), np.sin(PI * z)
    ana = sx * sy * sz
    if pid == 1:
        return ana, -(PI ** 2) * 3.0 * ana, 0.0, 1.0, 1.0, 1.0
    lam = 0.2 * x * y ** 2 * z ** 3
    if pid == 2:
        return ana, -((PI ** 2) * 3.0 + lam) * ana, lam, 1.0, 1.0, 1.0
    if pid == 3:
        return ana, -((PI ** 2) * (1.0 + 0.5 + 2.0) + lam) * ana, lam, 1.0, 0.5, 2.0
    raise ValueError(pid)
