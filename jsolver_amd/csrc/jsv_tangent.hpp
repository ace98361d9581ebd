// Tangent (forward-mode) pass of multigrid solves, FP64.
//
// The reference differentiates its solves forward with jax.jvp
// (examples/benchmark.py:1146-1182; tests/test_solver_2d.py:173-176, which
// runs check_grads(modes=["fwd"]) on the dynamic solve).  Every operation of a
// cycle is linear in (phi, rhs) for fixed coefficients, and the coefficient
// setup (solver_2d.py:326-421, solver_3d.py:338-429) is linear in (D, lambda),
// so the tangent of a cycle needs only:
//   * the tangent coefficients A' of every level: the same setup applied to
//     (D', lambda'), held by a second plan;
//   * per half sweep, at each swept point p with new primal value phi_p:
//       A_c phi'_p = sum_nb (A_nb phi'_nb + A'_nb phi_nb) - rhs'_p - A'_c phi_p
//   * the tangent residual r' = (L phi' + rhs') + L' phi restricted with the
//     primal's weights, and the primal's prolongation applied to phi'.
// The primal runs alongside (each kernel below also performs the primal
// update of its points with the primal's own device function), because the
// coefficient terms need the primal iterate at that point of the cycle.
// The part linear in (phi', rhs') is summed in the primal's term order and the
// coefficient part separately, so with A' = 0 the tangent carries exactly the
// bits of a primal solve run on (phi0', rhs').
#pragma once
#include "jsv_kernels.hpp"

namespace jsv {

// Tangent coefficient pointers of one level (the tangent plan's arrays).
struct TanCoef {
    const double *cx, *cy, *cz, *cc, *drr, *dlr;
};

// One RBGS half (5/7-point) with its tangent; threads laid out like k_rb3 /
// k_rb2 (one per point of the half).  dim 2: one problem, z = 0.
__global__ void __launch_bounds__(BX * BY)
k_rb_tan(double* __restrict__ phi, double* __restrict__ tphi, const double* __restrict__ rhs,
         const double* __restrict__ trhs, const double* __restrict__ cx, const double* __restrict__ cy,
         const double* __restrict__ cz, const double* __restrict__ cc, TanCoef t, Geo g, int par, int dim)
{
    const int z = (dim == 3) ? 2 + (int)blockIdx.z : 0;
    const int y = 2 + blockIdx.y * BY + threadIdx.y;
    if (y > g.my) return;
    const int x = 2 + ((par + y + z) & 1) + 2 * (blockIdx.x * BX + threadIdx.x);
    if (x > g.mx) return;
    const long p = (long)z * g.sz + (long)y * g.sy + x;
    const long sy = g.sy, sz = g.sz;
    double lin, cof;
    if (dim == 3) {
        rb3_point(phi, rhs, cx, cy, cz, cc, p, sy, sz);
        lin = cz[p] * tphi[p + sz] + cz[p - sz] * tphi[p - sz];
        lin = lin + cy[p] * tphi[p + sy];
        lin = lin + cy[p - sy] * tphi[p - sy];
        lin = lin + cx[p] * tphi[p + 1];
        lin = lin + cx[p - 1] * tphi[p - 1];
        lin = lin - trhs[p];
        cof = t.cz[p] * phi[p + sz] + t.cz[p - sz] * phi[p - sz];
        cof = cof + t.cy[p] * phi[p + sy];
        cof = cof + t.cy[p - sy] * phi[p - sy];
    } else {
        rb2_point(phi, rhs, cx, cy, cc, p, sy);
        lin = cy[p] * tphi[p + sy] + cy[p - sy] * tphi[p - sy];
        lin = lin + cx[p] * tphi[p + 1];
        lin = lin + cx[p - 1] * tphi[p - 1];
        lin = lin - trhs[p];
        cof = t.cy[p] * phi[p + sy] + t.cy[p - sy] * phi[p - sy];
    }
    cof = cof + t.cx[p] * phi[p + 1];
    cof = cof + t.cx[p - 1] * phi[p - 1];
    cof = cof - t.cc[p] * phi[p];
    tphi[p] = (lin + cof) / cc[p];
}

// One D_xy sub-colour (lx, ly) with its tangent (2D, one problem).
__global__ void __launch_bounds__(BX * BY)
k_rb2diag_tan(double* __restrict__ phi, double* __restrict__ tphi, const double* __restrict__ rhs,
              const double* __restrict__ trhs, const double* __restrict__ cx,
              const double* __restrict__ cy, const double* __restrict__ cc,
              const double* __restrict__ drr, const double* __restrict__ dlr, TanCoef t, Geo g, int lx,
              int ly)
{
    const int y = (ly ? 2 : 3) + 2 * (blockIdx.y * BY + threadIdx.y);
    if (y > g.my) return;
    const int x = (lx ? 2 : 3) + 2 * (blockIdx.x * BX + threadIdx.x);
    if (x > g.mx) return;
    const long sy = g.sy;
    const long p = (long)y * sy + x;
    rb2diag_point(phi, rhs, cx, cy, cc, drr, dlr, p, sy);
    double lin = cy[p] * tphi[p + sy] + cy[p - sy] * tphi[p - sy];
    lin = lin + cx[p] * tphi[p + 1];
    lin = lin + cx[p - 1] * tphi[p - 1];
    lin = lin - trhs[p];
    double off = drr[p - sy - 1] * tphi[p - sy - 1] + dlr[p - sy + 1] * tphi[p - sy + 1];
    off = off + dlr[p] * tphi[p + sy - 1];
    off = off + drr[p] * tphi[p + sy + 1];
    lin = lin + off;
    double cof = t.cy[p] * phi[p + sy] + t.cy[p - sy] * phi[p - sy];
    cof = cof + t.cx[p] * phi[p + 1];
    cof = cof + t.cx[p - 1] * phi[p - 1];
    cof = cof + t.drr[p - sy - 1] * phi[p - sy - 1];
    cof = cof + t.dlr[p - sy + 1] * phi[p - sy + 1];
    cof = cof + t.dlr[p] * phi[p + sy - 1];
    cof = cof + t.drr[p] * phi[p + sy + 1];
    cof = cof - t.cc[p] * phi[p];
    tphi[p] = (lin + cof) / cc[p];
}

// Tangent residual at fine point p: (rhs' - ...) part via the primal's
// residual function on (phi', rhs'), plus L' phi (no rhs term).
__device__ __forceinline__ double resid_tan(const double* __restrict__ phi, const double* __restrict__ tphi,
                                            const double* __restrict__ trhs, const double* __restrict__ cx,
                                            const double* __restrict__ cy, const double* __restrict__ cz,
                                            const double* __restrict__ cc, const double* __restrict__ drr,
                                            const double* __restrict__ dlr, const TanCoef& t, long p,
                                            long sy, long sz, int dim)
{
    double lin;
    if (dim == 3) lin = resid3<double, double>(tphi, trhs, cx, cy, cz, cc, p, sy, sz);
    else lin = resid2<double, double>(tphi, trhs, cx, cy, cc, drr, dlr, p, sy);
    double cof = t.cc[p] * phi[p];
    cof = cof - t.cx[p - 1] * phi[p - 1];
    cof = cof - t.cx[p] * phi[p + 1];
    cof = cof - t.cy[p - sy] * phi[p - sy];
    cof = cof - t.cy[p] * phi[p + sy];
    if (dim == 3) {
        cof = cof - t.cz[p - sz] * phi[p - sz];
        cof = cof - t.cz[p] * phi[p + sz];
    } else if (t.drr != nullptr) {
        double off = t.drr[p - sy - 1] * phi[p - sy - 1] + t.dlr[p - sy + 1] * phi[p - sy + 1];
        off = off + t.dlr[p] * phi[p + sy - 1];
        off = off + t.drr[p] * phi[p + sy + 1];
        cof = cof - off;
    }
    return lin + cof;
}

// Tangent of the defect + restriction: one thread per coarse unknown, the
// primal's weights and summation order (restrict2_at / restrict3_at); also
// zeroes the coarse tangent phi (with its rim) like the primal restriction.
__global__ void __launch_bounds__(BX * BY)
k_restrict_tan(double* __restrict__ trc, const double* __restrict__ phi, const double* __restrict__ tphi,
               const double* __restrict__ trhs, const double* __restrict__ cx, const double* __restrict__ cy,
               const double* __restrict__ cz, const double* __restrict__ cc, const double* __restrict__ drr,
               const double* __restrict__ dlr, TanCoef t, Geo gf, Geo gc, int dim,
               double* __restrict__ ztphi)
{
    const int I = (dim == 3) ? 2 + (int)blockIdx.z : 0;
    const int J = 2 + blockIdx.y * BY + threadIdx.y;
    const int K = 2 + blockIdx.x * BX + threadIdx.x;
    if (J > gc.my || K > gc.mx) return;
    const long sy = gf.sy, sz = gf.sz;
    auto R = [&](long o) -> double {
        return resid_tan(phi, tphi, trhs, cx, cy, cz, cc, drr, dlr, t, o, sy, sz, dim);
    };
    double out;
    if (dim == 3) {
        const long c = (long)(2 * I - 1) * sz + (long)(2 * J - 1) * sy + (2 * K - 1);
        double e = R(c + sy + 1) + R(c + sy - 1);
        e = e + R(c - sy + 1);
        e = e + R(c - sy - 1);
        e = e + R(c + sz + 1);
        e = e + R(c + sz - 1);
        e = e + R(c - sz + 1);
        e = e + R(c - sz - 1);
        e = e + R(c + sz + sy);
        e = e + R(c + sz - sy);
        e = e + R(c - sz + sy);
        e = e + R(c - sz - sy);
        const double rcen = R(c);
        out = 0.03125 * e + 0.125 * rcen;
    } else {
        const long c = (long)(2 * J - 1) * sy + (2 * K - 1);
        double s = R(c + sy + 1) + R(c + sy - 1);
        s = s + R(c - sy + 1);
        s = s + R(c - sy - 1);
        const double rcen = R(c);
        out = 0.0625 * s + 0.25 * rcen;
    }
    trc[(long)I * gc.sz + (long)J * gc.sy + K] = out;
    zero_with_rim(ztphi, gc, 0, dim, I, J, K);
}

}  // namespace jsv
