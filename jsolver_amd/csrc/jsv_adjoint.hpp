// Reverse pass (adjoint) of fixed-mode multigrid solves, FP64.
//
// The reference gets this from jax.vjp over its fori_loop of cycles
// (examples/benchmark.py:1184-1226, tests/test_solver_2d.py:200-207), which
// stores every intermediate of every cycle.  Here the transpose of each
// operation is written out by hand:
//   half sweep   phi'_S = (sum_nb A phi_nb - rhs_S) / A_c  on colour S
//                -> t = g_S / A_c ;  g_nb += A t ;  rhs_bar -= t ;
//                   A_c_bar -= t phi'_S ;  A_face_bar += t phi_nb ;  g_S = 0
//   restriction  rhs_c = R (A_c phi - sum A phi_nb + rhs)
//                -> r_bar = R^T rhs_c_bar on the first-half points, then the
//                   transposed residual into g, rhs_bar and the coefficient bars
//   prolongation phi += P phi_c  -> g_c = P^T g
// applied in reverse order of the forward cycle (post-smoothing first, then
// the coarse visits, then the restriction, then pre-smoothing).  The forward
// iterate is stored only at cycle boundaries; each cycle's interior states are
// recomputed into a tape (one snapshot per half sweep and before each
// restriction) right before that cycle is reversed.  Finally the coefficient
// bars flow back through compute_discretization (A_c = sum faces + lambda,
// faces = averaged D * idel^2, injection between levels) to D and lambda.
#pragma once
#include "jsv_kernels.hpp"

namespace jsv {

__device__ __forceinline__ bool is_unknown3(int z, int y, int x, const Geo& g, int dim)
{
    return x >= 2 && x <= g.mx && y >= 2 && y <= g.my && (dim == 2 || (z >= 2 && z <= g.mz));
}

// Kernel A of a half-sweep adjoint, over the points of the swept colour.
// S: forward state right after this half sweep.
__global__ void __launch_bounds__(BX * BY)
k_adj_half_a(double* __restrict__ g, double* __restrict__ t, const double* __restrict__ S,
             const double* __restrict__ cc, double* __restrict__ rbar, double* __restrict__ bx,
             double* __restrict__ by, double* __restrict__ bz, double* __restrict__ bc, Geo gd,
             int par, int dim)
{
    const int z = (dim == 3) ? 2 + (int)blockIdx.z : 0;
    const int y = 2 + blockIdx.y * BY + threadIdx.y;
    if (y > gd.my) return;
    const int x = 2 + ((par + y + z) & 1) + 2 * (blockIdx.x * BX + threadIdx.x);
    if (x > gd.mx) return;
    const long p = (long)z * gd.sz + (long)y * gd.sy + x;
    const double tp = g[p] / cc[p];
    t[p] = tp;
    rbar[p] -= tp;
    bc[p] -= tp * S[p];
    bx[p] += tp * S[p + 1];
    bx[p - 1] += tp * S[p - 1];
    by[p] += tp * S[p + gd.sy];
    by[p - gd.sy] += tp * S[p - gd.sy];
    if (dim == 3) {
        bz[p] += tp * S[p + gd.sz];
        bz[p - gd.sz] += tp * S[p - gd.sz];
    }
    g[p] = 0.0;
}

// Kernel B, over the other colour: g_q += sum_nb A(q, nb) t_nb.
__global__ void __launch_bounds__(BX * BY)
k_adj_half_b(double* __restrict__ g, const double* __restrict__ t, const double* __restrict__ cx,
             const double* __restrict__ cy, const double* __restrict__ cz, Geo gd, int par, int dim)
{
    const int z = (dim == 3) ? 2 + (int)blockIdx.z : 0;
    const int y = 2 + blockIdx.y * BY + threadIdx.y;
    if (y > gd.my) return;
    const int x = 2 + ((par + y + z) & 1) + 2 * (blockIdx.x * BX + threadIdx.x);
    if (x > gd.mx) return;
    const long p = (long)z * gd.sz + (long)y * gd.sy + x;
    double s = cx[p] * t[p + 1] + cx[p - 1] * t[p - 1];
    s += cy[p] * t[p + gd.sy] + cy[p - gd.sy] * t[p - gd.sy];
    if (dim == 3) s += cz[p] * t[p + gd.sz] + cz[p - gd.sz] * t[p - gd.sz];
    g[p] += s;
}

// Restriction adjoint, step 1: r_bar on the fine unknowns (first-half points
// get R^T rhs_c_bar, the others 0).  first = parity of the first RBGS half.
__global__ void __launch_bounds__(BX * BY)
k_adj_restrict_rbar(double* __restrict__ rb, const double* __restrict__ cbar, Geo gf, Geo gc,
                    int dim, int first)
{
    const int z = (dim == 3) ? 2 + (int)blockIdx.z : 0;
    const int y = 2 + blockIdx.y * BY + threadIdx.y;
    const int x = 2 + blockIdx.x * BX + threadIdx.x;
    if (y > gf.my || x > gf.mx) return;
    const long p = (long)z * gf.sz + (long)y * gf.sy + x;
    if (((x + y + z) & 1) != first) {
        rb[p] = 0.0;
        return;
    }
    // coarse index ranges per axis: odd i -> {(i+1)/2}, even i -> {i/2, i/2+1}
    int lo[3], hi[3], q[3] = {x, y, z};
    int neven = 0;
    for (int a = 0; a < dim; ++a) {
        if (q[a] & 1) {
            lo[a] = hi[a] = (q[a] + 1) >> 1;
        } else {
            lo[a] = q[a] >> 1;
            hi[a] = lo[a] + 1;
            ++neven;
        }
    }
    double w;
    if (dim == 2) w = neven == 0 ? 0.25 : 0.0625;
    else w = neven == 0 ? 0.125 : 0.03125;
    double s = 0.0;
    const int zl = dim == 3 ? lo[2] : 0, zh = dim == 3 ? hi[2] : 0;
    for (int K = zl; K <= zh; ++K)
        for (int J = lo[1]; J <= hi[1]; ++J)
            for (int I = lo[0]; I <= hi[0]; ++I) s += cbar[(long)K * gc.sz + (long)J * gc.sy + I];
    rb[p] = w * s;
}

// Restriction adjoint, step 2, over all fine unknowns: transposed residual.
__global__ void __launch_bounds__(BX * BY)
k_adj_restrict_apply(double* __restrict__ g, const double* __restrict__ rb,
                     const double* __restrict__ phi, const double* __restrict__ cx,
                     const double* __restrict__ cy, const double* __restrict__ cz,
                     const double* __restrict__ cc, double* __restrict__ rbar,
                     double* __restrict__ bx, double* __restrict__ by, double* __restrict__ bz,
                     double* __restrict__ bc, Geo gd, int dim, int first,
                     const double* __restrict__ drr, const double* __restrict__ dlr,
                     double* __restrict__ brr, double* __restrict__ blr)
{
    const int z = (dim == 3) ? 2 + (int)blockIdx.z : 0;
    const int y = 2 + blockIdx.y * BY + threadIdx.y;
    const int x = 2 + blockIdx.x * BX + threadIdx.x;
    if (y > gd.my || x > gd.mx) return;
    const long sy = gd.sy;
    const long p = (long)z * gd.sz + (long)y * sy + x;
    const double r = rb[p];
    double s = cc[p] * r;
    s -= cx[p] * rb[p + 1] + cx[p - 1] * rb[p - 1];
    s -= cy[p] * rb[p + sy] + cy[p - sy] * rb[p - sy];
    if (dim == 3) s -= cz[p] * rb[p + gd.sz] + cz[p - gd.sz] * rb[p - gd.sz];
    if (drr != nullptr)   // 2D D_xy: the residual subtracts the diagonal couplings
        s -= drr[p] * rb[p + sy + 1] + drr[p - sy - 1] * rb[p - sy - 1] + dlr[p] * rb[p + sy - 1] +
             dlr[p - sy + 1] * rb[p - sy + 1];
    g[p] += s;
    if (((x + y + z) & 1) == first) {
        rbar[p] += r;
        bc[p] += r * phi[p];
        bx[p] -= r * phi[p + 1];
        bx[p - 1] -= r * phi[p - 1];
        by[p] -= r * phi[p + sy];
        by[p - sy] -= r * phi[p - sy];
        if (dim == 3) {
            bz[p] -= r * phi[p + gd.sz];
            bz[p - gd.sz] -= r * phi[p - gd.sz];
        }
        if (drr != nullptr) {
            // the couplings this point owns; the ones it shares with the
            // first-half point (y-1, x-1) / (y-1, x+1) follow in
            // k_adj_restrict_diag (that point writes these same entries)
            brr[p] -= r * phi[p + sy + 1];
            blr[p] -= r * phi[p + sy - 1];
        }
    }
}

// Restriction adjoint with D_xy, second part over the first-half points: the
// diagonal couplings to (y-1, x-1) and (y-1, x+1).
__global__ void __launch_bounds__(BX * BY)
k_adj_restrict_diag(const double* __restrict__ rb, const double* __restrict__ phi,
                    double* __restrict__ brr, double* __restrict__ blr, Geo gd, int first)
{
    const int y = 2 + blockIdx.y * BY + threadIdx.y;
    const int x = 2 + blockIdx.x * BX + threadIdx.x;
    if (y > gd.my || x > gd.mx || ((x + y) & 1) != first) return;
    const long sy = gd.sy;
    const long p = (long)y * sy + x;
    const double r = rb[p];
    brr[p - sy - 1] -= r * phi[p - sy - 1];
    blr[p - sy + 1] -= r * phi[p - sy + 1];
}

// ---- 2D D_xy smoother: one of the four sub-colours (lx, ly) -------------
// Forward (rb2diag_point): phi_p = (sum_faces A phi_nb - rhs_p + sum_diag A phi_nb) / A_c
// on the points x = (lx ? 2 : 3) + 2i, y = (ly ? 2 : 3) + 2j.  Kernel A over
// those points (S: state after the sub-colour).
__global__ void __launch_bounds__(BX * BY)
k_adj_diag_a(double* __restrict__ g, double* __restrict__ t, const double* __restrict__ S,
             const double* __restrict__ cc, double* __restrict__ rbar, double* __restrict__ bx,
             double* __restrict__ by, double* __restrict__ bc, double* __restrict__ brr,
             double* __restrict__ blr, Geo gd, int lx, int ly)
{
    const int y = (ly ? 2 : 3) + 2 * (int)(blockIdx.y * BY + threadIdx.y);
    const int x = (lx ? 2 : 3) + 2 * (int)(blockIdx.x * BX + threadIdx.x);
    if (y > gd.my || x > gd.mx) return;
    const long sy = gd.sy;
    const long p = (long)y * sy + x;
    const double tp = g[p] / cc[p];
    t[p] = tp;
    rbar[p] -= tp;
    bc[p] -= tp * S[p];
    bx[p] += tp * S[p + 1];
    bx[p - 1] += tp * S[p - 1];
    by[p] += tp * S[p + sy];
    by[p - sy] += tp * S[p - sy];
    brr[p] += tp * S[p + sy + 1];
    brr[p - sy - 1] += tp * S[p - sy - 1];
    blr[p] += tp * S[p + sy - 1];
    blr[p - sy + 1] += tp * S[p - sy + 1];
    g[p] = 0.0;
}

// Kernel B over the unknowns outside the sub-colour: g_q += sum over q's
// neighbours in the sub-colour of A(q, nb) t_nb (x neighbours when q shares
// the sub-colour's row parity, y neighbours when it shares the column parity,
// the diagonals when it shares neither).
__global__ void __launch_bounds__(BX * BY)
k_adj_diag_b(double* __restrict__ g, const double* __restrict__ t, const double* __restrict__ cx,
             const double* __restrict__ cy, const double* __restrict__ drr,
             const double* __restrict__ dlr, Geo gd, int lx, int ly)
{
    const int y = 2 + blockIdx.y * BY + threadIdx.y;
    const int x = 2 + blockIdx.x * BX + threadIdx.x;
    if (y > gd.my || x > gd.mx) return;
    const bool xs = (x & 1) == (lx ? 0 : 1), ys = (y & 1) == (ly ? 0 : 1);
    if (xs && ys) return;
    const long sy = gd.sy;
    const long p = (long)y * sy + x;
    double s;
    if (ys) s = cx[p] * t[p + 1] + cx[p - 1] * t[p - 1];
    else if (xs) s = cy[p] * t[p + sy] + cy[p - sy] * t[p - sy];
    else
        s = drr[p] * t[p + sy + 1] + drr[p - sy - 1] * t[p - sy - 1] + dlr[p] * t[p + sy - 1] +
            dlr[p - sy + 1] * t[p - sy + 1];
    g[p] += s;
}

// Prolongation adjoint: g_c = P^T g over the coarse unknowns.
__global__ void __launch_bounds__(BX * BY)
k_adj_prolong(double* __restrict__ gc_out, const double* __restrict__ g, Geo gf, Geo gc, int dim,
              int full)
{
    const int I3 = (dim == 3) ? 2 + (int)blockIdx.z : 0;
    const int J = 2 + blockIdx.y * BY + threadIdx.y;
    const int K = 2 + blockIdx.x * BX + threadIdx.x;
    if (J > gc.my || K > gc.mx) return;
    const int fx = 2 * K - 1, fy = 2 * J - 1, fz = (dim == 3) ? 2 * I3 - 1 : 0;
    double s = 0.0;
    const int dzl = dim == 3 ? -1 : 0, dzh = dim == 3 ? 1 : 0;
    for (int dz = dzl; dz <= dzh; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int ne = (dx != 0) + (dy != 0) + (dz != 0);
                double w;
                if (ne == 0) w = 1.0;
                else if (ne == 1) w = 0.5;
                else if (ne == 2) { if (!full) continue; w = 0.25; }
                else w = 0.125;
                if (dim == 2 && ne == 2 && !full) continue;
                const long pf = (long)(fz + dz) * gf.sz + (long)(fy + dy) * gf.sy + (fx + dx);
                s += w * g[pf];
            }
    gc_out[(long)I3 * gc.sz + (long)J * gc.sy + K] = s;
}

// ---- coefficient chain --------------------------------------------------
// total bar of the face array of one axis: direct face bar plus the two
// diagonals the face enters; times idel^2 -> bar of the (averaged) D face.
__global__ void __launch_bounds__(256)
k_adj_face_total(double* __restrict__ out, const double* __restrict__ bface,
                 const double* __restrict__ bc, Geo g, long step, double sq, int accumulate)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= g.n) return;
    double v = bface[i] + bc[i];
    if (i + step < g.n) v += bc[i + step];
    v *= sq;
    out[i] = accumulate ? out[i] + v : v;
}

// transpose of the face average along `axis` over the window of `g`
// (2D: solver_2d.py:360-361; 3D: solver_3d.py:368-370): raw_bar[p] =
// 0.5 (avg_bar[p] + avg_bar[p - e]) for the parts inside the window.
__global__ void __launch_bounds__(256)
k_adj_average(double* __restrict__ raw_bar, const double* __restrict__ avg_bar, Geo g, int dim,
              int axis)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= g.n) return;
    const int x = (int)(i % g.nx);
    const int y = (int)((i / g.nx) % g.ny);
    const int z = (dim == 3) ? (int)(i / ((long)g.nx * g.ny)) : 0;
    const int q[3] = {x, y, z};
    const int m[3] = {g.mx, g.my, g.mz};
    const long st[3] = {1, g.sy, g.sz};
    // avg window: axis in [0, m+1], other axes in [1, m+1]
    auto in_win = [&](int qa) -> bool {
        for (int a = 0; a < dim; ++a) {
            const int v = (a == axis) ? qa : q[a];
            if (a == axis) {
                if (v < 0 || v > m[a] + 1) return false;
            } else if (v < 1 || v > m[a] + 1) {
                return false;
            }
        }
        return true;
    };
    double s = 0.0;
    if (in_win(q[axis])) s += 0.5 * avg_bar[i];
    if (in_win(q[axis] - 1)) s += 0.5 * avg_bar[i - st[axis]];
    raw_bar[i] = s;
}

// scatter a level-l array (window entries) into the fine array at stride 2^l
// (transpose of the reference's injection chain; 2D D, lambda in 2D and 3D).
__global__ void __launch_bounds__(256)
k_adj_inject_to_fine(double* __restrict__ fine, const double* __restrict__ lev, Geo g, Geo g0,
                     int stride, int dim)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= g.n) return;
    const int x = (int)(i % g.nx);
    const int y = (int)((i / g.nx) % g.ny);
    const int z = (dim == 3) ? (int)(i / ((long)g.nx * g.ny)) : 0;
    if (x < 1 || x > g.mx + 1 || y < 1 || y > g.my + 1) return;
    if (dim == 3 && (z < 1 || z > g.mz + 1)) return;
    const long zz = dim == 3 ? (long)(z - 1) * stride + 1 : 0;
    const long f = zz * g0.sz + ((long)(y - 1) * stride + 1) * g0.sy + (long)(x - 1) * stride + 1;
    fine[f] += lev[i];
}

// 3D: transpose of one injection step (level l from level l-1, odd indices)
__global__ void __launch_bounds__(256)
k_adj_inject_one(double* __restrict__ prev, const double* __restrict__ lev, Geo g, Geo gp)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= g.n) return;
    const int x = (int)(i % g.nx);
    const int y = (int)((i / g.nx) % g.ny);
    const int z = (int)(i / ((long)g.nx * g.ny));
    if (x < 1 || x > g.mx + 1 || y < 1 || y > g.my + 1 || z < 1 || z > g.mz + 1) return;
    prev[(long)(2 * z - 1) * gp.sz + (long)(2 * y - 1) * gp.sy + (2 * x - 1)] += lev[i];
}

// 2D D_xy: bar of level l's injected D_xy from the bars of its coupling
// arrays (k_setup2d: drr = (Dxy[y, x+1] + Dxy[y+1, x]) c, dlr = -(Dxy[y, x-1] +
// Dxy[y+1, x]) c on the window [1, m+1]^2; solver_2d.py:411-416), gathered
// per D_xy entry.
__global__ void __launch_bounds__(256)
k_adj_dxy_level(double* __restrict__ qbar, const double* __restrict__ brr,
                const double* __restrict__ blr, Geo g, double c)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= g.n) return;
    const int x = (int)(i % g.nx);
    const int y = (int)((i / g.nx) % g.ny);
    auto win = [&](int yy, int xx) { return yy >= 1 && yy <= g.my + 1 && xx >= 1 && xx <= g.mx + 1; };
    double s = 0.0;
    if (win(y, x - 1)) s += brr[i - 1];                       // Dxy[y, x] as the x+1 term of (y, x-1)
    if (win(y - 1, x)) s += brr[i - g.sy] - blr[i - g.sy];    // as the y+1 term of (y-1, x)
    if (win(y, x + 1)) s -= blr[i + 1];                       // as the x-1 term of (y, x+1)
    qbar[i] = s * c;
}

// Sum of n doubles into *out (set, or added when accumulate): one workgroup,
// fixed order, so repeated runs give the same bits.
__global__ void __launch_bounds__(1024)
k_sum_array(const double* __restrict__ a, long n, double* __restrict__ out, int accumulate)
{
    __shared__ double sh[16];
    double v = 0.0;
    for (long i = threadIdx.x; i < n; i += 1024) v += a[i];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < 16; ++w) s += sh[w];
        *out = accumulate ? *out + s : s;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_axpy(T* __restrict__ y, const T* __restrict__ x, long n, T a)
{
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) y[i] += a * x[i];
}

}  // namespace jsv
